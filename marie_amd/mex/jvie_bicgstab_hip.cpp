// [x, flag, relres, iter, resvec] = jvie_bicgstab_hip(b, fN, Mr, Mc, Gram, idx, tol, maxit, precond, x0)
//
// The GPU path of JVIE_Solver.m's default BiCGStab branch (JVIE_Solver.m:451-497):
// bicgstab(@(J) mv_AN(J, fN, Mr, Mc, Gram, 'notransp', idx, 1), b, tol,
// maxit[, pfun]) with pfun the preconditioner mv_AN(J, fN, 1./Mr, 1./Mr - 1,
// Gram, ...) (JVIE_Solver.m:278, :297) when `precond` is nonzero (left, as
// MATLAB's M argument), the whole iteration on the device
// (marie_vie_bicgstab). Outputs as MATLAB's bicgstab: flag 0 converged,
// 1 maxit reached, 4 breakdown, by the true residual; relres =
// ||b - A x|| / ||b||; iter counts half steps as 0.5; resvec the absolute
// residual norm of every half step. precond and x0 are optional.
// jvie_bicgstab_hip('clear') releases the cached operator.
//
// Build on a MATLAB host:
//   mex -O -I<repo>/include jvie_bicgstab_hip.cpp -L<repo>/marie_amd -lmarie_vie -Wl,-rpath,<repo>/marie_amd
#include "marie_mex.hpp"

static marie_mex::Cache g_cache;

static void release() { g_cache.clear(); }

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
    using namespace marie_mex;
    mexAtExit(release);
    if (is_clear(nrhs, prhs)) {
        g_cache.clear();
        return;
    }
    if (nrhs < 8 || nrhs > 10)
        mexErrMsgIdAndTxt("MARIE:hip:nargin",
                          "8 to 10 inputs required: b, fN, Mr, Mc, Gram, idx, tol, maxit[, precond[, x0]].");
    if (nlhs > 5) mexErrMsgIdAndTxt("MARIE:hip:nargout", "at most 5 outputs: x, flag, relres, iter, resvec.");
    g_cache.operator_for(prhs[1], prhs[2]);
    g_cache.material(prhs[2], prhs[3], mxGetScalar(prhs[4]));
    g_cache.index(prhs[5]);
    const double tol = mxGetScalar(prhs[6]);
    const int maxit = (int)mxGetScalar(prhs[7]);
    const int precond = nrhs >= 9 && mxGetNumberOfElements(prhs[8]) > 0 && mxGetScalar(prhs[8]) != 0.0 ? 1 : 0;
    if (maxit < 1 || !(tol > 0)) mexErrMsgIdAndTxt("MARIE:hip:args", "maxit >= 1, tol > 0.");
    const size_t n = (size_t)g_cache.nidx;
    if (mxGetNumberOfElements(prhs[0]) != n) mexErrMsgIdAndTxt("MARIE:hip:size", "b must have numel(idx) elements.");
    std::vector<double> b, x;
    interleave(prhs[0], b);
    if (nrhs >= 10 && mxGetNumberOfElements(prhs[9]) > 0) {
        if (mxGetNumberOfElements(prhs[9]) != n) mexErrMsgIdAndTxt("MARIE:hip:size", "x0 must match b.");
        interleave(prhs[9], x);
    } else {
        x.assign(2 * n, 0.0);
    }
    const int cap = 2 * maxit + 1;
    std::vector<double> resvec((size_t)cap);
    int flag = 0, nres = 0;
    double relres = 0.0, iter = 0.0;
    fail(marie_vie_bicgstab(g_cache.op, b.data(), x.data(), tol, maxit, precond, &flag, &relres, &iter,
                            resvec.data(), cap, &nres));
    plhs[0] = complex_column(x, n);
    if (nlhs > 1) plhs[1] = mxCreateDoubleScalar((double)flag);
    if (nlhs > 2) plhs[2] = mxCreateDoubleScalar(relres);
    if (nlhs > 3) plhs[3] = mxCreateDoubleScalar(iter);
    if (nlhs > 4) {
        const int m = nres < cap ? nres : cap;
        plhs[4] = mxCreateDoubleMatrix((mwSize)m, 1, mxREAL);
        for (int i = 0; i < m; ++i) mxGetPr(plhs[4])[i] = resvec[(size_t)i];
    }
}
