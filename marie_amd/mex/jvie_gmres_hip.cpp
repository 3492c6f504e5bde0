// [x, flag, relres, iter, resvec] = jvie_gmres_hip(b, fN, Mr, Mc, Gram, idx, restart, tol, maxit, x0, precision)
//
// The GPU path of JVIE_Solver.m's GMRES branch (JVIE_Solver.m:276, :309-377):
// pgmres(@(J) mv_AN(J, fN, Mr, Mc, Gram, 'notransp', idx, 1), b, restart, tol,
// maxit, [], [], [], [], x0) (src_numeric/src_iterative/pgmres.m:70) with the
// whole Krylov loop on the device (marie_vie_gmres). Outputs keep pgmres'
// meaning: flag 0/1 from the true residual (pgmres.m:175-180), relres =
// ||b - A x|| / ||b||, iter = [inner outer] (pgmres.m:172), resvec relative to
// ||b|| (pgmres.m:149, :160). x0 is optional (zeros, or [] when a precision
// follows). precision 'fp64' (the default) is pgmres itself; 'fp32-refined'
// runs complex64 products in block-orthogonalised cycles of `restart`
// products with FP64 refinement on the true residual (marie_vie_gmres mode
// 3, no counterpart in the reference): the same flag/relres meaning, resvec
// the cycles' least-squares residuals and each cycle's true one.
// jvie_gmres_hip('clear') releases the cached operator.
//
// Build on a MATLAB host:
//   mex -O -I<repo>/include jvie_gmres_hip.cpp -L<repo>/marie_amd -lmarie_vie -Wl,-rpath,<repo>/marie_amd
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
    if (nrhs < 9 || nrhs > 11)
        mexErrMsgIdAndTxt("MARIE:hip:nargin", "9 to 11 inputs required: b, fN, Mr, Mc, Gram, idx, restart, tol, "
                                              "maxit[, x0[, precision]].");
    int mode = 0;
    if (nrhs == 11) {
        char prec[32] = {0};
        if (!mxIsChar(prhs[10]) || mxGetString(prhs[10], prec, sizeof prec) != 0)
            mexErrMsgIdAndTxt("MARIE:hip:precision", "precision must be 'fp64' or 'fp32-refined'.");
        if (std::strcmp(prec, "fp32-refined") == 0)
            mode = 3;
        else if (std::strcmp(prec, "fp64") != 0)
            mexErrMsgIdAndTxt("MARIE:hip:precision", "precision must be 'fp64' or 'fp32-refined'.");
    }
    // the complex64 copy of the operator serves the refined precision
    g_cache.precisions = MARIE_PREC_F64 | MARIE_PREC_F32;
    if (nlhs > 5) mexErrMsgIdAndTxt("MARIE:hip:nargout", "at most 5 outputs: x, flag, relres, iter, resvec.");
    g_cache.operator_for(prhs[1], prhs[2]);
    g_cache.material(prhs[2], prhs[3], mxGetScalar(prhs[4]));
    g_cache.index(prhs[5]);
    const int restart = (int)mxGetScalar(prhs[6]);
    const double tol = mxGetScalar(prhs[7]);
    const int maxit = (int)mxGetScalar(prhs[8]);
    if (restart < 1 || maxit < 1 || !(tol > 0)) mexErrMsgIdAndTxt("MARIE:hip:args", "restart, maxit >= 1, tol > 0.");
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
    const int cap = (restart + 1) * maxit + 2;
    std::vector<double> resvec((size_t)cap);
    int flag = 0, iter[2] = {0, 0}, nres = 0;
    double relres = 0.0;
    fail(marie_vie_gmres(g_cache.op, b.data(), x.data(), restart, tol, maxit, mode, mode ? 1e-7 : 0.0, &flag,
                         &relres, iter, resvec.data(), cap, &nres));
    plhs[0] = complex_column(x, n);
    if (nlhs > 1) plhs[1] = mxCreateDoubleScalar((double)flag);
    if (nlhs > 2) plhs[2] = mxCreateDoubleScalar(relres);
    if (nlhs > 3) {
        plhs[3] = mxCreateDoubleMatrix(1, 2, mxREAL);
        mxGetPr(plhs[3])[0] = iter[0];
        mxGetPr(plhs[3])[1] = iter[1];
    }
    if (nlhs > 4) {
        const int m = nres < cap ? nres : cap;
        plhs[4] = mxCreateDoubleMatrix((mwSize)m, 1, mxREAL);
        for (int i = 0; i < m; ++i) mxGetPr(plhs[4])[i] = resvec[(size_t)i];
    }
}
