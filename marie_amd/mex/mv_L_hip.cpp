// JOut = mv_L_hip(JIn0, fN, r, Gram, idx)
//
// The GPU branch of src_vie/src_mvp/mv_L.m:1 (GPU_flag = 1): the same
// arguments and result (N JIn - Gram JIn, a column of numel(idx) complex values for
// JIn0 on the idx DOFs of the L x M x N x 3 grid of r), computed by
// marie_vie_apply(MARIE_OP_L) on a cached device operator. mv_L_hip('clear')
// releases the operator.
//
// Build on a MATLAB host:
//   mex -O -I<repo>/include mv_L_hip.cpp -L<repo>/marie_amd -lmarie_vie -Wl,-rpath,<repo>/marie_amd
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
    if (nrhs != 5) mexErrMsgIdAndTxt("MARIE:hip:nargin", "5 inputs required: JIn0, fN, r, Gram, idx.");
    if (nlhs > 1) mexErrMsgIdAndTxt("MARIE:hip:nargout", "1 output required.");
    g_cache.operator_on_grid(prhs[1], prhs[2], 6);
    g_cache.gram_only(mxGetScalar(prhs[3]));
    g_cache.index(prhs[4]);
    const size_t n = (size_t)g_cache.nidx;
    if (mxGetNumberOfElements(prhs[0]) != n)
        mexErrMsgIdAndTxt("MARIE:hip:size", "JIn0 must have numel(idx) elements.");
    std::vector<double> x, y(2 * n);
    interleave(prhs[0], x);
    if (n) fail(marie_vie_apply(g_cache.op, MARIE_OP_L, 0, x.data(), y.data(), 1));
    plhs[0] = complex_column(y, n);
}
