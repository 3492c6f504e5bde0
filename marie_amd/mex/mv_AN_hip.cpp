// JOut = mv_AN_hip(JIn0, fN, Mr, Mc, Gram, transp_flag, idx)
//
// The GPU branch of src_vie/src_mvp/mv_AN.m:1 (GPU_flag = 1): the same
// arguments and result (a column of numel(idx) complex values, 'notransp' or
// 'transp'), computed by marie_vie_apply on a cached device operator.
// mv_AN_hip('clear') releases the operator.
//
// Build on a MATLAB host:
//   mex -O -I<repo>/include mv_AN_hip.cpp -L<repo>/marie_amd -lmarie_vie -Wl,-rpath,<repo>/marie_amd
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
    if (nrhs != 7) mexErrMsgIdAndTxt("MARIE:hip:nargin", "7 inputs required: JIn0, fN, Mr, Mc, Gram, transp_flag, idx.");
    if (nlhs > 1) mexErrMsgIdAndTxt("MARIE:hip:nargout", "1 output required.");
    if (!mxIsChar(prhs[5])) mexErrMsgIdAndTxt("MARIE:hip:transp", "transp_flag must be 'notransp' or 'transp'.");
    char flag[16] = {0};
    mxGetString(prhs[5], flag, sizeof flag);
    const bool transp = std::strcmp(flag, "transp") == 0;
    if (!transp && std::strcmp(flag, "notransp") != 0)
        mexErrMsgIdAndTxt("MARIE:hip:transp", "transp_flag must be 'notransp' or 'transp'.");
    g_cache.operator_for(prhs[1], prhs[2]);
    g_cache.material(prhs[2], prhs[3], mxGetScalar(prhs[4]));
    g_cache.index(prhs[6]);
    const size_t n = (size_t)g_cache.nidx;
    if (mxGetNumberOfElements(prhs[0]) != n)
        mexErrMsgIdAndTxt("MARIE:hip:size", "JIn0 must have numel(idx) elements.");
    std::vector<double> x, y(2 * n);
    interleave(prhs[0], x);
    if (n) fail(marie_vie_apply(g_cache.op, MARIE_OP_AN, transp ? 1 : 0, x.data(), y.data(), 1));
    plhs[0] = complex_column(y, n);
}
