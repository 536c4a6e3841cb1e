// MEX gateway: FPWD = reWeigAdFiltr_GPU(single Proj[YL,ZL,ViewN], single PLC,
//   single ZLC, single dYL, single dZL, single SO)
// Same arguments as src/reWeigAdFiltr_GPU.cpp:36-63: sizes from Proj, PLC and
// ZLC already 0-based, output a new single array of Proj's size.  Wrong arity
// raises a MATLAB error (the reference prints and calls exit(-1)).
#include "gateway_common.h"

using arecon_mex::scalar;

extern "C" void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
    if (nlhs != 1 || nrhs != 6)
        mexErrMsgTxt("reWeigAdFiltr_GPU: requires 6 input parameters and provides one return "
                     "value.");
    if (mxGetNumberOfDimensions(prhs[0]) < 2 || mxGetClassID(prhs[0]) != mxSINGLE_CLASS)
        mexErrMsgTxt("reWeigAdFiltr_GPU: Proj must be a single array [YL, ZL, ViewN].");
    const mwSize* siz = mxGetDimensions(prhs[0]);
    const int YL = (int)siz[0];
    const int ZL = (int)siz[1];
    const int ViewN = mxGetNumberOfDimensions(prhs[0]) > 2 ? (int)siz[2] : 1;
    plhs[0] = mxCreateNumericArray(mxGetNumberOfDimensions(prhs[0]), siz, mxSINGLE_CLASS, mxREAL);
    float* fpwd = static_cast<float*>(mxGetData(plhs[0]));
    const float* Proj = static_cast<const float*>(mxGetData(prhs[0]));
    filtering(fpwd, Proj, YL, ZL, ViewN, scalar<float>(prhs[1]), scalar<float>(prhs[2]),
              scalar<float>(prhs[3]), scalar<float>(prhs[4]), scalar<float>(prhs[5]));
    arecon_mex::raise_if_failed("reWeigAdFiltr_GPU");
}
