// MEX gateway: P = ParallelRebinningCBCurve(single Proj[ZL,YL,ViewN], int32 YL,
//   int32 ZL, int32 ViewN, single YLC, single dYA, single DeltaTheta, single PLC,
//   single DeltaT, single DeltaFai, single SO, int32 useGPU)
// Same arguments and layout as src/ParallelRebinningCBCurve.cpp:101-129: YLC and
// PLC are 1-based and shifted by -1 here; the output is a new single array of
// Proj's size.  The GPU path is the only path (useGPU must be 1).
#include "gateway_common.h"

using arecon_mex::scalar;

extern "C" void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
    if (nrhs != 12) mexErrMsgTxt("ParallelRebinningCBCurve: 12 inputs are required.");
    if (nlhs > 1) mexErrMsgTxt("ParallelRebinningCBCurve: one output.");
    const float* Proj = static_cast<const float*>(mxGetData(prhs[0]));
    const int YL = scalar<int>(prhs[1]);
    const int ZL = scalar<int>(prhs[2]);
    const int ViewN = scalar<int>(prhs[3]);
    const float YLC = scalar<float>(prhs[4]);
    const float dYA = scalar<float>(prhs[5]);
    const float DeltaTheta = scalar<float>(prhs[6]);
    const float PLC = scalar<float>(prhs[7]);
    const float DeltaT = scalar<float>(prhs[8]);
    const float DeltaFai = scalar<float>(prhs[9]);
    const float SO = scalar<float>(prhs[10]);
    const int useGPU = scalar<int>(prhs[11]);
    if (useGPU != 1)
        mexErrMsgTxt("ParallelRebinningCBCurve: only useGPU == 1 is provided (no CPU path).");
    if (YL <= 0 || ZL <= 0 || ViewN <= 0 ||
        mxGetNumberOfElements(prhs[0]) != (size_t)YL * ZL * ViewN)
        mexErrMsgTxt("ParallelRebinningCBCurve: Proj does not hold ZL*YL*ViewN samples.");
    const mwSize dims[3] = {(mwSize)ZL, (mwSize)YL, (mwSize)ViewN};
    plhs[0] = mxCreateNumericArray(3, dims, mxSINGLE_CLASS, mxREAL);
    float* out = static_cast<float*>(mxGetData(plhs[0]));
    ParallelRebinningCBCurve_GPU(out, Proj, YL, ZL, ViewN, YLC - 1.0f, dYA, DeltaTheta,
                                 PLC - 1.0f, DeltaT, DeltaFai, SO);
    arecon_mex::raise_if_failed("ParallelRebinningCBCurve");
}
