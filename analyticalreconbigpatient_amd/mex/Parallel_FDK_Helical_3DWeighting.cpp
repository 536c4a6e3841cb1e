// MEX gateway: Parallel_FDK_Helical_3DWeighting(ScanGeom, double fp[YL,ZL,PN],
//   double ctimage[XN*YN*ZN], int32 useGPU)  -- no outputs, ctimage written in place.
// Same arguments as src/Parallel_FDK_Helical_3DWeighting.cpp:204-269.  ScanGeom
// fields are read by number: ScDet = [SO DO YL ZL DecWidth DecHeigh YLC ZLC h],
// Proj = [BetaS BetaE ViewN N_2pi], Obj = [ObjR XN YN ZN XC YC ZC],
// Rec = [delta HSCoef k1]; YLC is taken as (YL-1)/2 whatever the struct holds,
// centres are converted to 0-based, h = ScDet(9)*DecHeigh, k1 = Rec(3)*ScDet(9).
#include "gateway_common.h"

using arecon_mex::scalar;

extern "C" void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
    (void)plhs;
    if (nrhs != 4) mexErrMsgTxt("Backward projection needs 4 inputs.");
    if (nlhs != 0) mexErrMsgTxt("Backward projection does not need any output.");
    const mxArray* f[4];
    for (int i = 0; i < 4; ++i) {
        f[i] = mxGetFieldByNumber(prhs[0], 0, i);
        if (!f[i] || mxGetClassID(f[i]) != mxDOUBLE_CLASS)
            mexErrMsgTxt("Parallel_FDK_Helical_3DWeighting: ScanGeom must hold ScDet, Proj, "
                         "Obj, Rec as double vectors.");
    }
    if (mxGetNumberOfElements(f[0]) < 9 || mxGetNumberOfElements(f[1]) < 4 ||
        mxGetNumberOfElements(f[2]) < 7 || mxGetNumberOfElements(f[3]) < 3)
        mexErrMsgTxt("Parallel_FDK_Helical_3DWeighting: ScanGeom field too short.");
    const double* sd = mxGetPr(f[0]);
    const double* pr = mxGetPr(f[1]);
    const double* ob = mxGetPr(f[2]);
    const double* rc = mxGetPr(f[3]);
    const double SO = sd[0], DO = sd[1];
    const int YL = (int)(sd[2] + 0.5), ZL = (int)(sd[3] + 0.5);
    const double DecWidth = sd[4], DecHeigh = sd[5];
    const double YLC = (YL - 1) * 0.5;
    const double ZLC = sd[7] - 1.0;
    const double h = sd[8] * DecHeigh;
    const double BetaS = pr[0];
    const int PN = (int)pr[2], N_2pi = (int)pr[3];
    const double ObjR = ob[0];
    const int XN = (int)(ob[1] + 0.5), YN = (int)(ob[2] + 0.5), ZN = (int)(ob[3] + 0.5);
    const double XNC = ob[4] - 1.0, YNC = ob[5] - 1.0, ZNC = ob[6] - 1.0;
    const int delta = (int)(rc[0] + 0.5);
    const double HSCoef = rc[1];
    const double k1 = rc[2] * sd[8];
    if (mxGetClassID(prhs[1]) != mxDOUBLE_CLASS ||
        mxGetNumberOfElements(prhs[1]) != (size_t)YL * ZL * PN)
        mexErrMsgTxt("Parallel_FDK_Helical_3DWeighting: fp must be double [YL, ZL, PN].");
    if (mxGetClassID(prhs[2]) != mxDOUBLE_CLASS ||
        mxGetNumberOfElements(prhs[2]) != (size_t)XN * YN * ZN)
        mexErrMsgTxt("Parallel_FDK_Helical_3DWeighting: ctimage must be double with XN*YN*ZN "
                     "elements.");
    const double* fp = mxGetPr(prhs[1]);
    double* ctimage = mxGetPr(prhs[2]);   // written in place, as the reference does
    const int useGPU = scalar<int>(prhs[3]);
    ParallelFDKHelical3DWeighting(ctimage, fp, SO, DO, YL, ZL, DecWidth, DecHeigh, YLC, ZLC, h,
                                  BetaS, PN, N_2pi, ObjR, XN, YN, ZN, XNC, YNC, ZNC, delta,
                                  HSCoef, k1, useGPU);
    arecon_mex::raise_if_failed("Parallel_FDK_Helical_3DWeighting");
}
