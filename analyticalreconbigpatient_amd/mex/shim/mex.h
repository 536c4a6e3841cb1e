/* mex.h -- gateway-side API of the mx shim (see matrix.h). */
#ifndef ARECON_MX_SHIM_MEX_H_
#define ARECON_MX_SHIM_MEX_H_

#include "matrix.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Aborts the running gateway with an error message (the shim unwinds back to
 * mxshim_call, which reports the message to the caller). */
void mexErrMsgTxt(const char* msg);

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]);

#ifdef __cplusplus
}
#endif

#endif
