/*
 * matrix.h -- the subset of MATLAB's mx API that the arecon MEX gateways use,
 * implemented by mx_shim.cpp so the gateways can be built and tested without
 * MATLAB.  With a MATLAB installation, build the gateways against MATLAB's own
 * mex.h / matrix.h instead (the gateway sources are the same).
 */
#ifndef ARECON_MX_SHIM_MATRIX_H_
#define ARECON_MX_SHIM_MATRIX_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef size_t mwSize;
typedef struct mxArray_tag mxArray;

typedef enum {
    mxUNKNOWN_CLASS = 0,
    mxSTRUCT_CLASS = 2,
    mxDOUBLE_CLASS = 6,
    mxSINGLE_CLASS = 7,
    mxINT32_CLASS = 12
} mxClassID;

typedef enum { mxREAL = 0, mxCOMPLEX = 1 } mxComplexity;

double* mxGetPr(const mxArray* a);
void* mxGetData(const mxArray* a);
const mwSize* mxGetDimensions(const mxArray* a);
mwSize mxGetNumberOfDimensions(const mxArray* a);
size_t mxGetNumberOfElements(const mxArray* a);
mxClassID mxGetClassID(const mxArray* a);
mxArray* mxCreateNumericArray(mwSize ndim, const mwSize* dims, mxClassID cls,
                              mxComplexity cplx);
mxArray* mxGetFieldByNumber(const mxArray* a, size_t index, int field);
void mxDestroyArray(mxArray* a);

#ifdef __cplusplus
}
#endif

#endif
