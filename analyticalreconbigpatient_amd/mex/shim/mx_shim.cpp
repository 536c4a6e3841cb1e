// Minimal implementation of the mx API subset in matrix.h / mex.h, plus a
// C-ABI to build arrays and call a gateway's mexFunction from a test harness.
// MATLAB arrays are column-major; here an array is a flat buffer with MATLAB's
// dims, exactly what a gateway sees through mxGetPr/mxGetDimensions.
#include "mex.h"

#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

struct mxArray_tag {
    mxClassID cls = mxUNKNOWN_CLASS;
    std::vector<mwSize> dims;
    std::vector<unsigned char> data;     // numeric payload
    std::vector<mxArray*> fields;         // struct (1x1) fields, owned
};

namespace {
struct MexError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

size_t elem_size(mxClassID c) {
    switch (c) {
    case mxDOUBLE_CLASS: return 8;
    case mxSINGLE_CLASS: return 4;
    case mxINT32_CLASS: return 4;
    default: return 0;
    }
}
}  // namespace

extern "C" {

double* mxGetPr(const mxArray* a) {
    return a ? reinterpret_cast<double*>(const_cast<unsigned char*>(a->data.data())) : nullptr;
}
void* mxGetData(const mxArray* a) {
    return a ? const_cast<unsigned char*>(a->data.data()) : nullptr;
}
const mwSize* mxGetDimensions(const mxArray* a) { return a->dims.data(); }
mwSize mxGetNumberOfDimensions(const mxArray* a) { return a->dims.size(); }
size_t mxGetNumberOfElements(const mxArray* a) {
    size_t n = 1;
    for (mwSize d : a->dims) n *= d;
    return n;
}
mxClassID mxGetClassID(const mxArray* a) { return a->cls; }

mxArray* mxCreateNumericArray(mwSize ndim, const mwSize* dims, mxClassID cls, mxComplexity) {
    mxArray* a = new mxArray_tag();
    a->cls = cls;
    a->dims.assign(dims, dims + ndim);
    a->data.assign(mxGetNumberOfElements(a) * elem_size(cls), 0);   // zero-filled, as MATLAB
    return a;
}

mxArray* mxGetFieldByNumber(const mxArray* a, size_t index, int field) {
    if (!a || a->cls != mxSTRUCT_CLASS || index != 0 || field < 0 ||
        field >= (int)a->fields.size())
        return nullptr;
    return a->fields[field];
}

void mxDestroyArray(mxArray* a) {
    if (!a) return;
    for (mxArray* f : a->fields) mxDestroyArray(f);
    delete a;
}

void mexErrMsgTxt(const char* msg) { throw MexError(msg ? msg : "mexErrMsgTxt"); }

// ---- harness side ----------------------------------------------------------
mxArray* mxshim_numeric(int cls, int ndim, const long long* dims, const void* src) {
    std::vector<mwSize> d(dims, dims + ndim);
    mxArray* a = mxCreateNumericArray(ndim, d.data(), (mxClassID)cls, mxREAL);
    if (src && !a->data.empty()) std::memcpy(a->data.data(), src, a->data.size());
    return a;
}

mxArray* mxshim_struct(int nfields) {
    mxArray* a = new mxArray_tag();
    a->cls = mxSTRUCT_CLASS;
    a->dims = {1, 1};
    a->fields.assign(nfields, nullptr);
    return a;
}

int mxshim_set_field(mxArray* s, int field, mxArray* value) {
    if (!s || s->cls != mxSTRUCT_CLASS || field < 0 || field >= (int)s->fields.size()) return 1;
    mxDestroyArray(s->fields[field]);
    s->fields[field] = value;
    return 0;
}

size_t mxshim_bytes(const mxArray* a) { return a ? a->data.size() : 0; }

typedef void (*gateway_fn)(int, mxArray**, int, const mxArray**);

// Calls a gateway; returns 0, or 1 with the mexErrMsgTxt message in err.
int mxshim_call(gateway_fn fn, int nlhs, mxArray** plhs, int nrhs, const mxArray** prhs,
                char* err, size_t errlen) {
    try {
        fn(nlhs, plhs, nrhs, prhs);
        return 0;
    } catch (const std::exception& e) {
        if (err && errlen) {
            std::strncpy(err, e.what(), errlen - 1);
            err[errlen - 1] = '\0';
        }
        return 1;
    }
}

}  // extern "C"
