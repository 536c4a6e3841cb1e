// Shared helpers of the arecon MEX gateways.
#pragma once

#include <mex.h>

#include "arecon.h"

namespace arecon_mex {

// Turns a failure recorded by the last libarecon call into a MATLAB error.
inline void raise_if_failed(const char* who) {
    static char buf[1024];
    if (arecon_last_error(buf, sizeof(buf)) != ARECON_OK) {
        static char msg[1200];
        int n = 0;
        for (const char* p = who; *p && n < 150; ++p) msg[n++] = *p;
        msg[n++] = ':';
        msg[n++] = ' ';
        for (const char* p = buf; *p && n < (int)sizeof(msg) - 1; ++p) msg[n++] = *p;
        msg[n] = '\0';
        mexErrMsgTxt(msg);
    }
}

template <class T>
inline T scalar(const mxArray* a) {
    return *static_cast<const T*>(mxGetData(a));
}

}  // namespace arecon_mex
