"""Test harness for the MEX gateways: builds mxArrays through libmxshim.so and
calls a gateway's mexFunction the way MATLAB would."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

MEX = Path(__file__).resolve().parent.parent / "analyticalreconbigpatient_amd" / "mex"
CLS = {np.float64: 6, np.float32: 7, np.int32: 12}

_shim = None
_gw = {}


def shim() -> C.CDLL:
    global _shim
    if _shim is None:
        # RTLD_GLOBAL: the gateways resolve mx* from this copy
        _shim = C.CDLL(str(MEX / "libmxshim.so"), mode=C.RTLD_GLOBAL)
        _shim.mxshim_numeric.restype = C.c_void_p
        _shim.mxshim_numeric.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_longlong), C.c_void_p]
        _shim.mxshim_struct.restype = C.c_void_p
        _shim.mxshim_struct.argtypes = [C.c_int]
        _shim.mxshim_set_field.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        _shim.mxGetData.restype = C.c_void_p
        _shim.mxGetData.argtypes = [C.c_void_p]
        _shim.mxshim_bytes.restype = C.c_size_t
        _shim.mxshim_bytes.argtypes = [C.c_void_p]
        _shim.mxDestroyArray.argtypes = [C.c_void_p]
        _shim.mxshim_call.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_void_p), C.c_int,
                                      C.POINTER(C.c_void_p), C.c_char_p, C.c_size_t]
    return _shim


def gateway(name: str):
    if name not in _gw:
        shim()
        lib = C.CDLL(str(MEX / f"{name}.so"))
        _gw[name] = C.cast(lib.mexFunction, C.c_void_p)
        _gw[name + ".lib"] = lib
    return _gw[name]


def numeric(arr, matlab_dims=None):
    """mxArray from a numpy array; matlab_dims defaults to the reversed shape
    (a C-order array of shape (a, b, c) is MATLAB's column-major (c, b, a))."""
    arr = np.ascontiguousarray(arr)
    dims = list(matlab_dims) if matlab_dims is not None else list(arr.shape[::-1]) or [1, 1]
    if len(dims) == 1:
        dims = dims + [1]
    d = (C.c_longlong * len(dims))(*dims)
    return shim().mxshim_numeric(CLS[arr.dtype.type], len(dims), d, arr.ctypes.data)


def scalar(v, dtype):
    return numeric(np.array([v], dtype=dtype), [1, 1])


def struct(fields):
    s = shim().mxshim_struct(len(fields))
    for i, f in enumerate(fields):
        shim().mxshim_set_field(s, i, numeric(np.asarray(f, dtype=np.float64), [1, len(f)]))
    return s


def to_numpy(mx, dtype, shape):
    n = shim().mxshim_bytes(mx)
    buf = (C.c_char * n).from_address(shim().mxGetData(mx))
    return np.frombuffer(bytes(buf), dtype=dtype).reshape(shape)


def call(name, nlhs, args):
    """Returns (list of output mxArray handles, error message or None)."""
    plhs = (C.c_void_p * max(nlhs, 1))()
    prhs = (C.c_void_p * max(len(args), 1))(*args)
    err = C.create_string_buffer(1024)
    rc = shim().mxshim_call(gateway(name), nlhs, plhs, len(args), prhs, err, len(err))
    return list(plhs)[:nlhs], (err.value.decode() if rc else None)
