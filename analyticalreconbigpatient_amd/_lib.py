"""ctypes binding of libarecon.so (the gfx950 HIP library).

The library is loaded from the package directory only.  There is no fallback:
if ``libarecon.so`` is missing or does not export a declared symbol,
``lib()`` raises, so nothing can silently run a CPU substitute.
"""
from __future__ import annotations

import ctypes as C
import threading
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libarecon.so"

ARECON_OK = 0
ARECON_F32, ARECON_F16, ARECON_BF16 = 0, 1, 2
ARECON_LAYOUT_VYZ, ARECON_LAYOUT_VZY = 0, 1
ARECON_FP_ROWS, ARECON_FP_ROWPAIRS, ARECON_FP_QUADS, ARECON_FP_ROWDIFF = 0, 1, 2, 3

_f = C.c_float
_d = C.c_double
_i = C.c_int
_p = C.c_void_p
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)   # host int array (volume dims)
_dp = C.POINTER(C.c_double)


class BPGeom(C.Structure):
    """Mirror of ``arecon_bp_geom`` (include/arecon.h)."""

    _fields_ = [
        ("N_2pi", _i), ("XN", _i), ("YN", _i), ("ZN", _i), ("YL", _i), ("ZL", _i), ("PN", _i),
        ("XNC", _d), ("YNC", _d), ("ZNC", _d),
        ("dx", _d), ("dy", _d), ("dz", _d),
        ("YLC", _d), ("ZLC", _d),
        ("dYL", _d), ("dZL", _d),
        ("h", _d),
        ("BetaS", _d), ("DeltaFai", _d),
        ("N_pi", _d), ("HSCoef", _d),
        ("SO", _d), ("SD", _d), ("k1", _d),
        ("nz", _i),
        ("order", _i),
        ("variant", _i),
        ("group", _i),
    ]


# name -> (restype, argtypes); every symbol declared in include/arecon.h.
SIGNATURES = {
    "ParallelRebinningCBCurve_GPU": (None, [_fp, _fp, _i, _i, _i, _f, _f, _f, _f, _f, _f, _f]),
    "filtering": (None, [_fp, _fp, _i, _i, _i, _f, _f, _f, _f, _f]),
    "ParallekFDKHelical3DWeighting_GPU_float": (
        None, [_fp, _fp, _fp, _i, _i, _i, _i, _f, _f, _f, _f, _f, _f, _i, _i, _i,
               _f, _f, _f, _f, _f, _f, _f, _f, _f, _f, _f, _f]),
    "ParallekFDKHelical3DWeighting_GPU_double": (
        None, [_dp, _dp, _dp, _i, _i, _i, _i, _d, _d, _d, _d, _d, _d, _i, _i, _i,
               _d, _d, _d, _d, _d, _d, _d, _d, _d, _d, _d, _d]),
    "ParallelFDKHelical3DWeighting": (
        None, [_dp, _dp, _d, _d, _i, _i, _d, _d, _d, _d, _d, _d, _i, _i, _d, _i, _i, _i,
               _d, _d, _d, _i, _d, _d, _i]),
    "arecon_last_error": (_i, [C.c_char_p, C.c_size_t]),
    "arecon_version": (_i, []),
    "arecon_device_count": (_i, []),
    "arecon_rebin_dev": (_i, [_p, _p, _i, _i, _i, _f, _f, _f, _f, _f, _f, _f, _i, _p]),
    "arecon_filter_dev": (_i, [_p, _i, _i, _p, _i, _i, _i, _f, _f, _f, _f, _f, _p]),
    "arecon_filter_dev_noalloc": (_i, [_p, _i, _i, _p, _i, _i, _i, _f, _f, _f, _f, _f, _p]),
    "arecon_rebin_views_dev": (_i, [_p, _p, _i, _i, _i, _f, _f, _f, _f, _f, _f, _f, _i, _i, _p]),
    "arecon_bp_plan_create": (_i, [C.POINTER(BPGeom), _dp, C.POINTER(_p)]),
    "arecon_bp_plan_destroy": (_i, [_p]),
    "arecon_bp_plan_nz": (_i, [_p]),
    "arecon_bp_plan_execute": (_i, [_p, _p, _p, _i, _i, _p]),
    "arecon_bp_plan_nblocks": (_i, [_p]),
    "arecon_bp_plan_block_views": (_i, [_p, _i, C.POINTER(_i), C.POINTER(_i)]),
    "arecon_bp_plan_execute_blocks": (_i, [_p, _p, _p, _i, _i, _i, _i, _p]),
    "arecon_bp_plan_stats": (_i, [_p, C.POINTER(C.c_ulonglong), C.POINTER(C.c_ulonglong)]),
    "arecon_bp_plan_inside": (_i, [_p]),
    "arecon_bp_batch_create": (_i, [C.POINTER(_p), C.POINTER(_p), C.POINTER(_p), _i, _i, _i,
                                    C.POINTER(_p)]),
    "arecon_bp_batch_destroy": (_i, [_p]),
    "arecon_bp_batch_execute": (_i, [_p, _p]),
    "arecon_graph_capture_begin": (_i, [_p]),
    "arecon_graph_capture_end": (_i, [_p, C.POINTER(_p)]),
    "arecon_graph_launch": (_i, [_p, _p]),
    "arecon_graph_destroy": (_i, [_p]),
    # include/arecon_cnn.h
    "arecon_cnn_conv3": (_i, [_p, _p, _p, _p, _p, _ip, _i, _i, _i, _p]),
    "arecon_cnn_conv3_wgrad": (_i, [_p, _p, _p, _ip, _i, _i, _p]),
    "arecon_cnn_relu_bwd": (_i, [_p, _p, _p, _p, _p, _ip, _i, _i, _p]),
    "arecon_cnn_channel_major3": (_i, [_p, _p, _ip, _i, _p]),
    "arecon_cnn_pool": (_i, [_p, _p, _ip, _i, _p]),
    "arecon_cnn_pool_bwd": (_i, [_p, _p, _ip, _i, _i, _p]),
    "arecon_cnn_upsample": (_i, [_p, _p, _ip, _i, _p]),
    "arecon_cnn_upsample_bwd": (_i, [_p, _p, _ip, _i, _p]),
    "arecon_cnn_add": (_i, [_p, _p, C.c_longlong, _p]),
    "arecon_cnn_pack_weights": (_i, [_p, _p, _p, _i, _i, _p]),
    "arecon_cnn_mse_residual": (_i, [_p, _p, _p, _i, _p, _p, C.c_longlong, _p]),
    "arecon_cnn_adam": (_i, [_p, _p, _p, _p, C.c_longlong, _f, _f, _f, _f, _i, _p]),
    "arecon_cnn_to_bf16": (_i, [_p, _p, C.c_longlong, _i, _p]),
    "arecon_cnn_residual_out": (_i, [_p, _p, _i, _p, C.c_longlong, _p]),
    "arecon_bp_dev_f64": (_i, [_p, _p, C.POINTER(BPGeom), _dp, _p]),
    "arecon_fdk_weights": (_i, [_dp, _i, _d, _i]),
    "arecon_probe_lds": (_i, [_p, _i, _i, _i, _dp, _p]),
    "arecon_probe_l2": (_i, [_p, _p, C.c_longlong, _i, _i, _dp, _p]),
    "arecon_probe_copy": (_i, [_p, _p, C.c_longlong, _p]),
    "arecon_probe_valu": (_i, [_p, _i, _i, _dp, _p]),
    "arecon_project_ellipsoids_dev": (_i, [_p, _fp, _i, _i, _i, _i, _d, _d, _d, _d, _d, _d,
                                           _d, _d, _d, _p]),
}

_lock = threading.Lock()
_lib: C.CDLL | None = None


class AreconError(RuntimeError):
    """A call into libarecon.so reported a failure."""


def lib() -> C.CDLL:
    """Load libarecon.so (once) and bind every declared symbol."""
    global _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise AreconError(
                    f"{LIB_PATH} is missing: build it with "
                    "`python -c 'import __graft_entry__ as g; g.build()'`")
            handle = C.CDLL(str(LIB_PATH))
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(handle, name)
                fn.restype = res
                fn.argtypes = args
            _lib = handle
        return _lib


def last_error() -> tuple[int, str]:
    buf = C.create_string_buffer(2048)
    code = lib().arecon_last_error(buf, len(buf))
    return code, buf.value.decode(errors="replace")


def check(status: int | None, what: str) -> None:
    """Raise AreconError if ``status`` (or, for void drop-ins, the thread's
    last-error slot) reports a failure."""
    if status is None:
        code, msg = last_error()
    else:
        code = status
        msg = last_error()[1] if status != ARECON_OK else ""
    if code != ARECON_OK:
        raise AreconError(f"{what}: {msg} (code {code})")


def fptr(arr) -> C.POINTER(C.c_float):
    return arr.ctypes.data_as(_fp)


def dptr(arr) -> C.POINTER(C.c_double):
    return arr.ctypes.data_as(_dp)
