"""CPU oracle of the helical FDK path -- TEST INFRASTRUCTURE ONLY.

A ctypes wrapper around ``liboracle.so`` (built from ``arecon_oracle.c``, which
restates the reference's CPU code; see that file's header for the reference
lines each function follows and for how each function is pinned).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
may import this package, and only as the checker or the timed CPU baseline.
The product library never loads it.
"""
from __future__ import annotations

import ctypes as C
import math
from pathlib import Path

import numpy as np

ORACLE_LIB = Path(__file__).resolve().parent / "liboracle.so"

_f, _d, _i = C.c_float, C.c_double, C.c_int
_fp, _dp = C.POINTER(C.c_float), C.POINTER(C.c_double)

_SIGS = {
    "oracle_rebin_f32": [_fp, _fp, _i, _i, _i, _f, _f, _f, _f, _f, _f, _f],
    "oracle_rebin_f64": [_dp, _dp, _i, _i, _i, _d, _d, _d, _d, _d, _d, _d],
    "oracle_preweight_f64": [_dp, _i, _i, _i, _d, _d, _d, _d, _d],
    "oracle_filter_nn": [_i],
    "oracle_filter_kernel": [_dp, _i, _d, _i],
    "oracle_filter_f64": [_dp, _dp, _i, _i, _i, _d, _d, _d, _d, _d, _i],
    "oracle_fdk_weights": [_dp, _i, _d, _i],
    "oracle_fdk_core": [_dp, _dp, _dp, _i, _i, _i, _i, _d, _d, _d, _d, _d, _d, _i, _i, _i,
                        _d, _d, _d, _d, _d, _d, _d, _d, _d, _d, _d, _d, _i, _i, _i],
    "oracle_fdk": [_dp, _dp, _d, _d, _i, _i, _d, _d, _d, _d, _d, _d, _i, _i, _d, _i, _i, _i,
                   _d, _d, _d, _i, _d, _d, _i],
}

_lib = None


def build_variant(mutation: int, outdir) -> Path:
    """Compile arecon_oracle.c with -DORACLE_MUTATION=<mutation> (a planted
    mistake, used by the tests that show each pin can fail) into outdir."""
    import subprocess
    src = Path(__file__).resolve().parent / "arecon_oracle.c"
    out = Path(outdir) / f"liboracle_m{mutation}.so"
    subprocess.run(["gcc", "-O2", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
                    "-std=c99", "-D_GNU_SOURCE", f"-DORACLE_MUTATION={mutation}", str(src),
                    "-o", str(out), "-lm"], check=True)
    return out


class using:
    """Context manager that routes this module's functions to another build of
    the oracle (e.g. one from build_variant)."""

    def __init__(self, path):
        self.path = Path(path)

    def __enter__(self):
        global _lib
        self._saved = _lib
        _lib = _bind(C.CDLL(str(self.path)))
        return self

    def __exit__(self, *exc):
        global _lib
        _lib = self._saved
        return False


def _bind(h: C.CDLL) -> C.CDLL:
    for name, args in _SIGS.items():
        fn = getattr(h, name)
        fn.argtypes = args
        fn.restype = _i if name == "oracle_filter_nn" else None
    return h


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not ORACLE_LIB.exists():
            raise RuntimeError(f"{ORACLE_LIB} is missing: run __graft_entry__.build()")
        _lib = _bind(C.CDLL(str(ORACLE_LIB)))
    return _lib


def _f32(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(_fp)


def _f64(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def rebin_f32(proj, YLC0, dYA, DeltaTheta, PLC0, DeltaT, DeltaFai, SO):
    """proj [ViewN][YL][ZL] float32 -> rebinned, same layout (0-based YLC, PLC)."""
    proj, pin = _f32(proj)
    ViewN, YL, ZL = proj.shape
    out = np.empty_like(proj)
    lib().oracle_rebin_f32(out.ctypes.data_as(_fp), pin, YL, ZL, ViewN, YLC0, dYA,
                           DeltaTheta, PLC0, DeltaT, DeltaFai, SO)
    return out


def rebin_f64(proj, YLC0, dYA, DeltaTheta, PLC0, DeltaT, DeltaFai, SO):
    proj, pin = _f64(proj)
    ViewN, YL, ZL = proj.shape
    out = np.empty_like(proj)
    lib().oracle_rebin_f64(out.ctypes.data_as(_dp), pin, YL, ZL, ViewN, YLC0, dYA,
                           DeltaTheta, PLC0, DeltaT, DeltaFai, SO)
    return out


def preweight_f64(proj, PLC0, ZLC0, dYL, dZL, SO):
    """proj [ViewN][ZL][YL] -> preweighted copy (FP64)."""
    p = np.array(proj, dtype=np.float64, order="C", copy=True)
    ViewN, ZL, YL = p.shape
    lib().oracle_preweight_f64(p.ctypes.data_as(_dp), YL, ZL, ViewN, PLC0, ZLC0, dYL, dZL, SO)
    return p


def filter_nn(YL: int) -> int:
    return lib().oracle_filter_nn(YL)


def filter_kernel(YL: int, dYL: float, NN: int | None = None):
    NN = filter_nn(YL) if NN is None else NN
    HS = np.empty(NN, dtype=np.float64)
    lib().oracle_filter_kernel(HS.ctypes.data_as(_dp), YL, dYL, NN)
    return HS


def filter_f64(proj, PLC0, ZLC0, dYL, dZL, SO, nthreads: int = 1):
    """proj [ViewN][ZL][YL] -> preweighted + filtered (FP64)."""
    proj, pin = _f64(proj)
    ViewN, ZL, YL = proj.shape
    out = np.empty_like(proj)
    lib().oracle_filter_f64(out.ctypes.data_as(_dp), pin, YL, ZL, ViewN, PLC0, ZLC0, dYL, dZL,
                            SO, nthreads)
    return out


def fdk_weights(N_2pi: int, HSCoef: float, delta: int):
    w = np.empty(N_2pi, dtype=np.float64)
    lib().oracle_fdk_weights(w.ctypes.data_as(_dp), N_2pi, HSCoef, delta)
    return w


def fdk_core(fp, w, *, N_2pi, XN, YN, ZN, XNC, YNC, ZNC, dx, dy, dz, YLC, ZLC, dYL, dZL, h,
             BetaS, DeltaFai, N_pi, HSCoef, SO, SD, k1, zi_range=None, nthreads: int = 1,
             out=None):
    """Backprojection core (src/Parallel_FDK_Helical_3DWeighting.cpp:42-113).
    fp [PN][ZL][YL] -> volume [YN][XN][ZN] (FP64)."""
    fp, pfp = _f64(fp)
    PN, ZL, YL = fp.shape
    w, pw = _f64(w)
    img = np.zeros((YN, XN, ZN), dtype=np.float64) if out is None else out
    z0, z1 = (0, ZN) if zi_range is None else zi_range
    lib().oracle_fdk_core(img.ctypes.data_as(_dp), pfp, pw, N_2pi, XN, YN, ZN, XNC, YNC, ZNC,
                          dx, dy, dz, YL, ZL, PN, YLC, ZLC, dYL, dZL, h, BetaS, DeltaFai, N_pi,
                          HSCoef, SO, SD, k1, z0, z1, nthreads)
    return img


def fdk_from_params(fp, p, *, dz=None, zi_range=None, nthreads: int = 1, w=None):
    """Core backprojection for a ``geometry.BPParams`` (FP64 throughout)."""
    w = fdk_weights(p.N_2pi, p.HSCoef, p.delta) if w is None else w
    return fdk_core(fp, w, N_2pi=p.N_2pi, XN=p.XN, YN=p.YN, ZN=p.ZN, XNC=p.XNC, YNC=p.YNC,
                    ZNC=p.ZNC, dx=p.dx, dy=p.dy, dz=p.dz if dz is None else dz, YLC=p.YLC,
                    ZLC=p.ZLC, dYL=p.dYL, dZL=p.dZL, h=p.h, BetaS=p.BetaS,
                    DeltaFai=p.DeltaFai, N_pi=p.N_pi, HSCoef=p.HSCoef, SO=p.SO, SD=p.SD,
                    k1=p.k1, zi_range=zi_range, nthreads=nthreads)


def fdk_mid(fp, *, SO, DO, DecWidth, DecHeigh, YLC, ZLC, h, BetaS, N_2pi, ObjR, XN, YN, ZN,
            XNC, YNC, ZNC, delta, HSCoef, k1, nthreads: int = 1):
    """Mid-level entry (src/Parallel_FDK_Helical_3DWeighting.cpp:117-200, useGPU = 0)."""
    fp, pfp = _f64(fp)
    PN, ZL, YL = fp.shape
    img = np.zeros((YN, XN, ZN), dtype=np.float64)
    lib().oracle_fdk(img.ctypes.data_as(_dp), pfp, SO, DO, YL, ZL, DecWidth, DecHeigh, YLC, ZLC,
                     h, BetaS, PN, N_2pi, ObjR, XN, YN, ZN, XNC, YNC, ZNC, delta, HSCoef, k1,
                     nthreads)
    return img


# ---------------------------------------------------------------------------
# The reference's own CPU functions, compiled by build_ref.sh from the source
# files under the reference tree into _ref/libarecon_ref.so (git-ignored; it
# travels with the tree as a built library).  They pin the restatement above
# bit for bit and serve as the timed CPU baseline ("kind": "reference").
# ---------------------------------------------------------------------------
REF_LIB = Path(__file__).resolve().parent / "_ref" / "libarecon_ref.so"
REF_BUILD = Path(__file__).resolve().parent / "build_ref.sh"
_REF_SIGS = {
    "arecon_ref_rebin_f32": [_fp, _fp, _i, _i, _i, _f, _f, _f, _f, _f, _f, _f],
    "arecon_ref_rebin_f64": [_dp, _dp, _i, _i, _i, _d, _d, _d, _d, _d, _d, _d],
    "arecon_ref_fdk_cpu": [_dp, _dp, _dp, _i, _i, _i, _i, _d, _d, _d, _d, _d, _d, _i, _i, _i,
                           _d, _d, _d, _d, _d, _d, _d, _d, _d, _d, _d, _d],
}
_ref = None


def build_ref(reference_src: str | None = None) -> Path:
    """Run build_ref.sh (needs the reference sources; returns the library path)."""
    import os
    import subprocess
    env = dict(os.environ)
    if reference_src:
        env["AREF_REFERENCE_SRC"] = str(reference_src)
    subprocess.run(["sh", str(REF_BUILD)], check=True, env=env)
    return REF_LIB


def ref_available() -> bool:
    return REF_LIB.exists()


def ref_lib() -> C.CDLL:
    global _ref
    if _ref is None:
        if not REF_LIB.exists():
            raise RuntimeError(f"{REF_LIB} is missing: run oracle/build_ref.sh")
        h = C.CDLL(str(REF_LIB))
        for name, args in _REF_SIGS.items():
            fn = getattr(h, name)
            fn.argtypes = args
            fn.restype = None
        _ref = h
    return _ref


def ref_rebin_f32(proj, YLC0, dYA, DeltaTheta, PLC0, DeltaT, DeltaFai, SO):
    """The reference's ParallelRebinningCBCurve_CPU<float>
    (src/ParallelRebinningCBCurve.cpp:22-97); proj [ViewN][YL][ZL] float32."""
    proj, pin = _f32(proj)
    ViewN, YL, ZL = proj.shape
    out = np.empty_like(proj)
    ref_lib().arecon_ref_rebin_f32(out.ctypes.data_as(_fp), pin, YL, ZL, ViewN, YLC0, dYA,
                                   DeltaTheta, PLC0, DeltaT, DeltaFai, SO)
    return out


def ref_rebin_f64(proj, YLC0, dYA, DeltaTheta, PLC0, DeltaT, DeltaFai, SO):
    proj, pin = _f64(proj)
    ViewN, YL, ZL = proj.shape
    out = np.empty_like(proj)
    ref_lib().arecon_ref_rebin_f64(out.ctypes.data_as(_dp), pin, YL, ZL, ViewN, YLC0, dYA,
                                   DeltaTheta, PLC0, DeltaT, DeltaFai, SO)
    return out


def ref_fdk_core(fp, w, *, N_2pi, XN, YN, ZN, XNC, YNC, ZNC, dx, dy, dz, YLC, ZLC, dYL, dZL,
                 h, BetaS, DeltaFai, N_pi, HSCoef, SO, SD, k1, y_range=None,
                 nthreads: int = 1):
    """The reference's ParallelFDKHelical3DWeightingCPU
    (src/Parallel_FDK_Helical_3DWeighting.cpp:42-113) on fp [PN][ZL][YL] (FP64).

    Returns the rows y_range = (y0, y1) of the volume [YN][XN][ZN] (all rows by
    default).  nthreads > 1 runs runs of y rows on threads (ctypes releases the
    GIL); each call gets the row offset through YNC, which reproduces the serial
    run bit for bit."""
    from concurrent.futures import ThreadPoolExecutor
    fp, pfp = _f64(fp)
    PN, ZL, YL = fp.shape
    w, pw = _f64(w)
    y0, y1 = (0, YN) if y_range is None else y_range
    img = np.zeros((y1 - y0, XN, ZN), dtype=np.float64)
    n = max(1, min(int(nthreads), y1 - y0))
    cuts = np.linspace(0, y1 - y0, n + 1).round().astype(int)
    fn = ref_lib().arecon_ref_fdk_cpu

    def part(a: int, b: int) -> None:
        if b <= a:
            return
        dst = img[a:b].ctypes.data_as(_dp)
        fn(dst, pfp, pw, N_2pi, XN, b - a, ZN, XNC, YNC - (y0 + a), ZNC, dx, dy, dz, YL, ZL,
           PN, YLC, ZLC, dYL, dZL, h, BetaS, DeltaFai, N_pi, HSCoef, SO, SD, k1)

    if n == 1:
        part(0, y1 - y0)
    else:
        with ThreadPoolExecutor(n) as ex:
            list(ex.map(lambda i: part(cuts[i], cuts[i + 1]), range(n)))
    return img


def ref_fdk_from_params(fp, p, *, dz=None, w=None, y_range=None, nthreads: int = 1):
    """ref_fdk_core for a ``geometry.BPParams``."""
    w = fdk_weights(p.N_2pi, p.HSCoef, p.delta) if w is None else w
    return ref_fdk_core(fp, w, N_2pi=p.N_2pi, XN=p.XN, YN=p.YN, ZN=p.ZN, XNC=p.XNC,
                        YNC=p.YNC, ZNC=p.ZNC, dx=p.dx, dy=p.dy,
                        dz=p.dz if dz is None else dz, YLC=p.YLC, ZLC=p.ZLC, dYL=p.dYL,
                        dZL=p.dZL, h=p.h, BetaS=p.BetaS, DeltaFai=p.DeltaFai, N_pi=p.N_pi,
                        HSCoef=p.HSCoef, SO=p.SO, SD=p.SD, k1=p.k1, y_range=y_range,
                        nthreads=nthreads)


def sl_response(omega, dYL: float):
    """Frequency response of h(n) = -2/(pi^2 (4n^2-1))/dYL: |sin(w/2)| / (pi dYL)."""
    return np.abs(np.sin(np.asarray(omega) / 2.0)) / (math.pi * dYL)
