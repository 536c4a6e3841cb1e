"""The reference's MATLAB-level operator interface, over the C ABI.

Each function mirrors one MEX gateway or MATLAB helper of the reference, with
the same name, argument order and meaning, on numpy arrays whose C-order shape
is the MEX memory layout (last axis fastest):

=================================  ==========================================
``ParallelRebinningCBCurve``        src/ParallelRebinningCBCurve.cpp:101-129
``reWeigAdFiltr_GPU``               src/reWeigAdFiltr_GPU.cpp:36-63
``Parallel_FDK_Helical_3DWeighting``  src/Parallel_FDK_Helical_3DWeighting.cpp:204-269
``Parallel_Rebinning_CB_Curve``     src/Parallel_Rebinning_CB_Curve.m:1-25
``analyticalReconBigPatient``       src/analyticalReconBigPatient.m:10-150
=================================  ==========================================

MATLAB's ``Proj(ViewN, YL, ZL)`` is held as a numpy array of shape
``(ViewN, YL, ZL)`` with the same subscripts; the MEX layouts are C-order
``(ViewN, YL, ZL)`` (rebinning, ZL fastest), ``(ViewN, ZL, YL)`` (filter and
backprojection input, YL fastest) and ``(YN, XN, ZN)`` (volume, ZN fastest).

All compute runs on the GPU through ``libarecon.so``.  ``useGPU`` values other
than 1 are rejected: this package has no CPU implementation.
"""
from __future__ import annotations

import time

import numpy as np

from . import _lib
from .geometry import (GEHeader, ScanGeom, bp_params, bp_scan_geom, derive_setup, filter_args,
                       rebin_args, seg_count, trim_segment)


class MexArgumentError(ValueError):
    """Wrong arity or argument type at a gateway (the reference prints and
    calls exit(-1) or mexErrMsgTxt; this raises instead)."""


def _require_gpu(useGPU, who: str) -> None:
    if int(np.asarray(useGPU).reshape(-1)[0]) != 1:
        raise MexArgumentError(f"{who}: only useGPU == 1 is provided (no CPU implementation)")


def _scalar(v, dtype):
    return np.asarray(v, dtype=dtype).reshape(-1)[0]


def ParallelRebinningCBCurve(Proj, YL, ZL, ViewN, YLC, dYA, DeltaTheta, PLC, DeltaT, DeltaFai,
                             SO, useGPU):
    """Rebinning gateway.  ``YLC`` and ``PLC`` are 1-based (the gateway
    subtracts 1, :122); ``Proj`` is float32 ``(ViewN, YL, ZL)``."""
    _require_gpu(useGPU, "ParallelRebinningCBCurve")
    YL, ZL, ViewN = int(_scalar(YL, np.int32)), int(_scalar(ZL, np.int32)), int(_scalar(ViewN, np.int32))
    P = np.ascontiguousarray(Proj, dtype=np.float32)
    if P.size != YL * ZL * ViewN:
        raise MexArgumentError("ParallelRebinningCBCurve: Proj does not hold YL*ZL*ViewN samples")
    out = np.empty((ViewN, YL, ZL), dtype=np.float32)
    f = np.float32
    L = _lib.lib()
    L.ParallelRebinningCBCurve_GPU(_lib.fptr(out), _lib.fptr(P), YL, ZL, ViewN,
                                   f(_scalar(YLC, f) - f(1.0)), _scalar(dYA, f),
                                   _scalar(DeltaTheta, f), f(_scalar(PLC, f) - f(1.0)),
                                   _scalar(DeltaT, f), _scalar(DeltaFai, f), _scalar(SO, f))
    _lib.check(None, "ParallelRebinningCBCurve")
    return out


def reWeigAdFiltr_GPU(*args, nlhs: int = 1):
    """``FPWD = reWeigAdFiltr_GPU(Proj, PLC, ZLC, dYL, dZL, SO)``; ``PLC`` and
    ``ZLC`` are 0-based; ``Proj`` is float32 ``(ViewN, ZL, YL)``."""
    if nlhs != 1 or len(args) != 6:      # src/reWeigAdFiltr_GPU.cpp:41-46
        raise MexArgumentError("reWeigAdFiltr_GPU: requires 6 input parameters and provides "
                               "one return value")
    Proj, PLC, ZLC, dYL, dZL, SO = args
    P = np.ascontiguousarray(Proj, dtype=np.float32)
    if P.ndim != 3:
        raise MexArgumentError("reWeigAdFiltr_GPU: Proj must be 3-D (ViewN, ZL, YL)")
    ViewN, ZL, YL = P.shape
    out = np.empty_like(P)
    f = np.float32
    _lib.lib().filtering(_lib.fptr(out), _lib.fptr(P), YL, ZL, ViewN, _scalar(PLC, f),
                         _scalar(ZLC, f), _scalar(dYL, f), _scalar(dZL, f), _scalar(SO, f))
    _lib.check(None, "reWeigAdFiltr_GPU")
    return out


def Parallel_FDK_Helical_3DWeighting(*args, nlhs: int = 0):
    """``Parallel_FDK_Helical_3DWeighting(ScanGeom, fp, ctimage, useGPU)``:
    writes the volume into ``ctimage`` in place (float64, XN*YN*ZN elements,
    z fastest) and returns nothing."""
    if len(args) != 4:                   # src/Parallel_FDK_Helical_3DWeighting.cpp:211-216
        raise MexArgumentError("Backward projection needs 4 inputs.")
    if nlhs != 0:
        raise MexArgumentError("Backward projection does not need any output.")
    geom, fp, ctimage, useGPU = args
    _require_gpu(useGPU, "Parallel_FDK_Helical_3DWeighting")
    p = bp_params(geom)
    fp64 = np.ascontiguousarray(fp, dtype=np.float64)
    if fp64.size != p.YL * p.ZL * p.PN:
        raise MexArgumentError("Parallel_FDK_Helical_3DWeighting: fp does not hold YL*ZL*PN samples")
    if not (isinstance(ctimage, np.ndarray) and ctimage.dtype == np.float64
            and ctimage.flags.c_contiguous and ctimage.size == p.XN * p.YN * p.ZN):
        raise MexArgumentError("Parallel_FDK_Helical_3DWeighting: ctimage must be a contiguous "
                               "float64 array of XN*YN*ZN elements")
    _lib.lib().ParallelFDKHelical3DWeighting(
        _lib.dptr(ctimage), _lib.dptr(fp64), p.SO, p.DO, p.YL, p.ZL, p.DecWidth, p.DecHeigh,
        p.YLC, p.ZLC, p.h, p.BetaS, p.PN, p.N_2pi, p.ObjR, p.XN, p.YN, p.ZN, p.XNC, p.YNC,
        p.ZNC, p.delta, p.HSCoef, p.k1, 1)
    _lib.check(None, "Parallel_FDK_Helical_3DWeighting")


def Parallel_Rebinning_CB_Curve(Proj, ScanGeom_, useGPU):
    """MATLAB helper (src/Parallel_Rebinning_CB_Curve.m:1-25): ``Proj`` is
    ``(ViewN, YL, ZL)`` double; returns the rebinned ``(ViewN, YL, ZL)`` double."""
    a = rebin_args(ScanGeom_)
    # MATLAB permutes to (ZL, YL, ViewN) column-major = this C-order array.
    P = np.ascontiguousarray(Proj, dtype=np.float32)
    out = ParallelRebinningCBCurve(P, np.int32(a.YL), np.int32(a.ZL), np.int32(a.ViewN),
                                   np.float32(ScanGeom_.ScDet[6]), a.dYA, a.DeltaTheta,
                                   np.float32((1 + a.YL) * 0.5), a.DeltaT, a.DeltaFai, a.SO,
                                   useGPU)
    return out.astype(np.float64)


def analyticalReconBigPatient(Proj, CTGE: GEHeader, SegLength, SegOverLap, SegInd, DLength, objR,
                              xn, yn, zn, mWtInd=3, useGPU=1, mDetHei=None, mk1=5.0,
                              mdelta=60.0, mHSCoef=0.68, ZC=10.0):
    """One segment (src/analyticalReconBigPatient.m:10-150).  ``Proj`` is the
    segment's ``(DLength, YL, ZL)`` projections as read from the file (before
    the channel flip).  Returns ``(ctimage (YN, XN, ZN) float64, seconds)``.
    ``ZC`` is the driver's hard-coded value 10 (:65) unless given."""
    del mWtInd, mDetHei                     # dead parameters of the reference (:14-16, :52)
    _require_gpu(useGPU, "analyticalReconBigPatient")
    t0 = time.perf_counter()
    Proj = np.asarray(Proj)[:, ::-1, :]     # :30 channel flip
    s = derive_setup(CTGE, seg_length=SegLength, seg_overlap_views=SegOverLap, seg_ind=SegInd,
                     d_length=DLength, objR=objR, xn=xn, yn=yn, zn=zn, k1=mk1, delta=mdelta,
                     HSCoef=mHSCoef, ZC=ZC)
    Proj = Parallel_Rebinning_CB_Curve(Proj, s.geom, np.int32(1))          # :79
    seg_num = seg_count(CTGE, SegLength, SegOverLap)
    first, n, g = trim_segment(s.geom, SegInd, seg_num, SegOverLap, CTGE.NumberOfViewPerRotation)
    Proj = Proj[first:first + n]                                          # :80-93
    fa = filter_args(s)
    newProj = np.ascontiguousarray(np.transpose(Proj, (0, 2, 1)))          # permute [2 3 1]
    fpwd = reWeigAdFiltr_GPU(newProj.astype(np.float32), fa.PLC0, fa.ZLC0, fa.dYL, fa.dZL,
                             fa.SO).astype(np.float64)                     # :138-139
    geom_bp = bp_scan_geom(s, geom=g)                                      # :143
    ctimage = np.zeros((yn, xn, zn), dtype=np.float64)                     # :145
    Parallel_FDK_Helical_3DWeighting(geom_bp, fpwd, ctimage, np.int32(useGPU))
    return ctimage, time.perf_counter() - t0


__all__ = [
    "ParallelRebinningCBCurve", "reWeigAdFiltr_GPU", "Parallel_FDK_Helical_3DWeighting",
    "Parallel_Rebinning_CB_Curve", "analyticalReconBigPatient", "MexArgumentError", "ScanGeom",
]
