"""Whole-scan driver: overlapping view segments, reconstructed and summed.

Restates src/analyticalRecon.m:1-28 on the device-resident pipeline:
SegOverLap = ceil(atan(pitch/SDD/2) * YL * N_2pi / 2pi) + 1 views (:15-16), SegNum
segments of SegLength views read from the projection file (:20-22), each
reconstructed with its BetaS shift and overlap trimming
(src/analyticalReconBigPatient.m:69, :80-93) onto the same grid and summed into
RecImg (:23).  Each view's contribution depends only on (voxel, view), so the
sum over the disjoint trimmed view ranges is the reconstruction of the scan.
"""
from __future__ import annotations

import time

import numpy as np
import torch

from . import geometry as G
from .gedata import GetHeadofGeDataSet, readBigPatientData
from .pipeline import HelicalFDK


def reconstruct_segment(Proj_file_order: np.ndarray, CTGE: G.GEHeader, SegLength: int,
                        SegOverLap: int, SegInd: int, DLength: int, objR: float, xn: int,
                        yn: int, zn: int, *, ZC: float = 10.0, k1: float = 5.0,
                        delta: float = 60.0, HSCoef: float = 0.68, storage: str = "f32",
                        out: torch.Tensor | None = None, device="cuda") -> torch.Tensor:
    """One segment (src/analyticalReconBigPatient.m:10-150) on the device.
    Proj_file_order is (DLength, YL, ZL) as read (before the channel flip).
    The volume is added into ``out`` when given (RecImg += ctimage)."""
    s = G.derive_setup(CTGE, seg_length=SegLength, seg_overlap_views=SegOverLap, seg_ind=SegInd,
                       d_length=DLength, objR=objR, xn=xn, yn=yn, zn=zn, k1=k1, delta=delta,
                       HSCoef=HSCoef, ZC=ZC)
    seg_num = G.seg_count(CTGE, SegLength, SegOverLap)
    first, n, g = G.trim_segment(s.geom, SegInd, seg_num, SegOverLap,
                                 CTGE.NumberOfViewPerRotation)
    flipped = np.ascontiguousarray(Proj_file_order[:, ::-1, :], dtype=np.float32)  # :30
    host = torch.from_numpy(flipped).pin_memory()
    raw = host.to(device, non_blocking=True)
    fdk = HelicalFDK(s, storage=storage, first_view=first, n_views=n, bp_geom=g, device=device)
    vol = fdk.run(raw)
    if out is None:
        return vol.clone()
    out += vol
    return out


def analyticalRecon(GEProjectionDataName, GEProjHeadDataName, SegLength: int, objR: float,
                    Scale: int, useGPU: int = 1, *, ZC: float = 10.0, storage: str = "f32",
                    device="cuda"):
    """Returns (RecImg (Scale, Scale, Scale) volume on the device, seconds).
    Volume layout [YN][XN][ZN] as the backprojection writes it."""
    if int(useGPU) != 1:
        raise ValueError("analyticalRecon: only useGPU == 1 is provided (no CPU implementation)")
    t0 = time.perf_counter()
    xn = yn = zn = Scale                                               # :8-10
    CTGE = GetHeadofGeDataSet(GEProjHeadDataName)                      # :14
    SegOverLap = G.seg_overlap(CTGE)                                   # :15-16
    SegNum = G.seg_count(CTGE, SegLength, SegOverLap)                  # :17
    RecImg = torch.zeros((yn, xn, zn), dtype=torch.float32, device=device)
    for SegInd in range(1, SegNum + 1):                                # :20
        Proj, DLength = readBigPatientData(GEProjectionDataName, CTGE, SegLength, SegOverLap,
                                           SegInd)
        reconstruct_segment(np.asarray(Proj), CTGE, SegLength, SegOverLap, SegInd, DLength,
                            objR, xn, yn, zn, ZC=ZC, storage=storage, out=RecImg, device=device)
    torch.cuda.synchronize()
    return RecImg, time.perf_counter() - t0
