"""Scan geometry: the derivations the reference's MATLAB drivers perform.

Everything here is FP64 host arithmetic, restating

* ``src/analyticalReconBigPatient.m:30-101``  (header -> ScanGeom struct),
* ``src/Parallel_Rebinning_CB_Curve.m:8-18``  (rebinning arguments),
* ``src/analyticalReconBigPatient.m:95-105,139`` (filter arguments),
* ``src/Parallel_FDK_Helical_3DWeighting.cpp:222-263,145-154`` (backprojection
  arguments parsed from the struct),
* ``src/analyticalRecon.m:15-17`` (segment overlap and count).

``ScanGeom`` holds the four numeric fields of the MATLAB struct
(``ScDet``, ``Proj``, ``Obj``, ``Rec``) with MATLAB's 1-based centres.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field, replace

import numpy as np

from ._lib import BPGeom


@dataclass
class GEHeader:
    """The header fields the reference reads (src/GetHeadofGeDataSet.m:20-30)."""

    Source2DetDis: float           # SDD, mm
    Source2IsoDis: float           # SO, mm
    DetCellPitch: float            # channel pitch at the detector, mm
    BeginViewAngle: float          # degrees
    NumberOfViewPerRotation: int
    TotalViewNumber: int
    NumberOfDetChannelPerRow: int
    NumberOfDetectorRow: int
    IsoChannelLocation: float      # as stored in the header (before the flip)
    SliceHeigh: float              # row height at iso, mm
    HelicalPitch: float            # table travel per rotation in units of SliceHeigh


@dataclass
class ScanGeom:
    """The MATLAB ScanGeom struct (analyticalReconBigPatient.m:72-75, 98-101)."""

    ScDet: np.ndarray   # [SO DO YL ZL DecAngle|DecWidth DecHeigh YLC ZLC h]
    Proj: np.ndarray    # [BetaS BetaE ViewN N_2pi]
    Obj: np.ndarray     # [ObjR XN YN ZN XC YC ZC]
    Rec: np.ndarray     # [delta HSCoef k1]

    def copy(self) -> "ScanGeom":
        return ScanGeom(self.ScDet.copy(), self.Proj.copy(), self.Obj.copy(), self.Rec.copy())


@dataclass
class ReconSetup:
    """ScanGeom before rebinning plus the derived scalars of the driver."""

    geom: ScanGeom           # as passed to the rebinning (ScDet(5) = DecAngle)
    DecAngle: float
    DecWidth: float
    dYL: float
    dZL: float
    PLC: float               # 1-based, (1+YL)/2
    ZLC: float               # 1-based, (1+ZL)/2
    extras: dict = field(default_factory=dict)


def seg_overlap(hdr: GEHeader) -> int:
    """SegOverLap, src/analyticalRecon.m:15-16."""
    v = math.atan(hdr.DetCellPitch / hdr.Source2DetDis / 2) * hdr.NumberOfDetChannelPerRow
    return int(math.ceil(v * hdr.NumberOfViewPerRotation / (2 * math.pi)) + 1)


def seg_count(hdr: GEHeader, seg_length: int, overlap: int) -> int:
    """SegNum, src/analyticalRecon.m:17 and analyticalReconBigPatient.m:68."""
    return int(math.ceil((float(hdr.TotalViewNumber) - seg_length) / (seg_length - 2 * overlap)) + 1)


def derive_setup(hdr: GEHeader, *, seg_length: int, seg_overlap_views: int, seg_ind: int,
                 d_length: int, objR: float, xn: int, yn: int, zn: int,
                 k1: float = 5.0, delta: float = 60.0, HSCoef: float = 0.68,
                 ZC: float = 10.0) -> ReconSetup:
    """Geometry derivation of src/analyticalReconBigPatient.m:30-75.

    ``ZC`` defaults to the driver's hard-coded 10 (:65).  The channel flip of
    :30 is applied to the data by the caller; its effect on the iso channel
    (:31) is applied here.
    """
    iso = hdr.NumberOfDetChannelPerRow - hdr.IsoChannelLocation          # :31
    SO = hdr.Source2IsoDis                                              # :35
    DO = hdr.Source2DetDis - hdr.Source2IsoDis                          # :36
    YL = hdr.NumberOfDetChannelPerRow                                   # :37
    YLC = iso                                                           # :38
    ZL = hdr.NumberOfDetectorRow                                        # :39
    ZLC = (1 + ZL) * 0.5                                                # :40
    DetWid = hdr.DetCellPitch                                           # :42
    DetHei = hdr.SliceHeigh * (DO + SO) / SO                            # :43
    DecAngle = math.atan(DetWid / (SO + DO) / 2) * YL * 2               # :45
    DecHeigh = DetHei * ZL                                              # :46
    N_2pi = hdr.NumberOfViewPerRotation                                 # :47
    h = hdr.HelicalPitch / hdr.NumberOfDetectorRow * SO / (DO + SO)     # :48
    XC = (xn + 1) * 0.5                                                 # :63
    YC = (yn + 1) * 0.5
    ViewN = d_length                                                    # :67
    BetaS = (math.pi - hdr.BeginViewAngle * math.pi / 180.0
             + (seg_ind - 1) * (seg_length - 2 * seg_overlap_views) * 2 * math.pi / N_2pi)  # :69
    BetaE = BetaS + ViewN * 2 * math.pi / N_2pi                         # :70
    g = ScanGeom(
        ScDet=np.array([SO, DO, YL, ZL, DecAngle, DecHeigh, YLC, ZLC, h], dtype=np.float64),
        Proj=np.array([BetaS, BetaE, ViewN, N_2pi], dtype=np.float64),
        Obj=np.array([objR, xn, yn, zn, XC, YC, ZC], dtype=np.float64),
        Rec=np.array([delta, HSCoef, k1], dtype=np.float64),
    )
    DecWidth = math.tan(DecAngle * 0.5) * SO * 2 * 1.5                  # :95
    return ReconSetup(geom=g, DecAngle=DecAngle, DecWidth=DecWidth, dYL=DecWidth / YL,
                      dZL=DecHeigh / ZL, PLC=(1 + YL) * 0.5, ZLC=ZLC)


def trim_segment(geom: ScanGeom, seg_ind: int, seg_num: int, overlap: int, N_2pi: int):
    """Segment overlap trimming, src/analyticalReconBigPatient.m:80-93.

    Returns (first_view, n_views, new_geom).  The views kept are
    [first_view, first_view + n_views) of the rebinned array.
    """
    g = geom.copy()
    ViewN = int(g.Proj[2])
    step = overlap * 2 * math.pi / N_2pi
    if seg_ind == 1:
        first, n = 0, ViewN - overlap
        g.Proj[1] -= step
    elif seg_ind == seg_num:
        first, n = overlap, ViewN - overlap
        g.Proj[0] += step
    else:
        first, n = overlap, ViewN - 2 * overlap
        g.Proj[0] += step
        g.Proj[1] -= step
    g.Proj[2] = n
    return first, n, g


@dataclass
class RebinArgs:
    """Arguments of the rebinning MEX (src/Parallel_Rebinning_CB_Curve.m:8-18),
    as the single-precision values MATLAB passes (:21-22), with the MEX's
    0-based shift of YLC and PLC applied (src/ParallelRebinningCBCurve.cpp:122)."""

    YL: int
    ZL: int
    ViewN: int
    YLC0: np.float32
    dYA: np.float32
    DeltaTheta: np.float32
    PLC0: np.float32
    DeltaT: np.float32
    DeltaFai: np.float32
    SO: np.float32


def rebin_args(g: ScanGeom) -> RebinArgs:
    SO = g.ScDet[0]
    YL = int(g.ScDet[2])
    ZL = int(g.ScDet[3])
    dYA = g.ScDet[4] / YL                                   # :12
    YLC = g.ScDet[6]                                        # :13
    ViewN = int(g.Proj[2])                                  # :14
    DeltaFai = (g.Proj[1] - g.Proj[0]) / (ViewN - 1)        # :15
    DeltaTheta = DeltaFai                                   # :16
    DeltaT = math.tan(g.ScDet[4] * 0.5) * SO * 1.5 * 2 / YL  # :17
    PLC = (1 + YL) * 0.5                                    # :18
    f = np.float32
    # MEX: YLC - 1.0f and PLC - 1.0f in single precision.
    return RebinArgs(YL, ZL, ViewN, f(f(YLC) - f(1.0)), f(dYA), f(DeltaTheta),
                     f(f(PLC) - f(1.0)), f(DeltaT), f(DeltaFai), f(SO))


@dataclass
class FilterArgs:
    """Arguments of reWeigAdFiltr_GPU (src/analyticalReconBigPatient.m:139)."""

    PLC0: np.float32
    ZLC0: np.float32
    dYL: np.float32
    dZL: np.float32
    SO: np.float32


def filter_args(s: ReconSetup) -> FilterArgs:
    f = np.float32
    return FilterArgs(f(s.PLC - 1.0), f(s.ZLC - 1.0), f(s.dYL), f(s.dZL), f(s.geom.ScDet[0]))


def bp_scan_geom(s: ReconSetup, first_view: int = 0, n_views: int | None = None,
                 geom: ScanGeom | None = None) -> ScanGeom:
    """ScanGeom handed to the backprojection MEX: ScDet(5) = DecWidth
    (src/analyticalReconBigPatient.m:143)."""
    g = (geom or s.geom).copy()
    g.ScDet[4] = s.DecWidth
    if n_views is not None:
        g.Proj[2] = n_views
    return g


@dataclass
class BPParams:
    """Backprojection arguments as the MEX gateway derives them
    (src/Parallel_FDK_Helical_3DWeighting.cpp:222-263) together with the
    mid-level derivations (:145-154)."""

    SO: float
    DO: float
    YL: int
    ZL: int
    DecWidth: float
    DecHeigh: float
    YLC: float      # 0-based, (YL-1)/2 regardless of the struct (:229)
    ZLC: float      # 0-based (:230)
    h: float        # ScDet(9) * DecHeigh (:231)
    BetaS: float
    PN: int
    N_2pi: int
    ObjR: float
    XN: int
    YN: int
    ZN: int
    XNC: float
    YNC: float
    ZNC: float
    delta: int
    HSCoef: float
    k1: float       # Rec(3) * ScDet(9) (:250)

    @property
    def N_pi(self) -> float:
        return self.N_2pi / 2.0

    @property
    def dx(self) -> float:
        return 2.0 * self.ObjR / self.XN

    @property
    def dy(self) -> float:
        return 2.0 * self.ObjR / self.YN

    @property
    def dz(self) -> float:
        return 2.0 * self.ObjR / self.ZN

    @property
    def dYL(self) -> float:
        return self.DecWidth / self.YL

    @property
    def dZL(self) -> float:
        return self.DecHeigh / self.ZL

    @property
    def DeltaFai(self) -> float:
        return 2.0 * 3.141592653589793 / self.N_2pi

    @property
    def SD(self) -> float:
        return self.SO + self.DO

    def to_geom(self, *, dz: float | None = None, nz: int = 0, order: int = 0,
                variant: int = 0, group: int = 0) -> BPGeom:
        """FP64 ``arecon_bp_geom``; ``dz`` overrides the 2 ObjR / ZN coupling."""
        return BPGeom(
            N_2pi=self.N_2pi, XN=self.XN, YN=self.YN, ZN=self.ZN, YL=self.YL, ZL=self.ZL,
            PN=self.PN, XNC=self.XNC, YNC=self.YNC, ZNC=self.ZNC, dx=self.dx, dy=self.dy,
            dz=self.dz if dz is None else dz, YLC=self.YLC, ZLC=self.ZLC, dYL=self.dYL,
            dZL=self.dZL, h=self.h, BetaS=self.BetaS, DeltaFai=self.DeltaFai,
            N_pi=self.N_pi, HSCoef=self.HSCoef, SO=self.SO, SD=self.SD, k1=self.k1,
            nz=nz, order=order, variant=variant, group=group)


def bp_params(g: ScanGeom) -> BPParams:
    """Struct parse of src/Parallel_FDK_Helical_3DWeighting.cpp:222-250."""
    sd, pr, ob, rc = g.ScDet, g.Proj, g.Obj, g.Rec
    YL = int(sd[2] + 0.5)
    ZL = int(sd[3] + 0.5)
    return BPParams(
        SO=float(sd[0]), DO=float(sd[1]), YL=YL, ZL=ZL, DecWidth=float(sd[4]),
        DecHeigh=float(sd[5]), YLC=(YL - 1) * 0.5, ZLC=float(sd[7]) - 1.0,
        h=float(sd[8]) * float(sd[5]), BetaS=float(pr[0]), PN=int(pr[2]), N_2pi=int(pr[3]),
        ObjR=float(ob[0]), XN=int(ob[1] + 0.5), YN=int(ob[2] + 0.5), ZN=int(ob[3] + 0.5),
        XNC=float(ob[4]) - 1.0, YNC=float(ob[5]) - 1.0, ZNC=float(ob[6]) - 1.0,
        delta=int(rc[0] + 0.5), HSCoef=float(rc[1]), k1=float(rc[2]) * float(sd[8]))


def window_half_width(N_2pi: int, HSCoef: float) -> int:
    """W = ceil(N_pi * HSCoef); the window has L = 2W + 1 views
    (src/Parallel_FDK_Helical_3DWeighting.cpp:157)."""
    return int(math.ceil(N_2pi / 2.0 * HSCoef))


def voxel_updates(p: BPParams, dz: float | None = None) -> int:
    """Algorithmic voxel-updates: sum over voxels of the clamped window length
    (src/Parallel_FDK_Helical_3DWeighting.cpp:63-71)."""
    W = window_half_width(p.N_2pi, p.HSCoef)
    dzv = p.dz if dz is None else dz
    total = 0
    for zi in range(p.ZN):
        z = (zi - p.ZNC) * dzv
        beta0 = 2.0 * 3.14159265358979 * z / p.h
        s0 = math.ceil((beta0 - p.BetaS) / p.DeltaFai - 0.5)
        s1, s2 = s0 - W, s0 + W
        if s1 < p.PN and s2 > 0:
            total += min(s2, p.PN - 1) - max(s1, 0) + 1
    return total * p.XN * p.YN


# ---------------------------------------------------------------------------
# The synthetic workloads (configs are named by what they are).
# ---------------------------------------------------------------------------
@dataclass
class Workload:
    name: str
    hdr: GEHeader
    xn: int
    yn: int
    zn: int
    objR: float
    ZC: float           # 1-based volume centre in z
    k1: float = 5.0
    delta: float = 60.0
    HSCoef: float = 0.68

    def setup(self) -> ReconSetup:
        return derive_setup(self.hdr, seg_length=self.hdr.TotalViewNumber, seg_overlap_views=0,
                            seg_ind=1, d_length=self.hdr.TotalViewNumber, objR=self.objR,
                            xn=self.xn, yn=self.yn, zn=self.zn, k1=self.k1, delta=self.delta,
                            HSCoef=self.HSCoef, ZC=self.ZC)


def _header(YL: int, ZL: int, N_2pi: int, rotations: float, pitch: float) -> GEHeader:
    return GEHeader(Source2DetDis=949.0, Source2IsoDis=541.0, DetCellPitch=1.0239,
                    BeginViewAngle=0.0, NumberOfViewPerRotation=N_2pi,
                    TotalViewNumber=int(round(N_2pi * rotations)),
                    NumberOfDetChannelPerRow=YL, NumberOfDetectorRow=ZL,
                    IsoChannelLocation=(YL - 1) * 0.5, SliceHeigh=1.0963,
                    HelicalPitch=pitch * ZL)


def centred_ZC(hdr: GEHeader, zn: int, dz: float) -> float:
    """1-based ZC that centres a zn-slice slab of pitch dz on the helix travel
    (source z runs from h BetaS / 2pi over TotalViewNumber views)."""
    h = hdr.HelicalPitch * hdr.SliceHeigh
    BetaS = math.pi - hdr.BeginViewAngle * math.pi / 180.0
    z_first = h * BetaS / (2 * math.pi)
    z_mid = z_first + 0.5 * h * hdr.TotalViewNumber / hdr.NumberOfViewPerRotation
    # z(zi) = (zi - ZNC) dz with ZNC = ZC - 1; centre of slab at zi = (zn-1)/2.
    return (zn - 1) * 0.5 - z_mid / dz + 1.0


def shepp_logan_small() -> Workload:
    """368 x 16 curved detector, 720 views/rotation, pitch 1.0, 3 rotations;
    256 x 256 x 32 volume, 250 mm FOV (dz = 2 ObjR / ZN at the MEX level)."""
    hdr = _header(368, 16, 720, 3, 1.0)
    return Workload("shepp-logan 368x16x2160 -> 256x256x32", hdr, 256, 256, 32, 125.0, 10.0)


def body_clinical() -> Workload:
    """888 x 64 curved detector, 984 views/rotation, pitch 0.984, 12 rotations
    (11,808 views); 512 x 512 x 400 volume, 500 mm FOV, slab centred in the helix."""
    hdr = _header(888, 64, 984, 12, 0.984)
    zn, objR = 400, 250.0
    return Workload("body 888x64x11808 -> 512x512x400", hdr, 512, 512, zn, objR,
                    centred_ZC(hdr, zn, 2 * objR / zn))


def tiny(YL: int = 48, ZL: int = 8, N_2pi: int = 64, rotations: float = 3, pitch: float = 1.0,
         xn: int = 16, zn: int = 8, objR: float = 60.0) -> Workload:
    """Small geometry for fast parity tests (same scanner, fewer samples)."""
    hdr = _header(YL, ZL, N_2pi, rotations, pitch)
    # A wider channel pitch keeps the fan covering the test object.
    hdr = replace(hdr, DetCellPitch=1.0239 * 368 / YL)
    return Workload(f"tiny {YL}x{ZL}x{int(N_2pi * rotations)} -> {xn}^2x{zn}", hdr, xn, xn, zn,
                    objR, centred_ZC(hdr, zn, 2 * objR / zn))
