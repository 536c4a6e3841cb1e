"""GPU parity of each stage against the CPU oracle, through the C ABI.

Bars: rebinning bit-exact with the FP32 CPU path; filter within 1e-5 relative
L2 of the FP64 oracle (FP32 output), 2e-3 / 1e-2 for FP16 / BF16 output;
backprojection within 1e-4 relative L2 of the FP64 oracle (FP32).
"""
import ctypes as C
import math

import numpy as np
import pytest
import torch

import oracle
from analyticalreconbigpatient_amd import _lib, mexapi
from analyticalreconbigpatient_amd.pipeline import BPPlan

from .test_oracle_fdk import geom as bp_geom_dict

pytestmark = pytest.mark.gpu

SO = 541.0


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


# --------------------------------------------------------------------------- rebin
def rebin_args(YL, ViewN, N_2pi=64, YLC0=None):
    dYA = 2 * math.atan(1.0239 * 368 / YL / 949.0 / 2)
    DeltaFai = 2 * math.pi / N_2pi * ViewN / (ViewN - 1)
    f = np.float32
    return dict(YLC0=f((YL - 1) / 2.0 if YLC0 is None else YLC0), dYA=f(dYA),
                DeltaTheta=f(DeltaFai), PLC0=f((YL - 1) / 2.0),
                DeltaT=f(math.tan(YL * dYA * 0.5) * SO * 1.5 * 2 / YL), DeltaFai=f(DeltaFai),
                SO=f(SO))


# ZL 95 / 96: the switch from 128- to 64-channel blocks of the transposing
# kernel; ZL 242: its largest (64 KB of LDS with the static blend tables)
@pytest.mark.parametrize("ViewN,YL,ZL", [(120, 70, 8), (50, 33, 5), (2160, 368, 16), (64, 130, 3),
                                         (12, 200, 128), (10, 150, 95), (10, 150, 96),
                                         (6, 140, 242)])
@pytest.mark.parametrize("layout", [_lib.ARECON_LAYOUT_VYZ, _lib.ARECON_LAYOUT_VZY])
def test_rebin_bit_exact(gpu, ViewN, YL, ZL, layout):
    print(f"rebin ViewN={ViewN} YL={YL} ZL={ZL} layout={layout}", flush=True)
    rng = np.random.default_rng(ViewN + YL)
    P = rng.normal(size=(ViewN, YL, ZL)).astype(np.float32)
    a = rebin_args(YL, ViewN, YLC0=(YL - 1) / 2.0 + 0.25)
    want = oracle.rebin_f32(P, **a)
    din = torch.from_numpy(P).to(gpu)
    dout = torch.empty_like(din)
    _lib.check(_lib.lib().arecon_rebin_dev(
        C.c_void_p(dout.data_ptr()), C.c_void_p(din.data_ptr()), YL, ZL, ViewN, a["YLC0"],
        a["dYA"], a["DeltaTheta"], a["PLC0"], a["DeltaT"], a["DeltaFai"], a["SO"], layout,
        stream()), "rebin")
    got = dout.cpu().numpy()
    if layout == _lib.ARECON_LAYOUT_VZY:
        got = got.reshape(ViewN, ZL, YL).transpose(0, 2, 1)
    assert np.array_equal(got, want), f"max |diff| {np.abs(got - want).max()}"


def test_rebin_rejects_rows_beyond_the_lds_tile(gpu):
    """ZL 243 does not fit the transposing kernel's 64 KB of LDS: refused on the
    host with an error, before any launch."""
    ViewN, YL, ZL = 4, 64, 243
    a = rebin_args(YL, ViewN)
    din = torch.zeros(ViewN, YL, ZL, device=gpu)
    dout = torch.empty_like(din)
    rc = _lib.lib().arecon_rebin_dev(
        C.c_void_p(dout.data_ptr()), C.c_void_p(din.data_ptr()), YL, ZL, ViewN, a["YLC0"],
        a["dYA"], a["DeltaTheta"], a["PLC0"], a["DeltaT"], a["DeltaFai"], a["SO"],
        _lib.ARECON_LAYOUT_VZY, stream())
    assert rc == 1 and "ZL" in _lib.last_error()[1]


def test_rebin_drop_in_gateway(gpu):
    ViewN, YL, ZL = 96, 48, 8
    P = np.random.default_rng(5).normal(size=(ViewN, YL, ZL)).astype(np.float32)
    a = rebin_args(YL, ViewN)
    # gateway takes 1-based YLC, PLC and subtracts 1 (ParallelRebinningCBCurve.cpp:122)
    got = mexapi.ParallelRebinningCBCurve(P, np.int32(YL), np.int32(ZL), np.int32(ViewN),
                                          np.float32(a["YLC0"] + 1), a["dYA"], a["DeltaTheta"],
                                          np.float32(a["PLC0"] + 1), a["DeltaT"], a["DeltaFai"],
                                          a["SO"], np.int32(1))
    assert np.array_equal(got, oracle.rebin_f32(P, **a))


# --------------------------------------------------------------------------- filter
def run_filter(P, PLC0, ZLC0, dYL, dZL, dtype=_lib.ARECON_F32, inplace=False,
               layout=_lib.ARECON_FP_ROWS):
    """Filtered rows [ViewN][ZL][YL]; the row-pair layout is returned as stored,
    [ViewN][ZL-1][YL][2]."""
    ViewN, ZL, YL = P.shape
    din = torch.from_numpy(np.ascontiguousarray(P, dtype=np.float32)).cuda()
    tdt = {_lib.ARECON_F32: torch.float32, _lib.ARECON_F16: torch.float16,
           _lib.ARECON_BF16: torch.bfloat16}[dtype]
    shape = {_lib.ARECON_FP_ROWS: din.shape, _lib.ARECON_FP_ROWPAIRS: (ViewN, ZL - 1, YL, 2),
             _lib.ARECON_FP_QUADS: (ViewN, ZL - 1, YL - 1, 4),
             _lib.ARECON_FP_ROWDIFF: (ViewN, ZL - 1, YL, 2)}[layout]
    dout = din if inplace else torch.full(shape, float("nan"), dtype=tdt, device=din.device)
    _lib.check(_lib.lib().arecon_filter_dev(
        C.c_void_p(dout.data_ptr()), dtype, layout, C.c_void_p(din.data_ptr()), YL, ZL, ViewN,
        np.float32(PLC0), np.float32(ZLC0), np.float32(dYL), np.float32(dZL), np.float32(SO),
        stream()), "filter")
    return dout.float().cpu().numpy()


@pytest.mark.parametrize("ViewN,ZL,YL", [(40, 16, 368), (9, 64, 888), (7, 3, 20), (5, 5, 130), (6, 3, 100),
                                         (3, 1, 1024), (11, 4, 2048)])
def test_filter_matches_fp64_oracle(gpu, ViewN, ZL, YL):
    print(f"filter ViewN={ViewN} ZL={ZL} YL={YL}", flush=True)
    P = np.random.default_rng(YL).normal(size=(ViewN, ZL, YL)).astype(np.float32)
    args = ((YL - 1) / 2.0, (ZL - 1) / 2.0, 0.9493, 1.9231)
    want = oracle.filter_f64(P.astype(np.float64), *[float(np.float32(v)) for v in args], SO)
    got = run_filter(P, *args)
    assert rel_l2(got, want) < 1e-5


@pytest.mark.parametrize("ViewN,ZL,YL", [(9, 64, 888), (7, 3, 20), (5, 5, 130), (4, 2, 2048)])
@pytest.mark.parametrize("dtype", [_lib.ARECON_F32, _lib.ARECON_F16])
def test_filter_row_pair_layout(gpu, ViewN, ZL, YL, dtype):
    """Row-pair output holds rows r and r+1 of every channel, bit-identical to the
    rows layout of the same filter (odd ZL puts row pairs across transform pairs)."""
    print(f"filter row pairs ViewN={ViewN} ZL={ZL} YL={YL} dtype={dtype}", flush=True)
    P = np.random.default_rng(YL + ZL).normal(size=(ViewN, ZL, YL)).astype(np.float32)
    args = ((YL - 1) / 2.0, (ZL - 1) / 2.0, 0.9493, 1.9231)
    rows = run_filter(P, *args, dtype=dtype)
    pairs = run_filter(P, *args, dtype=dtype, layout=_lib.ARECON_FP_ROWPAIRS)
    assert np.array_equal(pairs[..., 0], rows[:, :-1, :])
    assert np.array_equal(pairs[..., 1], rows[:, 1:, :])


# YL <= 1024: fused kernel (filtered rows kept in LDS, one workgroup per view),
# every transform length 256..2048, odd ZL, ZL = 2 and YL = N/2; YL = 20 and
# 2048: filter in place, then expand_quads.
@pytest.mark.parametrize("ViewN,ZL,YL", [(9, 64, 888), (5, 5, 130), (4, 7, 200), (6, 3, 300),
                                         (3, 2, 1024), (5, 4, 128), (7, 3, 20), (4, 2, 2048)])
@pytest.mark.parametrize("dtype", [_lib.ARECON_F32, _lib.ARECON_F16, _lib.ARECON_BF16])
def test_filter_quad_layout(gpu, ViewN, ZL, YL, dtype):
    """Quad output = the differences of the filtered rows, rounded once to the
    output type.  With even ZL both filters transform the same row pairs
    (a + ib), so the quads are bit-identical to those formed from the rows
    output.  With odd ZL the rows filter pairs the last row of a view with the
    first of the next, the fused one with zeros; the rounding of a transform
    depends on its partner row, so the two agree to FP32 rounding only."""
    print(f"filter quads ViewN={ViewN} ZL={ZL} YL={YL} dtype={dtype}", flush=True)
    P = np.random.default_rng(YL + 7 * ZL).normal(size=(ViewN, ZL, YL)).astype(np.float32)
    args = ((YL - 1) / 2.0, (ZL - 1) / 2.0, 0.9493, 1.9231)
    rows = run_filter(P, *args)
    got = run_filter(P, *args, dtype=dtype, layout=_lib.ARECON_FP_QUADS)
    want = quads(rows)
    cast = {_lib.ARECON_F32: lambda a: a,
            _lib.ARECON_F16: lambda a: torch.from_numpy(a).half().float().numpy(),
            _lib.ARECON_BF16: lambda a: torch.from_numpy(a).bfloat16().float().numpy()}[dtype]
    if ZL % 2 == 0 or YL > 1024:
        assert np.array_equal(got, cast(want))
    else:
        tol = {_lib.ARECON_F32: 1e-6, _lib.ARECON_F16: 1e-3, _lib.ARECON_BF16: 8e-3}[dtype]
        assert rel_l2(got, want) < tol
        want64 = quads(oracle.filter_f64(P.astype(np.float64),
                                         *[float(np.float32(v)) for v in args], SO))
        assert rel_l2(got, want64) < max(tol, 1e-5)


@pytest.mark.parametrize("ViewN,ZL,YL", [(9, 64, 888), (5, 5, 130), (3, 2, 1024), (6, 4, 300),
                                         (7, 3, 20), (4, 2, 2048)])
@pytest.mark.parametrize("dtype", [_lib.ARECON_F32, _lib.ARECON_F16, _lib.ARECON_BF16])
def test_filter_row_difference_layout(gpu, ViewN, ZL, YL, dtype):
    """Row differences (f(r), f(r+1) - f(r)), formed in FP32 and rounded once to
    the output type: bit-identical to those formed from the rows output for
    even ZL (and for the in-place path, YL = 20 and 2048); to FP32 rounding of
    the filter for odd ZL (see the quad test)."""
    print(f"filter rowdiff ViewN={ViewN} ZL={ZL} YL={YL} dtype={dtype}", flush=True)
    P = np.random.default_rng(YL + 3 * ZL).normal(size=(ViewN, ZL, YL)).astype(np.float32)
    args = ((YL - 1) / 2.0, (ZL - 1) / 2.0, 0.9493, 1.9231)
    rows = run_filter(P, *args)
    got = run_filter(P, *args, dtype=dtype, layout=_lib.ARECON_FP_ROWDIFF)
    want = row_diffs(rows)
    cast = {_lib.ARECON_F32: lambda a: a,
            _lib.ARECON_F16: lambda a: torch.from_numpy(a).half().float().numpy(),
            _lib.ARECON_BF16: lambda a: torch.from_numpy(a).bfloat16().float().numpy()}[dtype]
    if ZL % 2 == 0 or YL > 1024 or YL < 128:
        assert np.array_equal(got, cast(want))
    else:
        tol = {_lib.ARECON_F32: 1e-6, _lib.ARECON_F16: 1e-3, _lib.ARECON_BF16: 8e-3}[dtype]
        assert rel_l2(got, want) < tol


def test_filter_quads_leave_input_unchanged(gpu):
    """The fused quad filter (YL <= 1024) only reads its input."""
    P = np.random.default_rng(5).normal(size=(6, 8, 400)).astype(np.float32)
    din = torch.from_numpy(P).cuda()
    dout = torch.empty((6, 7, 399, 4), dtype=torch.float16, device="cuda")
    _lib.check(_lib.lib().arecon_filter_dev(
        C.c_void_p(dout.data_ptr()), _lib.ARECON_F16, _lib.ARECON_FP_QUADS,
        C.c_void_p(din.data_ptr()), 400, 8, 6, np.float32(199.5), np.float32(3.5),
        np.float32(0.95), np.float32(1.92), np.float32(SO), stream()), "filter")
    torch.cuda.synchronize()
    assert np.array_equal(din.cpu().numpy(), P)


@pytest.mark.parametrize("layout", [_lib.ARECON_FP_ROWS, _lib.ARECON_FP_QUADS])
def test_filter_2048_point_schedules(gpu, layout, monkeypatch):
    """Both FFT schedules of the 2048-point transform (radix 16/16/8 on two
    wavefronts, radix 8/8/8/4 on four) match the FP64 oracle and each other."""
    P = np.random.default_rng(11).normal(size=(5, 6, 888)).astype(np.float32)
    args = (443.5, 2.5, 0.9493, 1.9231)
    want = oracle.filter_f64(P.astype(np.float64), *[float(np.float32(v)) for v in args], SO)
    if layout == _lib.ARECON_FP_QUADS:
        want = quads(want)
    got = {}
    for sched in ("0", "1"):
        monkeypatch.setenv("ARECON_FILTER_SCHEDULE", sched)
        got[sched] = run_filter(P, *args, layout=layout)
        assert rel_l2(got[sched], want) < 1e-5
    assert rel_l2(got["0"], got["1"]) < 1e-6


def test_filter_in_place_and_16_bit_outputs(gpu):
    P = np.random.default_rng(1).normal(size=(20, 8, 300)).astype(np.float32)
    args = (149.5, 3.5, 0.95, 1.92)
    want = oracle.filter_f64(P.astype(np.float64), *[float(np.float32(v)) for v in args], SO)
    assert rel_l2(run_filter(P, *args, inplace=True), want) < 1e-5
    assert rel_l2(run_filter(P, *args, dtype=_lib.ARECON_F16), want) < 2e-3
    assert rel_l2(run_filter(P, *args, dtype=_lib.ARECON_BF16), want) < 1e-2


def test_filter_drop_in_gateway(gpu):
    P = np.random.default_rng(2).normal(size=(6, 4, 100)).astype(np.float32)
    got = mexapi.reWeigAdFiltr_GPU(P, np.float32(49.5), np.float32(1.5), np.float32(0.9),
                                   np.float32(1.9), np.float32(SO))
    want = oracle.filter_f64(P.astype(np.float64), 49.5, 1.5, float(np.float32(0.9)),
                             float(np.float32(1.9)), SO)
    assert rel_l2(got, want) < 1e-5


# --------------------------------------------------------------------------- backprojection
class P_:
    """Minimal BPParams-like holder for BPPlan."""

    def __init__(self, g, YL, ZL, PN, delta=8):
        self.__dict__.update(g)
        self.YL, self.ZL, self.PN, self.delta = YL, ZL, PN, delta

    def to_geom(self, dz=None, nz=0, order=0, variant=0, group=0):
        return _lib.BPGeom(N_2pi=self.N_2pi, XN=self.XN, YN=self.YN, ZN=self.ZN, YL=self.YL,
                           ZL=self.ZL, PN=self.PN, XNC=self.XNC, YNC=self.YNC, ZNC=self.ZNC,
                           dx=self.dx, dy=self.dy, dz=self.dz if dz is None else dz,
                           YLC=self.YLC, ZLC=self.ZLC, dYL=self.dYL, dZL=self.dZL, h=self.h,
                           BetaS=self.BetaS, DeltaFai=self.DeltaFai, N_pi=self.N_pi,
                           HSCoef=self.HSCoef, SO=self.SO, SD=self.SD, k1=self.k1, nz=nz,
                           order=order, variant=variant, group=group)


def row_pairs(fp):
    """[PN][ZL][YL] -> the row-pair layout [PN][ZL-1][YL][2]."""
    return np.stack([fp[:, :-1, :], fp[:, 1:, :]], axis=-1)


def row_diffs(fp):
    """[PN][ZL][YL] -> the row-difference layout [PN][ZL-1][YL][2], in FP32."""
    f = np.asarray(fp, np.float32)
    return np.stack([f[:, :-1, :], f[:, 1:, :] - f[:, :-1, :]], axis=-1)


def quads(fp):
    """[PN][ZL][YL] -> the difference-quad layout [PN][ZL-1][YL-1][4], in FP32."""
    f = np.asarray(fp, np.float32)
    a, ac, b, bc = f[:, :-1, :-1], f[:, :-1, 1:], f[:, 1:, :-1], f[:, 1:, 1:]
    dr = b - a
    return np.stack([a, dr, ac - a, (bc - ac) - dr], axis=-1)


def gpu_bp(fp, g, nz=0, order=0, storage=torch.float32, variant=0, group=0, layout="rows"):
    PN, ZL, YL = fp.shape
    p = P_(g, YL, ZL, PN)
    w = oracle.fdk_weights(g["N_2pi"], g["HSCoef"], 8)
    plan = BPPlan(p, nz=nz, order=order, w=w, variant=variant, group=group)
    host = {"rows": lambda f: f, "rowpairs": row_pairs, "quads": quads,
            "rowdiff": row_diffs}[layout](fp)
    dfp = torch.from_numpy(np.ascontiguousarray(host, dtype=np.float32)).cuda().to(storage)
    img = torch.full((g["YN"], g["XN"], g["ZN"]), float("nan"), device="cuda")
    plan.execute(img, dfp, layout=layout)
    torch.cuda.synchronize()
    out = img.cpu().numpy()
    plan.close()
    return out, w


def realistic(**kw):
    """Clinical-like scanner scaled down: magnification, 1.5x parallel width,
    pitch ~1; slices inside the helix with partial windows at one end."""
    g = bp_geom_dict(N_2pi=96, XN=20, YN=18, ZN=13, XNC=9.7, YNC=8.3, dx=4.9, dy=4.9,
                     dz=1.5, ZNC=-1.37, YLC=63.5, ZLC=7.5, dYL=3.1, dZL=1.923, h=17.5,
                     k1=2.85)
    g.update(kw)
    g["N_pi"] = g["N_2pi"] / 2.0
    g["DeltaFai"] = 2 * math.pi / g["N_2pi"]
    return g


@pytest.mark.parametrize("nz", [1, 2, 4, 8, 16])
@pytest.mark.parametrize("order", [0, 1, 2])
@pytest.mark.parametrize("variant", [0, 1])
def test_bp_matches_fp64_oracle(gpu, nz, order, variant):
    print(f"bp nz={nz} order={order} variant={variant}", flush=True)
    g = realistic()
    PN, ZL, YL = 200, 16, 128
    fp = np.random.default_rng(nz + 10 * order).normal(size=(PN, ZL, YL))
    fp32 = fp.astype(np.float32)
    got, w = gpu_bp(fp32, g, nz=nz, order=order, variant=variant)
    want = oracle.fdk_core(fp32.astype(np.float64), w,
                           **{k: v for k, v in g.items() if k not in ("YL", "ZL")})
    assert not np.isnan(got).any()
    assert np.abs(want).max() > 0 and (want == 0).any()      # full, partial and empty voxels
    assert rel_l2(got, want) < 1e-4


@pytest.mark.parametrize("nz,group", [(8, 1), (8, 2), (8, 4), (8, 8), (16, 2), (16, 4), (16, 8),
                                      (32, 1), (32, 2), (32, 4)])
@pytest.mark.parametrize("ZNC", [-15.37, -68.41])
@pytest.mark.parametrize("layout", ["rows", "rowpairs", "quads", "rowdiff"])
def test_bp_z_column_matches_fp64_oracle(gpu, nz, group, ZNC, layout):
    """Z-column kernel: slice blocks whose windows overlap heavily (thin slices),
    a ragged last block, windows cut by the first (ZNC=-15.37) or the last
    (ZNC=-68.41) view of the helix."""
    print(f"bp z-column nz={nz} group={group} ZNC={ZNC} layout={layout}", flush=True)
    g = realistic(ZN=37, dz=0.45, ZNC=ZNC)
    PN, ZL, YL = 200, 16, 128
    fp = np.random.default_rng(nz + group).normal(size=(PN, ZL, YL)).astype(np.float32)
    got, w = gpu_bp(fp, g, nz=nz, variant=2, group=group, layout=layout)
    want = oracle.fdk_core(fp.astype(np.float64), w,
                           **{k: v for k, v in g.items() if k not in ("YL", "ZL")})
    assert not np.isnan(got).any()
    assert np.abs(want).max() > 0
    assert rel_l2(got, want) < 1e-4


@pytest.mark.parametrize("variant,nz", [(2, 8), (2, 16), (3, 8), (3, 16)])
@pytest.mark.parametrize("ZNC", [-15.37, -68.41])
def test_bp_row_differences_bit_identical_to_row_pairs(gpu, variant, nz, ZNC):
    """FP32 row differences hold the very differences the row-pair blend forms,
    so both z-column kernels give bit-identical volumes from either layout."""
    print(f"bp rowdiff variant={variant} nz={nz} ZNC={ZNC}", flush=True)
    g = realistic(ZN=37, dz=0.45, ZNC=ZNC)
    fp = np.random.default_rng(nz + 3).normal(size=(200, 16, 128)).astype(np.float32)
    a, w = gpu_bp(fp, g, nz=nz, variant=variant, group=8, layout="rowpairs")
    b, _ = gpu_bp(fp, g, nz=nz, variant=variant, group=8, layout="rowdiff")
    assert not np.isnan(b).any()
    assert np.array_equal(a, b)
    want = oracle.fdk_core(fp.astype(np.float64), w,
                           **{k: v for k, v in g.items() if k not in ("YL", "ZL")})
    assert rel_l2(b, want) < 1e-4


@pytest.mark.parametrize("storage,tol", [(torch.float16, 1e-3), (torch.bfloat16, 8e-3)])
@pytest.mark.parametrize("variant,group", [(2, 8), (2, 4), (3, 8)])
def test_bp_16_bit_row_differences(gpu, storage, tol, variant, group):
    """16-bit row differences: f and f(r+1) - f each rounded once, so the volume
    stays within the storage precision of the FP64 oracle, from the gathering
    kernel (split and grouped) and the LDS-staged one (4-byte records)."""
    print(f"bp 16-bit rowdiff {storage} variant={variant} group={group}", flush=True)
    g = realistic(ZN=37, dz=0.45, ZNC=-15.37)
    fp = np.random.default_rng(23).normal(size=(200, 16, 128)).astype(np.float32)
    got, w = gpu_bp(fp, g, storage=storage, variant=variant, nz=8, group=group, layout="rowdiff")
    want = oracle.fdk_core(fp.astype(np.float64), w,
                           **{k: v for k, v in g.items() if k not in ("YL", "ZL")})
    assert not np.isnan(got).any()
    err = rel_l2(got, want)
    print(f"16-bit row differences {storage}: rel L2 {err:.2e}")
    assert err < tol


@pytest.mark.parametrize("k1", [0.0, 2.85])
@pytest.mark.parametrize("variant", [0, 1, 2])
def test_bp_full_window_random(gpu, k1, variant):
    g = bp_geom_dict(k1=k1, XN=17, YN=21, ZN=9, XNC=8.2, YNC=10.1, dz=0.113, ZNC=-14.77)
    fp = np.random.default_rng(3).normal(size=(192, 16, 64)).astype(np.float32)
    got, w = gpu_bp(fp, g, variant=variant)
    want = oracle.fdk_core(fp.astype(np.float64), w,
                           **{k: v for k, v in g.items() if k not in ("YL", "ZL")})
    assert rel_l2(got, want) < 1e-4


@pytest.mark.parametrize("k1", [0.0, 5.0])
@pytest.mark.parametrize("variant", [0, 1, 2])
def test_bp_constant_projection_gives_pi_c(gpu, k1, variant):
    g = bp_geom_dict(h=1.0, k1=k1)
    c = 1.75
    got, _ = gpu_bp(np.full((192, 16, 64), c, dtype=np.float32), g, variant=variant)
    np.testing.assert_allclose(got, math.pi * c, rtol=2e-5)


@pytest.mark.parametrize("storage,tol", [(torch.float16, 1e-4), (torch.bfloat16, 1e-4)])
@pytest.mark.parametrize("variant,layout", [(0, "rows"), (2, "rows"), (2, "rowpairs")])
def test_bp_16_bit_storage(gpu, storage, tol, variant, layout):
    """The kernel reads 16-bit samples exactly: compare with the oracle fed the
    same rounded samples."""
    g = realistic()
    fp = np.random.default_rng(9).normal(size=(200, 16, 128)).astype(np.float32)
    rounded = torch.from_numpy(fp).to(storage).float().numpy().astype(np.float64)
    got, w = gpu_bp(fp, g, storage=storage, variant=variant, layout=layout)
    want = oracle.fdk_core(rounded, w, **{k: v for k, v in g.items() if k not in ("YL", "ZL")})
    assert rel_l2(got, want) < tol


@pytest.mark.parametrize("storage,tol", [(torch.float16, 1e-3), (torch.bfloat16, 8e-3)])
def test_bp_16_bit_quads(gpu, storage, tol):
    """16-bit difference quads: each of the four terms is rounded once, so the
    result stays within the storage precision of the FP64 oracle."""
    g = realistic(ZN=37, dz=0.45, ZNC=-15.37)
    fp = np.random.default_rng(19).normal(size=(200, 16, 128)).astype(np.float32)
    got, w = gpu_bp(fp, g, storage=storage, variant=2, nz=8, group=8, layout="quads")
    want = oracle.fdk_core(fp.astype(np.float64), w,
                           **{k: v for k, v in g.items() if k not in ("YL", "ZL")})
    err = rel_l2(got, want)
    print(f"16-bit quads {storage}: rel L2 {err:.2e}")
    assert err < tol


def test_bp_drop_in_float_and_double(gpu):
    g = realistic(XN=12, YN=10, ZN=7)
    PN, ZL, YL = 200, 16, 128
    fp = np.random.default_rng(11).normal(size=(PN, ZL, YL))
    w = oracle.fdk_weights(g["N_2pi"], g["HSCoef"], 8)
    want = oracle.fdk_core(fp, w, **{k: v for k, v in g.items() if k not in ("YL", "ZL")})
    L = _lib.lib()
    args = (g["N_2pi"], g["XN"], g["YN"], g["ZN"], g["XNC"], g["YNC"], g["ZNC"], g["dx"],
            g["dy"], g["dz"], YL, ZL, PN, g["YLC"], g["ZLC"], g["dYL"], g["dZL"], g["h"],
            g["BetaS"], g["DeltaFai"], g["N_pi"], g["HSCoef"], g["SO"], g["SD"], g["k1"])
    img64 = np.zeros(g["XN"] * g["YN"] * g["ZN"])
    L.ParallekFDKHelical3DWeighting_GPU_double(_lib.dptr(img64), _lib.dptr(fp), _lib.dptr(w),
                                                *args)
    _lib.check(None, "double drop-in")
    assert rel_l2(img64.reshape(want.shape), want) < 1e-10
    fp32 = fp.astype(np.float32)
    w32 = w.astype(np.float32)
    img32 = np.zeros(g["XN"] * g["YN"] * g["ZN"], dtype=np.float32)
    L.ParallekFDKHelical3DWeighting_GPU_float(_lib.fptr(img32), _lib.fptr(fp32), _lib.fptr(w32),
                                               *args)
    _lib.check(None, "float drop-in")
    assert rel_l2(img32.reshape(want.shape), want) < 1e-4


def test_bp_row_pairs_need_the_z_column_kernel(gpu):
    g = realistic()
    fp = np.zeros((200, 16, 128), dtype=np.float32)
    with pytest.raises(_lib.AreconError, match="z-column"):
        gpu_bp(fp, g, variant=0, layout="rowpairs")


def test_bp_rejects_oversized_window(gpu):
    g = realistic(HSCoef=1.2)
    with pytest.raises(_lib.AreconError, match="weight index"):
        BPPlan(P_(g, 128, 16, 200), w=np.ones(g["N_2pi"]))
