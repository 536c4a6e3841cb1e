"""LDS-staged z-column backprojector (variant 3) against the z-column kernel
(variant 2) and the FP64 oracle, through the C ABI.

Variant 3 reads each view's detector footprint from LDS but does the z-column
kernel's arithmetic step for step, so its volume is bit-identical to variant 2
on the same inputs; variant 2 is itself held to the oracle (1e-4 relative L2,
tests/test_gpu_stages.py).  The plan's workgroup counters show which path ran:
LDS-staged, or global gathers where a view's footprint does not fit.
"""
import math

import numpy as np
import pytest
import torch

import oracle
from analyticalreconbigpatient_amd import geometry as G
from analyticalreconbigpatient_amd.pipeline import BPPlan

from .test_gpu_stages import P_, quads, rel_l2, row_diffs, row_pairs
from .test_oracle_fdk import geom as bp_geom_dict

pytestmark = pytest.mark.gpu

STORAGE_LAYOUT = [(torch.float32, "rowpairs"), (torch.float32, "rowdiff"),
                  (torch.float16, "quads"), (torch.bfloat16, "quads"),
                  (torch.float16, "rowdiff"), (torch.bfloat16, "rowdiff")]


def lds_geom(**kw):
    """Scaled-down clinical-like geometry whose 16x16-column tiles span ~14
    channels (2 mm pixels, 3.1 mm channels): every footprint fits the LDS slots."""
    g = bp_geom_dict(N_2pi=96, XN=40, YN=36, ZN=37, XNC=19.3, YNC=17.6, dx=2.0, dy=2.0,
                     dz=0.45, ZNC=-15.37, YLC=63.5, ZLC=7.5, dYL=3.1, dZL=1.923, h=17.5,
                     k1=2.85)
    g.update(kw)
    g["N_pi"] = g["N_2pi"] / 2.0
    g["DeltaFai"] = 2 * math.pi / g["N_2pi"]
    return g


def run_bp(fp, g, variant, nz, storage, layout, order=0):
    PN, ZL, YL = fp.shape
    p = P_(g, YL, ZL, PN)
    w = oracle.fdk_weights(g["N_2pi"], g["HSCoef"], 8)
    plan = BPPlan(p, nz=nz, w=w, variant=variant, group=8, order=order)
    host = {"rowpairs": row_pairs, "quads": quads, "rowdiff": row_diffs}[layout](fp)
    dfp = torch.from_numpy(np.ascontiguousarray(host, dtype=np.float32)).cuda().to(storage)
    img = torch.full((g["YN"], g["XN"], g["ZN"]), float("nan"), device="cuda")
    plan.execute(img, dfp, layout=layout)
    torch.cuda.synchronize()
    stats = plan.stats()
    plan.close()
    return img.cpu().numpy(), w, stats


@pytest.mark.parametrize("nz", [8, 16])
@pytest.mark.parametrize("storage,layout", STORAGE_LAYOUT)
@pytest.mark.parametrize("ZNC", [-15.37, -68.41])
def test_lds_bit_identical_to_z_column(gpu, nz, storage, layout, ZNC):
    """Thin slices (heavily shared windows), ragged last block, windows cut by
    the first (ZNC=-15.37) or the last (ZNC=-68.41) view; partial tiles."""
    print(f"bp lds nz={nz} {storage} {layout} ZNC={ZNC}", flush=True)
    g = lds_geom(ZNC=ZNC)
    fp = np.random.default_rng(nz).normal(size=(200, 16, 128)).astype(np.float32)
    got, _, (staged, direct) = run_bp(fp, g, 3, nz, storage, layout)
    want, _, _ = run_bp(fp, g, 2, nz, storage, layout)
    assert not np.isnan(got).any()
    assert staged > 0 and direct == 0, (staged, direct)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("order,storage,layout", [
    (0, torch.float32, "rowdiff"), (1, torch.float32, "rowdiff"), (2, torch.float32, "rowdiff"),
    (0, torch.float32, "rowpairs"), (0, torch.float16, "quads"), (2, torch.bfloat16, "quads"),
    (0, torch.float16, "rowdiff"), (2, torch.bfloat16, "rowdiff")])
def test_lds_edges_and_workgroup_orders(gpu, order, storage, layout):
    """Columns projecting onto the detector's first and last channels and rows
    (the geometry is not INSIDE: the grid reaches past the detector's edges),
    every workgroup order (14 tile rows: one full band of 12 rows and a partial
    one for order 2; a partial tile column): bit-identical to the gathering
    kernel.  Staging moves 8-byte records in 16-byte pairs; a pair straddling
    the view's first or last record must not zero it (4-byte records load one
    per lane)."""
    print(f"bp lds edges order={order} {storage} {layout}", flush=True)
    g = lds_geom(XN=40, YN=213, YNC=105.6)
    fp = np.random.default_rng(21).normal(size=(200, 16, 128)).astype(np.float32)
    got, _, (staged, _) = run_bp(fp, g, 3, 8, storage, layout, order=order)
    want, _, _ = run_bp(fp, g, 2, 8, storage, layout, order=0)
    assert staged > 0
    assert not np.isnan(got).any()
    assert np.array_equal(got, want)


@pytest.mark.parametrize("nz", [8, 16])
def test_lds_matches_fp64_oracle(gpu, nz):
    g = lds_geom()
    fp = np.random.default_rng(11 + nz).normal(size=(200, 16, 128)).astype(np.float32)
    got, w, (staged, _) = run_bp(fp, g, 3, nz, torch.float32, "rowpairs")
    want = oracle.fdk_core(fp.astype(np.float64), w,
                           **{k: v for k, v in g.items() if k not in ("YL", "ZL")})
    assert staged > 0
    assert np.abs(want).max() > 0
    assert rel_l2(got, want) < 1e-4


def test_lds_constant_projection_gives_pi_c(gpu):
    """Constant projections: every voxel whose full window lies on the detector
    gets pi c (the 3D weights of conjugate views sum to one)."""
    g = bp_geom_dict(h=1.0, k1=5.0, XN=24, YN=20, XNC=11.5, YNC=9.5, dx=0.5, dy=0.5)
    c = 1.75
    got, _, (staged, _) = run_bp(np.full((192, 16, 64), c, dtype=np.float32), g, 3, 8,
                                 torch.float32, "rowpairs")
    assert staged > 0
    np.testing.assert_allclose(got, math.pi * c, rtol=2e-5)


def test_lds_falls_back_when_the_footprint_is_too_wide(gpu):
    """4.9 mm pixels on 3.1 mm channels: a tile spans ~34 channels, more than a
    staged row holds, so every workgroup gathers from global memory."""
    g = lds_geom(dx=4.9, dy=4.9, XN=20, YN=18, XNC=9.7, YNC=8.3)
    fp = np.random.default_rng(5).normal(size=(200, 16, 128)).astype(np.float32)
    got, _, (staged, direct) = run_bp(fp, g, 3, 8, torch.float32, "rowpairs")
    want, _, _ = run_bp(fp, g, 2, 8, torch.float32, "rowpairs")
    assert staged == 0 and direct > 0
    assert np.array_equal(got, want)


@pytest.mark.parametrize("nz,storage,layout", [(8, torch.float32, "rowpairs"),
                                               (8, torch.float32, "rowdiff"),
                                               (8, torch.float16, "quads"),
                                               (16, torch.float16, "quads"),
                                               (8, torch.float16, "rowdiff"),
                                               (16, torch.bfloat16, "rowdiff")])
@pytest.mark.parametrize("x0", [0.0, 200.0])
def test_lds_clinical_geometry(gpu, nz, storage, layout, x0):
    """The clinical scanner (888x64 detector, 984 views/rotation, pitch 0.984,
    1.25 mm slices): a 48x48x32 block of the 512x512 grid at the centre and at
    the edge of the 500 mm field (largest magnification, widest row span),
    backprojected from the views its slices need, random filtered data."""
    print(f"bp lds clinical nz={nz} {storage} x0={x0}", flush=True)
    wl = G.body_clinical()
    p = G.bp_params(G.bp_scan_geom(wl.setup()))
    n, zn, dx, dz = 48, 32, p.dx, p.dz
    z_start = p.ZN // 2 - zn // 2
    W = G.window_half_width(p.N_2pi, p.HSCoef)
    s0 = [math.ceil((2 * 3.14159265358979 * (z - p.ZNC) * dz / p.h - p.BetaS) / p.DeltaFai
                    - 0.5) for z in (z_start, z_start + zn - 1)]
    v_lo, v_hi = max(0, min(s0) - W - 2), min(p.PN, max(s0) + W + 3)
    # same pixel pitch (ObjR sets dx = 2 ObjR / XN), slices of the clinical slab
    sub = G.BPParams(**{**p.__dict__, "PN": v_hi - v_lo, "XN": n, "YN": n, "ZN": zn,
                        "ObjR": 0.5 * n * dx, "XNC": 0.5 * (n - 1) - x0 / dx,
                        "YNC": 0.5 * (n - 1), "ZNC": p.ZNC - z_start,
                        "BetaS": p.BetaS + v_lo * p.DeltaFai})
    assert abs(sub.dx - dx) < 1e-12
    shape = {"rowpairs": (sub.PN, p.ZL - 1, p.YL, 2), "rowdiff": (sub.PN, p.ZL - 1, p.YL, 2),
             "quads": (sub.PN, p.ZL - 1, p.YL - 1, 4)}
    gen = torch.Generator(device="cuda").manual_seed(7)
    fp = torch.randn(shape[layout], generator=gen, device="cuda").to(storage)
    out = {}
    for variant in (2, 3):
        plan = BPPlan(sub, dz=dz, nz=nz, variant=variant, group=8)
        img = torch.full((sub.YN, sub.XN, sub.ZN), float("nan"), device="cuda")
        plan.execute(img, fp, layout=layout)
        torch.cuda.synchronize()
        out[variant] = (img, plan.stats())
        plan.close()
    staged, direct = out[3][1]
    print(f"  staged workgroups {staged}, direct {direct}", flush=True)
    assert staged > 0
    assert not torch.isnan(out[3][0]).any()
    assert torch.equal(out[3][0], out[2][0])
    assert float(out[2][0].abs().max()) > 0


def test_lds_rejects_rows_layout(gpu):
    g = lds_geom()
    fp = np.zeros((200, 16, 128), dtype=np.float32)
    p = P_(g, 128, 16, 200)
    w = oracle.fdk_weights(g["N_2pi"], g["HSCoef"], 8)
    plan = BPPlan(p, nz=8, w=w, variant=3, group=8)
    img = torch.empty((g["YN"], g["XN"], g["ZN"]), device="cuda")
    with pytest.raises(RuntimeError, match="variant 3"):
        plan.execute(img, torch.from_numpy(fp).cuda(), layout="rows")
    plan.close()
