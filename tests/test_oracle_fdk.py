"""Pins the CPU oracle's 3D-weighted helical FDK backprojection
(oracle/arecon_oracle.c, restating src/Parallel_FDK_Helical_3DWeighting.cpp:42-200).

Independent checks:
* the weight table: cos^2 tapers in closed form (cos^2(pi/2 (d-k-1/2)/d) =
  sin^2(pi/2 (k+1/2)/d)), mirror symmetry, support [Shift, Shift+L);
* zero -> zero, linearity in fp, voxels whose window misses [0, PN) are 0;
* a constant projection c gives exactly pi*c in every fully covered voxel:
  conjugate views p, p+N_pi have complementary weights (any k1 when h = 1 mm,
  because the +-0.5 mm conjugate offset then equals h/2; k1 = 0 for any h) and
  views without a conjugate get weight 1, so the weights sum to N_pi and
  N_pi * DeltaFai = pi;
* the central ray (Yr = 0): the detector row is ZLC + (z - h V/2pi) SD/((SO-r) dZL),
  the textbook magnification of a point at distance SO - r from the source;
* the detector channel is YLC + t/dYL with t = -x sinV + y cosV, the textbook
  parallel-beam coordinate.
PARITY UNPINNED: off the central ray the row coordinate carries the factor
SO/sqrt(SO^2 - Yr^2) (:80); nothing independent fixes it.
"""
import math

import numpy as np
import pytest

import oracle

SO, SD = 541.0, 949.0


def geom(**kw):
    g = dict(N_2pi=64, XN=6, YN=6, ZN=4, XNC=2.5, YNC=2.5, ZNC=-6.5, dx=2.0, dy=2.0, dz=0.25,
             YLC=31.5, ZLC=7.5, dYL=0.5, dZL=0.5, h=1.0, BetaS=math.pi, HSCoef=0.68, SO=SO,
             SD=SD, k1=0.0)
    g.update(kw)
    g["N_pi"] = g["N_2pi"] / 2.0
    g["DeltaFai"] = 2 * math.pi / g["N_2pi"]
    return g


def run(fp, w, g, **kw):
    # YL, ZL and PN come from fp's shape
    return oracle.fdk_core(fp, w, **{k: v for k, v in g.items() if k not in ("YL", "ZL")}, **kw)


def test_weight_table_closed_form():
    for N_2pi, HS, delta in [(720, 0.68, 60), (984, 0.68, 60), (64, 0.68, 8), (65, 0.6, 5)]:
        w = oracle.fdk_weights(N_2pi, HS, delta)
        N_pi = N_2pi / 2.0
        W = math.ceil(N_pi * HS)
        L, Shift = 2 * W + 1, int(N_pi - W)
        nz = np.nonzero(w)[0]
        assert nz.min() == Shift and nz.max() == Shift + L - 1
        seg = w[Shift:Shift + L]
        np.testing.assert_allclose(seg, seg[::-1], rtol=1e-13)          # mirror symmetry
        k = np.arange(delta)
        np.testing.assert_allclose(seg[:delta], np.sin(math.pi / 2 * (k + 0.5) / delta) ** 2,
                                   rtol=1e-12)
        assert np.all(seg[delta:L - delta] == 1.0)


def test_zero_linearity_and_empty_windows():
    g = geom(k1=2.5)
    w = oracle.fdk_weights(g["N_2pi"], g["HSCoef"], 8)
    rng = np.random.default_rng(0)
    A = rng.normal(size=(192, 16, 64))
    B = rng.normal(size=(192, 16, 64))
    assert np.all(run(np.zeros_like(A), w, g) == 0)
    ra, rb, rab = run(A, w, g), run(B, w, g), run(2.0 * A - 3.0 * B, w, g)
    np.testing.assert_allclose(rab, 2.0 * ra - 3.0 * rb, rtol=0, atol=1e-10 * np.abs(rab).max())
    # z far below / above the helix: window misses [0, PN) -> exactly 0
    for ZNC in (400.0, -400.0):
        out = run(A, w, geom(k1=2.5, ZNC=ZNC))
        assert np.all(out == 0.0)


@pytest.mark.parametrize("h,k1", [(1.0, 0.0), (1.0, 2.8504), (1.0, 5.0), (3.0, 0.0)])
def test_constant_projection_gives_pi_c(h, k1):
    c = 1.75
    g = geom(h=h, k1=k1)
    if h != 1.0:   # keep the slab inside the helix travel
        g["dz"] = 0.25 * h
    w = oracle.fdk_weights(g["N_2pi"], g["HSCoef"], 8)
    fp = np.full((192, 16, 64), c)
    out = run(fp, w, g)
    np.testing.assert_allclose(out, math.pi * c, rtol=1e-12)


def _single_view(x, y, z, V0, h=10.0, **kw):
    g = geom(XN=1, YN=1, ZN=1, dx=1.0, dy=1.0, dz=1.0, XNC=-x, YNC=-y, ZNC=-z, BetaS=V0, h=h,
             **kw)
    return g


@pytest.mark.parametrize("r,dzs", [(0.0, 0.3), (120.0, -0.6), (-150.0, 0.45)])
def test_central_ray_row_is_textbook_magnification(r, dzs):
    V0, h = 0.3, 10.0
    z = h * V0 / (2 * math.pi) + dzs
    g = _single_view(r * math.cos(V0), r * math.sin(V0), z, V0, h=h, ZL=64, dZL=0.25,
                     ZLC=31.5)
    w = oracle.fdk_weights(g["N_2pi"], g["HSCoef"], 8)
    fp = np.broadcast_to(np.arange(64, dtype=np.float64)[:, None], (1, 64, 64)).copy()
    out = run(fp, w, g)[0, 0, 0]
    V1 = g["ZLC"] + dzs * SD / ((SO - r) * g["dZL"])
    assert out == pytest.approx(g["DeltaFai"] * V1, rel=1e-12)


@pytest.mark.parametrize("x,y", [(30.0, -12.0), (-7.0, 22.0), (0.0, 0.0)])
def test_channel_is_parallel_beam_offset(x, y):
    V0, h = 1.1, 10.0
    z = h * V0 / (2 * math.pi)
    g = _single_view(x, y, z, V0, h=h, dYL=1.5)
    w = oracle.fdk_weights(g["N_2pi"], g["HSCoef"], 8)
    fp = np.broadcast_to(np.arange(64, dtype=np.float64)[None, :], (1, 16, 64)).copy()
    out = run(fp, w, g)[0, 0, 0]
    t = -x * math.sin(V0) + y * math.cos(V0)
    assert out == pytest.approx(g["DeltaFai"] * (g["YLC"] + t / g["dYL"]), rel=1e-12)


def test_mid_level_derivations_match_core():
    """oracle_fdk (:117-200) == core with dx = dy = dz = 2 ObjR / N, dYL =
    DecWidth/YL, dZL = DecHeigh/ZL, DeltaFai = 2pi/N_2pi, SD = SO + DO."""
    rng = np.random.default_rng(7)
    YL, ZL, PN, N_2pi = 40, 8, 128, 64
    fp = rng.normal(size=(PN, ZL, YL))
    mid = dict(SO=SO, DO=SD - SO, DecWidth=40.0, DecHeigh=24.0, YLC=19.5, ZLC=3.5, h=16.0,
               BetaS=math.pi, N_2pi=N_2pi, ObjR=8.0, XN=5, YN=4, ZN=6, XNC=2.0, YNC=1.5,
               ZNC=-4.1, delta=8, HSCoef=0.68, k1=2.0)
    a = oracle.fdk_mid(fp, **mid)
    w = oracle.fdk_weights(N_2pi, 0.68, 8)
    b = oracle.fdk_core(fp, w, N_2pi=N_2pi, XN=5, YN=4, ZN=6, XNC=2.0, YNC=1.5, ZNC=-4.1,
                        dx=16.0 / 5, dy=16.0 / 4, dz=16.0 / 6, YLC=19.5, ZLC=3.5, dYL=1.0,
                        dZL=3.0, h=16.0, BetaS=math.pi, DeltaFai=2 * math.pi / N_2pi,
                        N_pi=N_2pi / 2.0, HSCoef=0.68, SO=SO, SD=SD, k1=2.0)
    assert np.array_equal(a, b, equal_nan=True)
    assert np.abs(a).max() > 0


def test_threads_do_not_change_results():
    g = geom(k1=2.0, XN=8, YN=7)
    w = oracle.fdk_weights(g["N_2pi"], g["HSCoef"], 8)
    fp = np.random.default_rng(3).normal(size=(192, 16, 64))
    assert np.array_equal(run(fp, w, g, nthreads=1), run(fp, w, g, nthreads=4))


def test_reference_nan_at_exact_conjugate_offset():
    """Quirk of :99-102: a view without a conjugate (w2 = 0) whose
    temp2 -+ 0.5 is exactly 0 gives Gc = 0 and weight = 0/0 = NaN.  The oracle
    keeps it; the GPU path uses the limit value (weight 1)."""
    g = geom(XN=1, YN=1, ZN=1, dx=1.0, dy=1.0, dz=1.0, XNC=0.0, YNC=0.0, ZNC=-16.0, h=16.0,
             k1=2.0)
    w = oracle.fdk_weights(64, 0.68, 8)
    out = run(np.ones((128, 16, 64)), w, g)
    assert np.isnan(out).all()


@pytest.mark.parametrize("x,y,k1", [(80.0, -40.0, 2.85), (-120.0, 65.0, 5.0), (30.0, 140.0, 1.5)])
def test_3d_weight_uses_geometric_cone_angles(x, y, k1):
    """One view p with a conjugate inside the window (w1, w2 > 0).  The 3D
    weight w1 Gc^k1 / (w2 G^k1 + w1 Gc^k1) must use the tangents of the cone
    angles measured from the sources: G = (z - z_src)/|V - S| for the source S
    on the circle behind the voxel along the parallel ray, and Gc with the
    conjugate source S' on the opposite side (distance |V - S'|) and the
    reference's +-0.5 mm conjugate offset.  Distances come from explicit 2-D
    vectors here, not from the oracle's temp1 +- UU."""
    V0, h, N_2pi = 0.7, 12.0, 64
    dF = 2 * math.pi / N_2pi
    s0 = 15                                       # p = 0 is 15 views before s0
    z = h * (V0 + s0 * dF) / (2 * math.pi) - 0.04     # |offset| < h/(2 N_2pi): s0 stays 15
    g = _single_view(x, y, z, V0, h=h, k1=k1, dYL=6.0, ZL=64, dZL=2.0, ZLC=31.5)
    w = oracle.fdk_weights(N_2pi, g["HSCoef"], 8)
    c = 2.0
    out = run(np.full((1, 64, 64), c), w, g)[0, 0, 0]
    # geometry of the parallel ray through (x, y) at view V0
    d = -np.array([math.cos(V0), math.sin(V0)])           # ray direction
    e = np.array([-math.sin(V0), math.cos(V0)])           # lateral axis
    vox = np.array([x, y])
    t = vox @ e
    s_near = t * e - math.sqrt(SO * SO - t * t) * d       # source behind the voxel
    s_far = t * e + math.sqrt(SO * SO - t * t) * d        # conjugate source
    temp2 = z - h * V0 / (2 * math.pi)
    G = abs(temp2) / np.linalg.norm(vox - s_near)
    Gc = abs(temp2 - 0.5) / np.linalg.norm(vox - s_far)   # p < s0: -0.5 (:99)
    N_pi = N_2pi / 2
    d1 = int(N_pi - s0)
    w1, w2 = w[d1], w[int(d1 + N_pi)]
    assert w1 > 0 and w2 > 0
    weight = w1 * Gc ** k1 / (w2 * G ** k1 + w1 * Gc ** k1)
    assert out == pytest.approx(dF * weight * c, rel=1e-10)
