"""Geometry derivations of the MATLAB drivers, checked against hand-derived
values of the two scanner configurations (closed forms of
src/analyticalReconBigPatient.m:35-101, src/Parallel_Rebinning_CB_Curve.m:8-18,
src/Parallel_FDK_Helical_3DWeighting.cpp:145-170,222-263, src/filtering.cu:65,
src/analyticalRecon.m:15-16)."""
import math

import pytest

import oracle
from analyticalreconbigpatient_amd import geometry as G


@pytest.mark.parametrize("wl,exp", [
    (G.shepp_logan_small, dict(DecAngle=0.39704, dYA=1.078925e-3, dYL=0.8872, dZL=1.9231,
                               hgeom=0.57007, h=17.541, k1=2.8504, W=245, L=491, NN=2048,
                               seg=24, samples=12.72e6)),
    (G.body_clinical, dict(DecAngle=0.95809, dYA=1.078925e-3, dYL=0.9493, dZL=1.9231,
                           hgeom=0.56095, h=69.041, k1=2.8048, W=335, L=671, NN=4096, seg=77,
                           samples=671.07e6)),
])
def test_derived_numbers(wl, exp):
    w = wl()
    s = w.setup()
    a = G.rebin_args(s.geom)
    p = G.bp_params(G.bp_scan_geom(s))
    assert s.DecAngle == pytest.approx(exp["DecAngle"], rel=2e-5)
    assert float(a.dYA) == pytest.approx(exp["dYA"], rel=1e-5)
    assert s.dYL == pytest.approx(exp["dYL"], rel=1e-4)
    assert float(a.DeltaT) == pytest.approx(exp["dYL"], rel=1e-4)      # DeltaT == dYL
    assert s.dZL == pytest.approx(exp["dZL"], rel=1e-4)
    assert s.geom.ScDet[8] == pytest.approx(exp["hgeom"], rel=1e-4)
    assert p.h == pytest.approx(exp["h"], rel=1e-4)
    assert p.k1 == pytest.approx(exp["k1"], rel=1e-4)
    assert G.window_half_width(p.N_2pi, p.HSCoef) == exp["W"]
    assert 2 * exp["W"] + 1 == exp["L"]
    assert oracle.filter_nn(p.YL) == exp["NN"]
    assert G.seg_overlap(w.hdr) == exp["seg"]
    assert p.YL * p.ZL * p.PN == pytest.approx(exp["samples"], rel=1e-3)
    # rebinning step is the segment's (BetaE-BetaS)/(ViewN-1) (analyticalRecon.m's segment arithmetic)
    assert float(a.DeltaFai) == pytest.approx(2 * math.pi / p.N_2pi * p.PN / (p.PN - 1), rel=1e-6)
    assert p.YLC == (p.YL - 1) / 2


def test_clinical_slab_is_fully_covered():
    w = G.body_clinical()
    p = G.bp_params(G.bp_scan_geom(w.setup()))
    assert p.dz == pytest.approx(1.25)
    assert G.voxel_updates(p) == 512 * 512 * 400 * 671                  # 7.036e10


def test_segment_trimming():
    hdr = G.body_clinical().hdr
    g = G.body_clinical().setup().geom
    ov = 77
    f, n, g1 = G.trim_segment(g, 1, 3, ov, 984)
    assert (f, n) == (0, 11808 - ov) and g1.Proj[0] == g.Proj[0]
    f, n, g3 = G.trim_segment(g, 3, 3, ov, 984)
    assert (f, n) == (ov, 11808 - ov)
    assert g3.Proj[0] == pytest.approx(g.Proj[0] + ov * 2 * math.pi / 984)
    f, n, g2 = G.trim_segment(g, 2, 3, ov, 984)
    assert (f, n) == (ov, 11808 - 2 * ov)
    assert G.seg_count(hdr, 4000, ov) == math.ceil((11808 - 4000) / (4000 - 2 * ov)) + 1
