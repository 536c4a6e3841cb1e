"""Pins the CPU oracle's rebinning (oracle/arecon_oracle.c, restating
src/ParallelRebinningCBCurve.cpp:22-97) against closed forms.

The reference ships no tests or golden vectors and its C++ cannot be compiled
here (every translation unit includes MATLAB's <mex.h>), so the oracle is
checked against what the mathematics fixes:

* the fan -> parallel coordinates of the textbook rebinning: the parallel ray
  (theta, t) is the fan ray of source angle beta = theta + asin(t/SO) at fan
  angle gamma = asin(t/SO);
* bilinear interpolation reproduces inputs affine in (fan view, channel) at
  those coordinates, and passes rows through untouched;
* the ">" edge clamps of :55-85 (last view and channel are repeated).
"""
import math

import numpy as np
import pytest

import oracle

SO = 541.0


def _args(YL=40, ZL=6, ViewN=50, N_2pi=64):
    dYA = 2 * math.atan(1.0239 * 368 / YL / 949.0 / 2)       # fan angle per channel
    YLC0 = (YL - 1) / 2.0
    DeltaFai = 2 * math.pi / N_2pi * ViewN / (ViewN - 1)      # :15 of the .m helper
    DeltaT = math.tan(YL * dYA * 0.5) * SO * 1.5 * 2 / YL
    PLC0 = (YL - 1) / 2.0
    return dict(YLC0=YLC0, dYA=dYA, DeltaTheta=DeltaFai, PLC0=PLC0, DeltaT=DeltaT,
                DeltaFai=DeltaFai, SO=SO), (YL, ZL, ViewN)


def _textbook_coords(a, YL, ViewN):
    """(fan view index, channel index) of every parallel (view, channel)."""
    i = np.arange(ViewN)[:, None]
    j = np.arange(YL)[None, :]
    t = (j - a["PLC0"]) * a["DeltaT"]
    gamma = np.arcsin(t / a["SO"])
    beta_rel = i * a["DeltaTheta"] + gamma
    return beta_rel / a["DeltaFai"], gamma / a["dYA"] + a["YLC0"]


def test_constant_is_reproduced():
    a, (YL, ZL, ViewN) = _args()
    P = np.full((ViewN, YL, ZL), 3.25, dtype=np.float64)
    out = oracle.rebin_f64(P, **a)
    np.testing.assert_allclose(out, 3.25, rtol=1e-14)
    out32 = oracle.rebin_f32(P.astype(np.float32), **a)
    np.testing.assert_allclose(out32, 3.25, rtol=2e-6)


def test_affine_input_lands_on_textbook_coordinates():
    a, (YL, ZL, ViewN) = _args()
    alpha, beta, gam, c0 = 0.7, -0.3, 0.11, 5.0
    i = np.arange(ViewN)[:, None, None]
    j = np.arange(YL)[None, :, None]
    k = np.arange(ZL)[None, None, :]
    P = alpha * i + beta * j + gam * k + c0 + 0 * (i + j + k)
    out = oracle.rebin_f64(P, **a)
    fi, ui = _textbook_coords(a, YL, ViewN)
    expect = alpha * fi[:, :, None] + beta * ui[:, :, None] + gam * k + c0
    # interior: both interpolation intervals inside the array (no clamping)
    inside = (np.ceil(fi) >= 1) & (np.ceil(fi) <= ViewN - 1) & (np.ceil(ui) >= 1) & (np.ceil(ui) <= YL - 1)
    assert inside.sum() > 0.5 * inside.size
    np.testing.assert_allclose(out[inside], expect[inside], rtol=0, atol=1e-9)
    out32 = oracle.rebin_f32(P.astype(np.float32), **{kk: np.float32(v) for kk, v in a.items()})
    np.testing.assert_allclose(out32[inside], expect[inside], rtol=0, atol=2e-3)


def test_rows_are_not_mixed():
    a, (YL, ZL, ViewN) = _args()
    rng = np.random.default_rng(0)
    rowvals = rng.normal(size=ZL)
    P = np.broadcast_to(rowvals, (ViewN, YL, ZL)).copy()
    out = oracle.rebin_f64(P, **a)
    np.testing.assert_allclose(out, np.broadcast_to(rowvals, out.shape), rtol=1e-13, atol=1e-13)


def test_edge_clamps_repeat_last_view_and_channel():
    a, (YL, ZL, ViewN) = _args()
    i = np.arange(ViewN, dtype=np.float64)[:, None, None]
    P = (i * i + 1.0) * np.ones((ViewN, YL, ZL))
    out = oracle.rebin_f64(P, **a)
    fi, ui = _textbook_coords(a, YL, ViewN)
    FI = np.ceil(fi)
    hi = FI > ViewN - 1                   # :60 -- both taps on the last view
    lo = FI <= 0                          # :55 -- both taps on view 0
    assert hi.any() and lo.any()
    np.testing.assert_allclose(out[hi], (ViewN - 1) ** 2 + 1.0, rtol=1e-13)
    np.testing.assert_allclose(out[lo], 1.0, rtol=1e-13)
    # channel clamps: input depends on channel only
    j = np.arange(YL, dtype=np.float64)[None, :, None]
    Q = (j * j + 2.0) * np.ones((ViewN, YL, ZL))
    outq = oracle.rebin_f64(Q, **a)
    UI = np.broadcast_to(np.ceil(ui), (ViewN, YL))
    assert (UI > YL - 1).any() and (UI <= 0).any()   # outer t clamp (1.5x wider grid)
    np.testing.assert_allclose(outq[UI > YL - 1], (YL - 1) ** 2 + 2.0, rtol=1e-13)
    np.testing.assert_allclose(outq[UI <= 0], 2.0, rtol=1e-13)


def test_float_and_double_agree():
    a, (YL, ZL, ViewN) = _args(YL=64, ZL=8, ViewN=80)
    rng = np.random.default_rng(1)
    P = rng.normal(size=(ViewN, YL, ZL))
    o64 = oracle.rebin_f64(P, **a)
    o32 = oracle.rebin_f32(P.astype(np.float32), **{kk: np.float32(v) for kk, v in a.items()})
    err = np.linalg.norm(o32 - o64) / np.linalg.norm(o64)
    assert err < 1e-5


@pytest.mark.parametrize("flip", [False, True])
def test_central_channel_copies_fan_channel(flip):
    """t = 0 (channel PLC) maps to fan angle 0, i.e. channel YLC, and to fan
    view theta/DeltaFai = i exactly: the output equals the input there."""
    a, (YL, ZL, ViewN) = _args(YL=41)
    a["PLC0"] = 20.0
    a["YLC0"] = 20.0
    rng = np.random.default_rng(2)
    P = rng.normal(size=(ViewN, YL, ZL))
    if flip:
        P = P[:, ::-1, :].copy()
    out = oracle.rebin_f64(P, **a)
    np.testing.assert_allclose(out[1:-1, 20, :], P[1:-1, 20, :], rtol=1e-12, atol=1e-12)
