"""Pins the CPU oracle's preweighting + ramp filter (oracle/arecon_oracle.c,
restating the MATLAB CPU branch src/analyticalReconBigPatient.m:107-134 and
src/filtering.cu:16-25,57-84,187-207).

Independent checks (the reference has no tests; its filter is GPU/MATLAB only):
* numpy's FFT run through the same MATLAB steps (fft(row, NN) .* fft(HS), ifft);
* direct O(YL^2) linear convolution;
* the telescoping closed form of a constant row:
      sum_{n=a}^{b} -2/(pi^2 (4n^2-1)) = -(1/pi^2) (1/(2a-1) - 1/(2b+1));
* the kernel's frequency response |sin(w/2)| / (pi dYL) (Shepp-Logan ramp);
* the preweight's special cases t = 0 (cone cosine SO/sqrt(SO^2+b^2)) and b = 0 (1).
"""
import math

import numpy as np
import pytest

import oracle

SO = 541.0


def h_taps(n, dYL):
    n = np.asarray(n, dtype=np.float64)
    return -2.0 / (math.pi ** 2 * (4.0 * n * n - 1.0)) / dYL


@pytest.mark.parametrize("YL,NN", [(368, 2048), (888, 4096), (48, 256), (64, 256), (100, 512)])
def test_transform_length(YL, NN):
    assert oracle.filter_nn(YL) == NN            # 2^ceil(log2(3 YL)), filtering.cu:65


def test_kernel_layout_wraps_negative_taps():
    YL, dYL = 20, 0.9
    NN = oracle.filter_nn(YL)
    HS = oracle.filter_kernel(YL, dYL)
    np.testing.assert_allclose(HS[:YL + 1], h_taps(np.arange(YL + 1), dYL), rtol=1e-15)
    np.testing.assert_allclose(HS[NN - YL:], h_taps(np.arange(-YL, 0), dYL), rtol=1e-15)
    assert np.all(HS[YL + 1:NN - YL] == 0)
    assert HS[0] == pytest.approx(2.0 / (math.pi ** 2 * dYL), rel=1e-15)


def _matlab_numpy(P, PLC0, ZLC0, dYL, dZL):
    """analyticalReconBigPatient.m:107-134 with numpy's FFT (P is [V][ZL][YL])."""
    V, ZL, YL = P.shape
    out = np.empty_like(P)
    j = np.arange(YL)
    t = (j - PLC0) * dYL
    NN = 2 ** math.ceil(math.log(YL * 3) / math.log(2))
    n = np.arange(-YL, YL + 1)
    hsl = -2.0 / ((math.pi ** 2) * (4 * n * n - 1)) / dYL
    HS = np.zeros(NN)
    HS[:YL + 1] = hsl[YL:]
    HS[NN - YL:] = hsl[:YL]
    FtS = np.fft.fft(HS)
    for k in range(ZL):
        b = (k - ZLC0) * dZL
        wgt = SO * SO / np.sqrt(SO ** 4 + (SO * b) ** 2 - (b * t) ** 2)
        rows = P[:, k, :] * wgt
        out[:, k, :] = np.real(np.fft.ifft(FtS * np.fft.fft(rows, NN, axis=1), axis=1))[:, :YL]
    return out


def test_matches_numpy_fft_restatement():
    rng = np.random.default_rng(3)
    V, ZL, YL = 5, 4, 96
    P = rng.normal(size=(V, ZL, YL))
    args = dict(PLC0=47.5, ZLC0=1.5, dYL=0.95, dZL=1.92)
    got = oracle.filter_f64(P, args["PLC0"], args["ZLC0"], args["dYL"], args["dZL"], SO)
    want = _matlab_numpy(P, **args)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * np.abs(want).max())


def test_matches_direct_convolution():
    rng = np.random.default_rng(4)
    V, ZL, YL = 3, 3, 70
    P = rng.normal(size=(V, ZL, YL))
    PLC0, ZLC0, dYL, dZL = 34.5, 1.0, 0.8, 2.0
    got = oracle.filter_f64(P, PLC0, ZLC0, dYL, dZL, SO)
    pw = oracle.preweight_f64(P, PLC0, ZLC0, dYL, dZL, SO)
    taps = h_taps(np.arange(-(YL - 1), YL), dYL)
    want = np.empty_like(P)
    for v in range(V):
        for k in range(ZL):
            full = np.convolve(pw[v, k], taps)            # index m + n + (YL-1)
            want[v, k] = full[YL - 1:2 * YL - 1]
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * np.abs(want).max())


def test_constant_row_telescoping_closed_form():
    YL, dYL, c = 128, 0.7, 2.5
    P = np.full((1, 1, YL), c)
    # ZL = 1, ZLC0 = 0 -> b = 0 -> the preweight is exactly 1
    got = oracle.filter_f64(P, 63.5, 0.0, dYL, 1.0, SO)[0, 0]
    j = np.arange(YL)
    a_, b_ = j - (YL - 1), j
    want = c * (-1.0 / (math.pi ** 2 * dYL)) * (1.0 / (2 * a_ - 1) - 1.0 / (2 * b_ + 1))
    np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-13)


def test_impulse_response_is_the_kernel():
    YL, dYL, m0 = 60, 1.1, 23
    P = np.zeros((1, 1, YL))
    P[0, 0, m0] = 1.0
    got = oracle.filter_f64(P, 29.5, 0.0, dYL, 1.0, SO)[0, 0]
    np.testing.assert_allclose(got, h_taps(np.arange(YL) - m0, dYL), rtol=1e-11, atol=1e-14)


def test_frequency_response_is_shepp_logan():
    """A cosine row far from the ends is scaled by |sin(w/2)|/(pi dYL); the
    residual is the kernel tail beyond the row, sum_{|n|>m} |h(n)| ~ 1/(pi^2 m dYL)."""
    YL, dYL = 2048, 1.0
    for w in (0.3, 1.1, 2.4):
        row = np.cos(w * np.arange(YL))
        got = oracle.filter_f64(row[None, None, :], 1023.5, 0.0, dYL, 1.0, SO)[0, 0]
        mid = slice(YL // 2 - 64, YL // 2 + 64)
        want = oracle.sl_response(w, dYL) * row[mid]
        tail = 2.0 / (math.pi ** 2 * (YL // 2 - 64) * dYL)
        np.testing.assert_allclose(got[mid], want, rtol=0, atol=tail)


def test_preweight_special_cases():
    YL, ZL = 9, 5
    PLC0, ZLC0, dYL, dZL = 4.0, 2.0, 1.3, 3.7
    P = np.ones((1, ZL, YL))
    w = oracle.preweight_f64(P, PLC0, ZLC0, dYL, dZL, SO)[0]
    b = (np.arange(ZL) - ZLC0) * dZL
    np.testing.assert_allclose(w[:, 4], SO / np.sqrt(SO * SO + b * b), rtol=1e-15)   # t = 0
    np.testing.assert_allclose(w[2, :], 1.0, rtol=1e-15)                            # b = 0


def test_threads_do_not_change_results():
    rng = np.random.default_rng(5)
    P = rng.normal(size=(6, 4, 50))
    a = oracle.filter_f64(P, 24.5, 1.5, 0.9, 1.9, SO, nthreads=1)
    b = oracle.filter_f64(P, 24.5, 1.5, 0.9, 1.9, SO, nthreads=4)
    assert np.array_equal(a, b)
