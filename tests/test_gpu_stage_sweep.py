"""Seeded sweeps of random shapes and geometries through the rebinning and the
filter, through the C ABI: rebinning bit-exact with the FP32 oracle (both
output layouts), the filter within 1e-5 relative L2 of the FP64 oracle (rows)
and its FP32 row differences against those formed from its rows.

Each case draws its sizes and geometry from its own seed and prints them before
it runs.  Rebinning: 1 to 242 detector rows (vector and scalar paths, 128- and
64-channel blocks of the transposing kernel), channel centres and parallel
widths that reach the clamps at the detector's and the scan's ends.  Filter:
1 to 64 rows (odd counts: the last row pair has one row), 8 to 1,500 channels
(one- to four-wavefront transforms and the general kernel beyond 1,024).
"""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle
from analyticalreconbigpatient_amd import _lib

from .test_gpu_stages import SO, rebin_args, rel_l2, row_diffs, run_filter, stream

pytestmark = pytest.mark.gpu


def draw_rebin(seed):
    r = np.random.default_rng(8000 + seed)
    while True:
        ViewN, YL, ZL = int(r.integers(3, 301)), int(r.integers(16, 421)), int(r.integers(1, 243))
        if ViewN * YL * ZL <= 8_000_000:
            break
    if r.uniform() < 0.5:
        ZL = max(4, ZL // 4 * 4)                   # the float4 path
    N_2pi = int(r.choice([48, 64, 96, 128]))
    a = rebin_args(YL, ViewN, N_2pi=N_2pi, YLC0=(YL - 1) / 2.0 + float(r.uniform(-3, 3)))
    a["PLC0"] = np.float32(float(a["PLC0"]) + float(r.uniform(-5, 5)))
    a["DeltaT"] = np.float32(float(a["DeltaT"]) * float(r.uniform(0.6, 1.4)))
    layout = [_lib.ARECON_LAYOUT_VYZ, _lib.ARECON_LAYOUT_VZY][seed % 2]
    return (ViewN, YL, ZL), a, layout


def rebin_on_gpu(P, a, layout, dev):
    ViewN, YL, ZL = P.shape
    din = torch.from_numpy(P).to(dev)
    dout = torch.empty_like(din)
    _lib.check(_lib.lib().arecon_rebin_dev(
        C.c_void_p(dout.data_ptr()), C.c_void_p(din.data_ptr()), YL, ZL, ViewN, a["YLC0"],
        a["dYA"], a["DeltaTheta"], a["PLC0"], a["DeltaT"], a["DeltaFai"], a["SO"], layout,
        stream()), "rebin")
    got = dout.cpu().numpy()
    if layout == _lib.ARECON_LAYOUT_VZY:
        got = got.reshape(ViewN, ZL, YL).transpose(0, 2, 1)
    return got


@pytest.mark.parametrize("case", range(12))
def test_random_rebin_bit_exact(gpu, case):
    shape, a, layout = draw_rebin(case)
    print(f"rebin sweep case {case}: {shape} layout={layout} "
          + " ".join(f"{k}={float(v):.6g}" for k, v in a.items()), flush=True)
    P = np.random.default_rng(case).normal(size=shape).astype(np.float32)
    want = oracle.rebin_f32(P, **a)
    got = rebin_on_gpu(P, a, layout, gpu)
    assert np.array_equal(got, want), f"max |diff| {np.abs(got - want).max()}"


def draw_filter(seed):
    r = np.random.default_rng(9000 + seed)
    ViewN, ZL = int(r.integers(1, 13)), int(r.integers(1, 65))
    YL = int(r.integers(1025, 1501)) if seed % 6 == 5 else int(r.integers(8, 1025))
    PLC0 = (YL - 1) / 2.0 + float(r.uniform(-20, 20))
    ZLC0 = (ZL - 1) / 2.0 + float(r.uniform(-2, 2))
    return (ViewN, ZL, YL), (PLC0, ZLC0, float(r.uniform(0.5, 1.5)), float(r.uniform(0.5, 2.0)))


@pytest.mark.parametrize("case", range(12))
def test_random_filter(gpu, case):
    shape, args = draw_filter(case)
    print(f"filter sweep case {case}: {shape} PLC0={args[0]:.3f} ZLC0={args[1]:.3f} "
          f"dYL={args[2]:.4f} dZL={args[3]:.4f}", flush=True)
    P = np.random.default_rng(case).normal(size=shape).astype(np.float32)
    want = oracle.filter_f64(P.astype(np.float64), *[float(np.float32(v)) for v in args], SO)
    rows = run_filter(P, *args)
    assert rel_l2(rows, want) < 1e-5
    if shape[1] >= 2:
        got = run_filter(P, *args, layout=_lib.ARECON_FP_ROWDIFF)
        assert rel_l2(got, row_diffs(rows)) < 1e-6
