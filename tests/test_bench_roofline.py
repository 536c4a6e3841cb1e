"""The bench's VALU-issue roofline, from synthetic counter values (no GPU).

The LDS-staged backprojector's roof counts issue slots: a quad-cycle that
dual-issues two FP32 FMAs spends one slot.  A kernel that keeps every slot of
every CU busy at the peak clock must land at exactly frac 1; the instruction
count alone (dual issue ignored) would put it above 1.
"""
from __future__ import annotations

import importlib.util
from pathlib import Path
from types import SimpleNamespace as NS

import pytest

ROOT = Path(__file__).resolve().parents[1]


def _bench():
    spec = importlib.util.spec_from_file_location("arecon_bench", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _fdk(storage="f32", nz=8):
    views, per_view = 11808, 512 * 64 * 4
    return NS(params=NS(XN=512, YN=512, ZN=400), updates=512 * 512 * 400 * 671,
              storage=storage, layout="rowdiff", plan=NS(inside=True), variant=3,
              plan_nz=nz, group=0, n_views=views,
              filtered=[NS(numel=lambda: per_view // 4)])


@pytest.mark.parametrize("dual_share", [0.0, 0.3])
def test_issue_bound_kernel_at_peak_clock_is_frac_one(dual_share):
    b = _bench()
    fdk = _fdk()
    wave_updates = fdk.updates / 64.0
    valu = 14.0 * wave_updates                 # non-transcendental VALU instructions
    trans = 0.5 * wave_updates
    dual = dual_share * valu                   # quad-cycles that issued two
    slots = valu - dual
    ms = slots / b.N_CU / b.CLOCK_HZ * 1e3     # every slot of every CU busy at the peak clock
    name = b.kernel_name(fdk)
    counters = {"kernels": {f"void arecon::{name}(params)": {
        "SQ_INSTS_VALU": valu + trans, "SQ_INSTS_VALU_TRANS_F32": trans,
        "SQ_ACTIVE_INST_VALU2": dual, "fetch_bytes": 4.0e10, "write_bytes": 4.2e8}}}
    roof = b.make_roofline(fdk, {"stages": [1.0, 2.0, ms]}, {}, {}, counters, None)
    assert roof["bound"] == "valu_issue"
    assert roof["frac"] == pytest.approx(1.0, abs=1e-3)
    assert roof["issue_slots_per_update"] == pytest.approx(14.0 * (1 - dual_share), abs=1e-3)
    assert roof["valu_per_update"] == pytest.approx(14.0, abs=1e-3)
    # the same time priced per instruction would claim more than the roof
    if dual_share:
        per_instr_peak = b.N_CU * b.CLOCK_HZ * 64 / 14.0 / 1e9
        assert roof["achieved"] / per_instr_peak > 1.2


def test_no_counters_falls_back_to_the_hbm_roof():
    b = _bench()
    roof = b.make_roofline(_fdk(), {"stages": [1.0, 2.0, 25.0]}, {}, {}, None, None)
    assert roof["bound"] == "hbm"
    assert 0 < roof["frac"] < 1
