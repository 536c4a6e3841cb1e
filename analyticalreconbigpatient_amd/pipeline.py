"""Device-resident helical FDK pipeline: rebin -> preweight + filter -> backproject.

The reference runs its three stages as separate MEX calls, each copying the
whole sinogram host -> device -> host, with host-side permutes in between
(src/analyticalReconBigPatient.m:79,138-146, src/Parallel_Rebinning_CB_Curve.m:20-23).
Here the sinogram stays in HBM: the rebinning writes the filter's
``[ViewN][ZL][YL]`` layout directly, the filter runs in place (FP32) or into a
16-bit buffer, and the backprojector reads the trimmed view range by pointer
offset.  torch supplies device memory and the stream; all compute is in
libarecon.so.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .geometry import (BPParams, ReconSetup, bp_params, bp_scan_geom, filter_args, rebin_args,
                       voxel_updates, window_half_width)

STORAGE = {"f32": (_lib.ARECON_F32, torch.float32), "f16": (_lib.ARECON_F16, torch.float16),
           "bf16": (_lib.ARECON_BF16, torch.bfloat16)}
# Filtered-sinogram layouts: the reference's rows [ViewN][ZL][YL]; row pairs
# [ViewN][ZL-1][YL][2] (rows r and r+1 of each channel side by side); difference
# quads [ViewN][ZL-1][YL-1][4] (sample, row, channel and cross differences).
LAYOUT = {"rows": _lib.ARECON_FP_ROWS, "rowpairs": _lib.ARECON_FP_ROWPAIRS,
          "quads": _lib.ARECON_FP_QUADS, "rowdiff": _lib.ARECON_FP_ROWDIFF}


def filtered_shape(ViewN: int, ZL: int, YL: int, layout: str) -> tuple:
    if layout == "rows":
        return (ViewN, ZL, YL)
    if layout in ("rowpairs", "rowdiff"):
        return (ViewN, ZL - 1, YL, 2)
    return (ViewN, ZL - 1, YL - 1, 4)


def _stream_ptr(stream: torch.cuda.Stream | None) -> C.c_void_p:
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def fdk_weights(N_2pi: int, HSCoef: float, delta: int) -> np.ndarray:
    """The reference's weight table (src/Parallel_FDK_Helical_3DWeighting.cpp:156-170),
    built by libarecon.so."""
    w = np.empty(N_2pi, dtype=np.float64)
    _lib.check(_lib.lib().arecon_fdk_weights(_lib.dptr(w), N_2pi, HSCoef, delta),
               "arecon_fdk_weights")
    return w


def choose_bp_kernel(params: BPParams, dz: float | None = None, storage: str = "f32") -> dict:
    """Backprojector launch shape for a geometry.  Consecutive slices' windows
    are N_2pi dz / h views apart; when a view lies in the windows of at least
    four consecutive slices, the LDS-staged z-column kernel (variant 3) shares
    the per-(column, view) work among a block of slices and reads each view's
    footprint from LDS, as row differences (f and f(r+1) - f per channel: four
    instructions per blend); a workgroup whose footprints do not fit runs the
    gathering z-column body (variant 2).  Blocks of 8 slices for FP32; in FP16 /
    BF16 a row difference takes 4 bytes, so 16-slice blocks (twice the slices
    per view's work) still leave seven workgroups per CU, where FP32's would
    leave three.  Otherwise (thick slices, few shared views) the view-table
    kernel runs on the rows layout."""
    dzv = params.dz if dz is None else dz
    L = 2 * window_half_width(params.N_2pi, params.HSCoef) + 1
    shift = abs(params.N_2pi * dzv / params.h)
    if shift * 4 <= L:
        return dict(variant=3, nz=8 if storage == "f32" else 16, group=8, order=2,
                    layout="rowdiff")
    return dict(variant=0, nz=0, group=0, order=0, layout="rows")


class BPPlan:
    """Owner of an ``arecon_bp_plan`` (tables built in FP64, stored FP32)."""

    def __init__(self, params: BPParams, *, dz: float | None = None, nz: int = 0,
                 order: int = 0, variant: int = 0, group: int = 0,
                 w: np.ndarray | None = None):
        self.params = params
        self.geom = params.to_geom(dz=dz, nz=nz, order=order, variant=variant, group=group)
        self.w = fdk_weights(params.N_2pi, params.HSCoef, params.delta) if w is None else \
            np.ascontiguousarray(w, dtype=np.float64)
        h = C.c_void_p()
        _lib.check(_lib.lib().arecon_bp_plan_create(C.byref(self.geom), _lib.dptr(self.w),
                                                    C.byref(h)), "arecon_bp_plan_create")
        self._h = h
        self.nz = _lib.lib().arecon_bp_plan_nz(h)

    def execute(self, img: torch.Tensor, fp: torch.Tensor, stream=None,
                layout: str = "rows") -> None:
        g = self.geom
        dtype = {torch.float32: _lib.ARECON_F32, torch.float16: _lib.ARECON_F16,
                 torch.bfloat16: _lib.ARECON_BF16}[fp.dtype]
        assert img.is_cuda and fp.is_cuda and img.dtype == torch.float32
        assert img.is_contiguous() and fp.is_contiguous()
        assert img.numel() == g.XN * g.YN * g.ZN, "volume size does not match the plan"
        per_view = math.prod(filtered_shape(1, g.ZL, g.YL, layout))
        assert fp.numel() >= g.PN * per_view, "fp holds fewer than PN views"
        _lib.check(_lib.lib().arecon_bp_plan_execute(self._h, C.c_void_p(img.data_ptr()),
                                                     C.c_void_p(fp.data_ptr()), dtype,
                                                     LAYOUT[layout], _stream_ptr(stream)),
                   "arecon_bp_plan_execute")

    @property
    def nblocks(self) -> int:
        """Slice blocks (of ``nz`` consecutive slices) of the plan."""
        return _lib.lib().arecon_bp_plan_nblocks(self._h)

    def block_views(self, b: int) -> tuple[int, int]:
        """Views [pmin, pmax] that slice block b reads (pmin > pmax: none)."""
        lo, hi = C.c_int(), C.c_int()
        _lib.check(_lib.lib().arecon_bp_plan_block_views(self._h, b, C.byref(lo), C.byref(hi)),
                   "arecon_bp_plan_block_views")
        return lo.value, hi.value

    def execute_blocks(self, img: torch.Tensor, fp: torch.Tensor, b0: int, b1: int, stream=None,
                       layout: str = "rows") -> None:
        """Backprojects only slice blocks [b0, b1)."""
        dtype = {torch.float32: _lib.ARECON_F32, torch.float16: _lib.ARECON_F16,
                 torch.bfloat16: _lib.ARECON_BF16}[fp.dtype]
        _lib.check(_lib.lib().arecon_bp_plan_execute_blocks(
            self._h, C.c_void_p(img.data_ptr()), C.c_void_p(fp.data_ptr()), dtype,
            LAYOUT[layout], b0, b1, _stream_ptr(stream)), "arecon_bp_plan_execute_blocks")

    @property
    def inside(self) -> bool:
        """Every column inside the source circle and the detector: the z-column
        kernels run their specialisation without per-column validity tests."""
        return bool(_lib.lib().arecon_bp_plan_inside(self._h))

    def stats(self) -> tuple[int, int]:
        """LDS-staged kernel (variant 3): workgroups that read their views from
        LDS and workgroups that fell back to global gathers, over every execute
        so far (both 0 for other variants)."""
        a, b = C.c_ulonglong(), C.c_ulonglong()
        _lib.check(_lib.lib().arecon_bp_plan_stats(self._h, C.byref(a), C.byref(b)),
                   "arecon_bp_plan_stats")
        return a.value, b.value

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().arecon_bp_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rebin_dev(out: torch.Tensor, raw: torch.Tensor, setup_geom, layout: int, stream=None):
    a = rebin_args(setup_geom)
    assert raw.dtype == torch.float32 and raw.is_contiguous() and out.is_contiguous()
    assert raw.numel() == a.YL * a.ZL * a.ViewN and out.numel() == raw.numel()
    _lib.check(_lib.lib().arecon_rebin_dev(
        C.c_void_p(out.data_ptr()), C.c_void_p(raw.data_ptr()), a.YL, a.ZL, a.ViewN, a.YLC0,
        a.dYA, a.DeltaTheta, a.PLC0, a.DeltaT, a.DeltaFai, a.SO, layout, _stream_ptr(stream)),
        "arecon_rebin_dev")


def rebin_views_dev(out: torch.Tensor, raw: torch.Tensor, setup_geom, v0: int, v1: int,
                    stream=None):
    """Rebinning of output views [v0, v1) into the full [ViewN][ZL][YL] buffer."""
    a = rebin_args(setup_geom)
    _lib.check(_lib.lib().arecon_rebin_views_dev(
        C.c_void_p(out.data_ptr()), C.c_void_p(raw.data_ptr()), a.YL, a.ZL, a.ViewN, a.YLC0,
        a.dYA, a.DeltaTheta, a.PLC0, a.DeltaT, a.DeltaFai, a.SO, v0, v1 - v0,
        _stream_ptr(stream)), "arecon_rebin_views_dev")


def filter_dev(out: torch.Tensor, inp: torch.Tensor, setup: ReconSetup, stream=None,
               layout: str = "rows"):
    fa = filter_args(setup)
    ViewN, ZL, YL = inp.shape
    dtype = {torch.float32: _lib.ARECON_F32, torch.float16: _lib.ARECON_F16,
             torch.bfloat16: _lib.ARECON_BF16}[out.dtype]
    assert inp.dtype == torch.float32 and inp.is_contiguous() and out.is_contiguous()
    assert out.numel() == math.prod(filtered_shape(ViewN, ZL, YL, layout))
    _lib.check(_lib.lib().arecon_filter_dev(
        C.c_void_p(out.data_ptr()), dtype, LAYOUT[layout], C.c_void_p(inp.data_ptr()), YL, ZL,
        ViewN, fa.PLC0, fa.ZLC0, fa.dYL, fa.dZL, fa.SO, _stream_ptr(stream)),
        "arecon_filter_dev")


@dataclass
class StageTimes:
    rebin_ms: float
    filter_ms: float
    bp_ms: float

    @property
    def total_ms(self) -> float:
        return self.rebin_ms + self.filter_ms + self.bp_ms


class HelicalFDK:
    """One segment's reconstruction, all stages on one device.

    ``first_view``/``n_views`` select the backprojected views of the rebinned
    sinogram (the segment trimming of src/analyticalReconBigPatient.m:80-93);
    ``bp_geom`` is the ScanGeom after trimming (BetaS shifted accordingly).
    ``storage`` is the element type of the filtered sinogram and ``layout`` its
    layout ("rows", or "rowpairs", "rowdiff" or "quads" for the z-column
    backprojectors, variants 2 and 3).
    With ``variant=None`` the backprojector's launch shape and the layout come
    from ``choose_bp_kernel``.
    """

    def __init__(self, setup: ReconSetup, *, storage: str = "f32", layout: str | None = None,
                 first_view: int = 0,
                 n_views: int | None = None, bp_geom=None, dz: float | None = None,
                 nz: int = 0, order: int | None = None, variant: int | None = None,
                 group: int = 0,
                 device: str = "cuda"):
        self.setup = setup
        self.ra = rebin_args(setup.geom)
        self.ViewN, self.YL, self.ZL = self.ra.ViewN, self.ra.YL, self.ra.ZL
        self.first_view = first_view
        self.n_views = self.ViewN - first_view if n_views is None else n_views
        g = bp_scan_geom(setup, geom=bp_geom, n_views=self.n_views)
        self.params = bp_params(g)
        self.dz = dz
        if variant is None:       # launch shape chosen from the geometry
            k = choose_bp_kernel(self.params, dz, storage)
            variant, nz, group = k["variant"], nz or k["nz"], group or k["group"]
            layout = layout or k["layout"]
            order = k["order"] if order is None else order
        elif variant == 3 and layout is None:   # not rows, which variant 3 cannot read
            layout = "rowdiff"
        order = 0 if order is None else order
        layout = layout or "rows"
        self.plan = BPPlan(self.params, dz=dz, nz=nz, order=order, variant=variant, group=group)
        self.variant, self.group = variant, group
        self.storage = storage
        self.layout = layout
        code, tdtype = STORAGE[storage]
        dev = torch.device(device)
        shape = (self.ViewN, self.ZL, self.YL)
        self.rebinned = torch.empty(shape, dtype=torch.float32, device=dev)
        if storage == "f32" and layout == "rows":
            self.filtered = self.rebinned          # in place
        else:
            self.filtered = torch.empty(filtered_shape(self.ViewN, self.ZL, self.YL, layout),
                                        dtype=tdtype, device=dev)
        p = self.params
        self.volume = torch.empty((p.YN, p.XN, p.ZN), dtype=torch.float32, device=dev)
        self.updates = voxel_updates(p, dz)
        self.plan_nz = self.plan.nz

    def rebin(self, raw: torch.Tensor, stream=None) -> torch.Tensor:
        rebin_dev(self.rebinned, raw, self.setup.geom, _lib.ARECON_LAYOUT_VZY, stream)
        return self.rebinned

    def filter(self, stream=None) -> torch.Tensor:
        filter_dev(self.filtered, self.rebinned, self.setup, stream, self.layout)
        return self.filtered

    def backproject(self, stream=None) -> torch.Tensor:
        fp = self.filtered[self.first_view:self.first_view + self.n_views]
        self.plan.execute(self.volume, fp, stream, self.layout)
        return self.volume

    def run(self, raw: torch.Tensor, stream=None) -> torch.Tensor:
        self.rebin(raw, stream)
        self.filter(stream)
        return self.backproject(stream)

    def run_overlapped(self, raw: torch.Tensor, nchunks: int = 8, ngroups: int = 8):
        """One pass with the stages overlapped: rebin + filter in ``nchunks``
        view chunks on the current stream, and the backprojection in
        ``ngroups`` launches of consecutive slice blocks on a second stream,
        each waiting only for the chunk that holds the last view its blocks
        read.  Same result as ``run``.  Needs the slice-major workgroup order."""
        main = torch.cuda.current_stream()
        if not hasattr(self, "_bp_stream"):
            self._bp_stream = torch.cuda.Stream()
            self._events = []
        bp = self._bp_stream
        V = self.ViewN
        bounds = [round(V * c / nchunks) for c in range(nchunks + 1)]
        while len(self._events) < nchunks:
            self._events.append(torch.cuda.Event())
        for c in range(nchunks):
            v0, v1 = bounds[c], bounds[c + 1]
            if v1 <= v0:
                continue
            rebin_views_dev(self.rebinned, raw, self.setup.geom, v0, v1, main)
            filter_dev(self.filtered[v0:v1], self.rebinned[v0:v1], self.setup, main, self.layout)
            self._events[c].record(main)
        nb = self.plan.nblocks
        gb = [round(nb * g / ngroups) for g in range(ngroups + 1)]
        fp = self.filtered[self.first_view:self.first_view + self.n_views]
        for g in range(ngroups):
            b0, b1 = gb[g], gb[g + 1]
            if b1 <= b0:
                continue
            need = max((self.plan.block_views(b)[1] for b in range(b0, b1)), default=-1)
            last_view = self.first_view + max(need, 0)   # in rebinned view numbering
            c = next((i for i in range(nchunks) if bounds[i + 1] > last_view), nchunks - 1)
            # waiting on chunk c also orders everything issued on `main` before it
            bp.wait_event(self._events[c])
            self.plan.execute_blocks(self.volume, fp, b0, b1, bp, self.layout)
        main.wait_stream(bp)
        return self.volume

    def fan_view_margin(self) -> int:
        """Fan views beyond a parallel view that its rebinning reads: parallel
        view i, channel j blends fan views ceil((i DeltaTheta + asin(t_j/SO)) /
        DeltaFai) - 1 and the next (src/ParallelRebinningCBCurve.cpp:36-69), with
        |t_j| at most max(PLC, YL-1-PLC) DeltaT."""
        a = self.ra
        t = max(abs(float(a.PLC0)), abs(a.YL - 1 - float(a.PLC0))) * abs(float(a.DeltaT))
        spread = math.asin(min(1.0, t / float(a.SO))) / abs(float(a.DeltaFai))
        return int(math.ceil(spread)) + 2

    def run_streamed(self, host_raw: torch.Tensor, raw: torch.Tensor, nchunks: int = 8,
                     ngroups: int = 8):
        """One pass that starts from the raw sinogram in pinned host memory: its
        views are copied into ``raw`` (on the device) in ``nchunks`` chunks on a
        copy stream, and each chunk of parallel views is rebinned and filtered
        on the current stream as soon as the fan views it reads have arrived,
        while the slice blocks whose views are filtered are backprojected on a
        third stream (as ``run_overlapped``).  The copy of the next pass waits
        for this pass's rebinning.  Same result as copying, then ``run``.
        Needs the slice-major workgroup order."""
        assert host_raw.is_pinned() and raw.is_cuda and host_raw.shape == raw.shape
        main = torch.cuda.current_stream()
        if not hasattr(self, "_bp_stream"):
            self._bp_stream = torch.cuda.Stream()
            self._events = []
        if not hasattr(self, "_copy_stream"):
            self._copy_stream = torch.cuda.Stream()
            self._copy_events = []
        cp, bp = self._copy_stream, self._bp_stream
        V = self.ViewN
        bounds = [round(V * c / nchunks) for c in range(nchunks + 1)]
        while len(self._events) < nchunks:
            self._events.append(torch.cuda.Event())
        while len(self._copy_events) < nchunks:
            self._copy_events.append(torch.cuda.Event())
        cp.wait_stream(main)            # the previous pass has read `raw`
        with torch.cuda.stream(cp):
            for c in range(nchunks):
                v0, v1 = bounds[c], bounds[c + 1]
                if v1 > v0:
                    raw[v0:v1].copy_(host_raw[v0:v1], non_blocking=True)
                self._copy_events[c].record(cp)
        margin = self.fan_view_margin()
        for c in range(nchunks):
            v0, v1 = bounds[c], bounds[c + 1]
            if v1 <= v0:
                continue
            need = min(V - 1, v1 - 1 + margin)
            k = next(i for i in range(nchunks) if bounds[i + 1] > need)
            main.wait_event(self._copy_events[k])
            rebin_views_dev(self.rebinned, raw, self.setup.geom, v0, v1, main)
            filter_dev(self.filtered[v0:v1], self.rebinned[v0:v1], self.setup, main, self.layout)
            self._events[c].record(main)
        nb = self.plan.nblocks
        gb = [round(nb * g / ngroups) for g in range(ngroups + 1)]
        fp = self.filtered[self.first_view:self.first_view + self.n_views]
        for g in range(ngroups):
            b0, b1 = gb[g], gb[g + 1]
            if b1 <= b0:
                continue
            need = max((self.plan.block_views(b)[1] for b in range(b0, b1)), default=-1)
            last_view = self.first_view + max(need, 0)
            c = next((i for i in range(nchunks) if bounds[i + 1] > last_view), nchunks - 1)
            bp.wait_event(self._events[c])
            self.plan.execute_blocks(self.volume, fp, b0, b1, bp, self.layout)
        main.wait_stream(bp)
        return self.volume

    def run_timed(self, raw: torch.Tensor) -> StageTimes:
        """One pass with HIP events around each stage on the current stream."""
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        ev[0].record()
        self.rebin(raw)
        ev[1].record()
        self.filter()
        ev[2].record()
        self.backproject()
        ev[3].record()
        ev[3].synchronize()
        return StageTimes(ev[0].elapsed_time(ev[1]), ev[1].elapsed_time(ev[2]),
                          ev[2].elapsed_time(ev[3]))
