"""Batched reconstruction: z-slabs of several helical scans in one pass.

The reference reconstructs one volume per MEX call (one segment at a time,
src/analyticalRecon.m:20-25).  Here a batch is a set of scans, each with its
raw sinogram on the device, and a set of z-slabs (any depth, any scan).  One
pass is:

* per scan: rebinning and preweight + filter, once, into the scan's filtered
  sinogram (every slab of the scan reads it);
* one batched backprojection launch over all slabs (``arecon_bp_batch``):
  each workgroup finds its slab in a descriptor table, so small slabs share
  the GPU with large ones instead of leaving it idle at the tail of their own
  launch.

The pass allocates nothing and never synchronises with the host, so
``BatchedRecon.capture()`` records it once as a HIP graph (``torch.cuda.CUDAGraph``
over the library's launches) and ``replay()`` reruns it with one call.
``autotune()`` times the launch shapes (slices per thread, gathers in flight,
workgroup order, slab order) on the actual batch and keeps the fastest; the
result is cached per batch signature in a JSON file.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import json
import math
import os
import time
from dataclasses import dataclass
from pathlib import Path

import numpy as np
import torch

from . import _lib
from .geometry import (BPParams, ReconSetup, Workload, bp_params, bp_scan_geom, rebin_args,
                       voxel_updates, window_half_width)
from .pipeline import LAYOUT, STORAGE, BPPlan, _stream_ptr, filter_dev, filtered_shape, rebin_dev


@dataclass(frozen=True)
class Slab:
    """``nz`` slices of scan ``scan`` starting at height ``z0`` (mm, the height
    of its first slice in the scan's coordinates), spaced ``dz``."""
    scan: int
    z0: float
    nz: int
    dz: float


@dataclass(frozen=True)
class Shape:
    """Launch shape of the batched backprojector."""
    nz: int = 8          # slices per thread
    group: int = 8       # slices whose gathers are in flight together (8: all)
    order: int = 0       # 0: slice-block-major workgroups, 1: column-tile-major,
                         # 2: slice-block-major, tiles in bands of 6 rows
    slab_order: str = "given"   # or "largest_first"
    variant: int = 2     # 2: gathering z-column kernel; 3: LDS-staged (nz 8 or 16)


def full_window_range(p: BPParams) -> tuple[float, float]:
    """Heights (mm) whose whole window [s0 - W, s0 + W] lies inside the scan
    (src/Parallel_FDK_Helical_3DWeighting.cpp:61-70)."""
    W = window_half_width(p.N_2pi, p.HSCoef)
    dphi = 2 * math.pi / p.N_2pi
    lo = p.h * (p.BetaS + (W + 1) * dphi) / (2 * math.pi)
    hi = p.h * (p.BetaS + (p.PN - 2 - W) * dphi) / (2 * math.pi)
    return lo, hi


def stack_slabs(p: BPParams, depths: list[int], dz: float, scan: int, gap: float = 0.0) -> list[Slab]:
    """Consecutive slabs of the given depths from the start of the scan's
    fully sampled range; raises if they do not fit."""
    lo, hi = full_window_range(p)
    out, z = [], lo
    for d in depths:
        out.append(Slab(scan, z, d, dz))
        z += d * dz + gap
    if z - gap > hi:
        raise ValueError(f"slabs of scan {scan} need {z - gap - lo:.1f} mm, "
                         f"the scan samples {hi - lo:.1f} mm fully")
    return out


def slab_params(p: BPParams, slab: Slab) -> BPParams:
    """The scan's backprojection arguments restricted to the slab:
    z(zi) = (zi - ZNC) dz = z0 + zi dz."""
    return dataclasses.replace(p, ZN=slab.nz, ZNC=-slab.z0 / slab.dz)


class Graph:
    """Owner of an ``arecon_graph``: a pass captured on a stream and
    instantiated by libarecon (hipStreamBeginCapture / EndCapture /
    GraphInstantiate), replayed with ``launch``."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle

    @classmethod
    def capture(cls, issue, stream: torch.cuda.Stream) -> "Graph":
        """Capture the calls ``issue()`` makes on ``stream``."""
        sp = _stream_ptr(stream)
        _lib.check(_lib.lib().arecon_graph_capture_begin(sp), "arecon_graph_capture_begin")
        h = C.c_void_p()
        try:
            issue()
        finally:
            # ends the capture even when issue() raised
            status = _lib.lib().arecon_graph_capture_end(sp, C.byref(h))
        _lib.check(status, "arecon_graph_capture_end")
        return cls(h)

    def launch(self, stream=None) -> None:
        _lib.check(_lib.lib().arecon_graph_launch(self._h, _stream_ptr(stream)),
                   "arecon_graph_launch")

    def close(self) -> None:
        if self._h is not None and self._h.value:
            _lib.lib().arecon_graph_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchedRecon:
    """Scans (setup + raw sinogram on the device) and slabs, reconstructed in
    one pass.  ``volumes[i]`` is slab i's [YN][XN][nz] FP32 volume."""

    def __init__(self, setups: list[ReconSetup], raws: list[torch.Tensor], slabs: list[Slab], *,
                 storage: str = "f16", layout: str | None = None, shape: Shape = Shape(),
                 device: str = "cuda"):
        assert len(setups) == len(raws)
        self.setups, self.raws, self.slabs = setups, raws, list(slabs)
        self.storage = storage
        self.layout = layout or "rowdiff"
        self.dev = torch.device(device)
        code, tdtype = STORAGE[storage]
        self.dtype_code = code
        self.params = [bp_params(bp_scan_geom(s)) for s in setups]
        self.rebinned, self.filtered = [], []
        for s, raw in zip(setups, raws):
            a = rebin_args(s.geom)
            self.rebinned.append(torch.empty((a.ViewN, a.ZL, a.YL), dtype=torch.float32,
                                             device=self.dev))
            if storage == "f32" and self.layout == "rows":
                self.filtered.append(self.rebinned[-1])
            else:
                self.filtered.append(torch.empty(filtered_shape(a.ViewN, a.ZL, a.YL, self.layout),
                                                 dtype=tdtype, device=self.dev))
        self.volumes = [torch.empty((self.params[sl.scan].YN, self.params[sl.scan].XN, sl.nz),
                                    dtype=torch.float32, device=self.dev) for sl in self.slabs]
        self.updates = sum(_slab_updates(self.params[sl.scan], sl) for sl in self.slabs)
        self.plans: list[BPPlan] = []
        self._batch = None
        self.shape = None
        self.graph = None
        self.set_shape(shape)

    # -- launch shape ---------------------------------------------------
    def set_shape(self, shape: Shape) -> None:
        self._release_batch()
        for pl in self.plans:
            pl.close()
        self.shape = shape
        self.plans = [BPPlan(slab_params(self.params[sl.scan], sl), dz=sl.dz, nz=shape.nz,
                             order=shape.order, variant=shape.variant, group=shape.group)
                      for sl in self.slabs]
        idx = list(range(len(self.slabs)))
        if shape.slab_order == "largest_first":
            idx.sort(key=lambda i: -self.slabs[i].nz)
        n = len(idx)
        plans = (C.c_void_p * n)(*[self.plans[i]._h.value for i in idx])
        imgs = (C.c_void_p * n)(*[self.volumes[i].data_ptr() for i in idx])
        fps = (C.c_void_p * n)(*[self.filtered[self.slabs[i].scan].data_ptr() for i in idx])
        h = C.c_void_p()
        _lib.check(_lib.lib().arecon_bp_batch_create(
            C.cast(plans, C.POINTER(C.c_void_p)), C.cast(imgs, C.POINTER(C.c_void_p)),
            C.cast(fps, C.POINTER(C.c_void_p)), n, self.dtype_code, LAYOUT[self.layout],
            C.byref(h)), "arecon_bp_batch_create")
        self._batch = h

    def _release_batch(self):
        if self.graph is not None:           # before the batch whose slabs it reads
            self.graph.close()
        self.graph = None
        if self._batch is not None and self._batch.value:
            _lib.lib().arecon_bp_batch_destroy(self._batch)
        self._batch = None

    def close(self):
        self._release_batch()
        for pl in self.plans:
            pl.close()
        self.plans = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- passes ----------------------------------------------------------
    def prepare(self, stream=None) -> None:
        """Rebinning and filter of every scan."""
        for s, raw, reb, filt in zip(self.setups, self.raws, self.rebinned, self.filtered):
            rebin_dev(reb, raw, s.geom, _lib.ARECON_LAYOUT_VZY, stream)
            filter_dev(filt, reb, s, stream, self.layout)

    def backproject(self, stream=None) -> None:
        """All slabs, one launch."""
        _lib.check(_lib.lib().arecon_bp_batch_execute(self._batch, _stream_ptr(stream)),
                   "arecon_bp_batch_execute")

    def backproject_sequential(self, stream=None) -> None:
        """All slabs, one launch each (the comparison point)."""
        for sl, pl, vol in zip(self.slabs, self.plans, self.volumes):
            pl.execute(vol, self.filtered[sl.scan], stream, self.layout)

    def run(self, stream=None) -> None:
        self.prepare(stream)
        self.backproject(stream)

    def capture(self) -> "Graph":
        """Record one pass as a HIP graph through the library's capture entry
        points (arecon_graph_capture_begin / _end).  Runs the pass once eagerly
        first so every cached table (filter spectra, rebinning coefficients,
        footprint tables) exists."""
        self.run()
        torch.cuda.synchronize()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        self.graph = None
        self.graph = Graph.capture(lambda: self.run(s), s)
        torch.cuda.synchronize()
        return self.graph

    def replay(self, stream=None) -> None:
        assert self.graph is not None, "capture() first"
        self.graph.launch(stream)

    # -- tuning ------------------------------------------------------------
    def signature(self) -> str:
        dims = [(p.XN, p.YN, p.YL, p.ZL, p.PN, p.N_2pi) for p in self.params]
        sl = [(s.scan, s.nz, round(s.dz, 6)) for s in self.slabs]
        name = torch.cuda.get_device_name(self.dev) if torch.cuda.is_available() else "cpu"
        return json.dumps({"device": name, "storage": self.storage, "layout": self.layout,
                           "scans": dims, "slabs": sl}, sort_keys=True)

    def time_backproject(self, reps: int = 3) -> float:
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        self.backproject()
        ts = []
        for _ in range(reps):
            ev[0].record()
            self.backproject()
            ev[1].record()
            ev[1].synchronize()
            ts.append(ev[0].elapsed_time(ev[1]))
        return float(np.median(ts))


def _slab_updates(p: BPParams, slab: Slab) -> int:
    return voxel_updates(slab_params(p, slab), slab.dz)


DEFAULT_CANDIDATES = [Shape(8, 8, 0, "given"), Shape(8, 8, 0, "largest_first"),
                      Shape(8, 8, 1, "given"), Shape(8, 8, 2, "given"), Shape(8, 4, 0, "given"),
                      Shape(16, 4, 0, "given"), Shape(8, 2, 0, "given"),
                      Shape(8, 8, 2, "given", 3), Shape(16, 8, 2, "given", 3)]


def autotune(recon: BatchedRecon, candidates=DEFAULT_CANDIDATES, cache: str | os.PathLike | None =
             None, reps: int = 3, log=None) -> tuple[Shape, dict]:
    """Times each launch shape on the batch (filtered sinograms must be
    prepared) and applies the fastest.  With ``cache``, results are stored and
    reused per ``recon.signature()``."""
    key = recon.signature()
    store = {}
    if cache is not None and Path(cache).exists():
        try:
            store = json.loads(Path(cache).read_text())
        except ValueError:
            store = {}
    if key in store:
        best = Shape(**store[key]["best"])
        recon.set_shape(best)
        return best, store[key]["timings"]
    timings = {}
    for sh in candidates:
        recon.set_shape(sh)
        t = recon.time_backproject(reps)
        timings[json.dumps(dataclasses.asdict(sh))] = t
        if log:
            log(f"autotune {sh}: {t:.3f} ms")
    best_key = min(timings, key=timings.get)
    best = Shape(**json.loads(best_key))
    recon.set_shape(best)
    if cache is not None:
        store[key] = {"best": dataclasses.asdict(best), "timings": timings,
                      "when": time.strftime("%Y-%m-%d %H:%M:%S")}
        Path(cache).parent.mkdir(parents=True, exist_ok=True)
        Path(cache).write_text(json.dumps(store, indent=1))
    return best, timings


def three_scan_workload(rotations=(4, 9, 15), depths=((16, 32, 48, 64), (24, 40, 56, 72, 88, 96),
                        (16, 28, 44, 60, 76, 92, 96, 80)), dz: float = 1.25, xn: int = 512,
                        objR: float = 250.0) -> tuple[list[Workload], list[Slab]]:
    """Three 888x64 scans (984 views/rotation, pitch 0.984) of 4, 9 and 15
    rotations and 18 z-slabs of 16..96 slices (512x512, dz 1.25 mm) placed in
    the fully sampled range of their scan."""
    from .geometry import _header
    wls, slabs = [], []
    for i, (rot, ds) in enumerate(zip(rotations, depths)):
        hdr = _header(888, 64, 984, rot, 0.984)
        wl = Workload(f"scan {i}: 888x64x{hdr.TotalViewNumber}", hdr, xn, xn, sum(ds), objR, 10.0)
        wls.append(wl)
        p = bp_params(bp_scan_geom(wl.setup()))
        slabs += stack_slabs(p, list(ds), dz, i, gap=2.0)
    return wls, slabs
