"""Benchmark: GVoxel-updates/s of the end-to-end helical FDK reconstruction
(rebin -> preweight + filter -> 3D-weighted backprojection) at 512 x 512 x 400
from an 888 x 64 x 11,808-view helical scan, on one MI355X.

One step = one pass of the three stages over the device-resident sinogram.
value = voxel-updates of one pass (512*512*400 voxels x 671-view windows =
7.036e10) / measured seconds per pass.  Prints one JSON line.

Also reported: per-stage times (HIP events on the launch stream); the chosen
form (FP16 filtered storage, the volume as uneven z-slabs in one batched
launch, the pass replayed from a HIP graph); the pass including the
host-to-device copy of the raw sinogram; the dominant kernel's roofline, from
hardware counters collected by rocprofv3 in two child passes of this run (VALU
instructions per voxel-update, HBM bytes calibrated on a copy) and from
bandwidth probes (LDS reads, L2 gathers, HBM copy); the HU RMSE against the
FP64 oracle on sampled slices; and the CPU baseline (the reference's own CPU
functions timed on this host on a bounded sample, scaled to the workload).

Other workloads: --workload batch (3 scans, 18 uneven slabs), cnn (the learned
correction), cpu (the reference CPU path on the 368 x 16 scan, in full, on one
thread and on all threads).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "GVoxel-updates/s end-to-end at 512²×400, chosen vs plain; HU RMSE vs FP64"
HBM_PEAK_GBS = 8000.0       # MI355X HBM3E
N_CU, CLOCK_HZ = 256, 2.4e9  # MI355X compute units and peak engine clock


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and HU RMSE")
    ap.add_argument("--cpu-slices", type=int, default=2, help="slices in the CPU sample")
    ap.add_argument("--nz", type=int, default=0, help="slices per thread (0: chosen)")
    ap.add_argument("--order", type=int, default=None,
                    help="workgroup order (default: chosen with the kernel)")
    ap.add_argument("--variant", type=int, default=None,
                    help="backprojector kernel variant (default: chosen from the geometry)")
    ap.add_argument("--group", type=int, default=0, help="z-column kernel slice grouping")
    ap.add_argument("--layout", default=None,
                    help="filtered layout: rows, rowpairs, rowdiff or quads")
    ap.add_argument("--storages", default="f32,f16",
                    help="filtered-sinogram storages to time (plain f32 first)")
    ap.add_argument("--no-counters", action="store_true",
                    help="skip the rocprofv3 counter passes (VALU per update, HBM bytes)")
    ap.add_argument("--no-slabs", action="store_true",
                    help="chosen form without slab batching (FP16 storage only)")
    ap.add_argument("--counter-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--workload", default="clinical", choices=["clinical", "batch", "cnn", "cpu", "probes"],
                    help="clinical: the 888x64x11808 -> 512x512x400 pass (default); batch: "
                         "3 scans of 4/9/15 rotations, 18 uneven slabs, batched + HIP graph; "
                         "cnn: train the learned correction and run it on 512x512x256; "
                         "cpu: the reference CPU path on the 368x16x2160 scan, 1 thread and "
                         "all threads")
    ap.add_argument("--cnn-pairs", type=int, default=200, help="training phantoms (cnn)")
    ap.add_argument("--cnn-clinical-pairs", type=int, default=192,
                    help="training crops of the 888x64 scan (cnn)")
    ap.add_argument("--cnn-steps", type=int, default=12000,
                    help="training steps of each network (cnn)")
    ap.add_argument("--cnn-seeds", type=int, default=3,
                    help="network seeds of the 888x64 correction (cnn)")
    ap.add_argument("--cnn-ema", type=float, default=0.0,
                    help="weight moving-average factor of the training (cnn; 0: last weights)")
    ap.add_argument("--cnn-batch", type=int, default=8, help="training patches per step (cnn)")
    ap.add_argument("--tune-cache", default=None, help="autotuner cache file (batch workload)")
    return ap.parse_args()


def kernel_name(fdk) -> str:
    """The backprojector's kernel as rocprofv3 names it."""
    lay = {"rows": 0, "rowpairs": 1, "quads": 2, "rowdiff": 3}[fdk.layout]
    inside = "true" if fdk.plan.inside else "false"
    t = {"f32": "float", "f16": "__half", "bf16": "__hip_bfloat16"}[fdk.storage]
    if fdk.variant == 2:
        g = 0 if fdk.group == 8 else fdk.group
        return f"bp_fdk3dw_zc<{fdk.plan_nz}, {g}, {t}, {lay}, {g == 0 and inside or 'false'}>"
    if fdk.variant == 3:
        rec4 = fdk.storage != "f32" and fdk.layout == "rowdiff"     # 4-byte records
        rmax = (21 if fdk.plan_nz == 8 else 48) if rec4 else (25 if fdk.plan_nz == 8 else 40)
        return f"bp_fdk3dw_zs<{fdk.plan_nz}, {t}, {lay}, {rmax}, {inside}>"
    return f"bp_fdk3dw{'_vt' if fdk.variant == 0 else ''}<{fdk.plan_nz}, {t}>"


def _kbase(k: str) -> tuple[str, str]:
    """(base name, template arguments) of a kernel name as rocprofv3 prints it."""
    import re
    m = re.search(r"::(\w+)(<[^()]*>)?\(", k) or re.search(r"(\w+)(<[^()]*>)?", k)
    return (m.group(1), m.group(2) or "") if m else (k, "")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_threads() -> int:
    try:
        n = int(os.environ.get("OMP_NUM_THREADS", "0"))
    except ValueError:
        n = 0
    return max(1, min(n or (os.cpu_count() or 1), 16))


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path "
              f"(--gpus {args.gpus} requested)", file=sys.stderr)
        sys.exit(2)

    if args.counter_child:
        counter_child(args)
        return
    if args.workload == "batch":
        bench_batch(args)
        return
    if args.workload == "cnn":
        bench_cnn(args)
        return
    if args.workload == "cpu":
        bench_cpu(args)
        return
    if args.workload == "probes":
        print(json.dumps({"probes": run_probes()}), flush=True)
        return
    bench_clinical(args)


def clinical_workload():
    """The 888x64 helical scan (984 views/rotation, pitch 0.984, 12 rotations)
    of a random-ellipsoid body (seed 2), exact line integrals, on the device."""
    import torch

    from analyticalreconbigpatient_amd import geometry as G
    from analyticalreconbigpatient_amd import phantom

    wl = G.body_clinical()
    s = wl.setup()
    p0 = G.bp_params(G.bp_scan_geom(s))
    z_mid = ((wl.zn - 1) * 0.5 - p0.ZNC) * p0.dz
    E = phantom.random_body(seed=2, fov_radius=250.0, z_centre=z_mid, z_half=250.0)
    raw = phantom.project(E, s, wl.hdr, device="cuda")
    torch.cuda.synchronize()
    return wl, s, raw


# Slabs of the 512 x 512 x 400 volume in the chosen form: uneven z-slabs (whole
# 16-slice blocks) in one batched launch.
CHOSEN_SLABS = (96, 112, 80, 112)


def bench_clinical(args):
    import numpy as np
    import torch

    from analyticalreconbigpatient_amd import batching as B
    from analyticalreconbigpatient_amd.pipeline import HelicalFDK

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    log("clinical workload: projecting the phantom")
    wl, s, raw = clinical_workload()

    def timed_stages(fdk, steps, warmup):
        for _ in range(warmup):
            fdk.run(raw)
        torch.cuda.synchronize()
        ev = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(steps)]
        t0 = time.perf_counter()
        for k in range(steps):
            e = ev[k]
            e[0].record()
            fdk.rebin(raw)
            e[1].record()
            fdk.filter()
            e[2].record()
            fdk.backproject()
            e[3].record()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        st = np.array([[e[0].elapsed_time(e[1]), e[1].elapsed_time(e[2]),
                        e[2].elapsed_time(e[3])] for e in ev])
        return wall * 1e3 / steps, st.mean(axis=0)

    results = {}
    for storage in [x for x in args.storages.split(",") if x]:
        fdk = HelicalFDK(s, storage=storage, nz=args.nz, order=args.order,
                         variant=args.variant, group=args.group, layout=args.layout)
        ms, st = timed_stages(fdk, args.steps, args.warmup)
        results[storage] = dict(fdk=fdk, ms=ms, stages=st)
        log(f"{storage}: {ms:.3f} ms/pass, stages (ms) {np.round(st, 3).tolist()}")
    plain = results.get("f32", next(iter(results.values())))
    fdk = plain["fdk"]
    p = fdk.params
    updates = fdk.updates
    value = updates / (plain["ms"] * 1e-3) / 1e9

    # chosen form: FP16 storage + slab batching, the whole pass replayed from a
    # HIP graph
    chosen = None
    if "f16" in results and not args.no_slabs:
        z, slabs = 0, []
        for d in CHOSEN_SLABS:
            slabs.append(B.Slab(0, (z - p.ZNC) * p.dz, d, p.dz))
            z += d
        assert z == p.ZN
        rb = B.BatchedRecon([s], [raw], slabs, storage="f16",
                            shape=B.Shape(16, 8, 2, "given", 3))
        rb.run()
        torch.cuda.synchronize()
        single = results["f16"]["fdk"].run(raw)
        torch.cuda.synchronize()
        joined = torch.cat(rb.volumes, dim=2)
        diff = float((joined - single).abs().max())
        rb.capture()
        for _ in range(args.warmup):
            rb.replay()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            rb.replay()
        torch.cuda.synchronize()
        gms = (time.perf_counter() - t0) * 1e3 / args.steps
        assert rb.updates == updates, (rb.updates, updates)
        chosen = dict(ms=gms, recon=rb, max_abs_diff_vs_single=diff,
                      bp_ms=rb.time_backproject(3))
        log(f"chosen (FP16, {len(slabs)} slabs {list(CHOSEN_SLABS)}, graph): {gms:.3f} ms/pass; "
            f"max |slabs - single plan| {diff:.3g}")
    if chosen is None:
        f16 = results.get("f16", plain)
        chosen = dict(ms=f16["ms"], recon=None, max_abs_diff_vs_single=0.0,
                      bp_ms=float(f16["stages"][2]))
    chosen_value = updates / (chosen["ms"] * 1e-3) / 1e9

    probes = run_probes()
    log(f"probes: {probes}")
    counters = None if args.no_counters else measure_counters(args, results)
    roofline = make_roofline(fdk, plain, results, probes, counters, wl)
    h2d = h2d_inclusive(raw, results, chosen, args)

    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "GVoxel-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(plain["ms"], 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: random-ellipsoid body phantom (seed 2), exact line integrals",
        "config": {"workload": wl.name, "detector": f"{p.YL}x{p.ZL}", "views": fdk.ViewN,
                   "volume": f"{p.XN}x{p.YN}x{p.ZN}", "voxel_updates": updates,
                   "storage": "f32", "path": "plain"},
        "stages_ms": {"rebin": round(float(plain["stages"][0]), 3),
                      "filter": round(float(plain["stages"][1]), 3),
                      "backproject": round(float(plain["stages"][2]), 3)},
        # algorithmic bytes: the rebinning reads and writes 4 B per sample; the
        # filter reads 4 B per sample and writes its output layout (FP32 row
        # differences: 8 B per record)
        "stage_hbm_GBs": {
            "rebin": round(8.0 * p.YL * p.ZL * fdk.ViewN / (plain["stages"][0] * 1e-3) / 1e9, 1),
            "filter": round((4.0 * p.YL * p.ZL + fdk.filtered[0].numel() * 4.0) * fdk.ViewN /
                            (plain["stages"][1] * 1e-3) / 1e9, 1)},
        "chosen": {"form": (f"FP16 filtered storage + {len(CHOSEN_SLABS)} z-slabs "
                            f"{list(CHOSEN_SLABS)} in one batched launch, pass replayed from a "
                            f"HIP graph") if chosen["recon"] is not None else "FP16 storage",
                   "value": round(chosen_value, 2), "ms_per_step": round(chosen["ms"], 3),
                   "backproject_ms": round(chosen["bp_ms"], 3),
                   "speedup_vs_plain": round(chosen_value / value, 3),
                   "max_abs_diff_vs_single_plan": chosen["max_abs_diff_vs_single"]},
        "per_storage": {k: {"ms_per_step": round(v["ms"], 3),
                            "value": round(updates / (v["ms"] * 1e-3) / 1e9, 2),
                            "stages_ms": [round(float(x), 3) for x in v["stages"]]}
                        for k, v in results.items()},
        "h2d_inclusive": h2d,
        "roofline": roofline,
        "probes": probes,
    }
    if not args.no_cpu:
        out["cpu_baseline"], out["hu_rmse"] = cpu_leg(s, fdk, results, chosen, raw, args)
    print(json.dumps(out), flush=True)


def run_probes():
    """Measured roofs of the backprojector's sample reads: LDS read bandwidth
    (8- and 16-byte lane reads, conflict-free) and L2 gather bandwidth (16-byte
    scattered reads of a 2 MiB buffer), plus an HBM streaming copy."""
    import ctypes as C

    import torch

    from analyticalreconbigpatient_amd import _lib
    lib = _lib.lib()
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    blocks = 256 * 8
    sink = torch.empty(blocks * 256, device="cuda")
    out = {}

    def timed(fn, reps=5):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        fn()
        ts = []
        for _ in range(reps):
            ev[0].record()
            nbytes = fn()
            ev[1].record()
            ev[1].synchronize()
            ts.append(ev[0].elapsed_time(ev[1]))
        return nbytes / (min(ts) * 1e-3) / 1e12

    for b in (8, 16):
        def lds(b=b):
            nb = C.c_double()
            _lib.check(lib.arecon_probe_lds(C.c_void_p(sink.data_ptr()), b, 4096, blocks,
                                            C.byref(nb), st), "arecon_probe_lds")
            return nb.value
        out[f"lds_read_b{8 * b}_TBs"] = round(timed(lds), 2)
    src = torch.randn(2 << 18, device="cuda")            # 2 MiB
    def l2():
        nb = C.c_double()
        _lib.check(lib.arecon_probe_l2(C.c_void_p(sink.data_ptr()), C.c_void_p(src.data_ptr()),
                                       src.numel() * 4, 1024, blocks, C.byref(nb), st),
                   "arecon_probe_l2")
        return nb.value
    out["l2_gather_16B_TBs"] = round(timed(l2), 2)
    a = torch.ones(1 << 28, device="cuda")                # 1 GiB
    b2 = torch.empty_like(a)
    def copy():
        _lib.check(lib.arecon_probe_copy(C.c_void_p(b2.data_ptr()), C.c_void_p(a.data_ptr()),
                                         a.numel() * 4, st), "arecon_probe_copy")
        return 2.0 * a.numel() * 4
    out["hbm_copy_TBs"] = round(timed(copy), 2)
    del a, b2, src
    # VALU issue rates (wavefront instructions per CU per ns) of the instruction
    # forms the backprojector is built from; at the ~2.26 GHz these kernels run
    # at, 2.26 per ns is one instruction per CU per clock
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    forms = ["fma_f32", "fma_f32_sgpr_operand", "fma_mix_f32", "fma_mixlo_f16", "dot2_f32_f16",
             "pk_fma_f16", "cvt_flr_i32_f32", "lshl_add_u32", "rcp_f32"]
    vb = cus * 8
    vsink = torch.empty(vb * 256, device="cuda")
    rates = {}
    for f, name in enumerate(forms):
        def valu(f=f):
            ni = C.c_double()
            _lib.check(lib.arecon_probe_valu(C.c_void_p(vsink.data_ptr()), f, vb, C.byref(ni), st),
                       "arecon_probe_valu")
            return ni.value
        rates[name] = round(timed(valu, 3) * 1e12 / cus / 1e9, 3)
    out["valu_wave_instr_per_cu_per_ns"] = rates
    del sink, vsink
    return out


def counter_child(args):
    """Child of measure_counters (runs under rocprofv3 --pmc): a 1 GiB
    calibration copy, then one pass of each storage's pipeline (the footprint
    table is built by a first pass, which the counters skip by kernel name)."""
    import ctypes as C

    import torch

    from analyticalreconbigpatient_amd import _lib
    from analyticalreconbigpatient_amd.pipeline import HelicalFDK
    a = torch.ones(1 << 28, device="cuda")
    b = torch.empty_like(a)
    _lib.check(_lib.lib().arecon_probe_copy(C.c_void_p(b.data_ptr()), C.c_void_p(a.data_ptr()),
                                            a.numel() * 4,
                                            C.c_void_p(torch.cuda.current_stream().cuda_stream)),
               "arecon_probe_copy")
    torch.cuda.synchronize()
    del a, b
    wl, s, raw = clinical_workload()
    for storage in [x for x in args.storages.split(",") if x]:
        fdk = HelicalFDK(s, storage=storage, nz=args.nz, order=args.order,
                         variant=args.variant, group=args.group, layout=args.layout)
        fdk.run(raw)
        torch.cuda.synchronize()
        print(f"counter child: {storage} kernel {kernel_name(fdk)}", flush=True)


def _read_counters(outdir: Path) -> list[dict]:
    import csv
    rows = []
    for f in outdir.rglob("*counter_collection.csv"):
        with open(f, newline="") as fh:
            rows += list(csv.DictReader(fh))
    return rows


def measure_counters(args, results):
    """Hardware counters of this run's kernels, collected by rocprofv3 in two
    child passes of the same workload (SQ instruction counts with HBM reads, then
    HBM writes): VALU instructions per voxel-update of each backprojector and
    the calibrated HBM bytes of each stage.  Returns None (and says why) when
    rocprofv3 is unavailable or fails."""
    import shutil
    import subprocess
    import tempfile

    exe = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not Path(exe).exists():
        log("counters: rocprofv3 not found")
        return None
    passes = {"sq": ["SQ_INSTS_VALU", "SQ_INSTS_VALU_TRANS_F32", "SQ_ACTIVE_INST_VALU2",
                     "SQ_WAVES", "GRBM_GUI_ACTIVE", "FETCH_SIZE"],
              "wr": ["WRITE_SIZE"]}
    per_kernel: dict[str, dict] = {}
    for name, ctrs in passes.items():
        d = Path(tempfile.mkdtemp(prefix=f"arecon_pmc_{name}_"))
        cmd = [exe, "--pmc", *ctrs, "--output-format", "csv", "-d", str(d), "-o", "run", "--",
               sys.executable, str(ROOT / "bench.py"), "--counter-child", "--storages",
               args.storages]
        for opt, val in (("--nz", args.nz), ("--order", args.order), ("--variant", args.variant),
                         ("--group", args.group), ("--layout", args.layout)):
            if val is not None:
                cmd += [opt, str(val)]
        log(f"counters ({name}): {' '.join(cmd[:4 + len(ctrs)])} ...")
        try:
            r = subprocess.run(cmd, timeout=240, stdout=subprocess.PIPE,
                               stderr=subprocess.STDOUT, cwd=str(ROOT))
        except subprocess.TimeoutExpired:
            log(f"counters ({name}): rocprofv3 timed out")
            return None
        rows = _read_counters(d)
        if not rows:
            log(f"counters ({name}): rocprofv3 exit {r.returncode}, no counter file; "
                f"{r.stdout.decode(errors='replace')[-400:]}")
            return None
        # the last dispatch of each kernel name (the footprint-table pass and the
        # pipeline run once per storage)
        last: dict[str, dict] = {}
        for row in rows:
            k = row.get("Kernel_Name", "")
            did = int(row.get("Dispatch_Id", 0))
            e = last.setdefault(k, {"id": -1, "c": {}})
            if did > e["id"]:
                e["id"], e["c"] = did, {}
            if did == e["id"]:
                e["c"][row["Counter_Name"]] = e["c"].get(row["Counter_Name"], 0.0) + \
                    float(row["Counter_Value"])
        for k, e in last.items():
            per_kernel.setdefault(k, {}).update(e["c"])
        shutil.rmtree(d, ignore_errors=True)
    # HBM byte counters calibrated on the 1 GiB copy (reads 1 GiB, writes 1 GiB)
    cal = next((v for k, v in per_kernel.items() if "probe_copy" in k), None)
    if cal is None or not cal.get("FETCH_SIZE") or not cal.get("WRITE_SIZE"):
        log("counters: no calibration copy in the counter files")
        return None
    fetch_f = float(1 << 30) / (cal["FETCH_SIZE"] * 1024.0)
    write_f = float(1 << 30) / (cal["WRITE_SIZE"] * 1024.0)
    out = {"fetch_factor": round(fetch_f, 4), "write_factor": round(write_f, 4), "kernels": {}}
    for k, c in per_kernel.items():
        if "probe_copy" in k:
            continue
        e = {}
        if "FETCH_SIZE" in c:
            e["fetch_bytes"] = c["FETCH_SIZE"] * 1024.0 * fetch_f
        if "WRITE_SIZE" in c:
            e["write_bytes"] = c["WRITE_SIZE"] * 1024.0 * write_f
        for n in ("SQ_INSTS_VALU", "SQ_INSTS_VALU_TRANS_F32", "SQ_ACTIVE_INST_VALU2", "SQ_WAVES",
                  "GRBM_GUI_ACTIVE"):
            if n in c:
                e[n] = c[n]
        out["kernels"][k] = e
    log(f"counters: {len(out['kernels'])} kernels, fetch x{fetch_f:.3f}, write x{write_f:.3f}")
    return out


def _match_kernel(counters, name):
    if not counters:
        return None
    for k, v in counters["kernels"].items():
        if k.split("(")[0].replace(" ", "").endswith(name.replace(" ", "")) or \
                name.replace(" ", "") in k.replace(" ", ""):
            return v
    return None


def _issue_slots(k):
    """VALU issue slots of a dispatch: a SIMD issues one wavefront VALU
    instruction per quad-cycle, and gfx950 can issue a second in the same
    quad-cycle (an FP32 FMA/add/mul whose operands are all VGPRs), so a slot
    holds one or two instructions.  Slots = SQ_INSTS_VALU - SQ_ACTIVE_INST_VALU2
    (quad-cycles that issued two).  Transcendentals run beside the VALU and are
    left out, which can only lower the count and so raise the roof."""
    return k["SQ_INSTS_VALU"] - k.get("SQ_INSTS_VALU_TRANS_F32", 0.0) - \
        k.get("SQ_ACTIVE_INST_VALU2", 0.0)


def make_roofline(fdk, plain, results, probes, counters, wl):
    """Roofline of the dominant kernel, the plain path's backprojector.  The
    LDS-staged kernel (variant 3) reads its samples from LDS and is bound by VALU
    issue: one slot per SIMD per quad-cycle, i.e. one slot per CU per clock,
    each slot issuing one wavefront instruction or a dual-issued pair.  At S
    slots per voxel-update the roof is N_CU * clock * 64 / S voxel-updates/s,
    with S measured in this run (rocprofv3 SQ_INSTS_VALU - SQ_INSTS_VALU_TRANS_F32
    - SQ_ACTIVE_INST_VALU2 over the launch's wavefront-updates), so the fraction
    cannot pass 1 at the peak clock.  Beside it: the VALU instructions per
    update, the compulsory HBM bytes (filtered sinogram once, volume once)
    against the 8 TB/s peak, and the algorithmic sample bytes (16 B per
    voxel-update) against the LDS read bandwidth the probe measured."""
    p = fdk.params
    updates = fdk.updates
    nvox = p.XN * p.YN * p.ZN
    bp_ms = float(plain["stages"][2])
    gvu = updates / (bp_ms * 1e-3) / 1e9
    kname = kernel_name(fdk)
    k = _match_kernel(counters, kname)
    elem = {"f32": 4, "f16": 2, "bf16": 2}[fdk.storage]
    per_view = fdk.filtered[0].numel() * elem
    compulsory = per_view * fdk.n_views + 4.0 * nvox
    traffic = None
    if k and "fetch_bytes" in k and "write_bytes" in k:
        traffic = round(k["fetch_bytes"] + k["write_bytes"])
    roof = {"kernel": kname, "kernel_ms": round(bp_ms, 3)}
    spu = vpu = None
    if k and "SQ_INSTS_VALU" in k and "SQ_ACTIVE_INST_VALU2" in k:
        vpu = (k["SQ_INSTS_VALU"] - k.get("SQ_INSTS_VALU_TRANS_F32", 0.0)) / (updates / 64.0)
        spu = _issue_slots(k) / (updates / 64.0)
    if fdk.variant == 3 and spu:
        peak = N_CU * CLOCK_HZ * 64 / spu / 1e9
        roof.update({"bound": "valu_issue", "achieved": round(gvu, 1), "peak": round(peak, 1),
                     "unit": "GVoxel-updates/s", "frac": round(gvu / peak, 4),
                     "traffic": traffic, "issue_slots_per_update": round(spu, 3),
                     "valu_per_update": round(vpu, 3),
                     "valu_source": "rocprofv3 --pmc SQ_INSTS_VALU - SQ_INSTS_VALU_TRANS_F32 - "
                                    "SQ_ACTIVE_INST_VALU2, this run",
                     "peak_basis": f"{N_CU} CUs x {CLOCK_HZ / 1e9} GHz x 1 issue slot x 64 "
                                   f"lanes / {spu:.3f} slots per voxel-update"})
        if k.get("GRBM_GUI_ACTIVE"):
            clk = k["GRBM_GUI_ACTIVE"] / 8.0 / (bp_ms * 1e-3)
            busy = _issue_slots(k) / N_CU / (k["GRBM_GUI_ACTIVE"] / 8.0)
            roof["measured_clock_GHz"] = round(clk / 1e9, 3)
            roof["issue_busy_at_measured_clock"] = round(busy, 4)
    else:
        # no counters: the HBM roof of the compulsory bytes is what can be stated
        roof.update({"bound": "hbm", "achieved": round(compulsory / (bp_ms * 1e-3) / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(compulsory / (bp_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "note": "VALU count not measured in this run (no rocprofv3 counters)"})
    roof["hbm_compulsory"] = {"bytes": round(compulsory),
                              "achieved_GBs": round(compulsory / (bp_ms * 1e-3) / 1e9, 1),
                              "peak_GBs": HBM_PEAK_GBS,
                              "frac": round(compulsory / (bp_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    if traffic:
        roof["hbm_measured"] = {"bytes": traffic,
                                "achieved_GBs": round(traffic / (bp_ms * 1e-3) / 1e9, 1),
                                "frac": round(traffic / (bp_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                                "x_compulsory": round(traffic / compulsory, 2)}
    lds_peak = probes.get("lds_read_b128_TBs") if elem == 4 else probes.get("lds_read_b64_TBs")
    samples = 4.0 * elem * updates
    if lds_peak:
        ach = samples / (bp_ms * 1e-3) / 1e12
        roof["lds_gather"] = {"bytes_per_update": 4 * elem, "achieved_TBs": round(ach, 2),
                              "peak_TBs": lds_peak, "frac": round(ach / lds_peak, 4),
                              "peak_source": "arecon_probe_lds, this run"}
    # every stage's HBM bytes from the counters
    stages = {}
    for kk, v in (counters or {}).get("kernels", {}).items():
        base, targs = _kbase(kk)
        if base in ("rebin_vzy", "filter_quads_wave", "filter_rows_wave", "bp_fdk3dw_zs",
                    "bp_zs_table") and "fetch_bytes" in v:
            stages[base + targs] = {"fetch_GB": round(v["fetch_bytes"] / 1e9, 3),
                                    "write_GB": round(v.get("write_bytes", 0.0) / 1e9, 3)}
    if stages:
        roof["stage_hbm_bytes"] = stages
    # the chosen (FP16) path's backprojector, same measurements
    f16 = results.get("f16")
    if f16 is not None and f16["fdk"] is not fdk:
        cf = f16["fdk"]
        cms = float(f16["stages"][2])
        cn = kernel_name(cf)
        ck = _match_kernel(counters, cn)
        c = {"kernel": cn, "kernel_ms": round(cms, 3),
             "achieved": round(cf.updates / (cms * 1e-3) / 1e9, 1)}
        if ck and "SQ_INSTS_VALU" in ck and "SQ_ACTIVE_INST_VALU2" in ck:
            cv = (ck["SQ_INSTS_VALU"] - ck.get("SQ_INSTS_VALU_TRANS_F32", 0.0)) / (cf.updates / 64)
            cs = _issue_slots(ck) / (cf.updates / 64)
            cpk = N_CU * CLOCK_HZ * 64 / cs / 1e9
            c.update({"bound": "valu_issue", "issue_slots_per_update": round(cs, 3),
                      "valu_per_update": round(cv, 3), "peak": round(cpk, 1),
                      "frac": round(c["achieved"] / cpk, 4)})
        if ck and "fetch_bytes" in ck and "write_bytes" in ck:
            c["traffic"] = round(ck["fetch_bytes"] + ck["write_bytes"])
        roof["chosen_kernel"] = c
    return roof


# View chunks of the streamed host-to-device pass.
STREAM_CHUNKS = 16


def h2d_inclusive(raw, results, chosen, args):
    """The pass including the host-to-device copy of the raw sinogram from
    pinned host memory (the reference copies the sinogram to the device in
    every stage)."""
    import torch

    host = raw.cpu().pin_memory()
    dev = torch.empty_like(raw)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    dev.copy_(host, non_blocking=True)
    ev[1].record()
    ev[1].synchronize()
    copy_ms = ev[0].elapsed_time(ev[1])
    out = {"h2d_ms": round(copy_ms, 3), "h2d_GBs": round(host.numel() * 4 / copy_ms / 1e6, 1)}
    steps = max(1, min(args.steps, 5))
    for name, res in results.items():
        fdk = res["fdk"]
        fdk.run(dev)
        torch.cuda.synchronize()
        ev[0].record()
        for _ in range(steps):
            dev.copy_(host, non_blocking=True)
            fdk.run(dev)
        ev[1].record()
        ev[1].synchronize()
        ms = ev[0].elapsed_time(ev[1]) / steps
        out[name] = {"ms_per_step": round(ms, 3),
                     "value": round(fdk.updates / (ms * 1e-3) / 1e9, 2)}
        # the copy streamed in view chunks, overlapped with rebin + filter of
        # the chunks that have arrived and the backprojection of ready blocks
        fdk.run_streamed(host, dev, nchunks=STREAM_CHUNKS, ngroups=STREAM_CHUNKS)
        torch.cuda.synchronize()
        ev[0].record()
        for _ in range(steps):
            fdk.run_streamed(host, dev, nchunks=STREAM_CHUNKS, ngroups=STREAM_CHUNKS)
        ev[1].record()
        ev[1].synchronize()
        ms = ev[0].elapsed_time(ev[1]) / steps
        out[name + "_streamed"] = {"ms_per_step": round(ms, 3), "chunks": STREAM_CHUNKS,
                                   "value": round(fdk.updates / (ms * 1e-3) / 1e9, 2)}
    rb = chosen.get("recon")
    if rb is not None:
        target = rb.raws[0]
        ev[0].record()
        for _ in range(steps):
            target.copy_(host, non_blocking=True)
            rb.replay()
        ev[1].record()
        ev[1].synchronize()
        ms = ev[0].elapsed_time(ev[1]) / steps
        out["chosen"] = {"ms_per_step": round(ms, 3),
                         "value": round(rb.updates / (ms * 1e-3) / 1e9, 2)}
    del host, dev
    return out


def _ref_rebin_threads(oracle, raw_np, a, nth):
    """The reference's float rebinning, run on nth threads over runs of views
    (each thread's run starts its view angle at 0: for timing only)."""
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np
    V = raw_np.shape[0]
    cuts = np.linspace(0, V, nth + 1).round().astype(int)

    def part(i):
        return oracle.ref_rebin_f32(raw_np[cuts[i]:cuts[i + 1]], a.YLC0, a.dYA, a.DeltaTheta,
                                    a.PLC0, a.DeltaT, a.DeltaFai, a.SO)
    with ThreadPoolExecutor(nth) as ex:
        list(ex.map(part, range(nth)))


def cpu_leg(s, fdk, results, chosen, raw, args):
    """CPU baseline and HU RMSE on a bounded sample: the views that 2 slices
    need and those 2 slices of 512 x 512.

    * Timed (``cpu_baseline``): the reference's own CPU functions (oracle/_ref,
      compiled from the reference sources) for the rebinning and the
      backprojection, on this host's threads (views split for the rebinning,
      runs of y rows for the backprojection), and the oracle's port of the
      MATLAB CPU filter (the reference has no C/C++ CPU filter), each scaled per
      view / per voxel-update to the whole workload.  Without oracle/_ref the
      oracle's port of all three stages is timed instead (kind "port").
    * HU RMSE: the GPU volumes' sampled slices against the FP64 oracle chain."""
    import numpy as np
    import torch

    import oracle
    from analyticalreconbigpatient_amd import geometry as G

    nth = host_threads()
    p = fdk.params
    W = G.window_half_width(p.N_2pi, p.HSCoef)
    zs = np.linspace(p.ZN * 0.45, p.ZN * 0.55, args.cpu_slices).astype(int)
    s0 = [math.ceil((2 * 3.14159265358979 * (z - p.ZNC) * p.dz / p.h - p.BetaS) / p.DeltaFai - 0.5)
          for z in zs]
    margin = int(math.ceil(math.asin(min(1.0, 1.5 * math.tan(s.DecAngle / 2))) / p.DeltaFai)) + 4
    v_lo = max(0, min(s0) - W - margin)
    v_hi = min(fdk.ViewN, max(s0) + W + margin + 1)
    a = G.rebin_args(s.geom)
    raw_np = raw[v_lo:v_hi].cpu().numpy()
    use_ref = oracle.ref_available()
    log(f"cpu leg: views [{v_lo}, {v_hi}), slices {zs.tolist()}, {nth} threads, "
        f"{'reference library' if use_ref else 'oracle port'}")

    # the FP64 oracle chain (HU reference): port, OpenMP
    reb = oracle.rebin_f32(raw_np, a.YLC0, a.dYA, a.DeltaTheta, a.PLC0, a.DeltaT, a.DeltaFai,
                           a.SO)
    t = time.perf_counter()
    if use_ref:
        _ref_rebin_threads(oracle, raw_np, a, nth)
    else:
        oracle.rebin_f32(raw_np, a.YLC0, a.dYA, a.DeltaTheta, a.PLC0, a.DeltaT, a.DeltaFai, a.SO)
    t_reb = time.perf_counter() - t
    fa = G.filter_args(s)
    t = time.perf_counter()
    filt = oracle.filter_f64(np.ascontiguousarray(reb.transpose(0, 2, 1)), float(fa.PLC0),
                             float(fa.ZLC0), float(fa.dYL), float(fa.dZL), float(fa.SO),
                             nthreads=nth)
    t_filt = time.perf_counter() - t
    sub = G.BPParams(**{**p.__dict__, "PN": v_hi - v_lo,
                        "BetaS": p.BetaS + v_lo * p.DeltaFai})
    w = oracle.fdk_weights(p.N_2pi, p.HSCoef, p.delta)
    ref = np.zeros((p.YN, p.XN, p.ZN))
    for z in zs:
        oracle.fdk_core(filt, w, N_2pi=sub.N_2pi, XN=sub.XN, YN=sub.YN, ZN=sub.ZN, XNC=sub.XNC,
                        YNC=sub.YNC, ZNC=sub.ZNC, dx=sub.dx, dy=sub.dy, dz=sub.dz, YLC=sub.YLC,
                        ZLC=sub.ZLC, dYL=sub.dYL, dZL=sub.dZL, h=sub.h, BetaS=sub.BetaS,
                        DeltaFai=sub.DeltaFai, N_pi=sub.N_pi, HSCoef=sub.HSCoef, SO=sub.SO,
                        SD=sub.SD, k1=sub.k1, zi_range=(int(z), int(z) + 1), nthreads=nth,
                        out=ref)
    # timed backprojection of the sampled slices
    ref_check = None
    t = time.perf_counter()
    for z in zs:
        if use_ref:
            one = oracle.ref_fdk_core(filt, w, N_2pi=sub.N_2pi, XN=sub.XN, YN=sub.YN, ZN=1,
                                      XNC=sub.XNC, YNC=sub.YNC, ZNC=sub.ZNC - int(z), dx=sub.dx,
                                      dy=sub.dy, dz=sub.dz, YLC=sub.YLC, ZLC=sub.ZLC,
                                      dYL=sub.dYL, dZL=sub.dZL, h=sub.h, BetaS=sub.BetaS,
                                      DeltaFai=sub.DeltaFai, N_pi=sub.N_pi, HSCoef=sub.HSCoef,
                                      SO=sub.SO, SD=sub.SD, k1=sub.k1, nthreads=nth)
            d = float(np.abs(one[:, :, 0] - ref[:, :, z]).max() /
                      max(np.abs(ref[:, :, z]).max(), 1e-300))
            ref_check = d if ref_check is None else max(ref_check, d)
        else:
            oracle.fdk_core(filt, w, N_2pi=sub.N_2pi, XN=sub.XN, YN=sub.YN, ZN=sub.ZN,
                            XNC=sub.XNC, YNC=sub.YNC, ZNC=sub.ZNC, dx=sub.dx, dy=sub.dy,
                            dz=sub.dz, YLC=sub.YLC, ZLC=sub.ZLC, dYL=sub.dYL, dZL=sub.dZL,
                            h=sub.h, BetaS=sub.BetaS, DeltaFai=sub.DeltaFai, N_pi=sub.N_pi,
                            HSCoef=sub.HSCoef, SO=sub.SO, SD=sub.SD, k1=sub.k1,
                            zi_range=(int(z), int(z) + 1), nthreads=nth, out=np.zeros_like(ref))
    t_bp = time.perf_counter() - t
    nv = v_hi - v_lo
    upd_sample = len(zs) * p.XN * p.YN * (2 * W + 1)
    t_total = (t_reb / nv * fdk.ViewN + t_filt / nv * fdk.ViewN
               + t_bp / upd_sample * fdk.updates)
    kind = "reference" if use_ref else "port"
    stages_desc = ("reference CPU rebinning and backprojection (oracle/_ref, compiled from the "
                   "reference sources) with the port of the MATLAB CPU filter" if use_ref
                   else "oracle port of all three stages")
    cpu = {"value": round(fdk.updates / t_total / 1e9, 5), "unit": "GVoxel-updates/s",
           "cores": nth, "kind": kind,
           "sample": (f"{stages_desc} on views [{v_lo},{v_hi}) and {len(zs)} slices of "
                      f"512x512, {nth} threads; measured {t_reb:.2f}+{t_filt:.2f}+{t_bp:.2f} s, "
                      f"scaled per view / per voxel-update to the whole workload "
                      f"({t_total:.0f} s); single-thread figure: python bench.py --workload cpu"),
           "stage_s_scaled": {"rebin": round(t_reb / nv * fdk.ViewN, 2),
                              "filter": round(t_filt / nv * fdk.ViewN, 2),
                              "backproject": round(t_bp / upd_sample * fdk.updates, 2)}}
    if ref_check is not None:
        cpu["reference_vs_oracle_max_rel"] = ref_check
    yy, xx = np.meshgrid((np.arange(p.YN) - p.YNC) * p.dy, (np.arange(p.XN) - p.XNC) * p.dx,
                         indexing="ij")
    mask = (xx ** 2 + yy ** 2) <= (0.98 * 250.0) ** 2
    hu = {}
    vols = {name: res["fdk"].run(raw) for name, res in results.items()}
    if chosen.get("recon") is not None:
        chosen["recon"].replay()
        vols["chosen"] = torch.cat(chosen["recon"].volumes, dim=2)
    torch.cuda.synchronize()
    for name, vol in vols.items():
        g = vol[:, :, torch.as_tensor(zs, device=vol.device)].cpu().numpy().astype(np.float64)
        r = ref[:, :, zs]
        diff = (g - r)[mask]
        hu[name] = round(float(1000.0 / 0.02 * np.sqrt(np.mean(diff ** 2))), 4)
        hu[name + "_max_abs"] = round(float(1000.0 / 0.02 * np.abs(diff).max()), 4)
        hu[name + "_rel_l2"] = float(np.linalg.norm(g - r) / np.linalg.norm(r))
    log(f"cpu leg done: {cpu['value']} GVU/s ({kind}), HU RMSE {hu}")
    return cpu, hu


def bench_cpu(args):
    """The reference's CPU path on the 368 x 16 Shepp-Logan scan (720
    views/rotation, pitch 1.0, 3 rotations) into 256 x 256 x 32, in full:
    single-threaded (as the reference ships) and on this host's threads.
    Rebinning and backprojection are the reference's own CPU functions
    (oracle/_ref); the filter is the oracle's port of the MATLAB CPU branch."""
    import numpy as np

    import oracle
    from analyticalreconbigpatient_amd import geometry as G
    from analyticalreconbigpatient_amd import phantom

    wl = G.shepp_logan_small()
    s = wl.setup()
    dz = 1.0            # 1 mm slices centred in the helix (the config's 32 slices)
    s.geom.Obj[6] = G.centred_ZC(wl.hdr, wl.zn, dz) + 0.137
    p = G.bp_params(G.bp_scan_geom(s))
    z_mid = ((wl.zn - 1) * 0.5 - p.ZNC) * dz
    E = phantom.shepp_logan(100.0, 30.0, z_mid)
    raw = phantom.project(E, s, wl.hdr).cpu().numpy()
    a = G.rebin_args(s.geom)
    fa = G.filter_args(s)
    w = oracle.fdk_weights(p.N_2pi, p.HSCoef, p.delta)
    use_ref = oracle.ref_available()
    nth = host_threads()
    updates = G.voxel_updates(p, dz)
    out = {"metric": "GVoxel-updates/s of the CPU path (368x16x2160 -> 256x256x32, 1 mm)",
           "unit": "GVoxel-updates/s", "workload": wl.name, "voxel_updates": updates,
           "kind": "reference" if use_ref else "port", "runs": {}}
    for threads in sorted({1, nth}):
        t = time.perf_counter()
        if use_ref:
            if threads == 1:
                reb = oracle.ref_rebin_f32(raw, a.YLC0, a.dYA, a.DeltaTheta, a.PLC0, a.DeltaT,
                                           a.DeltaFai, a.SO)
            else:
                _ref_rebin_threads(oracle, raw, a, threads)
                reb = None
        else:
            reb = oracle.rebin_f32(raw, a.YLC0, a.dYA, a.DeltaTheta, a.PLC0, a.DeltaT,
                                   a.DeltaFai, a.SO)
        t_reb = time.perf_counter() - t
        if reb is None:
            reb = oracle.ref_rebin_f32(raw, a.YLC0, a.dYA, a.DeltaTheta, a.PLC0, a.DeltaT,
                                       a.DeltaFai, a.SO)
        t = time.perf_counter()
        filt = oracle.filter_f64(np.ascontiguousarray(reb.transpose(0, 2, 1)), float(fa.PLC0),
                                 float(fa.ZLC0), float(fa.dYL), float(fa.dZL), float(fa.SO),
                                 nthreads=threads)
        t_filt = time.perf_counter() - t
        t = time.perf_counter()
        if use_ref:
            vol = oracle.ref_fdk_from_params(filt, p, dz=dz, w=w, nthreads=threads)
        else:
            vol = oracle.fdk_from_params(filt, p, dz=dz, w=w, nthreads=threads)
        t_bp = time.perf_counter() - t
        total = t_reb + t_filt + t_bp
        out["runs"][f"{threads}_threads"] = {
            "threads": threads, "s": round(total, 2),
            "stages_s": {"rebin": round(t_reb, 3), "filter": round(t_filt, 3),
                         "backproject": round(t_bp, 3)},
            "value": round(updates / total / 1e9, 5), "volume_max": float(np.abs(vol).max())}
        log(f"cpu {threads} threads: {out['runs'][f'{threads}_threads']}")
    out["host"] = {"os_cpu_count": os.cpu_count(), "OMP_NUM_THREADS":
                   os.environ.get("OMP_NUM_THREADS")}
    print(json.dumps(out), flush=True)


def bench_batch(args):
    """3 scans (888x64, 984 views/rotation, 4/9/15 rotations, random-ellipsoid
    bodies, seeds 41-43) and 18 z-slabs of 16..96 slices (512x512, 1.25 mm):
    per-scan rebin + filter, then the slabs' backprojection.  Times the
    sequential form (one launch per slab) and the batched form (one launch for
    all slabs, the whole pass replayed from a HIP graph), after the autotuner
    has picked the batched launch shape."""
    import numpy as np
    import torch

    from analyticalreconbigpatient_amd import batching as B
    from analyticalreconbigpatient_amd import geometry as G
    from analyticalreconbigpatient_amd import phantom

    wls, slabs = B.three_scan_workload()
    setups = [wl.setup() for wl in wls]
    raws = []
    for i, (wl, s) in enumerate(zip(wls, setups)):
        p = G.bp_params(G.bp_scan_geom(s))
        lo, hi = B.full_window_range(p)
        E = phantom.random_body(seed=41 + i, fov_radius=250.0, z_centre=0.5 * (lo + hi),
                                z_half=0.5 * (hi - lo))
        raws.append(phantom.project(E, s, wl.hdr))
        log(f"scan {i}: {wl.hdr.TotalViewNumber} views projected")
    torch.cuda.synchronize()
    r = B.BatchedRecon(setups, raws, slabs, storage="f16")
    r.prepare()
    torch.cuda.synchronize()
    best, timings = B.autotune(r, cache=args.tune_cache, log=log)
    log(f"autotuned shape {best}")

    def timed(fn, steps):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / steps

    def sequential():
        r.prepare()
        r.backproject_sequential()

    seq_ms = timed(sequential, args.steps)
    seq_vols = [v.clone() for v in r.volumes]
    r.capture()
    graph_ms = timed(r.replay, args.steps)
    same = all(torch.equal(a, b) for a, b in zip(r.volumes, seq_vols))
    bp_batched_ms = r.time_backproject(3)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    r.backproject_sequential()
    ev[1].record()
    ev[1].synchronize()
    bp_seq_ms = ev[0].elapsed_time(ev[1])
    upd = r.updates
    out = {"metric": "GVoxel-updates/s, 3 scans / 18 uneven z-slabs, batched (HIP graph) vs "
                     "per-slab launches",
           "value": round(upd / (graph_ms * 1e-3) / 1e9, 2), "unit": "GVoxel-updates/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(graph_ms, 3), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f16 storage, f32 accumulate",
           "data": "synthetic: random-ellipsoid bodies (seeds 41-43), exact line integrals",
           "config": {"workload": "3 scans 888x64 x {3936, 8856, 14760} views -> 18 slabs "
                                  "512x512x{16..96}", "slices": sum(s.nz for s in slabs),
                      "voxel_updates": upd, "shape": dataclasses_asdict(best)},
           "sequential_ms_per_step": round(seq_ms, 3),
           "speedup_vs_sequential": round(seq_ms / graph_ms, 3),
           "backproject_ms": {"batched": round(bp_batched_ms, 3), "per_slab": round(bp_seq_ms, 3)},
           "batched_equals_sequential": bool(same),
           "autotune_ms": {k: round(v, 3) for k, v in timings.items()}}
    print(json.dumps(out), flush=True)


def bench_cnn(args):
    """The learned correction (512x512x256 workload): the 3-level 32/64/128 network
    with its BF16 MFMA kernels, trained on 64^3 patches of pairs the method
    makes, one network per scanner:

    * 368x16: seeded random-ellipsoid phantoms (720 views/rotation, pitch
      U[0.5, 1.5]) reconstructed by the plain pipeline into 192x192x64 at 1 mm,
      evaluated on held-out phantoms;
    * 888x64: 192x192x64 crops of seeded random bodies (984 views/rotation, pitch
      U[0.5, 1.5], 0.977 x 0.977 x 1.25 mm), trained from --cnn-seeds network
      seeds, each applied with its plain-FDK fallback to a 512x512x256
      reconstruction of another body (pitch 0.984).

    Reported per network: RMSE against the phantom over the field, and the
    artifact RMSE over the voxels away from the phantom's edges, with two edge
    margins: phantom within 20 HU of its 3x3x3 neighbourhood mean ("artifact"),
    and additionally more than 2 voxels from any voxel that fails it
    ("artifact_far": past the reach of the reconstruction's blur).  The 512 x
    512 x 256 figures are given per seed and as mean and sample standard
    deviation over the seeds.  Training is deterministic for a given seed."""
    import statistics

    import torch

    from analyticalreconbigpatient_amd import cnn, learned
    from analyticalreconbigpatient_amd import geometry as G
    from analyticalreconbigpatient_amd import phantom
    from analyticalreconbigpatient_amd.pipeline import HelicalFDK

    def rnd(d):
        return {k: round(v, 4) for k, v in d.items()}

    spec = learned.PairSpec()
    cspec = learned.clinical_spec()
    t0 = time.perf_counter()
    def make(seed0, n, sp, what):
        out = []
        for i in range(n):
            out.append(learned.make_pair(seed0 + i, sp))
            if (i + 1) % 50 == 0:
                log(f"{what}: {i + 1} of {n} pairs ({time.perf_counter() - t0:.1f} s)")
        return out

    pairs = make(1000, args.cnn_pairs, spec, "368x16")
    cpairs = make(3000, args.cnn_clinical_pairs, cspec, "888x64")
    held = [learned.make_pair(2000 + i, spec) for i in range(10)]
    cheld = [learned.make_pair(4000 + i, cspec) for i in range(10)]
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    log(f"{len(pairs)} + {len(cpairs)} training pairs, {len(held)} + {len(cheld)} held out "
        f"in {t_data:.1f} s")

    # the 512 x 512 x 256 volume (888x64 scan, 984 views/rotation, pitch 0.984)
    zn, dz = 256, 1.25
    hdr = G._header(888, 64, 984, 6, 0.984)
    wl = G.Workload("body 888x64x5904 -> 512x512x256", hdr, 512, 512, zn, 250.0,
                    G.centred_ZC(hdr, zn, dz) + 0.137)
    s = wl.setup()
    p = G.bp_params(G.bp_scan_geom(s))
    z_mid = ((zn - 1) * 0.5 - p.ZNC) * dz
    E = phantom.random_body(seed=2, fov_radius=250.0, z_centre=z_mid, z_half=0.5 * zn * dz)
    raw = phantom.project(E, s, hdr)
    fdk = HelicalFDK(s, dz=dz)
    vol = fdk.run(raw)
    x0 = (vol / learned.MU_W - 1.0)[None].contiguous()
    dev = x0.device
    xs = (torch.arange(p.XN, device=dev, dtype=torch.float32) - p.XNC) * p.dx
    ys = (torch.arange(p.YN, device=dev, dtype=torch.float32) - p.YNC) * p.dy
    zs = (torch.arange(p.ZN, device=dev, dtype=torch.float32) - p.ZNC) * dz
    target = learned.voxelize_gpu(E, xs, ys, zs, ss=2) / learned.MU_W - 1.0
    field = (xs[None, :] ** 2 + ys[:, None] ** 2 <= (0.95 * 250.0) ** 2)[:, :, None]
    field = field.expand_as(target)
    masks = {"": field, "artifact_": field & learned.smooth_mask(target),
             "artifact_far_": field & learned.far_mask(target)}

    def rmse(a, m):
        return 1000 * float(((a - target)[m] ** 2).mean().sqrt())

    def big_volume(net):
        y, frac = net.correct(x0, bound=0.3)          # first call sizes the workspace
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        y, frac = net.correct(x0, bound=0.3)
        e1.record()
        e1.synchronize()
        out = {"ms": e0.elapsed_time(e1), "fallback_fraction": frac}
        for tag, m in masks.items():
            f, c = rmse(x0[0], m), rmse(y[0], m)
            out.update({f"{tag}rmse_fdk_hu": f, f"{tag}rmse_cnn_hu": c,
                        f"{tag}reduction": 1 - c / f})
        return out

    t0 = time.perf_counter()
    net = cnn.Net(seed=0)
    hist = learned.train(net, pairs, steps=args.cnn_steps, batch=args.cnn_batch, patch=64,
                         lr=1e-3, log=log, log_every=1000, ema=args.cnn_ema)
    torch.cuda.synchronize()
    t368 = time.perf_counter() - t0
    ev = learned.evaluate(net, held, spec, bound=0.3)
    log(f"368x16 held out: {rnd(ev)}")
    runs = []
    for seed in range(args.cnn_seeds):
        net = cnn.Net(seed=seed)
        t0 = time.perf_counter()
        h = learned.train(net, cpairs, steps=args.cnn_steps, batch=args.cnn_batch, patch=64,
                          lr=1e-3, seed=seed, log=log, log_every=1000, ema=args.cnn_ema)
        torch.cuda.synchronize()
        tt = time.perf_counter() - t0
        cev = learned.evaluate(net, cheld, cspec, bound=0.3)
        inf = big_volume(net)
        runs.append({"seed": seed, "train_s": round(tt, 2), "loss_first_last": [h[0][1], h[-1][1]],
                     "held_out_crops": rnd(cev), "volume_512x512x256": rnd(inf)})
        log(f"888x64 seed {seed}: crops {rnd(cev)}; 512x512x256 {rnd(inf)}")

    def stat(key):
        v = [r["volume_512x512x256"][key] for r in runs]
        return {"mean": round(statistics.mean(v), 4),
                "std": round(statistics.stdev(v), 4) if len(v) > 1 else 0.0,
                "min": round(min(v), 4), "max": round(max(v), 4)}

    summary = {k: stat(k) for k in ("reduction", "artifact_reduction", "artifact_far_reduction",
                                    "rmse_cnn_hu", "artifact_rmse_cnn_hu",
                                    "artifact_far_rmse_cnn_hu", "ms")}
    base = runs[0]["volume_512x512x256"]
    out = {"metric": "artifact RMSE vs phantom on a 512x512x256 reconstruction, learned "
                     "correction vs plain FDK (mean over network seeds)",
           "value": summary["artifact_far_reduction"]["mean"], "unit": "fraction",
           "higher_is_better": True, "n_gpus": 1,
           "dtype": "bf16 weights/activations, f32 accumulate",
           "data": "synthetic: seeded random-ellipsoid phantoms and bodies, exact projections, "
                   "plain-FDK reconstructions",
           "config": {"workload": f"3-level 32/64/128 network, 64^3 patches, batch "
                                  f"{args.cnn_batch}; 368x16 phantoms -> 192x192x64; 888x64 "
                                  "body crops -> 192x192x64; infer 512x512x256 (888x64, "
                                  "pitch 0.984)",
                      "train_pairs_368x16": len(pairs), "train_pairs_888x64": len(cpairs),
                      "steps": args.cnn_steps, "seeds": args.cnn_seeds,
                      "weight_ema": args.cnn_ema,
                      "artifact_voxels": "phantom within 20 HU of its 3x3x3 neighbourhood mean",
                      "artifact_far_voxels": "artifact voxels more than 2 voxels from any "
                                             "voxel that is not one"},
           "volume_512x512x256_plain_fdk": {k: round(v, 4) for k, v in base.items()
                                            if "fdk" in k},
           "volume_512x512x256_over_seeds": summary,
           "runs_888x64": runs,
           "held_out_368x16": rnd(ev), "loss_first_last_368x16": [hist[0][1], hist[-1][1]],
           "train_s_368x16": round(t368, 2),
           "ms_per_train_step": round(1e3 * t368 / args.cnn_steps, 2),
           "data_s": round(t_data, 2)}
    print(json.dumps(out), flush=True)


def dataclasses_asdict(x):
    import dataclasses
    return dataclasses.asdict(x)


if __name__ == "__main__":
    main()
