"""Benchmark: GVoxel-updates/s of the end-to-end helical FDK reconstruction
(rebin -> preweight + filter -> 3D-weighted backprojection) at 512 x 512 x 400
from an 888 x 64 x 11,808-view helical scan, on one MI355X.

One step = one pass of the three stages over the device-resident sinogram.
value = voxel-updates of one pass (512*512*400 voxels x 671-view windows =
7.036e10) / measured seconds per pass.  Prints one JSON line.

Also reported: per-stage times (HIP events on the launch stream), the FP16
filtered-storage variant, the dominant kernel's roofline, the HU RMSE against
the FP64 oracle on sampled slices, and the CPU baseline (the oracle, timed on
this host on a bounded sample and scaled to the whole workload).
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
    ap.add_argument("--overlap", type=int, default=1,
                    help="view chunks / slice-block groups of the overlapped pass (1: off)")
    ap.add_argument("--layout", default=None,
                    help="filtered layout: rows, rowpairs, rowdiff or quads")
    ap.add_argument("--storages", default="f32,f16",
                    help="filtered-sinogram storages to time (plain f32 first)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 2 GiB device copy (calibrates HBM byte counters)")
    ap.add_argument("--workload", default="clinical", choices=["clinical", "batch", "cnn"],
                    help="clinical: the 888x64x11808 -> 512x512x400 pass (default); batch: "
                         "3 scans of 4/9/15 rotations, 18 uneven slabs, batched + HIP graph; "
                         "cnn: train the learned correction and run it on 512x512x256")
    ap.add_argument("--cnn-pairs", type=int, default=200, help="training phantoms (cnn)")
    ap.add_argument("--cnn-clinical-pairs", type=int, default=300,
                    help="training crops of the 888x64 scan (cnn)")
    ap.add_argument("--cnn-steps", type=int, default=10000,
                    help="training steps of each network (cnn)")
    ap.add_argument("--tune-cache", default=None, help="autotuner cache file (batch workload)")
    return ap.parse_args()


def kernel_name(fdk) -> str:
    """The backprojector's kernel as rocprofv3 names it (FP32 storage)."""
    lay = {"rows": 0, "rowpairs": 1, "quads": 2, "rowdiff": 3}[fdk.layout]
    inside = "true" if fdk.plan.inside else "false"
    if fdk.variant == 2:
        g = 0 if fdk.group == 8 else fdk.group
        return f"bp_fdk3dw_zc<{fdk.plan_nz}, {g}, float, {lay}, {g == 0 and inside or 'false'}>"
    if fdk.variant == 3:
        rmax = 21 if fdk.plan_nz == 8 else 40
        return f"bp_fdk3dw_zs<{fdk.plan_nz}, float, {lay}, {rmax}, {inside}>"
    return f"bp_fdk3dw{'_vt' if fdk.variant == 0 else ''}<{fdk.plan_nz}, float>"


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

    import numpy as np
    import torch

    if args.copy_probe:
        n = 1 << 29                                   # 2 GiB of FP32
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": 4 * n, "copy_probe_bytes_written": 4 * n}))
        return

    if args.workload == "batch":
        bench_batch(args)
        return
    if args.workload == "cnn":
        bench_cnn(args)
        return

    from analyticalreconbigpatient_amd import geometry as G
    from analyticalreconbigpatient_amd import phantom
    from analyticalreconbigpatient_amd.pipeline import HelicalFDK

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    wl = G.body_clinical()
    s = wl.setup()
    p0 = G.bp_params(G.bp_scan_geom(s))
    z_mid = ((wl.zn - 1) * 0.5 - p0.ZNC) * p0.dz
    log(f"workload {wl.name}; projecting phantom")
    E = phantom.random_body(seed=2, fov_radius=250.0, z_centre=z_mid, z_half=250.0)
    raw = phantom.project(E, s, wl.hdr, device=str(dev))
    torch.cuda.synchronize()

    results = {}
    storages = [x for x in args.storages.split(",") if x]
    assert storages, "no storage to time"
    for storage in storages:
        fdk = HelicalFDK(s, storage=storage, nz=args.nz, order=args.order,
                         variant=args.variant, group=args.group, layout=args.layout)
        for _ in range(args.warmup):
            fdk.run(raw)
        torch.cuda.synchronize()
        log(f"{storage}: warmup done, timing {args.steps} steps")
        ev = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(args.steps)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(args.steps):
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
        seq_ms = wall * 1e3 / args.steps
        ovl_ms = None
        if args.overlap > 1 and fdk.plan.nblocks > 1:
            for _ in range(args.warmup):
                fdk.run_overlapped(raw, args.overlap, args.overlap)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                fdk.run_overlapped(raw, args.overlap, args.overlap)
            torch.cuda.synchronize()
            ovl_ms = (time.perf_counter() - t0) * 1e3 / args.steps
        ms = min(seq_ms, ovl_ms) if ovl_ms else seq_ms
        results[storage] = dict(fdk=fdk, wall=wall, stages=st.mean(axis=0), ms=ms,
                                seq_ms=seq_ms, ovl_ms=ovl_ms)
        log(f"{storage}: {ms:.2f} ms/step (stages back to back {seq_ms:.2f}, overlapped "
            f"{ovl_ms}), stages (ms) {np.round(st.mean(axis=0), 3).tolist()}")

    plain = results.get("f32", results[storages[0]])
    fdk = plain["fdk"]
    updates = fdk.updates
    value = updates / (plain["ms"] * 1e-3) / 1e9
    chosen = results.get("f16", plain)
    chosen_value = updates / (chosen["ms"] * 1e-3) / 1e9

    # Roofline of the dominant kernel (backprojection): algorithmic bytes per
    # launch = 16 B of gathered FP32 samples per voxel-update + 4 B per voxel.
    bp_ms = float(plain["stages"][2])
    p = fdk.params
    nvox = p.XN * p.YN * p.ZN
    alg_bytes = 16.0 * updates + 4.0 * nvox
    achieved = alg_bytes / (bp_ms * 1e-3) / 1e9
    kname = kernel_name(fdk)
    traffic, entry = None, None
    tfile = ROOT / "profiles" / "traffic.json"
    if tfile.exists():
        try:
            entry = json.loads(tfile.read_text()).get("kernels", {}).get(kname)
            if entry and entry.get("workload") == wl.name:
                traffic = entry["fetch_bytes"] + entry["write_bytes"]
            else:
                entry = None
        except (ValueError, OSError, KeyError, TypeError):
            traffic, entry = None, None
    gvu = updates / (bp_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": kname, "kernel_ms": round(bp_ms, 3), "bytes_per_update": 16}
    if fdk.variant == 3:
        # Samples come from LDS: each view's detector footprint is copied there once
        # per workgroup.  The bound is VALU issue: one wavefront instruction per CU
        # per clock (transcendentals issue on their own pipeline), so at V
        # instructions per voxel-update the roof is N_CU * clock * 64 / V.
        roofline["note"] = ("samples are read from LDS, where each view's detector footprint is "
                            "staged once per workgroup; compulsory HBM bytes are the filtered "
                            "sinogram and the volume; VALU issue bounds this kernel")
        vpu = entry.get("valu_per_update") if entry else None
        if vpu:
            peak = N_CU * CLOCK_HZ * 64 / vpu / 1e9
            roofline["valu_issue"] = {
                "achieved": round(gvu, 1), "peak": round(peak, 1), "unit": "GVoxel-updates/s",
                "frac": round(gvu / peak, 4), "valu_per_update": vpu,
                "trans_per_update": entry.get("trans_per_update"),
                "source": "profiles/sq_counters_bp_fdk3dw_zs_clinical.md"}
    else:
        # Issue bound of the gathers: the texture addresser takes four lane
        # addresses per clock, so one 64-lane gather instruction occupies it for 16
        # clocks; the row-pair, row-difference and quad layouts need one gather per
        # voxel-update, the rows layout two.
        gathers = 2 if fdk.layout == "rows" else 1
        gather_peak = N_CU * CLOCK_HZ / 16.0 * 64 / gathers / 1e9
        roofline["note"] = ("gathers are served from L1/L2; compulsory HBM bytes are the filtered "
                            "sinogram once and the volume once, so the gather issue rate, not HBM, "
                            "bounds this kernel")
        roofline["gather_issue"] = {"achieved": round(gvu, 1), "peak": round(gather_peak, 1),
                                    "unit": "GVoxel-updates/s", "frac": round(gvu / gather_peak, 4),
                                    "gathers_per_update": gathers}

    stage_roof = {
        "rebin_GBs": round(8.0 * p.YL * p.ZL * fdk.ViewN / (plain["stages"][0] * 1e-3) / 1e9, 1),
        "filter_GBs": round(8.0 * p.YL * p.ZL * fdk.ViewN / (plain["stages"][1] * 1e-3) / 1e9, 1),
    }

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
                      "backproject": round(bp_ms, 3)},
        "pass_ms": {"stages_back_to_back": round(plain["seq_ms"], 3),
                    "overlapped": round(plain["ovl_ms"], 3) if plain["ovl_ms"] else None,
                    "overlap_chunks": args.overlap},
        "stage_hbm_GBs": stage_roof,
        "chosen": {"storage": "f16" if "f16" in results else "f32",
                   "value": round(chosen_value, 2),
                   "ms_per_step": round(chosen["ms"], 3),
                   "ms_back_to_back": round(chosen["seq_ms"], 3),
                   "stages_ms": [round(float(x), 3) for x in chosen["stages"]],
                   "speedup_vs_plain": round(chosen_value / value, 3)},
        "roofline": roofline,
    }

    if not args.no_cpu:
        out["cpu_baseline"], out["hu_rmse"] = cpu_leg(s, fdk, results, raw, args)
    print(json.dumps(out), flush=True)


def cpu_leg(s, fdk, results, raw, args):
    """Time the oracle on this host on a bounded sample and compare sampled
    slices of the GPU volumes with the FP64 oracle chain."""
    import numpy as np
    import torch

    import oracle
    from analyticalreconbigpatient_amd import geometry as G

    nth = host_threads()
    p = fdk.params
    W = G.window_half_width(p.N_2pi, p.HSCoef)
    zs = np.linspace(p.ZN * 0.45, p.ZN * 0.55, args.cpu_slices).astype(int)
    # views needed: windows of the sampled slices plus the rebinning's fan margin
    s0 = [math.ceil((2 * 3.14159265358979 * (z - p.ZNC) * p.dz / p.h - p.BetaS) / p.DeltaFai - 0.5)
          for z in zs]
    margin = int(math.ceil(math.asin(min(1.0, 1.5 * math.tan(s.DecAngle / 2))) / p.DeltaFai)) + 4
    v_lo = max(0, min(s0) - W - margin)
    v_hi = min(fdk.ViewN, max(s0) + W + margin + 1)
    a = G.rebin_args(s.geom)
    raw_np = raw[v_lo:v_hi].cpu().numpy()
    log(f"cpu leg: views [{v_lo}, {v_hi}), slices {zs.tolist()}, {nth} threads")

    t = time.perf_counter()
    reb = oracle.rebin_f32(raw_np, a.YLC0, a.dYA, a.DeltaTheta, a.PLC0, a.DeltaT, a.DeltaFai,
                           a.SO)
    t_reb = time.perf_counter() - t
    fa = G.filter_args(s)
    t = time.perf_counter()
    filt = oracle.filter_f64(np.ascontiguousarray(reb.transpose(0, 2, 1)), float(fa.PLC0),
                             float(fa.ZLC0), float(fa.dYL), float(fa.dZL), float(fa.SO),
                             nthreads=nth)
    t_filt = time.perf_counter() - t
    # backproject the sampled slices from the view sub-range: shift BetaS and
    # the slice window by v_lo (views are independent).
    sub = G.BPParams(**{**p.__dict__, "PN": v_hi - v_lo,
                        "BetaS": p.BetaS + v_lo * p.DeltaFai})
    w = oracle.fdk_weights(p.N_2pi, p.HSCoef, p.delta)
    ref = np.zeros((p.YN, p.XN, p.ZN))
    t = time.perf_counter()
    for z in zs:
        oracle.fdk_core(filt, w, N_2pi=sub.N_2pi, XN=sub.XN, YN=sub.YN, ZN=sub.ZN, XNC=sub.XNC,
                        YNC=sub.YNC, ZNC=sub.ZNC, dx=sub.dx, dy=sub.dy, dz=sub.dz, YLC=sub.YLC,
                        ZLC=sub.ZLC, dYL=sub.dYL, dZL=sub.dZL, h=sub.h, BetaS=sub.BetaS,
                        DeltaFai=sub.DeltaFai, N_pi=sub.N_pi, HSCoef=sub.HSCoef, SO=sub.SO,
                        SD=sub.SD, k1=sub.k1, zi_range=(int(z), int(z) + 1), nthreads=nth,
                        out=ref)
    t_bp = time.perf_counter() - t
    nv = v_hi - v_lo
    upd_sample = len(zs) * p.XN * p.YN * (2 * W + 1)
    t_total = (t_reb / nv * fdk.ViewN + t_filt / nv * fdk.ViewN
               + t_bp / upd_sample * fdk.updates)
    cpu = {"value": round(fdk.updates / t_total / 1e9, 5), "unit": "GVoxel-updates/s",
           "cores": nth, "kind": "port",
           "sample": (f"oracle (rebin FP32, filter FP64, backprojection FP64) on views "
                      f"[{v_lo},{v_hi}) and {len(zs)} slices of 512x512; measured "
                      f"{t_reb:.2f}+{t_filt:.2f}+{t_bp:.2f} s, scaled per view / per "
                      f"voxel-update to the whole workload ({t_total:.0f} s)")}
    # HU RMSE (1000/mu_w) RMSE(mu) inside the FOV cylinder on the sampled slices
    yy, xx = np.meshgrid((np.arange(p.YN) - p.YNC) * p.dy, (np.arange(p.XN) - p.XNC) * p.dx,
                         indexing="ij")
    mask = (xx ** 2 + yy ** 2) <= (0.98 * 250.0) ** 2
    hu = {}
    for name, res in results.items():
        vol = res["fdk"].run(raw)
        torch.cuda.synchronize()
        g = vol[:, :, torch.as_tensor(zs, device=vol.device)].cpu().numpy().astype(np.float64)
        r = ref[:, :, zs]
        diff = (g - r)[mask]
        hu[name] = round(float(1000.0 / 0.02 * np.sqrt(np.mean(diff ** 2))), 4)
        hu[name + "_rel_l2"] = float(np.linalg.norm(g - r) / np.linalg.norm(r))
    log(f"cpu leg done: {cpu['value']} GVU/s, HU RMSE {hu}")
    return cpu, hu


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
    """The learned correction, one network per scanner, each trained on pairs
    the method makes for that scanner: (1) seeded random-ellipsoid phantoms on
    the 368x16 scan (720 views/rotation, pitch U[0.5, 1.5]) reconstructed by
    the plain pipeline into 192x192x64 at 1 mm; (2) 192x192x64 crops of seeded
    random bodies on the 888x64 scan (984 views/rotation, pitch U[0.5, 1.5],
    0.977 x 0.977 x 1.25 mm).  64^3 patches, batch 4, Adam.  Reported: RMSE
    against the phantom on held-out pairs of each scanner, over all voxels and
    over the voxels away from phantom edges (artifact RMSE), and inference of
    the 888x64 network on a 512x512x256 reconstruction of a random body
    (pitch 0.984) with the same two RMSEs."""
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
    pairs = [learned.make_pair(1000 + i, spec) for i in range(args.cnn_pairs)]
    cpairs = [learned.make_pair(3000 + i, cspec) for i in range(args.cnn_clinical_pairs)]
    held = [learned.make_pair(2000 + i, spec) for i in range(10)]
    cheld = [learned.make_pair(4000 + i, cspec) for i in range(10)]
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    log(f"{len(pairs)} + {len(cpairs)} training pairs, {len(held)} + {len(cheld)} held out "
        f"in {t_data:.1f} s")
    nets, hists, t_train = {}, {}, {}
    for name, train_pairs in (("368x16", pairs), ("888x64", cpairs)):
        net = cnn.Net(seed=0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        hists[name] = learned.train(net, train_pairs, steps=args.cnn_steps, batch=4, patch=64,
                                    lr=1e-3, log=log, log_every=1000)
        torch.cuda.synchronize()
        t_train[name] = time.perf_counter() - t0
        nets[name] = net
    ev = learned.evaluate(nets["368x16"], held, spec, bound=0.3)
    cev = learned.evaluate(nets["888x64"], cheld, cspec, bound=0.3)
    log(f"held-out 368x16: {rnd(ev)}; held-out 888x64 crops: {rnd(cev)}")

    # inference on 512x512x256 (888x64 scan, 984 views/rotation, pitch 0.984)
    net = nets["888x64"]
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
    y, frac = net.correct(x0, bound=0.3)          # first call sizes the workspace
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    y, frac = net.correct(x0, bound=0.3)
    ev1.record()
    ev1.synchronize()
    t_inf = ev0.elapsed_time(ev1)
    r2 = (xs[None, :] ** 2 + ys[:, None] ** 2 <= (0.95 * 250.0) ** 2)[:, :, None].expand_as(target)
    sm = r2 & learned.smooth_mask(target)

    def rmse(a, m):
        return 1000 * float(((a - target)[m] ** 2).mean().sqrt())

    inf = {"ms": round(t_inf, 2), "rmse_fdk_hu": rmse(x0[0], r2), "rmse_cnn_hu": rmse(y[0], r2),
           "artifact_rmse_fdk_hu": rmse(x0[0], sm), "artifact_rmse_cnn_hu": rmse(y[0], sm)}
    inf["reduction"] = 1 - inf["rmse_cnn_hu"] / inf["rmse_fdk_hu"]
    inf["artifact_reduction"] = 1 - inf["artifact_rmse_cnn_hu"] / inf["artifact_rmse_fdk_hu"]
    inf = rnd(inf)
    inf.update({"fallback_fraction": round(frac, 4), "bp_voxel_updates": fdk.updates})
    log(f"512x512x256: {inf}")
    out = {"metric": "artifact RMSE vs phantom, learned correction vs plain FDK (held-out "
                     "phantoms)", "value": round(ev["artifact_reduction"], 4),
           "unit": "fraction", "higher_is_better": True, "n_gpus": 1,
           "dtype": "bf16 weights/activations, f32 accumulate",
           "data": "synthetic: seeded random-ellipsoid phantoms, exact projections, plain-FDK "
                   "reconstructions",
           "config": {"workload": "one network per scanner: 368x16 scans (pitch U[0.5,1.5]) -> "
                                  "192x192x64; 888x64 crops -> 192x192x64; 64^3 patches; infer "
                                  "512x512x256 (888x64)",
                      "train_pairs_368x16": len(pairs), "train_pairs_888x64": len(cpairs),
                      "steps": args.cnn_steps, "batch": 4,
                      "artifact_voxels": "phantom within 20 HU of its 3x3x3 neighbourhood mean"},
           "held_out_368x16": rnd(ev), "held_out_888x64_crops": rnd(cev),
           "loss_first_last": {k: [h[0][1], h[-1][1]] for k, h in hists.items()},
           "train_s": {k: round(v, 2) for k, v in t_train.items()},
           "ms_per_train_step": round(1e3 * t_train["368x16"] / args.cnn_steps, 2),
           "data_s": round(t_data, 2), "inference_512x512x256": inf}
    print(json.dumps(out), flush=True)


def dataclasses_asdict(x):
    import dataclasses
    return dataclasses.asdict(x)


if __name__ == "__main__":
    main()
