"""Training data, training and evaluation of the learned correction.

Pairs are made by the method itself: a seeded random-ellipsoid phantom, its
exact helical projections, reconstructed by the plain GPU pipeline; the target
is the phantom sampled on the same grid (2x2x2 supersampled).  Values are
HU / 1000.  Two scanners:

* ``PairSpec()``: the 368x16 curved detector (720 views/rotation, pitch drawn
  from [0.5, 1.5]) into a 192x192x64 volume of 1 mm voxels;
* ``clinical_spec()``: the 888x64 detector (984 views/rotation, pitch drawn
  from [0.5, 1.5]) into a 192x192x64 crop of the 500 mm field at
  0.977 x 0.977 x 1.25 mm, centred at a random point of the body.

The network (``cnn.Net``) learns the residual target - FDK on 64^3 patches;
the plain FDK is both its input and the yardstick of its error.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import geometry as G
from . import phantom
from .cnn import Net
from .pipeline import HelicalFDK

MU_W = phantom.MU_WATER


@dataclass
class PairSpec:
    nx: int = 192
    nz: int = 64
    dx: float = 1.0
    radius: float = 88.0          # phantom support (inside the fan's 106.7 mm FOV)
    scanner: str = "368x16"       # or "888x64" (crops of the 500 mm field)
    dz: float | None = None       # slice pitch (default dx)
    field: float = 250.0          # 888x64: radius of the body's field
    body_z_half: float = 160.0    # 888x64: half length of the random body's ellipsoid range


def clinical_spec(nx: int = 192, nz: int = 64) -> PairSpec:
    """Crops of the 888x64 scanner's 512x512 grid (500 mm field, 1.25 mm slices)."""
    return PairSpec(nx=nx, nz=nz, dx=500.0 / 512, radius=None, scanner="888x64", dz=1.25)


def scan_setup(pitch: float, spec: PairSpec):
    """The 368x16 scan with enough rotations to sample nz slices fully."""
    h = pitch * 16 * 1.0963
    rotations = math.ceil(spec.nz * spec.dx / h + 1.0) + 1
    hdr = G._header(368, 16, 720, rotations, pitch)
    wl = G.Workload(f"368x16 pitch {pitch:.3f}", hdr, spec.nx, spec.nx, spec.nz,
                    spec.nx * spec.dx / 2, 10.0)
    s = wl.setup()
    s.geom.Obj[6] = G.centred_ZC(hdr, spec.nz, spec.dx) + 0.137   # off 0/0 ties
    return wl, s


def random_phantom(seed: int, z_mid: float, spec: PairSpec) -> np.ndarray:
    """Water-like body inside the support plus 10..30 random ellipsoids of
    HU in [-1000, 1500] (rows as ``phantom.shepp_logan``)."""
    rng = np.random.default_rng(seed)
    R = spec.radius
    rows = [(0.0, 0.0, z_mid, R * rng.uniform(0.8, 0.98), R * rng.uniform(0.65, 0.9),
             spec.nz * spec.dx, rng.uniform(0, math.pi), MU_W * (1 + rng.uniform(-0.1, 0.1)))]
    for _ in range(int(rng.integers(10, 31))):
        r = 0.6 * R * math.sqrt(rng.uniform())
        ang = rng.uniform(0, 2 * math.pi)
        rows.append((r * math.cos(ang), r * math.sin(ang),
                     z_mid + rng.uniform(-0.5, 0.5) * spec.nz * spec.dx,
                     rng.uniform(3, 25), rng.uniform(3, 25), rng.uniform(3, 30),
                     rng.uniform(0, math.pi), MU_W * rng.uniform(-1.0, 1.5)))
    return np.asarray(rows, dtype=np.float32)


def voxelize_gpu(E: np.ndarray, xs: torch.Tensor, ys: torch.Tensor, zs: torch.Tensor,
                 ss: int = 2) -> torch.Tensor:
    """Mean of mu over ss^3 sub-samples per voxel, [len(ys)][len(xs)][len(zs)]."""
    dx = float(xs[1] - xs[0]) if len(xs) > 1 else 1.0
    dz = float(zs[1] - zs[0]) if len(zs) > 1 else 1.0
    off = (torch.arange(ss, device=xs.device, dtype=torch.float32) + 0.5) / ss - 0.5
    out = torch.zeros((len(ys), len(xs), len(zs)), device=xs.device)
    for oy in off:
        for ox in off:
            for oz in off:
                Y = (ys + oy * dx)[:, None, None]
                X = (xs + ox * dx)[None, :, None]
                Z = (zs + oz * dz)[None, None, :]
                for cx, cy, cz, a, b, c, phi, mu in E.astype(np.float64):
                    cp, sp = math.cos(phi), math.sin(phi)
                    u = (cp * (X - cx) + sp * (Y - cy)) / a
                    v = (-sp * (X - cx) + cp * (Y - cy)) / b
                    w = (Z - cz) / c
                    out += torch.where(u * u + v * v + w * w <= 1.0, float(mu), 0.0)
    return out / ss ** 3


def make_clinical_pair(seed: int, spec: PairSpec, pitch: float | None = None):
    """A crop of a random body on the 888x64 scanner: the scan covers the crop's
    slices (plus the windows and the rebinning margin), the crop's centre is
    drawn so that it stays inside the body."""
    rng = np.random.default_rng(seed + 7919)
    pitch = float(rng.uniform(0.5, 1.5)) if pitch is None else pitch
    dz = spec.dz or spec.dx
    h = pitch * 64 * 1.0963
    rotations = math.ceil(spec.nz * dz / h + 1.0) + 1
    hdr = G._header(888, 64, 984, rotations, pitch)
    half = 0.5 * spec.nx * spec.dx
    # crop centre uniform in the body's cross-section (0.98 x 0.8 of the field),
    # so crops also hold the body's edge and the air around it
    r, ang = 0.98 * spec.field * math.sqrt(rng.uniform()), rng.uniform(0, 2 * math.pi)
    xc, yc = r * math.cos(ang), 0.8 / 0.98 * r * math.sin(ang)
    wl = G.Workload(f"888x64 crop pitch {pitch:.3f}", hdr, spec.nx, spec.nx, spec.nz, half,
                    G.centred_ZC(hdr, spec.nz, dz) + 0.137)
    s = wl.setup()
    s.geom.Obj[4] = (spec.nx - 1) * 0.5 - xc / spec.dx + 1.0     # 1-based centres
    s.geom.Obj[5] = (spec.nx - 1) * 0.5 - yc / spec.dx + 1.0
    p = G.bp_params(G.bp_scan_geom(s))
    z_mid = ((spec.nz - 1) * 0.5 - p.ZNC) * dz
    # a long body as in the clinical scans (the body runs far beyond the scanned
    # slab; a body that ends near the slab biases the plain FDK differently)
    E = phantom.random_body(seed, fov_radius=spec.field, z_centre=z_mid, z_half=spec.body_z_half)
    raw = phantom.project(E, s, hdr)
    fdk = HelicalFDK(s, dz=dz)
    vol = fdk.run(raw)
    x0 = vol / MU_W - 1.0
    dev = vol.device
    xs = (torch.arange(p.XN, device=dev, dtype=torch.float32) - p.XNC) * p.dx
    ys = (torch.arange(p.YN, device=dev, dtype=torch.float32) - p.YNC) * p.dy
    zs = (torch.arange(p.ZN, device=dev, dtype=torch.float32) - p.ZNC) * dz
    target = voxelize_gpu(E, xs, ys, zs) / MU_W - 1.0
    return x0.contiguous(), target.contiguous(), pitch


def make_pair(seed: int, spec: PairSpec = PairSpec(), pitch: float | None = None):
    """(plain FDK, phantom) in HU / 1000, FP32 [nx][nx][nz] on the device."""
    if spec.scanner == "888x64":
        return make_clinical_pair(seed, spec, pitch)
    rng = np.random.default_rng(seed + 7919)
    pitch = float(rng.uniform(0.5, 1.5)) if pitch is None else pitch
    wl, s = scan_setup(pitch, spec)
    p = G.bp_params(G.bp_scan_geom(s))
    z_mid = ((spec.nz - 1) * 0.5 - p.ZNC) * spec.dx
    E = random_phantom(seed, z_mid, spec)
    raw = phantom.project(E, s, wl.hdr)
    fdk = HelicalFDK(s, dz=spec.dx)
    vol = fdk.run(raw)
    x0 = vol / MU_W - 1.0
    dev = vol.device
    xs = (torch.arange(p.XN, device=dev, dtype=torch.float32) - p.XNC) * p.dx
    ys = (torch.arange(p.YN, device=dev, dtype=torch.float32) - p.YNC) * p.dy
    zs = (torch.arange(p.ZN, device=dev, dtype=torch.float32) - p.ZNC) * spec.dx
    target = voxelize_gpu(E, xs, ys, zs) / MU_W - 1.0
    return x0.contiguous(), target.contiguous(), pitch


def fov_mask(spec: PairSpec, device) -> torch.Tensor:
    """The phantom's support (368x16 pairs); every voxel of a clinical crop."""
    if spec.radius is None:
        return torch.ones((spec.nx, spec.nx, spec.nz), dtype=torch.bool, device=device)
    c = (torch.arange(spec.nx, device=device, dtype=torch.float32) - (spec.nx - 1) / 2) * spec.dx
    r2 = c[:, None] ** 2 + c[None, :] ** 2
    return (r2 <= spec.radius ** 2)[:, :, None].expand(spec.nx, spec.nx, spec.nz)


def sample_batch(pairs, batch: int, patch: int, gen: torch.Generator):
    """Random patch^3 crops (random flips) from random pairs, [B][p][p][p]."""
    xs, ts = [], []
    for _ in range(batch):
        i = int(torch.randint(len(pairs), (1,), generator=gen))
        x0, t = pairs[i][0], pairs[i][1]
        ny, nx, nz = x0.shape
        oy = int(torch.randint(ny - patch + 1, (1,), generator=gen))
        ox = int(torch.randint(nx - patch + 1, (1,), generator=gen))
        oz = int(torch.randint(nz - patch + 1, (1,), generator=gen))
        a = x0[oy:oy + patch, ox:ox + patch, oz:oz + patch]
        b = t[oy:oy + patch, ox:ox + patch, oz:oz + patch]
        flips = [d for d in range(3) if bool(torch.randint(2, (1,), generator=gen))]
        if flips:
            a, b = a.flip(flips), b.flip(flips)
        xs.append(a)
        ts.append(b)
    return torch.stack(xs).contiguous(), torch.stack(ts).contiguous()


def train(net: Net, pairs, steps: int, batch: int = 4, patch: int = 64, lr: float = 1e-3,
          seed: int = 0, log=None, log_every: int = 100, warmup: int = 300,
          ema: float = 0.0):
    """Adam with a linear learning-rate warmup over ``warmup`` steps, then a
    cosine decay; returns the loss history.  (Without the warmup one of three
    seeds of a 10,000-step, 192-crop run never left the plain FDK: its loss
    stayed flat from the first steps.)  ``ema`` > 0: the network ends with the
    exponential moving average of its weights over the steps after the warmup
    (w_avg += (1 - ema) (w - w_avg) after every step), which smooths the noise
    of small-batch updates."""
    gen = torch.Generator().manual_seed(seed)
    net.pack()
    hist = []
    avg = None
    t0 = time.perf_counter()
    for k in range(steps):
        x, t = sample_batch(pairs, batch, patch, gen)
        lr_k = lr * 0.5 * (1 + math.cos(math.pi * k / max(1, steps)))
        lr_k = max(lr_k, lr * 0.02) * min(1.0, (k + 1) / max(1, warmup))
        loss = net.train_step(x, t, lr=lr_k)
        if ema > 0.0 and k + 1 >= warmup:
            if avg is None:
                avg = {n: (L.w.clone(), L.b.clone()) for n, L in net.layers.items()}
            else:
                for n, L in net.layers.items():
                    avg[n][0].lerp_(L.w, 1.0 - ema)
                    avg[n][1].lerp_(L.b, 1.0 - ema)
        if k % 10 == 0 or k == steps - 1:
            hist.append((k, float(loss)))
        if log and (k % log_every == 0 or k == steps - 1):
            log(f"step {k}: loss {hist[-1][1]:.5f} ({time.perf_counter() - t0:.1f} s)")
    if avg is not None:
        net.load_state_dict(avg)
    return hist


def smooth_mask(target: torch.Tensor, tol: float = 0.02) -> torch.Tensor:
    """Voxels whose phantom value is within ``tol`` (HU / 1000) of its 3x3x3
    neighbourhood mean: away from the phantom's edges, where the error of a
    reconstruction is artifact (streaks, bias, cupping) rather than the edge
    blur of its resolution."""
    t = target[None, None]
    mean = torch.nn.functional.avg_pool3d(t, 3, 1, 1, count_include_pad=False)[0, 0]
    return (target - mean).abs() <= tol


def far_mask(target: torch.Tensor, tol: float = 0.02, margin: int = 2) -> torch.Tensor:
    """Voxels of ``smooth_mask`` more than ``margin`` voxels from every voxel
    that is not (no phantom edge in the (2 margin + 1)^3 neighbourhood).  The
    3x3x3 criterion of smooth_mask still admits voxels one or two voxels from an
    edge, where the reconstruction's error is mostly the blur of its resolution
    (its point spread reaches 2-3 voxels); past that margin the error is
    artifact: streaks, cone-beam (windmill) patterns, bias and cupping."""
    sm = smooth_mask(target, tol)
    edge = (~sm).float()[None, None]
    near = torch.nn.functional.max_pool3d(edge, 2 * margin + 1, 1, margin)[0, 0] > 0
    return sm & ~near


def evaluate(net: Net, pairs, spec: PairSpec, bound: float = 0.3):
    """RMSE in HU against the phantom inside the support, for the plain FDK
    and the corrected volume (with the fallback), summed over the pairs; the
    ``artifact`` figures cover only the voxels of ``smooth_mask``."""
    m = fov_mask(spec, pairs[0][0].device)
    se = {"fdk": 0.0, "cnn": 0.0, "fdk_s": 0.0, "cnn_s": 0.0, "fdk_f": 0.0, "cnn_f": 0.0}
    n = ns = nf = 0
    fb = []
    for x0, t, *_ in pairs:
        y, frac = net.correct(x0[None], bound=bound)
        y = y[0]
        ms = m & smooth_mask(t)
        mf = m & far_mask(t)
        for tag, mk in (("", m), ("_s", ms), ("_f", mf)):
            se["fdk" + tag] += float(((x0 - t)[mk] ** 2).sum())
            se["cnn" + tag] += float(((y - t)[mk] ** 2).sum())
        n += int(m.sum())
        ns += int(ms.sum())
        nf += int(mf.sum())
        fb.append(frac)
    cnt = {"": n, "_s": ns, "_f": nf}
    r = {k: 1000.0 * math.sqrt(v / max(cnt[k[3:]], 1)) for k, v in se.items()}
    return {"rmse_fdk_hu": r["fdk"], "rmse_cnn_hu": r["cnn"], "reduction": 1 - r["cnn"] / r["fdk"],
            "artifact_rmse_fdk_hu": r["fdk_s"], "artifact_rmse_cnn_hu": r["cnn_s"],
            "artifact_reduction": 1 - r["cnn_s"] / r["fdk_s"],
            "artifact_far_rmse_fdk_hu": r["fdk_f"], "artifact_far_rmse_cnn_hu": r["cnn_f"],
            "artifact_far_reduction": 1 - r["cnn_f"] / r["fdk_f"],
            "fallback_fraction": float(np.mean(fb))}
