"""Training data, training and evaluation of the learned correction.

Pairs are made by the method itself: a seeded random-ellipsoid phantom, its
exact helical projections on the 368x16 curved detector (720 views/rotation,
pitch drawn from [0.5, 1.5]), reconstructed by the plain GPU pipeline into a
192x192x64 volume of 1 mm voxels; the target is the phantom sampled on the
same grid (2x2x2 supersampled).  Values are HU / 1000.  The network
(``cnn.Net``) learns the residual target - FDK on 64^3 patches; the plain FDK
is both its input and the yardstick of its error.
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


def make_pair(seed: int, spec: PairSpec = PairSpec(), pitch: float | None = None):
    """(plain FDK, phantom) in HU / 1000, FP32 [nx][nx][nz] on the device."""
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
          seed: int = 0, log=None, log_every: int = 100):
    """Adam with a cosine learning-rate decay; returns the loss history."""
    gen = torch.Generator().manual_seed(seed)
    net.pack()
    hist = []
    t0 = time.perf_counter()
    for k in range(steps):
        x, t = sample_batch(pairs, batch, patch, gen)
        lr_k = lr * 0.5 * (1 + math.cos(math.pi * k / max(1, steps)))
        loss = net.train_step(x, t, lr=max(lr_k, lr * 0.02))
        if k % 10 == 0 or k == steps - 1:
            hist.append((k, float(loss)))
        if log and (k % log_every == 0 or k == steps - 1):
            log(f"step {k}: loss {hist[-1][1]:.5f} ({time.perf_counter() - t0:.1f} s)")
    return hist


def evaluate(net: Net, pairs, spec: PairSpec, bound: float = 0.3):
    """RMSE in HU against the phantom inside the support, for the plain FDK
    and the corrected volume (with the fallback), summed over the pairs."""
    m = fov_mask(spec, pairs[0][0].device)
    se_fdk = se_cnn = 0.0
    n = 0
    fb = []
    for x0, t, *_ in pairs:
        y, frac = net.correct(x0[None], bound=bound)
        y = y[0]
        se_fdk += float(((x0 - t)[m] ** 2).sum())
        se_cnn += float(((y - t)[m] ** 2).sum())
        n += int(m.sum())
        fb.append(frac)
    r_fdk = 1000.0 * math.sqrt(se_fdk / n)
    r_cnn = 1000.0 * math.sqrt(se_cnn / n)
    return {"rmse_fdk_hu": r_fdk, "rmse_cnn_hu": r_cnn, "reduction": 1 - r_cnn / r_fdk,
            "fallback_fraction": float(np.mean(fb))}
