"""Synthetic phantoms and their exact helical projections.

The reference ships no data, so every workload is an analytic ellipsoid
phantom projected by ``arecon_project_ellipsoids_dev`` (libarecon.so) onto the
post-flip rebinning input layout ``[ViewN][YL][ZL]``.

Attenuation is mu = MU_WATER * (1 + HU/1000) with MU_WATER = 0.02 / mm.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import _lib
from .geometry import GEHeader, ReconSetup

MU_WATER = 0.02  # 1/mm

# 3-D Shepp-Logan (Kak & Slaney): a, b, c, x0, y0, z0, phi [deg], additive density.
_SHEPP_LOGAN_3D = [
    (0.69, 0.92, 0.90, 0.0, 0.0, 0.0, 0.0, 2.0),
    (0.6624, 0.874, 0.88, 0.0, -0.0184, 0.0, 0.0, -0.98),
    (0.41, 0.16, 0.21, -0.22, 0.0, -0.25, 108.0, -0.02),
    (0.31, 0.11, 0.22, 0.22, 0.0, -0.25, 72.0, -0.02),
    (0.21, 0.25, 0.50, 0.0, 0.35, -0.25, 0.0, 0.01),
    (0.046, 0.046, 0.046, 0.0, 0.1, -0.25, 0.0, 0.01),
    (0.046, 0.023, 0.02, -0.08, -0.65, -0.25, 0.0, 0.01),
    (0.046, 0.023, 0.02, 0.06, -0.65, -0.25, 90.0, 0.01),
    (0.056, 0.04, 0.1, 0.06, -0.105, 0.625, 90.0, 0.02),
    (0.056, 0.056, 0.1, 0.0, 0.1, 0.625, 0.0, -0.02),
]


def shepp_logan(scale_xy: float, scale_z: float, z_centre: float) -> np.ndarray:
    """10 ellipsoids as rows (cx, cy, cz, a, b, c, phi_rad, mu).  Densities are
    the standard additive values in units of MU_WATER (background 0, skull
    2.0 -> +1000 HU, brain 1.02 -> +20 HU)."""
    rows = []
    for a, b, c, x0, y0, z0, phi, d in _SHEPP_LOGAN_3D:
        rows.append((x0 * scale_xy, y0 * scale_xy, z_centre + z0 * scale_z, a * scale_xy,
                     b * scale_xy, c * scale_z, math.radians(phi), d * MU_WATER))
    return np.asarray(rows, dtype=np.float32)


def random_body(seed: int, fov_radius: float, z_centre: float, z_half: float,
                n_range=(40, 80)) -> np.ndarray:
    """Water body filling the FOV plus U[40, 80] random ellipsoids with HU
    drawn from U[-1000, 1500] (each adds its HU on top of water)."""
    rng = np.random.default_rng(seed)
    rows = [(0.0, 0.0, z_centre, fov_radius * 0.98, fov_radius * 0.80, z_half * 1.5, 0.0,
             MU_WATER)]
    n = int(rng.integers(n_range[0], n_range[1] + 1))
    for _ in range(n):
        r = fov_radius * 0.7 * math.sqrt(rng.uniform())
        ang = rng.uniform(0, 2 * math.pi)
        cx, cy = r * math.cos(ang), r * 0.8 * math.sin(ang)
        cz = z_centre + rng.uniform(-z_half, z_half)
        a, b, c = rng.uniform(5, 40), rng.uniform(5, 40), rng.uniform(5, 60)
        phi = rng.uniform(0, math.pi)
        hu = rng.uniform(-1000, 1500)
        rows.append((cx, cy, cz, a, b, c, phi, MU_WATER * hu / 1000.0))
    return np.asarray(rows, dtype=np.float32)


def project(ellipsoids: np.ndarray, setup: ReconSetup, hdr: GEHeader,
            device: str = "cuda", stream=None) -> torch.Tensor:
    """Exact projections ``[ViewN][YL][ZL]`` FP32 on the device, in the
    post-flip channel order the rebinning receives."""
    g = setup.geom
    SO, DO, YL, ZL = g.ScDet[0], g.ScDet[1], int(g.ScDet[2]), int(g.ScDet[3])
    DecAngle, DecHeigh, YLC1, ZLC1, hgeom = g.ScDet[4], g.ScDet[5], g.ScDet[6], g.ScDet[7], g.ScDet[8]
    BetaS, ViewN, N_2pi = g.Proj[0], int(g.Proj[2]), int(g.Proj[3])
    h = hgeom * DecHeigh
    out = torch.empty((ViewN, YL, ZL), dtype=torch.float32, device=device)
    E = np.ascontiguousarray(ellipsoids, dtype=np.float32)
    s = stream if stream is not None else torch.cuda.current_stream()
    _lib.check(_lib.lib().arecon_project_ellipsoids_dev(
        C.c_void_p(out.data_ptr()), E.ctypes.data_as(C.POINTER(C.c_float)), E.shape[0], YL, ZL,
        ViewN, BetaS, 2 * math.pi / N_2pi, SO, SO + DO, h, YLC1 - 1.0, DecAngle / YL,
        ZLC1 - 1.0, DecHeigh / ZL, C.c_void_p(s.cuda_stream)), "arecon_project_ellipsoids_dev")
    return out


def voxelize(ellipsoids: np.ndarray, xs: np.ndarray, ys: np.ndarray, zs: np.ndarray) -> np.ndarray:
    """Phantom values on a grid, returned as (len(ys), len(xs), len(zs)) --
    the volume layout [YN][XN][ZN]."""
    Y, X, Z = np.meshgrid(ys, xs, zs, indexing="ij")
    out = np.zeros(Y.shape, dtype=np.float64)
    for cx, cy, cz, a, b, c, phi, mu in ellipsoids.astype(np.float64):
        cp, sp = math.cos(phi), math.sin(phi)
        u = (cp * (X - cx) + sp * (Y - cy)) / a
        v = (-sp * (X - cx) + cp * (Y - cy)) / b
        w = (Z - cz) / c
        out += np.where(u * u + v * v + w * w <= 1.0, mu, 0.0)
    return out
