"""Synthetic phantoms for the measured workloads (seeded, deterministic)."""
from __future__ import annotations

import math

import numpy as np

def ellipsoids(grid, seed=0, n_ell=8, background=0.1):
    """Seeded random ellipsoids (activity U(0.5, 2)) on a constant background, OMEGA image order."""
    rng = np.random.default_rng(seed)
    xc, yc, zc = grid.centers()
    Z, Y, X = np.meshgrid(zc, yc, xc, indexing="ij")
    img = np.full(X.shape, background, dtype=np.float64)
    rx, ry, rz = grid.Nx * grid.dx / 2, grid.Ny * grid.dy / 2, grid.Nz * grid.dz / 2
    for _ in range(n_ell):
        cx, cy, cz = rng.uniform(-0.5, 0.5) * rx, rng.uniform(-0.5, 0.5) * ry, rng.uniform(-0.5, 0.5) * rz
        ax, ay, az = rng.uniform(0.1, 0.4) * rx, rng.uniform(0.1, 0.4) * ry, rng.uniform(0.2, 0.6) * rz
        th = rng.uniform(0, math.pi)
        u = (X - cx) * math.cos(th) + (Y - cy) * math.sin(th)
        v = -(X - cx) * math.sin(th) + (Y - cy) * math.cos(th)
        inside = (u / ax) ** 2 + (v / ay) ** 2 + ((Z - cz) / az) ** 2 <= 1.0
        img[inside] += rng.uniform(0.5, 2.0)
    return np.ascontiguousarray(img.astype(np.float32).reshape(-1))
