"""Synthetic phantoms for the measured workloads (seeded, deterministic)."""
from __future__ import annotations

import math

import numpy as np

def ellipsoid_masks(grid, seed=0, n_ell=8):
    """The seeded ellipsoids as (flattened boolean mask, activity U(0.5, 2)) pairs, OMEGA image order."""
    rng = np.random.default_rng(seed)
    xc, yc, zc = grid.centers()
    Z, Y, X = np.meshgrid(zc, yc, xc, indexing="ij")
    rx, ry, rz = grid.Nx * grid.dx / 2, grid.Ny * grid.dy / 2, grid.Nz * grid.dz / 2
    out = []
    for _ in range(n_ell):
        cx, cy, cz = rng.uniform(-0.5, 0.5) * rx, rng.uniform(-0.5, 0.5) * ry, rng.uniform(-0.5, 0.5) * rz
        ax, ay, az = rng.uniform(0.1, 0.4) * rx, rng.uniform(0.1, 0.4) * ry, rng.uniform(0.2, 0.6) * rz
        th = rng.uniform(0, math.pi)
        u = (X - cx) * math.cos(th) + (Y - cy) * math.sin(th)
        v = -(X - cx) * math.sin(th) + (Y - cy) * math.cos(th)
        inside = (u / ax) ** 2 + (v / ay) ** 2 + ((Z - cz) / az) ** 2 <= 1.0
        out.append((inside.reshape(-1), rng.uniform(0.5, 2.0)))
    return out


def ellipsoids(grid, seed=0, n_ell=8, background=0.1):
    """Seeded random ellipsoids (activity U(0.5, 2)) on a constant background, OMEGA image order."""
    img = np.full(grid.N, background, dtype=np.float64)
    for inside, act in ellipsoid_masks(grid, seed, n_ell):
        img[inside] += act
    return np.ascontiguousarray(img.astype(np.float32))


# Frame schedule of the dynamic study: 24 frames, 55 min (durations in s).
FRAME_DURATIONS = [15.0] * 4 + [30.0] * 4 + [60.0] * 4 + [120.0] * 4 + [300.0] * 8


def one_tissue_tac(t_min, K1, k2, A=10.0, lam=0.8):
    """Tissue curve of the two-compartment (plasma + tissue) model, C_T = K1 * (C_p conv exp(-k2 t)),
    with a gamma-variate input C_p(t) = A t exp(-lam t), in closed form (t in minutes)."""
    t = np.asarray(t_min, dtype=np.float64)
    a = k2 - lam
    if abs(a) < 1e-9:
        return K1 * A * np.exp(-k2 * t) * t * t / 2.0
    return K1 * A * np.exp(-k2 * t) * (np.exp(a * t) * (a * t - 1.0) + 1.0) / (a * a)


def dynamic_ellipsoids(grid, seed=0, n_ell=8, background=0.1, durations=None):
    """Frame images [N, F] (frame fastest) of the ellipsoid phantom whose ellipsoids follow seeded
    one-tissue-compartment curves (K1 ~ U(0.05, 0.5), k2 ~ U(0.02, 0.3) per min), evaluated at the frame
    mid-times and scaled by each ellipsoid's activity; the background follows K1 = 0.1, k2 = 0.1."""
    durations = FRAME_DURATIONS if durations is None else durations
    ends = np.cumsum(durations)
    mid = (ends - np.asarray(durations) / 2.0) / 60.0
    rng = np.random.default_rng(seed + 7919)
    bg = one_tissue_tac(mid, 0.1, 0.1)
    bg = background * bg / bg.max()
    out = np.repeat(bg[None, :], grid.N, axis=0)
    for inside, act in ellipsoid_masks(grid, seed, n_ell):
        tac = one_tissue_tac(mid, rng.uniform(0.05, 0.5), rng.uniform(0.02, 0.3))
        out[inside] += act * tac / tac.max()
    return np.ascontiguousarray(out.astype(np.float32))


def attenuation_ellipsoids(grid, seed=0, n_ell=12, lo=0.01, hi=0.03):
    """Attenuation phantom: seeded ellipsoids (the geometry of ``ellipsoid_masks``) with
    mu ~ U(lo, hi) per mm, later ellipsoids overwriting earlier ones, zero elsewhere."""
    rng = np.random.default_rng(seed + 15485863)
    img = np.zeros(grid.N, dtype=np.float32)
    for inside, _ in ellipsoid_masks(grid, seed, n_ell):
        img[inside] = rng.uniform(lo, hi)
    return img
