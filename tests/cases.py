"""Shared test geometries (all deterministic).

* ``tiny_grid()``: an asymmetric 8x7x5 grid with unequal voxel sizes;
* ``list_rays(grid, n, seed)``: random LORs through/around the grid plus every special branch of
  the Siddon set-up: in-plane axis-parallel LORs (both axes), axial LORs parallel to the x or y
  axis, LORs starting or ending inside the FOV, and LORs that miss;
* ``c0()``: the 2-D plane PET scanner/grid of the OSEM-Siddon workload (128 radial x 160 angles x
  63 planes, 128x128x63 image at 2.5 mm, ring radius 400 mm);
* ``c2_small()``: a reduced 3-D ring-pair PET scanner for oblique-LOR checks;
* ``ct_small()``: a small cone-beam flat-panel CT trajectory.
"""
from __future__ import annotations

import math

import numpy as np

from omega_amd import geometry as G


def tiny_grid():
    return G.Grid(8, 7, 5, 1.5, 2.0, 2.5, -6.0, -7.0, -6.25)


def list_rays(grid, n=400, seed=0):
    rng = np.random.default_rng(seed)
    c = np.array([grid.bx + grid.Nx * grid.dx / 2, grid.by + grid.Ny * grid.dy / 2, grid.bz + grid.Nz * grid.dz / 2])
    ext = np.array([grid.Nx * grid.dx, grid.Ny * grid.dy, grid.Nz * grid.dz])
    R = 1.5 * np.linalg.norm(ext)
    a = rng.normal(size=(n, 3)); a /= np.linalg.norm(a, axis=1, keepdims=True)
    b = rng.normal(size=(n, 3)); b /= np.linalg.norm(b, axis=1, keepdims=True)
    # pull endpoints towards the FOV so most rays intersect it
    s = c + R * a
    d = c + R * b * 0.6 + rng.uniform(-0.4, 0.4, size=(n, 3)) * ext
    rays = [np.concatenate([s, d], axis=1)]
    special = []
    lo = np.array([grid.bx, grid.by, grid.bz]); hi = lo + ext
    for k in range(12):
        # voxel-centre and off-centre offsets
        fx = lo[0] + (k % grid.Nx + 0.5 + 0.3 * (k % 3 - 1)) * grid.dx
        fy = lo[1] + (k % grid.Ny + 0.5 - 0.25 * (k % 2)) * grid.dy
        fz = lo[2] + (k % grid.Nz + 0.5) * grid.dz
        special.append([lo[0] - 20, fy, fz, hi[0] + 20, fy, fz])            # parallel to x, in plane
        special.append([hi[0] + 15, fy, fz, lo[0] - 25, fy, fz])            # parallel to x, reversed
        special.append([fx, lo[1] - 20, fz, fx, hi[1] + 20, fz])            # parallel to y, in plane
        special.append([fx, hi[1] + 20, fz, fx, lo[1] - 20, fz])            # parallel to y, reversed
        special.append([lo[0] - 20, fy, lo[2] - 3, hi[0] + 20, fy, hi[2] + 2])  # y constant, oblique in z
        special.append([fx, lo[1] - 20, hi[2] + 3, fx, hi[1] + 20, lo[2] - 2])  # x constant, oblique in z
        special.append([lo[0] - 9, lo[1] - 8, fz, hi[0] + 7, hi[1] + 6, fz])  # in plane, oblique
        special.append([c[0] + 0.3 * k - 1.5, c[1] - 0.2 * k, c[2], hi[0] + 20, hi[1] + 10, hi[2] + 5])  # starts inside
        special.append([lo[0] - 20, lo[1] - 10, lo[2] - 5, c[0] + 0.1 * k, c[1], c[2] - 0.2 * k])     # ends inside
        special.append([lo[0] - 30, hi[1] + 30, fz, lo[0] - 10, hi[1] + 50, fz])  # misses
        special.append([lo[0] + 0.5 * grid.dx, fy, fz, lo[0] + 0.5 * grid.dx, fy, fz])  # degenerate (zero length)
    rays.append(np.asarray(special))
    return np.ascontiguousarray(np.concatenate(rays).astype(np.float32).reshape(-1))


def c0():
    grid = G.Grid.from_fov(128, 128, 63, 320.0, 320.0, 157.5)
    sc = G.pet_2d_planes(grid, radius=400.0, n_radial=128, n_angles=160, radial_pitch=2.5)
    return grid, sc


def c2_small(n_rings=6, crystals=96, max_ring_diff=5):
    grid = G.Grid.from_fov(48, 48, 2 * n_rings - 1, 200.0, 200.0, 4.0 * n_rings)
    sc = G.pet_3d_rings(radius=150.0, n_rings=n_rings, ring_pitch=4.0, crystals=crystals, max_ring_diff=max_ring_diff)
    return grid, sc


def ct_small(n_views=24, rows=40, cols=24):
    grid = G.Grid.from_fov(32, 32, 24, 32.0, 32.0, 24.0)
    x, uv = G.cbct_tables(n_views, sod=120.0, sdd=200.0, pitch_u=1.6)
    return grid, x, uv, rows, cols, 1.6


def square_grid():
    """8x8x5 grid with unequal voxel sizes (the orthogonal projector needs Nx == Ny)."""
    return G.Grid(8, 8, 5, 1.5, 2.0, 2.5, -6.0, -8.0, -6.25)


def orth_pet2d():
    """Small 2-D plane PET case for the orthogonal projector: 32x32x9 image of 2.5 mm voxels,
    32 radial bins x 40 angles at 2.5 mm radial pitch on a 100 mm ring, one plane per slice."""
    grid = G.Grid.from_fov(32, 32, 9, 80.0, 80.0, 22.5)
    sc = G.pet_2d_planes(grid, radius=100.0, n_radial=32, n_angles=40, radial_pitch=2.5)
    return grid, sc


def tube_brute_force(grid, s, d, width, plane=None):
    """Every voxel whose centre lies closer than 0.99*width to the infinite line through s and d,
    with weight 1 - dist/width, in float64 (optionally only slice `plane`). Returns {index: (w, dist/width)}."""
    Nx, Ny, Nz = grid.Nx, grid.Ny, grid.Nz
    xc = grid.bx + (np.arange(Nx) + 0.5) * grid.dx
    yc = grid.by + (np.arange(Ny) + 0.5) * grid.dy
    zc = grid.bz + (np.arange(Nz) + 0.5) * grid.dz
    Z, Y, X = np.meshgrid(zc, yc, xc, indexing="ij")
    P = np.stack([X, Y, Z], -1).reshape(-1, 3)
    s = np.asarray(s, np.float64)
    v = np.asarray(d, np.float64) - s
    dist = np.linalg.norm(np.cross(v[None, :], P - s[None, :]), axis=1) / np.linalg.norm(v) / width
    keep = dist < 0.99 + 1e-4
    if plane is not None:
        keep &= (np.arange(P.shape[0]) // (Nx * Ny)) == plane
    ind = np.nonzero(keep)[0]
    return {int(i): (1.0 - dist[i], dist[i]) for i in ind}
