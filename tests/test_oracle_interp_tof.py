"""Pins the oracle's time-of-flight interpolation projector (projector type 4 with TOF, the reference GPU
kernel's -DTOF semantics: projectorType4.cl:402-425 TOFDis at the FOV entry, :463-466 the Joseph shift,
:566-600 TOFLoop per sample with the weight-sum floor 1e-6, denominator / rhs).

Checked against what the geometry and the Gaussian fix, not against the restated loop:
* the bins of the TOF forward projection add up to the non-TOF forward projection (the per-sample weights
  are normalised), for the Joseph variant and a fixed interpolation length;
* the bins of the TOF backprojection of data equal in every bin add up to the non-TOF backprojection of
  that data, and BP of ones gives the sensitivity;
* Joseph rays deposit one sample per voxel of the dominant axis; backprojecting one bin at a time, each
  such voxel's share of the bins is the float64 trapezoid of the Gaussians over its sample's stretch
  [D_i, D_i - stepLen s], D_i = L (t_i - 1/2) for samples at t_i = t0 + i tStep on rays whose detectors lie
  outside the FOV (the entry before the midpoint, so s = -1).
(The type-4 PET backprojection deposits each sample into its nearest voxel, the forward projection samples
trilinearly, so the pair is not adjoint: no adjoint test.)
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests import cases
from tests.test_oracle_volume_tof import _rays

SIGMA = 5.0
CENTRES = np.array([0.0, 12.5, -12.5, 25.0, -25.0, 37.5, -37.5, 50.0, -50.0], np.float32)
BINS = CENTRES.size
TOF = (BINS, SIGMA, CENTRES)


def _case(grid, x, dL):
    return O.InterpCase(grid, 1, 1, x, np.zeros(2, np.float32), dL, mode=O.MODE_LIST)


def _finite_rays(x):
    r = x.reshape(-1, 6)
    return np.linalg.norm(r[:, 3:] - r[:, :3], axis=1) > 1e-3


@pytest.mark.parametrize("dL", [0.0, 0.7])
def test_bins_add_up_to_the_non_tof_projections(dL):
    grid = cases.square_grid()
    x = np.concatenate([_rays(grid, 120, 1), cases.list_rays(grid, 200, seed=3)])
    n = x.size // 6
    ok = _finite_rays(x)
    case = _case(grid, x, dL)
    rng = np.random.default_rng(2)
    img = rng.random(grid.N).astype(np.float32)
    y = case.fp(img, n).astype(np.float64)
    yt = case.tof_fp(img, n, TOF).astype(np.float64).reshape(BINS, n)
    assert np.count_nonzero(y[ok]) > 150
    np.testing.assert_allclose(yt.sum(0)[ok], y[ok], rtol=2e-5, atol=1e-7 * y[ok].max())
    r = rng.random(n).astype(np.float32)
    r[~ok] = 0
    bp = case.bp(r, n).astype(np.float64)
    bpt, sens = case.tof_bp(np.tile(r, BINS) / np.float32(1.0), n, TOF, sens=True)
    np.testing.assert_allclose(bpt, bp, rtol=2e-5, atol=1e-6 * bp.max())
    ones = np.ones(n, np.float32)
    ones[~ok] = 0
    np.testing.assert_allclose(sens, case.bp(ones, n), rtol=2e-5, atol=1e-6 * sens.max())


def test_joseph_sample_distribution():
    grid = cases.square_grid()
    x = _rays(grid, 40, 9)
    n = x.size // 6
    rays = x.reshape(-1, 6).astype(np.float64)
    lo = np.array([grid.bx, grid.by, grid.bz])
    N = np.array([grid.Nx, grid.Ny, grid.Nz])
    dv = np.array([grid.dx, grid.dy, grid.dz])
    checked = 0
    for i in range(n):
        # one ray at a time: its own per-bin backprojections
        xi = np.ascontiguousarray(x[6 * i:6 * i + 6])
        ci = _case(grid, xi, 0.0)
        s, d = rays[i, :3], rays[i, 3:]
        v = d - s
        L = float(np.linalg.norm(v.astype(np.float32).astype(np.float64)))
        tb = (lo - s) / v
        tf = (lo + N * dv - s) / v
        t_start = np.max(np.minimum(tb, tf))
        t_end = np.min(np.maximum(tb, tf))
        cr = np.abs(v) / (N * dv) * N
        ax = int(np.argmax(cr))
        t_step = 1.0 / cr[ax]
        s_ind = (s[ax] + v[ax] * t_start - lo[ax]) / (N[ax] * dv[ax]) * N[ax] - 0.5
        t0 = t_start + ((math.floor(s_ind) - s_ind) if v[ax] > 0 else (s_ind - math.floor(s_ind))) * t_step
        step_len = t_step * L
        D0 = float(np.linalg.norm(v * t_start)) - L / 2
        if D0 >= 0:
            continue
        bins = [ci.tof_bp(np.eye(BINS, dtype=np.float32)[b], 1, TOF).astype(np.float64) for b in range(BINS)]
        tot = sum(bins)
        k = 0
        while t0 + k * t_step <= t_end:
            t = t0 + k * t_step
            p = s + v * t
            vox = np.floor((p - lo) / dv).astype(int)
            k += 1
            if np.any(vox <= 0) or np.any(vox >= N - 1):
                continue   # image edges: a skipped sample would shift the later ones' distance
            ind = vox[0] + vox[1] * grid.Nx + vox[2] * grid.Nx * grid.Ny
            if tot[ind] <= 0:
                continue
            D = L * (t - 0.5)
            dX = step_len / 3
            want = np.array([(_p(D, c) + 2 * _p(D + dX, c) + 2 * _p(D + 2 * dX, c) + _p(D + step_len, c)) * dX
                             for c in CENTRES.astype(np.float64)])
            got = np.array([bins[b][ind] for b in range(BINS)])
            np.testing.assert_allclose(got / got.sum(), want / want.sum(), atol=3e-4, err_msg=f"ray {i} voxel {ind}")
            checked += 1
    assert checked > 100


def _p(xv, mu):
    return math.exp(-0.5 * ((xv - mu) / SIGMA) ** 2)
