"""Pins the oracle's interpolation projector with several rays per LOR (projectorType4.cl -DN_RAYS, :336-400,
:612-617, :633-660).

Each ray is the LOR shifted as multirayCoordinateShiftZ then XY (projector_functions.h:439-461): the shifts
are computed here in float64, independently of the oracle, and the shifted rays projected one at a time with
the single-ray oracle. With the LOR weight 1 / (L nRays) per ray:
* the multi-ray forward projection equals the sum over rays of the single-ray projections / nRays;
* the multi-ray backprojection of y equals the single-ray backprojection of y / nRays over the shifted rays;
* image attenuation factors are per ray (ray r's at r * n + lo) and equal the shifted rays' single factors;
* with TOF, the bins of the multi-ray forward projection add up to the non-TOF multi-ray projection.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import cases
from tests.test_oracle_volume_tof import _rays

RAYS = (3, 2, 2.4, 3.2)   # transaxial, axial, pitch_xy, pitch_z


def _shifted(x, rays):
    """Every LOR's rays in ray order (axial index outer), as list-mode rays, in float64 then float32."""
    nxy, nz, cxy, cz = rays
    r = x.reshape(-1, 6).astype(np.float64)
    out = []
    for lo in range(r.shape[0]):
        for kz in range(nz):
            for kxy in range(nxy):
                q = r[lo].copy()
                if nz > 1:
                    iz = cz / (2 * nz)
                    q[[2, 5]] += (iz - cz / 2) + 2 * iz * kz
                if nxy > 1:
                    ixy = cxy / (2 * nxy)
                    q[[0, 1, 3, 4]] += (ixy - cxy / 2) + 2 * ixy * kxy
                out.append(q)
    return np.ascontiguousarray(np.asarray(out, np.float32).ravel())


@pytest.mark.parametrize("dL", [0.0, 0.6])
def test_multiray_equals_shifted_single_rays(dL):
    grid = cases.square_grid()
    x = _rays(grid, 60, 11)
    n = x.size // 6
    nr = RAYS[0] * RAYS[1]
    multi = O.InterpCase(grid, 1, 1, x, np.zeros(2, np.float32), dL, mode=O.MODE_LIST, rays=RAYS)
    xs = _shifted(x, RAYS)
    single = O.InterpCase(grid, 1, 1, xs, np.zeros(2, np.float32), dL, mode=O.MODE_LIST)
    rng = np.random.default_rng(4)
    img = rng.random(grid.N).astype(np.float32)
    y = multi.fp(img, n).astype(np.float64)
    ys = single.fp(img, n * nr).astype(np.float64).reshape(n, nr).sum(1) / nr
    assert np.count_nonzero(y) > 40
    np.testing.assert_allclose(y, ys, rtol=2e-5, atol=1e-7 * y.max())
    r = rng.random(n).astype(np.float32)
    bp, sens = multi.bp(r, n, sens=True)
    bps, senss = single.bp(np.repeat(r / np.float32(nr), nr), n * nr, sens=True)
    np.testing.assert_allclose(bp, bps, rtol=2e-5, atol=1e-6 * bp.max())
    np.testing.assert_allclose(sens, senss / nr, rtol=2e-5, atol=1e-6 * sens.max())
    mu = (rng.random(grid.N) * 0.02).astype(np.float32)
    f = multi.atten_factors(mu, n).reshape(nr, n)
    fs = single.atten_factors(mu, n * nr).reshape(n, nr).T
    np.testing.assert_allclose(f, fs, rtol=2e-6)


def test_multiray_with_tof_bins_add_up():
    grid = cases.square_grid()
    x = _rays(grid, 60, 12)
    n = x.size // 6
    tof = (9, 5.0, np.array([0.0, 12.5, -12.5, 25.0, -25.0, 37.5, -37.5, 50.0, -50.0], np.float32))
    multi = O.InterpCase(grid, 1, 1, x, np.zeros(2, np.float32), 0.0, mode=O.MODE_LIST, rays=RAYS)
    img = np.random.default_rng(5).random(grid.N).astype(np.float32)
    y = multi.fp(img, n).astype(np.float64)
    yt = multi.tof_fp(img, n, tof).astype(np.float64).reshape(9, n)
    np.testing.assert_allclose(yt.sum(0), y, rtol=2e-5, atol=1e-7 * y.max())
    r = np.random.default_rng(6).random(n).astype(np.float32)
    np.testing.assert_allclose(multi.tof_bp(np.tile(r, 9), n, tof), multi.bp(r, n), rtol=2e-5,
                               atol=1e-6 * multi.bp(r, n).max())
