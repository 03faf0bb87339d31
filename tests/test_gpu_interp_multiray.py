"""GPU parity of the interpolation projector (projector type 4) with several rays per LOR against the oracle
(its multi-ray semantics pinned in test_oracle_interp_multiray.py).

Forward projections without TOF bit-identical to the oracle (the rays' sums added in ray order, as there);
backprojections within 1e-5 relative L2 (unordered float atomics); an attenuation image gives per-ray factors
within 1e-6 of the oracle's and projections within 1e-6 relative per element; with TOF, forward projections
within 2e-6 relative L2 and backprojections within 1e-5.
"""
import numpy as np
import pytest

from omega_amd import _lib
from oracle import oracle as O
from tests import cases
from tests.hip_helpers import bp, fp, make_geom, rel_l2
from tests.test_oracle_interp import _outside_rays

pytestmark = pytest.mark.gpu

RAYS = (3, 2, 2.4, 3.2)


def _rays_x(grid):
    rays = np.concatenate([_outside_rays(grid, 300, seed=7), cases.list_rays(grid, 150, seed=8).reshape(-1, 6)])
    rays = rays[np.linalg.norm(rays[:, 3:] - rays[:, :3], axis=1) > 0]
    return np.ascontiguousarray(rays.ravel()), len(rays)


@pytest.mark.parametrize("dl", [0.5, -1.0])
def test_multiray_fp_bitexact_bp(cuda, dl):
    import torch
    grid = cases.tiny_grid()
    x, n = _rays_x(grid)
    print(f"[interp multiray] dL={dl} n={n} rays={RAYS[:2]}", flush=True)
    case = O.InterpCase(grid, 1, 1, x, np.zeros(2, np.float32), dl, mode=O.MODE_LIST, rays=RAYS)
    geom = make_geom(grid, x, None, 1, 1, mode=_lib.MODE_LIST6, projector_type=4, interp=(np.float32(dl), 0.0),
                     rays=RAYS)
    rng = np.random.default_rng(3)
    img = (rng.random(grid.N) + 0.05).astype(np.float32)
    y = fp(geom, 0, n, torch.from_numpy(img).to(cuda)).cpu().numpy()
    yo = case.fp(img, n)
    assert np.array_equal(y, yo), f"max |diff| {np.abs(y - yo).max()}"
    yin = (rng.random(n) + 0.05).astype(np.float32)
    rhs, sens = bp(geom, 0, n, torch.from_numpy(yin).to(cuda), grid.N, sens=True)
    ro, so = case.bp(yin, n, sens=True)
    assert rel_l2(rhs.cpu().numpy(), ro) < 1e-5 and rel_l2(sens.cpu().numpy(), so) < 1e-5


def test_multiray_attenuation_image(cuda):
    import torch
    grid = cases.tiny_grid()
    x, n = _rays_x(grid)
    nr = RAYS[0] * RAYS[1]
    rng = np.random.default_rng(4)
    mu = (rng.random(grid.N) * 0.02).astype(np.float32)
    fac = O.InterpCase(grid, 1, 1, x, np.zeros(2, np.float32), 0.5, mode=O.MODE_LIST, rays=RAYS).atten_factors(mu, n)
    assert fac.size == nr * n
    case = O.InterpCase(grid, 1, 1, x, np.zeros(2, np.float32), 0.5, mode=O.MODE_LIST, rays=RAYS, atten=fac,
                        atten_image=True)
    geom = make_geom(grid, x, None, 1, 1, mode=_lib.MODE_LIST6, projector_type=4, interp=(np.float32(0.5), 0.0),
                     rays=RAYS, atten=mu, atten_mode=1)
    print(f"[interp multiray atten] n={n}", flush=True)
    img = (rng.random(grid.N) + 0.05).astype(np.float32)
    y = fp(geom, 0, n, torch.from_numpy(img).to(cuda)).cpu().numpy()
    yo = case.fp(img, n)
    assert np.all(np.abs(y - yo) <= 1e-6 * np.abs(yo) + 1e-12)
    yin = (rng.random(n) + 0.05).astype(np.float32)
    rhs = bp(geom, 0, n, torch.from_numpy(yin).to(cuda), grid.N)
    assert rel_l2(rhs.cpu().numpy(), case.bp(yin, n)) < 1e-5


def test_multiray_tof(cuda):
    import torch
    grid = cases.tiny_grid()
    x, n = _rays_x(grid)
    tof = (7, 3.0, np.array([0.0, 4.0, -4.0, 8.0, -8.0, 12.0, -12.0], np.float32), 1e-9)
    case = O.InterpCase(grid, 1, 1, x, np.zeros(2, np.float32), -1.0, mode=O.MODE_LIST, rays=RAYS)
    geom = make_geom(grid, x, None, 1, 1, mode=_lib.MODE_LIST6, projector_type=4, interp=(np.float32(-1.0), 0.0),
                     rays=RAYS, tof=tof)
    print(f"[interp multiray tof] n={n} bins={tof[0]}", flush=True)
    rng = np.random.default_rng(5)
    img = (rng.random(grid.N) + 0.05).astype(np.float32)
    y = fp(geom, 0, n, torch.from_numpy(img).to(cuda)).cpu().numpy()
    yo = case.tof_fp(img, n, tof)
    assert rel_l2(y, yo) <= 2e-6 and np.abs(y - yo).max() <= 1e-5 * np.abs(yo).max()
    yin = (rng.random(n * tof[0]) + 0.05).astype(np.float32)
    rhs, sens = bp(geom, 0, n, torch.from_numpy(yin).to(cuda), grid.N, sens=True)
    ro, so = case.tof_bp(yin, n, tof, sens=True)
    assert rel_l2(rhs.cpu().numpy(), ro) < 1e-5 and rel_l2(sens.cpu().numpy(), so) < 1e-5
