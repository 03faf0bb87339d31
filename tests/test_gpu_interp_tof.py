"""GPU parity of the time-of-flight interpolation projector (projector type 4 with TOF) against the oracle
(its TOF semantics pinned in test_oracle_interp_tof.py).

The bin weights take exponentials (the device's expf against the host's), so the bars are relative: forward
projections within 2e-6 relative L2 and 1e-5 of the largest value per element, backprojection and sensitivity
(unordered float atomics) within 1e-5 relative L2, the fused data step (FP -> ratio -> BP through the work
buffer) within 1e-5 of the oracle's; projectorClass with projector_type 4 and TOF data projects as the
geometry does.
"""
import numpy as np
import pytest

from omega_amd import _lib
from oracle import oracle as O
from tests import cases
from tests.hip_helpers import bp, fp, make_geom, ptr, rel_l2
from tests.test_oracle_interp import _outside_rays

pytestmark = pytest.mark.gpu

EPPS = 1e-9


def _tof(bins, step, sigma):
    c = [0.0]
    for k in range(1, bins // 2 + 1):
        c += [k * step, -k * step]
    return bins, float(sigma), np.asarray(c, np.float32), EPPS


def _close_fp(y, y_ref):
    assert rel_l2(y, y_ref) <= 2e-6, rel_l2(y, y_ref)
    assert np.abs(y - y_ref).max() <= 1e-5 * np.abs(y_ref).max()


@pytest.mark.parametrize("dl", [0.4, -1.0])
def test_listmode_fp_bp(cuda, dl):
    import torch
    grid = cases.tiny_grid()
    rays = np.concatenate([_outside_rays(grid, 300, seed=1), cases.list_rays(grid, 200, seed=2).reshape(-1, 6)])
    rays = rays[np.linalg.norm(rays[:, 3:] - rays[:, :3], axis=1) > 0]
    x = np.ascontiguousarray(rays.ravel())
    n = len(rays)
    tof = _tof(7, 4.0, 3.0)
    print(f"[interp tof list] dL={dl} n={n} bins={tof[0]}", flush=True)
    case = O.InterpCase(grid, 1, 1, x, np.zeros(2, np.float32), dl, mode=O.MODE_LIST)
    geom = make_geom(grid, x, None, 1, 1, mode=_lib.MODE_LIST6, projector_type=4, interp=(np.float32(dl), 0.0),
                     tof=tof)
    rng = np.random.default_rng(3)
    img = (rng.random(grid.N) + 0.05).astype(np.float32)
    _close_fp(fp(geom, 0, n, torch.from_numpy(img).to(cuda)).cpu().numpy(), case.tof_fp(img, n, tof))
    yin = (rng.random(n * tof[0]) + 0.05).astype(np.float32)
    rhs, sens = bp(geom, 0, n, torch.from_numpy(yin).to(cuda), grid.N, sens=True)
    ro, so = case.tof_bp(yin, n, tof, sens=True)
    assert rel_l2(rhs.cpu().numpy(), ro) < 1e-5 and rel_l2(sens.cpu().numpy(), so) < 1e-5


@pytest.mark.parametrize("randoms", [False, True])
def test_sinogram_fused_data_step(cuda, randoms):
    import torch
    grid, sc = cases.c2_small()
    x, z = sc.xy_table(), sc.z
    n = sc.n_radial * sc.n_angles * (z.size // 2)
    rng = np.random.default_rng(5)
    norm = (rng.random(n) + 0.5).astype(np.float32)
    tof = _tof(9, 15.0, 12.0)
    case = O.InterpCase(grid, sc.n_radial, sc.n_angles, x, z, 0.8, norm=norm, global_factor=1.3)
    geom = make_geom(grid, x, z, sc.n_radial, sc.n_angles, projector_type=4, interp=(np.float32(0.8), 0.0),
                     norm=norm, global_factor=1.3, tof=tof)
    nb = tof[0]
    print(f"[interp tof sinogram] n={n} bins={nb} randoms={randoms}", flush=True)
    img = (rng.random(grid.N) * 2 + 0.2).astype(np.float32)
    _close_fp(fp(geom, 0, n, torch.from_numpy(img).to(cuda)).cpu().numpy(), case.tof_fp(img, n, tof))
    y = rng.poisson(case.tof_fp(img, n, tof) * 30.0).astype(np.float32)
    add = (rng.random(n) * 0.5 + 0.05).astype(np.float32) if randoms else None
    xs = (rng.random(grid.N) + 0.5).astype(np.float32)
    xt, yt = torch.from_numpy(xs).to(cuda), torch.from_numpy(y).to(cuda)
    rhs = torch.zeros(grid.N, device=cuda)
    sens = torch.zeros(grid.N, device=cuda)
    eps = 1e-5
    if add is None:
        _lib.call("omg_osem_data_step", geom.handle, 0, n, ptr(xt), ptr(yt), ptr(rhs), ptr(sens), eps, None)
    else:
        at = torch.from_numpy(add).to(cuda)
        _lib.call("omg_osem_data_step_ex", geom.handle, 0, n, ptr(xt), ptr(yt), ptr(at), ptr(rhs), ptr(sens), eps,
                  None)
    yhat = case.tof_fp(xs, n, tof).reshape(nb, n).astype(np.float32)
    den = (yhat + np.float32(eps)) + (add if add is not None else np.float32(0))
    ratio = (y.reshape(nb, n) / den).astype(np.float32)
    ratio[~np.isfinite(ratio)] = np.float32(eps)
    ro, so = case.tof_bp(ratio.ravel(), n, tof, sens=True)
    assert rel_l2(rhs.cpu().numpy(), ro) <= 1e-5
    assert rel_l2(sens.cpu().numpy(), so) <= 1e-5


def test_projector_class_interp_tof(cuda):
    import torch
    from omega_amd.projector import projectorClass
    grid, sc = cases.orth_pet2d()
    p = projectorClass()
    p.Nx, p.Ny, p.Nz = grid.Nx, grid.Ny, grid.Nz
    p.FOVa_x, p.FOVa_y, p.axial_fov = grid.Nx * grid.dx, grid.Ny * grid.dy, grid.Nz * grid.dz
    p.projector_type = 4
    p.x, p.z = sc.xy_table(), sc.z
    p.Ndist, p.Nang = sc.n_radial, sc.n_angles
    p.subsets, p.subsetType = 3, 8
    p.TOF_bins, p.TOF_bins_used, p.TOF_width, p.TOF_FWHM = 7, 7, 100e-12, 300e-12
    p.addProjector()
    p.initProj()
    assert p.nBins == 7
    dl = float(p._interp_scalars()[0])
    geom = make_geom(p.grid, p.x, p.z, sc.n_radial, sc.n_angles, projector_type=4, interp=(np.float32(dl), 0.0),
                     tof=(7, float(p.sigma_x), p.TOFCenter, float(p.epps)))
    print(f"[interp tof projectorClass] subsets={p.subsets} dL={dl}", flush=True)
    img = torch.rand(grid.N, device=cuda) + 0.1
    for k in range(p.subsets):
        y = p.forwardProject(img, k)
        n = y.numel() // 7
        assert float(y.sum()) > 0
        assert np.array_equal(y.cpu().numpy(), fp(geom, int(p.nMeas[k]), n, img).cpu().numpy())
        b = p.backwardProject(y, k)
        assert rel_l2(b.cpu().numpy(), bp(geom, int(p.nMeas[k]), n, y, grid.N).cpu().numpy()) < 1e-6
