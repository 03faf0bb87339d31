"""GPU parity of the time-of-flight volume-of-intersection projector (projector type 3 with TOF) against the
oracle (its TOF semantics pinned in test_oracle_volume_tof.py).

The bin weights take exponentials (the device's expf against the host's), so the bars are relative: forward
projections within 2e-6 relative L2 and 1e-5 of the largest value per element, backprojection and sensitivity
(unordered float atomics) within 1e-5 relative L2, the fused data step within 1e-5 of the oracle's
FP -> ratio -> BP; projectorClass with projector_type 3 and TOF data projects as the geometry does.
"""
import numpy as np
import pytest

from omega_amd import _lib
from omega_amd import geometry as G
from oracle import oracle as O
from tests import cases
from tests.hip_helpers import bp, fp, make_geom, ptr, rel_l2
from tests.test_oracle_volume_tof import _rays

pytestmark = pytest.mark.gpu

TUBE_R, VOXEL_R = 2.0, 1.0
EPPS = 1e-9


def _volume(grid, tr=TUBE_R):
    V, bmin, bmax, vmax, _ = G.voxel_volume_table(tr, VOXEL_R, grid.dx, grid.dy, grid.dz)
    return V, np.float32(bmin), np.float32(bmax), np.float32(vmax)


def _tof(bins, step, sigma):
    c = [0.0]
    for k in range(1, bins // 2 + 1):
        c += [k * step, -k * step]
    return bins, float(sigma), np.asarray(c, np.float32), EPPS


def _close_fp(y, y_ref):
    assert rel_l2(y, y_ref) <= 2e-6, rel_l2(y, y_ref)
    assert np.abs(y - y_ref).max() <= 1e-5 * np.abs(y_ref).max()


def _pet2d(flags=3):
    grid, sc = cases.orth_pet2d()
    vol = _volume(grid)
    n = sc.n_radial * sc.n_angles * (sc.z.size // 2)
    norm = np.random.default_rng(3).uniform(0.5, 1.5, n).astype(np.float32)
    tof = _tof(7, 20.0, 15.0)
    case = O.OrthCase(grid, sc.n_radial, sc.n_angles, sc.xy_table(), sc.z, TUBE_R, flags=flags, volume=vol,
                      norm=norm, global_factor=0.8)
    geom = make_geom(grid, sc.xy_table(), sc.z, sc.n_radial, sc.n_angles, projector_type=3, orth_width=TUBE_R,
                     orth_flags=flags, volume=vol, norm=norm, global_factor=0.8, tof=tof)
    return grid, n, tof, case, geom


@pytest.mark.parametrize("flags", [1, 3])
def test_pet2d_fp_bp(cuda, flags):
    import torch
    grid, n, tof, case, geom = _pet2d(flags)
    nb = tof[0]
    print(f"[volume tof pet2d] flags={flags} n={n} bins={nb}", flush=True)
    rng = np.random.default_rng(6)
    img = rng.random(grid.N).astype(np.float32) + 0.05
    y = fp(geom, 0, n, torch.from_numpy(img).to(cuda)).cpu().numpy()
    y_ref = case.tof_fp(img, n, tof)
    _close_fp(y, y_ref)
    r = rng.random(n * nb).astype(np.float32)
    rhs, s = bp(geom, 0, n, torch.from_numpy(r).to(cuda), grid.N, sens=True)
    rhs_o, s_o = case.tof_bp(r, n, tof, sens=True)
    assert rel_l2(rhs.cpu().numpy(), rhs_o) <= 1e-5
    assert rel_l2(s.cpu().numpy(), s_o) <= 1e-5
    # one sinogram plane from a plane offset: the same values as the full call's slice
    plane, m = 2, 32 * 40
    ys = fp(geom, plane, m, torch.from_numpy(img).to(cuda)).cpu().numpy().reshape(nb, m)
    np.testing.assert_array_equal(ys, y.reshape(nb, n)[:, plane * m:(plane + 1) * m])


def test_list_mode_all_directions(cuda):
    import torch
    grid = cases.square_grid()
    x = np.concatenate([_rays(grid, 200, 17), cases.list_rays(grid, 300, seed=12)])
    n = x.size // 6
    vol = _volume(grid, 1.5)
    tof = _tof(9, 6.0, 5.0)
    case = O.OrthCase(grid, 1, 1, x, np.zeros(2, np.float32), 1.5, flags=3, mode=O.MODE_LIST, volume=vol)
    geom = make_geom(grid, x, None, 1, 1, mode=_lib.MODE_LIST6, projector_type=3, orth_width=1.5, orth_flags=3,
                     volume=vol, tof=tof)
    print(f"[volume tof list-mode] n={n} bins={tof[0]}", flush=True)
    img = np.random.default_rng(13).random(grid.N).astype(np.float32)
    _close_fp(fp(geom, 0, n, torch.from_numpy(img).to(cuda)).cpu().numpy(), case.tof_fp(img, n, tof))
    r = np.random.default_rng(14).random(n * tof[0]).astype(np.float32)
    rhs = bp(geom, 0, n, torch.from_numpy(r).to(cuda), grid.N)
    assert rel_l2(rhs.cpu().numpy(), case.tof_bp(r, n, tof)) <= 1e-5


@pytest.mark.parametrize("randoms", [False, True])
def test_fused_data_step(cuda, randoms):
    import torch
    grid, n, tof, case, geom = _pet2d()
    nb = tof[0]
    rng = np.random.default_rng(21)
    img = (rng.random(grid.N) * 2 + 0.2).astype(np.float32)
    y = rng.poisson(case.tof_fp(img, n, tof) * 40.0).astype(np.float32)
    add = (rng.random(n) * 0.5 + 0.05).astype(np.float32) if randoms else None
    x = (rng.random(grid.N) + 0.5).astype(np.float32)
    print(f"[volume tof data step] n={n} bins={nb} randoms={randoms}", flush=True)
    xt = torch.from_numpy(x).to(cuda)
    yt = torch.from_numpy(y).to(cuda)
    rhs = torch.zeros(grid.N, device=cuda)
    sens = torch.zeros(grid.N, device=cuda)
    eps = 1e-5
    if add is None:
        _lib.call("omg_osem_data_step", geom.handle, 0, n, ptr(xt), ptr(yt), ptr(rhs), ptr(sens), eps, None)
    else:
        at = torch.from_numpy(add).to(cuda)
        _lib.call("omg_osem_data_step_ex", geom.handle, 0, n, ptr(xt), ptr(yt), ptr(at), ptr(rhs), ptr(sens), eps,
                  None)
    yhat = case.tof_fp(x, n, tof).reshape(nb, n).astype(np.float32)
    den = (yhat + np.float32(eps)) + (add if add is not None else np.float32(0))
    ratio = (y.reshape(nb, n) / den).astype(np.float32)
    ratio[~np.isfinite(ratio)] = np.float32(eps)
    rhs_o, s_o = case.tof_bp(ratio.ravel(), n, tof, sens=True)
    assert rel_l2(rhs.cpu().numpy(), rhs_o) <= 1e-5
    assert rel_l2(sens.cpu().numpy(), s_o) <= 1e-5


def test_projector_class_volume_tof(cuda):
    import torch
    from omega_amd.projector import projectorClass
    grid, sc = cases.orth_pet2d()
    p = projectorClass()
    p.Nx, p.Ny, p.Nz = grid.Nx, grid.Ny, grid.Nz
    p.FOVa_x, p.FOVa_y, p.axial_fov = grid.Nx * grid.dx, grid.Ny * grid.dy, grid.Nz * grid.dz
    p.projector_type = 3
    p.tube_radius, p.voxel_radius = TUBE_R, VOXEL_R
    p.x, p.z = sc.xy_table(), sc.z
    p.Ndist, p.Nang = sc.n_radial, sc.n_angles
    p.subsets, p.subsetType = 3, 8
    p.TOF_bins, p.TOF_bins_used, p.TOF_width, p.TOF_FWHM = 7, 7, 100e-12, 300e-12
    p.addProjector()
    p.initProj()
    assert p.nBins == 7
    vol = (p.V, np.float32(p.bmin), np.float32(p.bmax), np.float32(p.Vmax))
    geom = make_geom(p.grid, p.x, p.z, sc.n_radial, sc.n_angles, projector_type=3,
                     orth_width=float(np.float32(TUBE_R)), orth_flags=3, volume=vol,
                     tof=(7, float(p.sigma_x), p.TOFCenter, float(p.epps)))
    print(f"[volume tof projectorClass] subsets={p.subsets}", flush=True)
    img = torch.rand(grid.N, device=cuda) + 0.1
    for k in range(p.subsets):
        y = p.forwardProject(img, k)
        n = y.numel() // 7
        assert float(y.sum()) > 0
        assert np.array_equal(y.cpu().numpy(), fp(geom, int(p.nMeas[k]), n, img).cpu().numpy())
        b = p.backwardProject(y, k)
        assert rel_l2(b.cpu().numpy(), bp(geom, int(p.nMeas[k]), n, y, grid.N).cpu().numpy()) < 1e-6
