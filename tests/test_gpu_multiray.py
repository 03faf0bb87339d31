"""GPU multi-ray Siddon projections against the reference's own CPU outputs
(tests/golden/siddon_multiray.npz, written by tests/golden/make_multiray.py from the reference headers)
and the oracle.

Bars: forward projections bit-identical to the reference's (the same float operations: shifted
endpoints, per-ray weight 1 / (L nRays), rays summed in order); backprojections and sensitivity within
1e-6 relative L2 (float atomics against the reference's sequential sums); fused OSEM subset iterations
with several rays per LOR within 1e-5 of the oracle run.
"""
import os

import numpy as np
import pytest

from omega_amd import _lib
from omega_amd.recon import OSEMRunner
from oracle import oracle as O
from tests import cases
from tests.hip_helpers import bp, fp, make_geom, rel_l2

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load():
    return np.load(os.path.join(GOLD, "siddon_multiray.npz"), allow_pickle=False)


def test_listmode_against_reference(cuda):
    import torch
    f = _load()
    grid = cases.tiny_grid()
    x = f["list_x"]
    n = x.size // 6
    print(f"[multi-ray list-mode] n={n} rays=3x2", flush=True)
    geom = make_geom(grid, x, None, 1, 1, mode=_lib.MODE_LIST6, rays=(3, 2, 0.9, 1.1))
    img = torch.from_numpy(f["list_img"]).to(cuda)
    y = fp(geom, 0, n, img).cpu().numpy()
    assert np.array_equal(y, f["list_fp32"]), np.abs(y - f["list_fp32"]).max()
    yin = torch.from_numpy(f["list_yin"]).to(cuda)
    rhs, sens = bp(geom, 0, n, yin, grid.N, sens=True)
    assert rel_l2(rhs.cpu().numpy(), f["list_bp32"]) < 1e-6
    assert rel_l2(sens.cpu().numpy(), f["list_sens32"]) < 1e-6
    geom = make_geom(grid, x, None, 1, 1, mode=_lib.MODE_LIST6, rays=(3, 2, 0.9, 1.1), norm=f["list_norm"],
                     atten=f["list_fac"], atten_mode=2, scat=f["list_scat"], global_factor=0.9)
    assert np.array_equal(fp(geom, 0, n, img).cpu().numpy(), f["list_all_fp32"])
    assert rel_l2(bp(geom, 0, n, yin, grid.N).cpu().numpy(), f["list_all_bp32"]) < 1e-6


@pytest.mark.parametrize("x_fastest_only", [False, True])
def test_sinograms_against_reference(cuda, x_fastest_only):
    import torch
    f = _load()
    g2, sc = cases.orth_pet2d()
    planes = int(f["pet2d_planes"])
    n = sc.n_radial * sc.n_angles * planes
    print(f"[multi-ray pet2d] n={n} rays=5x1 x_fastest_only={x_fastest_only}", flush=True)
    geom = make_geom(g2, sc.xy_table(), sc.z[: 2 * planes], sc.n_radial, sc.n_angles, rays=(5, 1, 2.5, 0.0),
                     x_fastest_only=x_fastest_only)
    assert np.array_equal(fp(geom, 0, n, torch.from_numpy(f["pet2d_img"]).to(cuda)).cpu().numpy(), f["pet2d_fp32"])
    rhs, sens = bp(geom, 0, n, torch.from_numpy(f["pet2d_yin"]).to(cuda), g2.N, sens=True)
    assert rel_l2(rhs.cpu().numpy(), f["pet2d_bp32"]) < 1e-6
    assert rel_l2(sens.cpu().numpy(), f["pet2d_sens32"]) < 1e-6
    g3, sc3 = cases.c2_small()
    planes3 = int(f["pet3d_planes"])
    n3 = sc3.n_radial * sc3.n_angles * planes3
    geom = make_geom(g3, sc3.xy_table(), sc3.z[: 2 * planes3], sc3.n_radial, sc3.n_angles, rays=(2, 3, 2.0, 4.0),
                     x_fastest_only=x_fastest_only)
    assert np.array_equal(fp(geom, 0, n3, torch.from_numpy(f["pet3d_img"]).to(cuda)).cpu().numpy(), f["pet3d_fp32"])


def test_osem_with_multiray(cuda):
    import torch
    from omega_amd.projector import projectorClass
    grid, sc = cases.orth_pet2d()
    p = projectorClass()
    p.Nx, p.Ny, p.Nz = grid.Nx, grid.Ny, grid.Nz
    p.FOVa_x, p.FOVa_y, p.axial_fov = grid.Nx * grid.dx, grid.Ny * grid.dy, grid.Nz * grid.dz
    p.x, p.z = sc.xy_table(), sc.z
    p.Ndist, p.Nang = sc.n_radial, sc.n_angles
    p.subsets, p.subsetType = 3, 8
    p.n_rays_transaxial, p.cr_pz, p.cr_p = 3, 2.5, 2.5
    p.addProjector()
    p.initProj()
    rng = np.random.default_rng(22)
    xt = torch.from_numpy((rng.random(grid.N) * 2 + 0.2).astype(np.float32)).to(cuda)
    y = torch.cat([p.forwardProject(xt, k) for k in range(p.subsets)]).cpu().numpy()
    y = rng.poisson(y * 50.0).astype(np.float32)
    print(f"[multi-ray osem] rays=3 LORs={y.size}", flush=True)
    run = OSEMRunner(p, torch.from_numpy(y).to(cuda))
    x = torch.ones(grid.N, device=cuda)
    n_sub, done = 5, 0
    for it in range(2):
        for k in range(p.subsets):
            if done < n_sub:
                run.subset_step(x, k)
                done += 1
        run.sens_done = True
    torch.cuda.synchronize()
    nmeas = np.diff(p.nMeas).astype(np.int64)
    xo = O.osem(p.grid, p.nRowsD, p.nColsD, p.x, p.z, nmeas, y, 2, np.ones(grid.N, np.float32), epps=1e-5,
                subsets_to_run=n_sub, rays=(3, 1), ray_pitch=(2.5, 2.5))
    one = O.osem(p.grid, p.nRowsD, p.nColsD, p.x, p.z, nmeas, y, 2, np.ones(grid.N, np.float32), epps=1e-5,
                 subsets_to_run=n_sub)
    e = rel_l2(x.cpu().numpy(), xo)
    print(f"  rel-L2 vs oracle {e:.2e}; one ray vs three {rel_l2(one, xo):.2e}", flush=True)
    assert e < 1e-5
    assert rel_l2(one, xo) > 1e-3
