"""Measurement-level subsets (subset types 1-5) through projectorClass and the C ABI's index table.

Each subset's LORs are the sinogram bins of the reference's index vectors (checked against the
reference's own indexMaker in test_subset_indices.py). The oracle runs the same LORs as list-mode
rays built from the same coordinate tables, so the GPU forward projection must equal it bit for bit;
backprojections (float atomics) within 1e-5; a 2-iteration OSEM and an OSL-OSEM(RDP) run through
OSEMRunner within 1e-5 of the oracle's step sequence.
"""
import numpy as np
import pytest

from omega_amd import geometry as G
from omega_amd.projector import projectorClass
from omega_amd.recon import OSEMRunner
from oracle import oracle as O
from tests.hip_helpers import rel_l2

pytestmark = pytest.mark.gpu


def _projector(subset_type, projector_type=1, subsets=4):
    grid = G.Grid.from_fov(32, 32, 4, 80.0, 80.0, 10.0)
    sc = G.pet_2d_planes(grid, radius=100.0, n_radial=32, n_angles=40, radial_pitch=2.5)
    p = projectorClass()
    p.Nx, p.Ny, p.Nz = grid.Nx, grid.Ny, grid.Nz
    p.FOVa_x, p.FOVa_y, p.axial_fov = 80.0, 80.0, 10.0
    p.projector_type = projector_type
    if projector_type == 2:
        p.tube_width_xy = p.tube_width_z = 2.2
    p.x, p.z = sc.xy_table(), sc.z
    p.Ndist, p.Nang = sc.n_radial, sc.n_angles
    p.subsets = subsets
    p.subsetType = subset_type
    p.seed = 11
    p.addProjector()
    p.initProj()
    return p, grid, sc


def _list_rays(p, k):
    """The subset's LORs as list-mode rays [xs, ys, zs, xd, yd, zd] from the same float tables."""
    per = int(p.nRowsD) * int(p.nColsD)
    bins = p.index[int(p.nMeas[k]):int(p.nMeas[k + 1])]
    xy = p.x.reshape(-1, 4)[bins % per]
    zz = p.z.reshape(-1, 2)[bins // per]
    return np.ascontiguousarray(np.stack([xy[:, 0], xy[:, 1], zz[:, 0], xy[:, 2], xy[:, 3], zz[:, 1]], -1)
                                .astype(np.float32).ravel())


def _case(p, grid, k):
    rays = _list_rays(p, k)
    if p.projector_type == 2:
        return O.OrthCase(grid, 1, 1, rays, np.zeros(2, np.float32), 2.2, flags=3, mode=O.MODE_LIST), rays.size // 6
    return O.SiddonCase(grid, 1, 1, rays, np.zeros(2, np.float32), mode=O.MODE_LIST), rays.size // 6


@pytest.mark.parametrize("subset_type", [1, 2, 3, 4, 5])
def test_siddon_subsets_fp_bp(cuda, subset_type):
    import torch
    p, grid, sc = _projector(subset_type)
    rng = np.random.default_rng(subset_type)
    img = (rng.random(grid.N) + 0.05).astype(np.float32)
    for k in range(p.subsets):
        case, n = _case(p, grid, k)
        assert n == int(p.nMeasSubset[k, 0])
        print(f"[subsets] type={subset_type} subset={k} n={n}", flush=True)
        y = p.forwardProject(torch.from_numpy(img).to(cuda), k).cpu().numpy()
        assert np.array_equal(y, case.fp(img, n))
        r = rng.random(n).astype(np.float32)
        f = p.backwardProject(torch.from_numpy(r).to(cuda), k).cpu().numpy()
        assert rel_l2(f, case.bp(r, n)) < 1e-5


def test_orth_subsets_fp(cuda):
    import torch
    p, grid, sc = _projector(1, projector_type=2)
    img = (np.random.default_rng(5).random(grid.N) + 0.05).astype(np.float32)
    for k in range(p.subsets):
        case, n = _case(p, grid, k)
        y = p.forwardProject(torch.from_numpy(img).to(cuda), k).cpu().numpy()
        assert np.array_equal(y, case.fp(img, n))


# beta = 0.01 for OSL: at beta = 0.1, S_k + beta grad R changes sign at edge voxels of this small
# geometry (endpoints inside the image, 4 subsets), and a 1e-7 relative perturbation of one
# backprojection moves the oracle's own result by 1.4e-3: no FP32 path can be compared there.
@pytest.mark.parametrize("subset_type,beta", [(1, 0.0), (5, 0.0), (2, 0.01)])
def test_osem_with_measurement_subsets(cuda, subset_type, beta):
    import torch
    p, grid, sc = _projector(subset_type)
    rng = np.random.default_rng(3)
    ph = (rng.random(grid.N) * 2 + 0.1).astype(np.float32)
    full = O.SiddonCase(grid, sc.n_radial, sc.n_angles, p.x, p.z).fp(ph, sc.n_radial * sc.n_angles * (p.z.size // 2))
    y = p.subset_order(rng.poisson(full * 50).astype(np.float32))
    run = OSEMRunner(p, torch.from_numpy(y).to(cuda), beta=beta, gamma=2.0)
    x = torch.ones(grid.N, device=cuda)
    n_iter = 2
    for it in range(n_iter):
        for k in range(p.subsets):
            run.subset_step(x, k)
        run.sens_done = True
    xg = x.cpu().numpy()
    xo = np.ones(grid.N, np.float32)
    S = [None] * p.subsets
    for it in range(n_iter):
        for k in range(p.subsets):
            case, n = _case(p, grid, k)
            off = int(p.nTotMeas[k])
            rk = O.ratio(y[off:off + n], case.fp(xo, n), 1e-5)
            if it == 0:
                rhs, s = case.bp(rk, n, sens=True)
                S[k] = O.sens_clamp(s, 1e-5)
            else:
                rhs = case.bp(rk, n)
            if beta:
                dU = O.rdp_gradient(xo, (grid.Nx, grid.Ny, grid.Nz), 2.0, 1e-5, beta)
                xo = O.osl_update(xo, S[k], dU, rhs, 1e-5)
            else:
                xo = O.osem_update(xo, S[k], rhs, 1e-5)
    e = rel_l2(xg, xo)
    print(f"[subsets osem] type={subset_type} beta={beta} rel-L2 vs oracle {e:.2e}", flush=True)
    assert e < 1e-5


@pytest.mark.parametrize("subset_type", [9, 11])
def test_projection_level_orders_osem(cuda, subset_type):
    """Subset types 9 (seeded permutation) and 11 (prime-factor order) reorder the planes like type 8:
    per-subset forward projections bit-identical to the oracle on the reordered plane table, and a
    2-iteration OSEM within 1e-5 of the oracle's run."""
    import torch
    grid = G.Grid.from_fov(32, 32, 12, 80.0, 80.0, 30.0)
    sc = G.pet_2d_planes(grid, radius=100.0, n_radial=32, n_angles=40, radial_pitch=2.5)
    p = projectorClass()
    p.Nx, p.Ny, p.Nz = grid.Nx, grid.Ny, grid.Nz
    p.FOVa_x, p.FOVa_y, p.axial_fov = 80.0, 80.0, 30.0
    p.x, p.z = sc.xy_table(), sc.z
    p.Ndist, p.Nang = sc.n_radial, sc.n_angles
    p.subsets, p.subsetType, p.seed = 3, subset_type, 11
    p.addProjector()
    p.initProj()
    idx, cnt = G.subset_index_projections(subset_type, sc.n_planes, 3, seed=11)
    assert np.array_equal(np.asarray(p.index), idx)
    rng = np.random.default_rng(4)
    img = (rng.random(grid.N) + 0.1).astype(np.float32)
    it = torch.from_numpy(img).to(cuda)
    per = sc.n_radial * sc.n_angles
    ys = []
    for k in range(p.subsets):
        y = p.forwardProject(it, k).cpu().numpy()
        case = O.SiddonCase(grid, sc.n_radial, sc.n_angles, p.x, p.z, plane_offset=int(p.nMeas[k]))
        assert np.array_equal(y, case.fp(img, int(cnt[k]) * per))
        ys.append(y)
    meas = rng.poisson(np.concatenate(ys) * 30.0).astype(np.float32)
    print(f"[subsets type {subset_type}] planes per subset {cnt.tolist()}", flush=True)
    run = OSEMRunner(p, torch.from_numpy(meas).to(cuda))
    x = torch.ones(grid.N, device=cuda)
    for _ in range(2):
        for k in range(p.subsets):
            run.subset_step(x, k)
        run.sens_done = True
    torch.cuda.synchronize()
    xo = O.osem(grid, sc.n_radial, sc.n_angles, p.x, p.z, cnt, meas, 2, np.ones(grid.N, np.float32), epps=1e-5)
    assert rel_l2(x.cpu().numpy(), xo) < 1e-5
