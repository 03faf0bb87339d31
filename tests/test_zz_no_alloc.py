"""omg_osem_subset_iteration_noalloc against omg_osem_subset_iteration_ex, through the C ABI.

The no-allocation entry point runs the same kernels in the same order with its work buffers in a fixed
device array, so on the deterministic paths (the direct-plane LDS kernels of the Siddon and orthogonal
projectors: 64-bit fixed-point sums, partial slices reduced in workgroup order) its image and sensitivity
are bit-identical to the allocating entry point's; on the one-lane-per-LOR path (float atomics in any
order) they agree within 1e-5 relative L2. Each case makes two calls in a row, the second on a smaller
measurement range than the first. A call whose work buffers exceed the fixed array fails with
OMG_ERR_ARG and allocates nothing.
"""
import numpy as np
import pytest

from omega_amd import _lib, workloads
from tests import cases
from tests.hip_helpers import make_geom, ptr, rel_l2

pytestmark = pytest.mark.gpu

def _noalloc():
    return _lib.load().omg_osem_subset_iteration_noalloc


def _both(handle, first, n, x0, meas, S0, compute_sens, beta, additive=None, epps=1e-5):
    """(x, S) after one subset iteration through _ex and through _noalloc, from the same inputs."""
    import torch
    out = []
    for name in ("omg_osem_subset_iteration_ex", "noalloc"):
        x = x0.clone()
        S = S0.clone()
        args = (handle, first, n, ptr(x), ptr(meas), ptr(additive), ptr(S), int(compute_sens), epps, beta, 2.0, 1,
                None)
        if name == "noalloc":
            rc = _noalloc()(*args)
            assert rc == 0, _lib.last_error()
        else:
            _lib.call(name, *args)
        torch.cuda.synchronize()
        out.append((x.cpu().numpy(), S.cpu().numpy()))
    return out


def _pet2d(cuda):
    import torch
    p = workloads.pet2d_projector(subsets=8)
    p.initProj()
    _, y = workloads.pet2d_data(p, seed=5)
    return p, y


@pytest.mark.parametrize("beta", [0.0, 0.1])
def test_planar_siddon_bitexact_two_calls(cuda, beta):
    """2-D plane PET (128x128x63 image, 128x160x63 sinogram): subset 0 (8 planes), then 3 planes."""
    import torch
    p, y = _pet2d(cuda)
    per = p.nRowsD * p.nColsD
    N = int(p.N[0])
    x0 = torch.from_numpy(np.random.default_rng(1).random(N).astype(np.float32) + 0.5).to(cuda)
    S0 = torch.zeros(N, device=cuda)
    for first, n_planes, sens in ((int(p.nMeas[0]), 8, 1), (int(p.nMeas[3]), 3, 1), (int(p.nMeas[3]), 3, 0)):
        off = first * per
        meas = y[off: off + n_planes * per]
        print(f"[noalloc planar siddon] beta={beta} first={first} planes={n_planes} sens={sens}", flush=True)
        (xa, Sa), (xb, Sb) = _both(p.geometry.handle, first, n_planes * per, x0, meas,
                                   S0 if sens else torch.ones(N, device=cuda), sens, beta)
        assert np.array_equal(xa, xb)
        assert np.array_equal(Sa, Sb)
        assert rel_l2(xb, x0.cpu().numpy()) > 1e-3   # the iteration did change the image


def test_voxel_driven_bp_bitexact_two_calls(cuda):
    """The opt-in voxel-driven backprojection (per-LOR q buffers in the fixed array): subset 1, then 2 planes."""
    import torch
    from omega_amd.projector import Geometry
    p, y = _pet2d(cuda)
    d, keep = p._descriptor()
    d.reserved[2] = 1
    geom = Geometry(d, keep)
    per = p.nRowsD * p.nColsD
    N = int(p.N[0])
    x0 = torch.from_numpy(np.random.default_rng(4).random(N).astype(np.float32) + 0.5).to(cuda)
    for first, n_planes, beta in ((int(p.nMeas[1]), 8, 0.0), (int(p.nMeas[5]), 2, 0.1)):
        print(f"[noalloc voxel bp] first={first} planes={n_planes} beta={beta}", flush=True)
        meas = y[first * per:(first + n_planes) * per]
        (xa, Sa), (xb, Sb) = _both(geom.handle, first, n_planes * per, x0, meas, torch.zeros(N, device=cuda), 1,
                                   beta)
        assert np.array_equal(xa, xb)
        assert np.array_equal(Sa, Sb)


def test_planar_orth_bitexact_two_calls(cuda):
    """Orthogonal projector, direct-plane kernels (32x32x9 image, 32x40x9 sinogram): 9 planes, then 4."""
    import torch
    grid, sc = cases.orth_pet2d()
    geom = make_geom(grid, sc.xy_table(), sc.z, sc.n_radial, sc.n_angles, projector_type=2, orth_width=3.0,
                     orth_flags=1)
    per = sc.n_radial * sc.n_angles
    rng = np.random.default_rng(2)
    y = torch.from_numpy(rng.poisson(3.0, per * 9).astype(np.float32)).to(cuda)
    x0 = torch.from_numpy(rng.random(grid.N).astype(np.float32) + 0.5).to(cuda)
    for first, n_planes in ((0, 9), (2, 4)):
        print(f"[noalloc planar orth] first={first} planes={n_planes}", flush=True)
        meas = y[first * per:(first + n_planes) * per]
        (xa, Sa), (xb, Sb) = _both(geom.handle, first, n_planes * per, x0, meas, torch.zeros(grid.N, device=cuda),
                                   1, 0.0)
        assert np.array_equal(xa, xb)
        assert np.array_equal(Sa, Sb)


def test_generic_3d_two_calls(cuda):
    """One-lane-per-LOR kernels on a reduced 3-D ring-pair scanner (oblique LORs, float atomics, randoms
    term): 1e-5 relative L2; the second call covers a quarter of the planes."""
    import torch
    p = workloads.pet3d_projector(subsets=1, n_rings=6, crystals=96, max_ring_diff=5, radius=150.0,
                                  image=(48, 48, 11), fov_xy=200.0)
    p.initProj()
    n_all = int(p.nMeasSubset[0, 0])
    per = p.nRowsD * p.nColsD
    rng = np.random.default_rng(3)
    y = torch.from_numpy(rng.poisson(2.0, n_all).astype(np.float32)).to(cuda)
    add = torch.from_numpy((rng.random(n_all) * 0.2).astype(np.float32)).to(cuda)
    N = int(p.N[0])
    x0 = torch.from_numpy(rng.random(N).astype(np.float32) + 0.5).to(cuda)
    planes = n_all // per
    for first, n_planes, beta in ((0, planes, 0.0), (planes // 2, planes // 4, 0.1)):
        print(f"[noalloc generic 3d] first={first} planes={n_planes} beta={beta}", flush=True)
        sl = slice(first * per, (first + n_planes) * per)
        (xa, Sa), (xb, Sb) = _both(p.geometry.handle, first, n_planes * per, x0, y[sl],
                                   torch.zeros(N, device=cuda), 1, beta, additive=add[sl])
        assert rel_l2(xb, xa) < 1e-5
        assert rel_l2(Sb, Sa) < 1e-5


def test_too_large_is_rejected(cuda):
    """A 256x256x256 image needs 2 x 16.8 M floats of image work buffers, more than the 64 MiB array:
    the call fails with OMG_ERR_ARG before any launch and leaves x untouched."""
    import torch
    from omega_amd import geometry as G
    grid = G.Grid.from_fov(256, 256, 256, 320.0, 320.0, 320.0)
    x = cases.list_rays(cases.tiny_grid(), 20, seed=1)
    geom = make_geom(grid, x, None, 1, 1, mode=_lib.MODE_LIST6)
    n = x.size // 6
    xi = torch.ones(grid.N, device=cuda)
    S = torch.zeros(grid.N, device=cuda)
    meas = torch.ones(n, device=cuda)
    rc = _noalloc()(geom.handle, 0, n, ptr(xi), ptr(meas), None, ptr(S), 1, 1e-5, 0.0, 2.0, 1, None)
    assert rc == _lib.OMG_ERR_ARG
    assert "no-allocation scratch" in _lib.last_error()
    torch.cuda.synchronize()
    assert bool((xi == 1).all())
