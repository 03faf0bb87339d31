"""GPU parity of the orthogonal distance-based projector (projector type 2) against the oracle.

Bar: forward projections and per-LOR voxel lists (indices in visiting order, tube weights)
bit-identical to the oracle's restatement of the reference kernel; backprojection and sensitivity
(unordered float atomics, or exact fixed-point sums in the direct-plane LDS kernels) within 1e-5
relative L2; the fused OSEM subset iteration within 1e-5 of the oracle's FP -> ratio -> BP -> update
sequence; FP/BP adjoint to 1e-5. Geometries whose tubes stay in their slice run the direct-plane
kernels; the same cases through the one-lane-per-LOR kernels are checked too.
"""
import numpy as np
import pytest

from omega_amd import _lib
from oracle import oracle as O
from tests import cases
from tests.hip_helpers import bp, fp, make_geom, ptr, rel_l2, trace

pytestmark = pytest.mark.gpu

W = 3.0


def _pet2d(flags, norm=None, global_factor=1.0):
    grid, sc = cases.orth_pet2d()
    case = O.OrthCase(grid, sc.n_radial, sc.n_angles, sc.xy_table(), sc.z, W, flags=flags, norm=norm,
                      global_factor=global_factor)
    geom = make_geom(grid, sc.xy_table(), sc.z, sc.n_radial, sc.n_angles, projector_type=2, orth_width=W,
                     orth_flags=flags, norm=norm, global_factor=global_factor)
    return grid, sc, case, geom


@pytest.mark.parametrize("flags", [1, 3])
def test_pet2d_fp_trace_bitexact(cuda, flags):
    import torch
    grid, sc, case, geom = _pet2d(flags)
    n = sc.n_radial * sc.n_angles * (sc.z.size // 2)
    print(f"[orth pet2d] flags={flags} n={n}", flush=True)
    img = np.random.default_rng(4).random(grid.N).astype(np.float32) + 0.05
    y = fp(geom, 0, n, torch.from_numpy(img).to(cuda)).cpu().numpy()
    y_ref = case.fp(img, n)
    assert np.array_equal(y, y_ref), f"max |diff| {np.abs(y - y_ref).max()}"
    c_g, i_g, w_g = trace(geom, 0, n, cuda)
    c_o, i_o, w_o, _ = case.trace(n)
    assert np.array_equal(c_g, c_o)
    assert np.array_equal(i_g.astype(np.int64), i_o)
    assert np.array_equal(w_g, w_o)


@pytest.mark.parametrize("flags", [1, 3])
def test_list_mode_fp_bitexact_bp(cuda, flags):
    import torch
    grid = cases.square_grid()
    x = cases.list_rays(grid, 400, seed=6)
    n = x.size // 6
    print(f"[orth list-mode] flags={flags} n={n}", flush=True)
    case = O.OrthCase(grid, 1, 1, x, np.zeros(2, np.float32), 2.2, flags=flags, mode=O.MODE_LIST)
    geom = make_geom(grid, x, None, 1, 1, mode=_lib.MODE_LIST6, projector_type=2, orth_width=2.2, orth_flags=flags)
    img = np.random.default_rng(7).random(grid.N).astype(np.float32) + 0.05
    y = fp(geom, 0, n, torch.from_numpy(img).to(cuda)).cpu().numpy()
    assert np.array_equal(y, case.fp(img, n))
    r = np.random.default_rng(8).random(n).astype(np.float32)
    rhs, s = bp(geom, 0, n, torch.from_numpy(r).to(cuda), grid.N, sens=True)
    rhs_o, s_o = case.bp(r, n, sens=True)
    assert rel_l2(rhs.cpu().numpy(), rhs_o) <= 1e-5
    assert rel_l2(s.cpu().numpy(), s_o) <= 1e-5


def test_subset_bp_sens_and_adjoint(cuda):
    import torch
    rng = np.random.default_rng(9)
    grid, sc = cases.orth_pet2d()
    n_all = sc.n_radial * sc.n_angles * (sc.z.size // 2)
    norm = rng.uniform(0.5, 1.5, n_all).astype(np.float32)
    _, _, _, geom = _pet2d(3, norm=norm, global_factor=1.2)
    per = sc.n_radial * sc.n_angles
    first, planes = 3, 4                       # planes 3..6, a subset-like slab of the table
    n = per * planes
    case = O.OrthCase(grid, sc.n_radial, sc.n_angles, sc.xy_table(), sc.z, W, flags=3, plane_offset=first,
                      norm=norm[first * per:first * per + n], global_factor=1.2)
    print(f"[orth subset bp] first={first} n={n}", flush=True)
    r = rng.random(n).astype(np.float32)
    rhs, s = bp(geom, first, n, torch.from_numpy(r).to(cuda), grid.N, sens=True)
    rhs_o, s_o = case.bp(r, n, sens=True)
    assert rel_l2(rhs.cpu().numpy(), rhs_o) <= 1e-5
    assert rel_l2(s.cpu().numpy(), s_o) <= 1e-5
    img = rng.random(grid.N).astype(np.float32)
    y = fp(geom, first, n, torch.from_numpy(img).to(cuda)).cpu().numpy()
    assert np.array_equal(y, case.fp(img, n))
    lhs = float(np.dot(y.astype(np.float64), r))
    rr = float(np.dot(img.astype(np.float64), rhs.cpu().numpy().astype(np.float64)))
    assert abs(lhs - rr) <= 1e-5 * abs(lhs)


def test_osem_subset_iteration_orth(cuda):
    import torch
    grid, sc, case_all, geom = _pet2d(3)
    per = sc.n_radial * sc.n_angles
    nplanes = sc.z.size // 2
    rng = np.random.default_rng(10)
    x0 = (rng.random(grid.N).astype(np.float32) + 0.5)
    meas = rng.poisson(case_all.fp(x0, per * nplanes) * 20).astype(np.float32)
    first, planes = 0, nplanes
    n = per * planes
    epps = 1e-5
    print(f"[orth osem] n={n}", flush=True)
    xg = torch.from_numpy(np.ones(grid.N, np.float32)).to(cuda)
    S = torch.zeros(grid.N, dtype=torch.float32, device=cuda)
    mt = torch.from_numpy(meas).to(cuda)
    _lib.call("omg_osem_subset_iteration", geom.handle, first, n, ptr(xg), ptr(mt), ptr(S), 1, epps, 0.0, 2.0, 1,
              None)
    x = np.ones(grid.N, np.float32)
    yhat = case_all.fp(x, n)
    r_out = yhat.copy()
    O.lib().oracle_ratio_f32(meas, r_out, n, epps)
    rhs, sens = case_all.bp(r_out, n, sens=True)
    O.lib().oracle_sens_clamp_f32(sens, grid.N, epps)
    O.lib().oracle_osem_update_f32(x, sens, rhs, grid.N, epps, 1)
    assert rel_l2(S.cpu().numpy(), sens) <= 1e-5
    assert rel_l2(xg.cpu().numpy(), x) <= 1e-5


def test_projector_class_orth(cuda):
    import torch
    from omega_amd import geometry as G
    from omega_amd.projector import projectorClass
    grid, sc = cases.orth_pet2d()
    p = projectorClass()
    p.Nx, p.Ny, p.Nz = 32, 32, 9
    p.FOVa_x, p.FOVa_y, p.axial_fov = 80.0, 80.0, 22.5
    p.projector_type = 2
    p.tube_width_xy = W
    p.tube_width_z = W
    p.x = sc.xy_table()
    p.z = sc.z
    p.Ndist, p.Nang = sc.n_radial, sc.n_angles
    p.subsets = 3
    p.subsetType = 8
    p.addProjector()
    p.initProj()
    img = torch.rand(grid.N, device=cuda) + 0.1
    for k in range(p.subsets):
        y = p.forwardProject(img, k)
        case = O.OrthCase(grid, sc.n_radial, sc.n_angles, p.x, p.z, W, flags=3, plane_offset=int(p.nMeas[k]))
        assert np.array_equal(y.cpu().numpy(), case.fp(img.cpu().numpy(), y.numel()))


@pytest.mark.parametrize("flags,width", [(1, 3.0), (3, 2.2)])
@pytest.mark.parametrize("path", ["lists", "traced", "generic"])
def test_planar_tube_in_slice(cuda, flags, width, path):
    """Direct planes whose tube stays in its slice (transaxial sweep; or axial sweep with 0.99 * width
    below the 2.5 mm slice pitch): the direct-plane LDS kernels fed by the kept tube lists (default),
    the same kernels tracing the tubes (lists off) and the one-lane-per-LOR kernels give FP bit-identical
    to the oracle, BP / sensitivity / the fused data step within 1e-5, OSEM and OSL-OSEM subset
    iterations within 1e-5; both LDS paths are bit-identical run to run (64-bit fixed-point sums)."""
    generic = path == "generic"
    import torch
    grid, sc = cases.orth_pet2d()
    per = sc.n_radial * sc.n_angles
    nplanes = sc.z.size // 2
    n = per * nplanes
    print(f"[orth planar] flags={flags} width={width} path={path} n={n}", flush=True)
    case = O.OrthCase(grid, sc.n_radial, sc.n_angles, sc.xy_table(), sc.z, width, flags=flags)
    geom = make_geom(grid, sc.xy_table(), sc.z, sc.n_radial, sc.n_angles, projector_type=2, orth_width=width,
                     orth_flags=flags, generic=generic, tube_budget_mib=-1 if path == "traced" else 0)
    rng = np.random.default_rng(20 + flags)
    img = rng.random(grid.N).astype(np.float32) + 0.05
    img_t = torch.from_numpy(img).to(cuda)
    y = fp(geom, 0, n, img_t).cpu().numpy()
    assert np.array_equal(y, case.fp(img, n))
    r = rng.random(n).astype(np.float32)
    rhs, s = bp(geom, 0, n, torch.from_numpy(r).to(cuda), grid.N, sens=True)
    rhs_o, s_o = case.bp(r, n, sens=True)
    assert rel_l2(rhs.cpu().numpy(), rhs_o) <= 1e-5
    assert rel_l2(s.cpu().numpy(), s_o) <= 1e-5
    meas = rng.poisson(case.fp(img, n) * 20).astype(np.float32)
    mt = torch.from_numpy(meas).to(cuda)
    rhs_t = torch.zeros(grid.N, device=cuda)
    sens_t = torch.zeros(grid.N, device=cuda)
    _lib.call("omg_osem_data_step", geom.handle, 0, n, ptr(img_t), ptr(mt), ptr(rhs_t), ptr(sens_t), 1e-5, None)
    r_o = O.ratio(meas, case.fp(img, n), 1e-5)
    rhs_o, s_o = case.bp(r_o, n, sens=True)
    assert rel_l2(rhs_t.cpu().numpy(), rhs_o) <= 1e-5
    assert rel_l2(sens_t.cpu().numpy(), s_o) <= 1e-5
    # two subset iterations (the first computes S), OSEM and OSL-OSEM(RDP), against the oracle's steps
    half = (nplanes // 2) * per
    for beta in (0.0, 0.1):
        outs = []
        for _ in range(2):
            xg = torch.ones(grid.N, device=cuda)
            S = torch.zeros(grid.N, device=cuda)
            for k, (f, m) in enumerate(((0, half), (0, half))):
                _lib.call("omg_osem_subset_iteration", geom.handle, f, m, ptr(xg), ptr(mt), ptr(S), int(k == 0),
                          1e-5, beta, 2.0, 1, None)
            outs.append(xg.cpu().numpy())
        if not generic:
            assert np.array_equal(outs[0], outs[1])
        x = np.ones(grid.N, np.float32)
        S_o = None
        for k in range(2):
            ck = O.OrthCase(grid, sc.n_radial, sc.n_angles, sc.xy_table(), sc.z, width, flags=flags)
            rk = O.ratio(meas[:half], ck.fp(x, half), 1e-5)
            if k == 0:
                rhs_k, s_k = ck.bp(rk, half, sens=True)
                S_o = O.sens_clamp(s_k, 1e-5)
            else:
                rhs_k = ck.bp(rk, half)
            if beta:
                dU = O.rdp_gradient(x, (grid.Nx, grid.Ny, grid.Nz), 2.0, 1e-5, beta)
                x = O.osl_update(x, S_o, dU, rhs_k, 1e-5)
            else:
                x = O.osem_update(x, S_o, rhs_k, 1e-5)
        e = rel_l2(outs[0], x)
        print(f"  beta={beta} rel-L2 vs oracle {e:.2e}", flush=True)
        assert e <= 1e-5
