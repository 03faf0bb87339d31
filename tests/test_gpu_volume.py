"""GPU parity of the volume-of-intersection projector (projector type 3) against the oracle.

Bars as for projector 2: forward projections and per-LOR voxel lists (indices in visiting order, table
weights) bit-identical to the oracle; backprojection and sensitivity (unordered float atomics) within
1e-5 relative L2; projectorClass with projector_type 3 builds the reference's table and projects as the
oracle does; an OSEM subset iteration through it within 1e-5 of the oracle's FP -> ratio -> BP -> update.
"""
import numpy as np
import pytest

from omega_amd import _lib
from omega_amd import geometry as G
from oracle import oracle as O
from tests import cases
from tests.hip_helpers import bp, fp, make_geom, ptr, rel_l2, trace

pytestmark = pytest.mark.gpu

TUBE_R, VOXEL_R = 2.0, 1.0


def _volume(grid, tr=TUBE_R):
    V, bmin, bmax, vmax, _ = G.voxel_volume_table(tr, VOXEL_R, grid.dx, grid.dy, grid.dz)
    return V, np.float32(bmin), np.float32(bmax), np.float32(vmax)


@pytest.mark.parametrize("flags", [1, 3])
def test_pet2d_fp_trace_bp(cuda, flags):
    import torch
    grid, sc = cases.orth_pet2d()
    vol = _volume(grid)
    rng = np.random.default_rng(6)
    n = sc.n_radial * sc.n_angles * (sc.z.size // 2)
    norm = rng.uniform(0.5, 1.5, n).astype(np.float32)
    case = O.OrthCase(grid, sc.n_radial, sc.n_angles, sc.xy_table(), sc.z, TUBE_R, flags=flags, volume=vol,
                      norm=norm, global_factor=0.8)
    geom = make_geom(grid, sc.xy_table(), sc.z, sc.n_radial, sc.n_angles, projector_type=3, orth_width=TUBE_R,
                     orth_flags=flags, volume=vol, norm=norm, global_factor=0.8)
    print(f"[volume pet2d] flags={flags} n={n} table={vol[0].size}", flush=True)
    img = rng.random(grid.N).astype(np.float32) + 0.05
    y = fp(geom, 0, n, torch.from_numpy(img).to(cuda)).cpu().numpy()
    y_ref = case.fp(img, n)
    assert np.array_equal(y, y_ref), f"max |diff| {np.abs(y - y_ref).max()}"
    c_g, i_g, w_g = trace(geom, 0, n, cuda)
    c_o, i_o, w_o, _ = case.trace(n)
    assert np.array_equal(c_g, c_o)
    assert np.array_equal(i_g.astype(np.int64), i_o)
    assert np.array_equal(w_g, w_o)
    r = rng.random(n).astype(np.float32)
    rhs, s = bp(geom, 0, n, torch.from_numpy(r).to(cuda), grid.N, sens=True)
    rhs_o, s_o = case.bp(r, n, sens=True)
    assert rel_l2(rhs.cpu().numpy(), rhs_o) <= 1e-5
    assert rel_l2(s.cpu().numpy(), s_o) <= 1e-5


def test_list_mode_all_directions(cuda):
    import torch
    grid = cases.square_grid()
    x = cases.list_rays(grid, 400, seed=12)
    n = x.size // 6
    vol = _volume(grid, 1.5)
    case = O.OrthCase(grid, 1, 1, x, np.zeros(2, np.float32), 1.5, flags=3, mode=O.MODE_LIST, volume=vol)
    geom = make_geom(grid, x, None, 1, 1, mode=_lib.MODE_LIST6, projector_type=3, orth_width=1.5, orth_flags=3,
                     volume=vol)
    print(f"[volume list-mode] n={n}", flush=True)
    img = np.random.default_rng(13).random(grid.N).astype(np.float32)
    assert np.array_equal(fp(geom, 0, n, torch.from_numpy(img).to(cuda)).cpu().numpy(), case.fp(img, n))
    r = np.random.default_rng(14).random(n).astype(np.float32)
    rhs = bp(geom, 0, n, torch.from_numpy(r).to(cuda), grid.N)
    assert rel_l2(rhs.cpu().numpy(), case.bp(r, n)) <= 1e-5


def _projector():
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
    p.addProjector()
    p.initProj()
    return grid, sc, p


def test_projector_class_and_osem(cuda):
    import torch
    grid, sc, p = _projector()
    vol = (p.V, np.float32(p.bmin), np.float32(p.bmax), np.float32(p.Vmax))
    assert np.array_equal(p.V, _volume(p.grid)[0])
    img = torch.rand(grid.N, device=cuda) + 0.1
    for k in range(p.subsets):
        y = p.forwardProject(img, k)
        case = O.OrthCase(p.grid, sc.n_radial, sc.n_angles, p.x, p.z, TUBE_R, flags=3, plane_offset=int(p.nMeas[k]),
                          volume=vol)
        assert np.array_equal(y.cpu().numpy(), case.fp(img.cpu().numpy(), y.numel()))
    n = int(p.nMeasSubset[0, 0])   # measurements of subset 0 (nMeas counts planes here)
    case = O.OrthCase(p.grid, sc.n_radial, sc.n_angles, p.x, p.z, TUBE_R, flags=3, volume=vol)
    rng = np.random.default_rng(15)
    meas = rng.poisson(case.fp(rng.random(grid.N).astype(np.float32) + 0.5, n) * 50).astype(np.float32)
    print(f"[volume osem] n={n}", flush=True)
    xg = torch.ones(grid.N, device=cuda)
    S = torch.zeros(grid.N, dtype=torch.float32, device=cuda)
    mt = torch.from_numpy(meas).to(cuda)
    _lib.call("omg_osem_subset_iteration", p._geom.handle, 0, n, ptr(xg),
              ptr(mt), ptr(S), 1, 1e-5, 0.0, 2.0, 1, None)
    x = np.ones(grid.N, np.float32)
    r_out = case.fp(x, n)
    O.lib().oracle_ratio_f32(meas, r_out, n, 1e-5)
    rhs, sens = case.bp(r_out, n, sens=True)
    O.lib().oracle_sens_clamp_f32(sens, grid.N, 1e-5)
    O.lib().oracle_osem_update_f32(x, sens, rhs, grid.N, 1e-5, 1)
    assert rel_l2(S.cpu().numpy(), sens) <= 1e-5
    assert rel_l2(xg.cpu().numpy(), x) <= 1e-5


@pytest.mark.parametrize("flags", [1, 3])
def test_fused_steps_after_the_first(cuda, flags):
    """The fused subset steps, the sensitivity-building first one and one after it, on the three paths of
    projector 3: the tube lists kept in device memory (default), the tracing slice-accumulator kernel
    (lists off; slices tk and tk + 1 in LDS, tk - 1 through global atomics) and the one-lane-per-LOR
    kernel. Each against the one-lane-per-LOR kernel within 1e-6 (only the order of the float additions
    differs) and against the oracle's FP -> ratio -> BP -> update twice (1e-5)."""
    import torch
    grid, sc = cases.orth_pet2d()
    vol = _volume(grid)
    n = sc.n_radial * sc.n_angles * (sc.z.size // 2)
    case = O.OrthCase(grid, sc.n_radial, sc.n_angles, sc.xy_table(), sc.z, TUBE_R, flags=flags, volume=vol)
    rng = np.random.default_rng(16)
    meas = rng.poisson(case.fp(rng.random(grid.N).astype(np.float32) + 0.5, n) * 50).astype(np.float32)
    mt = torch.from_numpy(meas).to(cuda)
    print(f"[volume fused steps] flags={flags} n={n} planes={sc.z.size // 2} Nz={grid.Nz}", flush=True)
    out = {}
    for path, kw in (("lists", {}), ("traced", dict(tube_budget_mib=-1)), ("generic", dict(generic=True))):
        geom = make_geom(grid, sc.xy_table(), sc.z, sc.n_radial, sc.n_angles, projector_type=3, orth_width=TUBE_R,
                         orth_flags=flags, volume=vol, **kw)
        xg = torch.ones(grid.N, device=cuda)
        S = torch.zeros(grid.N, dtype=torch.float32, device=cuda)
        for step in range(2):
            _lib.call("omg_osem_subset_iteration", geom.handle, 0, n, ptr(xg), ptr(mt), ptr(S), int(step == 0), 1e-5,
                      0.0, 2.0, 1, None)
        out[path] = xg.cpu().numpy()
    x = np.ones(grid.N, np.float32)
    sens = None
    for step in range(2):
        r = case.fp(x, n)
        O.lib().oracle_ratio_f32(meas, r, n, 1e-5)
        if step == 0:
            rhs, sens = case.bp(r, n, sens=True)
            O.lib().oracle_sens_clamp_f32(sens, grid.N, 1e-5)
        else:
            rhs = case.bp(r, n)
        O.lib().oracle_osem_update_f32(x, sens, rhs, grid.N, 1e-5, 1)
    for path in ("lists", "traced"):
        e_gen = rel_l2(out[path], out["generic"])
        e_or = rel_l2(out[path], x)
        print(f"  {path} vs one-lane-per-LOR {e_gen:.2e}, vs oracle {e_or:.2e}", flush=True)
        assert e_gen <= 1e-6
        assert e_or <= 1e-5
    assert rel_l2(out["generic"], x) <= 1e-5


def test_tube_lists_forward_bitexact_and_budget_off(cuda):
    """Forward projections from the kept tube lists are bit-identical to the oracle (the same pairs in the
    same order), repeatedly (the lists are built on the first call), on every subset; the backprojection
    within 1e-5; with the budget off the tracing kernels give the same forward projection."""
    import torch
    grid, sc = cases.orth_pet2d()
    vol = _volume(grid)
    rng = np.random.default_rng(17)
    n_planes = sc.z.size // 2
    per = sc.n_radial * sc.n_angles
    img = torch.from_numpy(rng.random(grid.N).astype(np.float32) + 0.05).to(cuda)
    on = make_geom(grid, sc.xy_table(), sc.z, sc.n_radial, sc.n_angles, projector_type=3, orth_width=TUBE_R,
                   orth_flags=3, volume=vol)
    off = make_geom(grid, sc.xy_table(), sc.z, sc.n_radial, sc.n_angles, projector_type=3, orth_width=TUBE_R,
                    orth_flags=3, volume=vol, tube_budget_mib=-1)
    print(f"[volume tube lists] planes={n_planes}", flush=True)
    for first, planes in ((0, 2), (2, n_planes - 2), (0, 2)):
        n = planes * per
        case = O.OrthCase(grid, sc.n_radial, sc.n_angles, sc.xy_table(), sc.z, TUBE_R, flags=3, volume=vol,
                          plane_offset=first)
        y_ref = case.fp(img.cpu().numpy(), n)
        assert np.array_equal(fp(on, first, n, img).cpu().numpy(), y_ref)
        assert np.array_equal(fp(off, first, n, img).cpu().numpy(), y_ref)
        r = rng.random(n).astype(np.float32)
        rhs, s = bp(on, first, n, torch.from_numpy(r).to(cuda), grid.N, sens=True)
        rhs_o, s_o = case.bp(r, n, sens=True)
        assert rel_l2(rhs.cpu().numpy(), rhs_o) <= 1e-5
        assert rel_l2(s.cpu().numpy(), s_o) <= 1e-5
