"""GPU parity of the improved Siddon projector (projector type 1) against the CPU oracle.

Bar: forward projections and traversal lists (voxel indices, lengths) bit-identical to the
oracle, which is itself bit-identical to the reference CPU build (test_oracle_ref.py,
test_golden.py). Backprojections (unordered float atomics in the one-lane-per-LOR kernels, as in
the reference; exact 64-bit fixed-point sums of the float terms in the direct-plane LDS kernels)
within 1e-5 relative L2 of the FP32 oracle; the FP32 error against the FP64 oracle is printed.
"""
import numpy as np
import pytest

from omega_amd import _lib
from omega_amd import geometry as G
from oracle import oracle as O
from tests import cases
from tests.hip_helpers import bp, fp, make_geom, ptr, rel_l2, trace

pytestmark = pytest.mark.gpu

BP_TOL = 1e-5


def _img(N, seed):
    return np.random.default_rng(seed).random(N).astype(np.float32) + 0.05


def test_list_mode_tiny_grid_bitexact(cuda):
    import torch
    grid = cases.tiny_grid()
    x = cases.list_rays(grid, 400, seed=1)
    n = x.size // 6
    print(f"[siddon list-mode tiny] n={n}", flush=True)
    case = O.SiddonCase(grid, 1, 1, x, np.zeros(2, np.float32), mode=O.MODE_LIST)
    img = _img(grid.N, 2)
    geom = make_geom(grid, x, None, 1, 1, mode=_lib.MODE_LIST6)
    y = fp(geom, 0, n, torch.from_numpy(img).to(cuda)).cpu().numpy()
    y_ref = case.fp(img, n)
    assert np.array_equal(y, y_ref), f"max |diff| {np.abs(y - y_ref).max()}"
    c_g, i_g, l_g = trace(geom, 0, n, cuda)
    c_o, i_o, l_o, _ = case.trace(n)
    assert np.array_equal(c_g, c_o)
    assert np.array_equal(i_g, i_o)
    assert np.array_equal(l_g, l_o)
    yin = _img(n, 3)
    rhs, sens = bp(geom, 0, n, torch.from_numpy(yin).to(cuda), grid.N, sens=True)
    rhs_o, sens_o = case.bp(yin, n, sens=True)
    assert rel_l2(rhs.cpu().numpy(), rhs_o) < BP_TOL
    assert rel_l2(sens.cpu().numpy(), sens_o) < BP_TOL


@pytest.mark.parametrize("subset", [0, 3, 7])
def test_c0_subset_fp_bitexact_bp_close(cuda, subset):
    import torch
    grid, sc = cases.c0()
    xy = sc.xy_table()
    idx, nmeas = G.subset_index_type8(sc.n_planes, 8)
    zs = np.ascontiguousarray(sc.z.reshape(-1, 2)[idx].reshape(-1))
    per = sc.n_radial * sc.n_angles
    first = int(nmeas[:subset].sum())
    n = int(nmeas[subset]) * per
    print(f"[siddon c0] subset={subset} first={first} n={n}", flush=True)
    case = O.SiddonCase(grid, sc.n_radial, sc.n_angles, xy, zs, plane_offset=first)
    geom = make_geom(grid, xy, zs, sc.n_radial, sc.n_angles)
    img = _img(grid.N, 10 + subset)
    y = fp(geom, first, n, torch.from_numpy(img).to(cuda)).cpu().numpy()
    y_ref = case.fp(img, n)
    assert np.array_equal(y, y_ref), f"{np.count_nonzero(y != y_ref)} of {n} differ"
    yin = _img(n, 20 + subset)
    rhs, sens = bp(geom, first, n, torch.from_numpy(yin).to(cuda), grid.N, sens=True)
    rhs_o, sens_o = case.bp(yin, n, sens=True)
    e = rel_l2(rhs.cpu().numpy(), rhs_o)
    case64 = O.SiddonCase(grid, sc.n_radial, sc.n_angles, xy, zs, plane_offset=first, dtype=np.float64)
    rhs64 = case64.bp(yin.astype(np.float64), n)
    print(f"  bp rel-L2 vs fp32 oracle {e:.2e}; vs fp64 oracle {rel_l2(rhs.cpu().numpy(), rhs64):.2e}; "
          f"fp vs fp64 {rel_l2(y, case64.fp(img.astype(np.float64), n)):.2e}")
    assert e < BP_TOL
    assert rel_l2(sens.cpu().numpy(), sens_o) < BP_TOL


def test_c0_trace_lists_bitexact(cuda):
    grid, sc = cases.c0()
    xy = sc.xy_table()
    per = sc.n_radial * sc.n_angles
    case = O.SiddonCase(grid, sc.n_radial, sc.n_angles, xy, sc.z, plane_offset=31)
    geom = make_geom(grid, xy, sc.z, sc.n_radial, sc.n_angles)
    c_g, i_g, l_g = trace(geom, 31, per, cuda)
    c_o, i_o, l_o, _ = case.trace(per)
    assert np.array_equal(c_g, c_o)
    assert np.array_equal(i_g, i_o)
    assert np.array_equal(l_g, l_o)


@pytest.mark.parametrize("x_fastest_only", [False, True])
def test_3d_ring_pairs_fp_bitexact(cuda, x_fastest_only):
    """Oblique ring-pair planes through the one-lane-per-LOR kernels, with LORs closer to the x axis
    backprojecting into y-fastest buffers (default) or everything in the x-fastest image: FP
    bit-identical to the oracle either way, BP within 1e-5, and the fused data step within 1e-5."""
    import torch
    grid, sc = cases.c2_small()
    xy = sc.xy_table()
    per = sc.n_radial * sc.n_angles
    n = sc.n_planes * per
    print(f"[siddon 3d rings] planes={sc.n_planes} n={n} x_fastest_only={x_fastest_only}", flush=True)
    case = O.SiddonCase(grid, sc.n_radial, sc.n_angles, xy, sc.z)
    geom = make_geom(grid, xy, sc.z, sc.n_radial, sc.n_angles, x_fastest_only=x_fastest_only)
    img = _img(grid.N, 5)
    y = fp(geom, 0, n, torch.from_numpy(img).to(cuda)).cpu().numpy()
    assert np.array_equal(y, case.fp(img, n))
    yin = _img(n, 6)
    rhs, sens = bp(geom, 0, n, torch.from_numpy(yin).to(cuda), grid.N, sens=True)
    rhs_o, sens_o = case.bp(yin, n, sens=True)
    assert rel_l2(rhs.cpu().numpy(), rhs_o) < BP_TOL
    assert rel_l2(sens.cpu().numpy(), sens_o) < BP_TOL
    meas = (yin * 5).astype(np.float32)
    rhs_t = torch.zeros(grid.N, device=cuda)
    sens_t = torch.zeros(grid.N, device=cuda)
    img_t = torch.from_numpy(img).to(cuda)
    meas_t = torch.from_numpy(meas).to(cuda)
    _lib.call("omg_osem_data_step", geom.handle, 0, n, ptr(img_t), ptr(meas_t), ptr(rhs_t), ptr(sens_t), 1e-5, None)
    r_o = O.ratio(meas, case.fp(img, n), 1e-5)
    rhs_o, sens_o = case.bp(r_o, n, sens=True)
    assert rel_l2(rhs_t.cpu().numpy(), rhs_o) < BP_TOL
    assert rel_l2(sens_t.cpu().numpy(), sens_o) < BP_TOL


def test_3d_rings_non_square_grid(cuda):
    """Nx != Ny (40 x 24 x 12 voxels): the y-fastest accumulators and their transposed adds cover
    rectangular slices; FP bit-identical to the oracle, BP and sensitivity within 1e-5."""
    import torch
    grid = G.Grid.from_fov(40, 24, 12, 100.0, 60.0, 24.0)
    sc = G.pet_3d_rings(radius=150.0, n_rings=6, ring_pitch=4.0, crystals=96, max_ring_diff=5)
    xy = sc.xy_table()
    per = sc.n_radial * sc.n_angles
    n = sc.n_planes * per
    print(f"[siddon 3d non-square] n={n}", flush=True)
    case = O.SiddonCase(grid, sc.n_radial, sc.n_angles, xy, sc.z)
    geom = make_geom(grid, xy, sc.z, sc.n_radial, sc.n_angles)
    img = _img(grid.N, 15)
    y = fp(geom, 0, n, torch.from_numpy(img).to(cuda)).cpu().numpy()
    assert np.array_equal(y, case.fp(img, n))
    yin = _img(n, 16)
    rhs, sens = bp(geom, 0, n, torch.from_numpy(yin).to(cuda), grid.N, sens=True)
    rhs_o, sens_o = case.bp(yin, n, sens=True)
    assert rel_l2(rhs.cpu().numpy(), rhs_o) < BP_TOL
    assert rel_l2(sens.cpu().numpy(), sens_o) < BP_TOL


def test_ct_flat_panel_fp_bitexact(cuda):
    import torch
    grid, x, uv, rows, cols, pitch = cases.ct_small()
    nv = x.size // 6
    n = nv * rows * cols
    case = O.SiddonCase(grid, rows, cols, x, uv, mode=O.MODE_CT, dPitchZ=pitch, n_proj=nv)
    geom = make_geom(grid, x, uv, rows, cols, mode=_lib.MODE_CT_FLAT, n_proj=nv, dPitchZ=pitch)
    img = _img(grid.N, 7)
    y = fp(geom, 0, n, torch.from_numpy(img).to(cuda)).cpu().numpy()
    assert np.array_equal(y, case.fp(img, n))
    yin = _img(n, 8)
    rhs = bp(geom, 0, n, torch.from_numpy(yin).to(cuda), grid.N)
    assert rel_l2(rhs.cpu().numpy(), case.bp(yin, n)) < BP_TOL


def test_adjoint_identity_c0_subset(cuda):
    """<A x, y> == <x, A^T y> in FP64 accumulation (FP and BP trace identical rays)."""
    import torch
    grid, sc = cases.c0()
    xy = sc.xy_table()
    per = sc.n_radial * sc.n_angles
    geom = make_geom(grid, xy, sc.z, sc.n_radial, sc.n_angles)
    n = 8 * per
    for seed in range(3):
        x = _img(grid.N, 100 + seed)
        yv = _img(n, 200 + seed)
        ax = fp(geom, 4, n, torch.from_numpy(x).to(cuda)).cpu().numpy().astype(np.float64)
        aty = bp(geom, 4, n, torch.from_numpy(yv).to(cuda), grid.N).cpu().numpy().astype(np.float64)
        lhs = float(ax @ yv.astype(np.float64))
        rhs = float(x.astype(np.float64) @ aty)
        assert abs(lhs - rhs) / abs(lhs) < 1e-5


@pytest.mark.parametrize("radius", [100.0, 400.0])
def test_planar_endpoints_inside_and_outside(cuda, radius):
    """Direct-plane sinograms whose LOR endpoints lie inside the image (ring radius 100 mm, image
    half-width 80 mm: the planar kernels keep the reference's full length expression) and outside it
    (400 mm: every crossing parameter is in [0, 1] and the short expression applies). FP and the
    traversal lists bit-identical to the oracle either way; BP, sensitivity and the fused
    FP -> ratio -> BP step within 1e-5."""
    import torch
    grid = G.Grid.from_fov(64, 64, 4, 160.0, 160.0, 10.0)
    sc = G.pet_2d_planes(grid, radius=radius, n_radial=64, n_angles=48, radial_pitch=2.5)
    xy = sc.xy_table()
    per = sc.n_radial * sc.n_angles
    n = 4 * per
    print(f"[planar] radius={radius} LORs={n}", flush=True)
    case = O.SiddonCase(grid, sc.n_radial, sc.n_angles, xy, sc.z)
    geom = make_geom(grid, xy, sc.z, sc.n_radial, sc.n_angles)
    img = _img(grid.N, 31)
    y = fp(geom, 0, n, torch.from_numpy(img).to(cuda)).cpu().numpy()
    assert np.array_equal(y, case.fp(img, n))
    c_g, i_g, l_g = trace(geom, 0, n, cuda)
    c_o, i_o, l_o, _ = case.trace(n)
    assert np.array_equal(c_g, c_o) and np.array_equal(i_g, i_o) and np.array_equal(l_g, l_o)
    yin = _img(n, 32)
    rhs, sens = bp(geom, 0, n, torch.from_numpy(yin).to(cuda), grid.N, sens=True)
    rhs_o, sens_o = case.bp(yin, n, sens=True)
    assert rel_l2(rhs.cpu().numpy(), rhs_o) < BP_TOL
    assert rel_l2(sens.cpu().numpy(), sens_o) < BP_TOL
    meas = (yin * 3).astype(np.float32)
    rhs_t = torch.zeros(grid.N, device=cuda)
    sens_t = torch.zeros(grid.N, device=cuda)
    img_t, meas_t = torch.from_numpy(img).to(cuda), torch.from_numpy(meas).to(cuda)
    _lib.call("omg_osem_data_step", geom.handle, 0, n, ptr(img_t), ptr(meas_t), ptr(rhs_t), ptr(sens_t), 1e-5, None)
    r_o = O.ratio(meas, case.fp(img, n), 1e-5)
    rhs_o, sens_o = case.bp(r_o, n, sens=True)
    assert rel_l2(rhs_t.cpu().numpy(), rhs_o) < BP_TOL
    assert rel_l2(sens_t.cpu().numpy(), sens_o) < BP_TOL


def test_norm_and_global_factor(cuda):
    import torch
    grid, sc = cases.c0()
    xy = sc.xy_table()
    per = sc.n_radial * sc.n_angles
    norm = np.random.default_rng(4).uniform(0.5, 1.5, size=sc.n_planes * per).astype(np.float32)
    geom = make_geom(grid, xy, sc.z, sc.n_radial, sc.n_angles, norm=norm, global_factor=1.7)
    case = O.SiddonCase(grid, sc.n_radial, sc.n_angles, xy, sc.z, plane_offset=10, norm=norm[10 * per:12 * per],
                        global_factor=1.7)
    img = _img(grid.N, 9)
    y = fp(geom, 10, 2 * per, torch.from_numpy(img).to(cuda)).cpu().numpy()
    assert np.array_equal(y, case.fp(img, 2 * per))


def test_errors_are_reported(cuda):
    import torch
    grid, sc = cases.c0()
    geom = make_geom(grid, sc.xy_table(), sc.z, sc.n_radial, sc.n_angles)
    img = torch.zeros(grid.N, device=cuda)
    with pytest.raises(_lib.OmegaHipError) as e:
        fp(geom, 62, 2 * sc.n_radial * sc.n_angles, img)   # runs past the last plane
    assert e.value.code == _lib.OMG_ERR_ARG
    with pytest.raises(_lib.OmegaHipError):
        fp(geom, 0, 100, img)                              # not a whole plane


def test_c2_geometry_fp32_vs_fp64(cuda):
    """Full 3-D ring geometry (224 x 224 x 1024 sinogram, 128 x 128 x 63 image), 8 planes of it: the FP32
    forward projection bit-identical to the FP32 oracle; its error against the FP64 oracle is stated."""
    import torch
    from omega_amd import workloads
    p = workloads.pet3d_projector(subsets=1)
    p.initProj()
    per = int(p.nRowsD) * int(p.nColsD)
    first, planes = 500, 8
    n = per * planes
    img = _img(p.grid.N, 9)
    y = fp(p.geometry, first, n, torch.from_numpy(img).to(cuda)).cpu().numpy()
    c32 = O.SiddonCase(p.grid, int(p.nRowsD), int(p.nColsD), p.x, p.z, plane_offset=first)
    c64 = O.SiddonCase(p.grid, int(p.nRowsD), int(p.nColsD), p.x, p.z, plane_offset=first, dtype=np.float64)
    assert np.array_equal(y, c32.fp(img, n))
    e = rel_l2(y, c64.fp(img.astype(np.float64), n))
    print(f"[c2 fp] FP32 vs FP64 relative L2 {e:.2e} over {n} LORs", flush=True)
    assert e < 1e-5
