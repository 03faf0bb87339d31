"""Time of flight in the CPU oracle, pinned against the reference's own CPU code and invariants.

Fixtures: tests/golden/make_tof.py (reference headers compiled as oracle/_ref with paramStruct::TOF;
single-threaded). Bar: forward projections bit-identical in float and double, single-threaded
backprojections bit-identical, also with normalisation, attenuation, scatter and global factor.
Invariants the TOF weighting must keep (each voxel's bin weights are normalised by their sum):
  * the bins of a forward projection add up to the non-TOF forward projection (1e-5, float sums);
  * with a standard deviation far beyond the image (3000 mm) every bin carries 1/bins of it (1e-3,
    the Gaussian's own variation across the image);
  * in float64 the TOF forward and backprojection are adjoint (1e-12).
"""
import os

import numpy as np

from omega_amd import geometry as G
from oracle import oracle as O
from tests import cases

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
EPPS = 1e-5


def _load():
    return np.load(os.path.join(GOLD, "siddon_tof.npz"), allow_pickle=False)


def _list(f, dtype=np.float32, **kw):
    grid = cases.tiny_grid()
    return grid, O.SiddonCase(grid, 1, 1, f["list_x"], np.zeros(2, np.float32), mode=O.MODE_LIST, dtype=dtype,
                              tof_bins=int(f["list_bins"]), tof_sigma=dtype(f["list_sigma"]),
                              tof_centers=f["list_centers"].astype(dtype), tof_epps=EPPS, **kw)


def test_listmode_matches_reference_code():
    f = _load()
    n = f["list_x"].size // 6
    grid, c32 = _list(f)
    assert np.array_equal(c32.fp(f["list_img"], n, 1), f["list_fp32"])
    _, c64 = _list(f, np.float64)
    assert np.array_equal(c64.fp(f["list_img"].astype(np.float64), n, 1), f["list_fp64"])
    bp, s = c32.bp(f["list_yin"], n, sens=True, nthreads=1)
    assert np.array_equal(bp, f["list_bp32"]) and np.array_equal(s, f["list_sens32"])
    _, a32 = _list(f, norm=f["list_norm"], atten=f["list_fac"], atten_mode=2, scat=f["list_scat"], global_factor=0.9)
    assert np.array_equal(a32.fp(f["list_img"], n, 1), f["list_all_fp32"])
    assert np.array_equal(a32.bp(f["list_yin"], n, nthreads=1), f["list_all_bp32"])


def test_sinograms_match_reference_code():
    f = _load()
    g2, sc = cases.orth_pet2d()
    planes = int(f["pet2d_planes"])
    n2 = sc.n_radial * sc.n_angles * planes
    p32 = O.SiddonCase(g2, sc.n_radial, sc.n_angles, sc.xy_table(), sc.z[: 2 * planes], tof_bins=int(f["pet2d_bins"]),
                       tof_sigma=np.float32(f["pet2d_sigma"]), tof_centers=f["pet2d_centers"], tof_epps=EPPS)
    assert np.array_equal(p32.fp(f["pet2d_img"], n2, 1), f["pet2d_fp32"])
    bp, s = p32.bp(f["pet2d_yin"], n2, sens=True, nthreads=1)
    assert np.array_equal(bp, f["pet2d_bp32"]) and np.array_equal(s, f["pet2d_sens32"])
    g3, sc3 = cases.c2_small()
    planes3 = int(f["pet3d_planes"])
    n3 = sc3.n_radial * sc3.n_angles * planes3
    q32 = O.SiddonCase(g3, sc3.n_radial, sc3.n_angles, sc3.xy_table(), sc3.z[: 2 * planes3],
                       tof_bins=int(f["pet3d_bins"]), tof_sigma=np.float32(f["pet3d_sigma"]),
                       tof_centers=f["pet3d_centers"], tof_epps=EPPS)
    assert np.array_equal(q32.fp(f["pet3d_img"], n3, 1), f["pet3d_fp32"])


def _pet2d(bins, sigma, width_s):
    g2, sc = cases.orth_pet2d()
    n = sc.n_radial * sc.n_angles * 2
    tc = G.tof_centers(bins, width_s)
    tof = O.SiddonCase(g2, sc.n_radial, sc.n_angles, sc.xy_table(), sc.z[:4], tof_bins=bins, tof_sigma=sigma,
                       tof_centers=tc, tof_epps=EPPS)
    plain = O.SiddonCase(g2, sc.n_radial, sc.n_angles, sc.xy_table(), sc.z[:4])
    return g2, n, tof, plain


def test_bins_add_up_to_the_plain_projection():
    g2, n, tof, plain = _pet2d(15, np.float32(G.tof_sigma(200e-12)), 100e-12)   # 15 bins of 15 mm cover the FOV
    img = (np.random.default_rng(3).random(g2.N) + 0.1).astype(np.float32)
    y = tof.fp(img, n).reshape(15, n).astype(np.float64)
    y0 = plain.fp(img, n).astype(np.float64)
    np.testing.assert_allclose(y.sum(0), y0, rtol=1e-5, atol=1e-6 * y0.max())
    assert (y > 0.05 * y0[None, :]).sum(0).max() < 15   # the bins do separate positions along the LOR


def test_wide_kernel_spreads_evenly():
    # sigma = 3000 mm: across the image and the bin centres the Gaussian varies by ~5e-4, and a voxel's
    # weight sum (~1.3e-3 per mm of length) falls under the 1e-5 floor only for clips shorter than
    # 8 um (wider kernels put whole voxels under the floor, which then scales them by sum / epps)
    g2, n, tof, plain = _pet2d(5, np.float32(3000.0), 100e-12)
    img = (np.random.default_rng(4).random(g2.N) + 0.1).astype(np.float32)
    y = tof.fp(img, n).reshape(5, n)
    y0 = plain.fp(img, n)
    np.testing.assert_allclose(y, np.broadcast_to(y0 / 5.0, y.shape), rtol=1e-3, atol=1e-7 * y0.max())


def test_tof_adjoint_float64():
    f = _load()
    n = f["list_x"].size // 6
    grid, c64 = _list(f, np.float64)
    rng = np.random.default_rng(8)
    x = rng.random(grid.N)
    y = rng.random(n * int(f["list_bins"]))
    lhs = float(c64.fp(x, n, 1) @ y)
    rhs = float(x @ c64.bp(y, n, nthreads=1))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_weight_sum_floor_damps_wide_kernels():
    """The floor of a voxel's weight sum at epps (projector_functions.h:583-584): with a 1e7 mm kernel
    every voxel's sum (~1e-7 per mm) is floored, so the bins carry far less than the projection."""
    g2, n, tof, plain = _pet2d(5, np.float32(1e7), 100e-12)
    img = (np.random.default_rng(5).random(g2.N) + 0.1).astype(np.float32)
    y = tof.fp(img, n).reshape(5, n).sum(0)
    y0 = plain.fp(img, n)
    hit = y0 > 0
    assert np.all(y[hit] < 0.2 * y0[hit])
