"""Several rays per LOR in the CPU oracle, pinned against the reference's own CPU code and invariants.

Fixtures: tests/golden/make_multiray.py (reference headers compiled as oracle/_ref with nRays2D /
nRays3D and dPitchXY / dPitchZ; single-threaded). Bar: forward projections bit-identical in float and
double, single-threaded backprojections bit-identical, also with normalisation, attenuation factors,
scatter and global factor. Invariants: with no spread (pitch 0) every ray is the single ray, so the
projection equals the one-ray projection (1e-6, the n-fold sum of value / n); a spread moves the
projection (it is not the one-ray one); float64 forward and backprojection are adjoint (1e-12).
"""
import os

import numpy as np

from oracle import oracle as O
from tests import cases

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load():
    return np.load(os.path.join(GOLD, "siddon_multiray.npz"), allow_pickle=False)


def _list(f, dtype=np.float32, **kw):
    grid = cases.tiny_grid()
    return grid, O.SiddonCase(grid, 1, 1, f["list_x"], np.zeros(2, np.float32), mode=O.MODE_LIST, dtype=dtype,
                              rays=(3, 2), ray_pitch=(0.9, 1.1), **kw)


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
    p32 = O.SiddonCase(g2, sc.n_radial, sc.n_angles, sc.xy_table(), sc.z[: 2 * planes], rays=(5, 1),
                       ray_pitch=(2.5, 0.0))
    assert np.array_equal(p32.fp(f["pet2d_img"], n2, 1), f["pet2d_fp32"])
    bp, s = p32.bp(f["pet2d_yin"], n2, sens=True, nthreads=1)
    assert np.array_equal(bp, f["pet2d_bp32"]) and np.array_equal(s, f["pet2d_sens32"])
    g3, sc3 = cases.c2_small()
    planes3 = int(f["pet3d_planes"])
    n3 = sc3.n_radial * sc3.n_angles * planes3
    q32 = O.SiddonCase(g3, sc3.n_radial, sc3.n_angles, sc3.xy_table(), sc3.z[: 2 * planes3], rays=(2, 3),
                       ray_pitch=(2.0, 4.0))
    assert np.array_equal(q32.fp(f["pet3d_img"], n3, 1), f["pet3d_fp32"])


def test_no_spread_is_the_single_ray():
    g2, sc = cases.orth_pet2d()
    n = sc.n_radial * sc.n_angles * 2
    img = (np.random.default_rng(1).random(g2.N) + 0.1).astype(np.float32)
    one = O.SiddonCase(g2, sc.n_radial, sc.n_angles, sc.xy_table(), sc.z[:4]).fp(img, n)
    same = O.SiddonCase(g2, sc.n_radial, sc.n_angles, sc.xy_table(), sc.z[:4], rays=(4, 1),
                        ray_pitch=(0.0, 0.0)).fp(img, n)
    spread = O.SiddonCase(g2, sc.n_radial, sc.n_angles, sc.xy_table(), sc.z[:4], rays=(4, 1),
                          ray_pitch=(2.5, 0.0)).fp(img, n)
    np.testing.assert_allclose(same, one, rtol=1e-6, atol=0)
    assert np.abs(spread - one).max() > 1e-3 * one.max()


def test_multiray_adjoint_float64():
    f = _load()
    n = f["list_x"].size // 6
    grid, c64 = _list(f, np.float64)
    rng = np.random.default_rng(9)
    x = rng.random(grid.N)
    y = rng.random(n)
    lhs = float(c64.fp(x, n, 1) @ y)
    rhs = float(x @ c64.bp(y, n, nthreads=1))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
