"""Attenuation correction in the CPU oracle, pinned against the reference's own CPU code and closed forms.

Fixtures: tests/golden/make_attenuation.py (reference headers compiled as oracle/_ref, attenuation
image via paramStruct::CTAttenuation and per-measurement factors, alone and with normalisation,
multiplicative scatter correction and global factor; single-threaded). Bar: forward
projections bit-identical in float and double, single-threaded backprojections bit-identical; a
uniform attenuation image gives exp(-mu * chord) with the chord from slab clipping in float64
(2e-5 relative: float sums along ~50 voxels); mu = 0 changes nothing. The reference's
backprojection integrates mu over the unclipped walk (t_next - t_c) * L while its forward projection
uses the clipped lengths (projector_functions.h:1754-1775 vs :1938-1939): the two are adjoint for LORs
whose endpoints lie outside the image and not otherwise, which is checked here too.
"""
import os

import numpy as np

from oracle import oracle as O
from tests import cases

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load():
    return np.load(os.path.join(GOLD, "siddon_attenuation.npz"), allow_pickle=False)


def _list_case(f, dtype=np.float32, mode=1):
    grid = cases.tiny_grid()
    atn = f["list_mu"] if mode == 1 else f["list_fac"]
    return grid, O.SiddonCase(grid, 1, 1, f["list_x"], np.zeros(2, np.float32), mode=O.MODE_LIST,
                              atten=atn.astype(dtype), atten_mode=mode, dtype=dtype)


def test_listmode_matches_reference_code():
    f = _load()
    grid, c32 = _list_case(f)
    n = f["list_x"].size // 6
    assert np.array_equal(c32.fp(f["list_img"], n, 1), f["list_fp32"])
    _, c64 = _list_case(f, np.float64)
    assert np.array_equal(c64.fp(f["list_img"].astype(np.float64), n, 1), f["list_fp64"])
    bp, s = c32.bp(f["list_yin"], n, sens=True, nthreads=1)
    assert np.array_equal(bp, f["list_bp32"]) and np.array_equal(s, f["list_sens32"])
    _, l32 = _list_case(f, mode=2)
    assert np.array_equal(l32.fp(f["list_img"], n, 1), f["list_lor_fp32"])
    bp, s = l32.bp(f["list_yin"], n, sens=True, nthreads=1)
    assert np.array_equal(bp, f["list_lor_bp32"]) and np.array_equal(s, f["list_lor_sens32"])
    # with norm, multiplicative scatter and global factor (their order with the attenuation factor)
    kw = dict(mode=O.MODE_LIST, norm=f["list_norm"], scat=f["list_scat"], global_factor=0.9)
    z0 = np.zeros(2, np.float32)
    a32 = O.SiddonCase(grid, 1, 1, f["list_x"], z0, atten=f["list_fac"], atten_mode=2, **kw)
    i32 = O.SiddonCase(grid, 1, 1, f["list_x"], z0, atten=f["list_mu"], atten_mode=1, **kw)
    assert np.array_equal(a32.fp(f["list_img"], n, 1), f["list_all_fp32"])
    assert np.array_equal(i32.fp(f["list_img"], n, 1), f["list_img_all_fp32"])
    assert np.array_equal(a32.bp(f["list_yin"], n, nthreads=1), f["list_all_bp32"])
    assert np.array_equal(i32.bp(f["list_yin"], n, nthreads=1), f["list_img_all_bp32"])
    # the attenuation changes the projections (the fixture is not the uncorrected one)
    plain = O.SiddonCase(grid, 1, 1, f["list_x"], np.zeros(2, np.float32), mode=O.MODE_LIST)
    assert not np.array_equal(plain.fp(f["list_img"], n, 1), f["list_fp32"])


def test_sinograms_match_reference_code():
    f = _load()
    g2, sc = cases.orth_pet2d()
    n2 = sc.n_radial * sc.n_angles * sc.n_planes
    p32 = O.SiddonCase(g2, sc.n_radial, sc.n_angles, sc.xy_table(), sc.z, atten=f["pet2d_mu"], atten_mode=1)
    assert np.array_equal(p32.fp(f["pet2d_img"], n2, 1), f["pet2d_fp32"])
    bp, s = p32.bp(f["pet2d_yin"], n2, sens=True, nthreads=1)
    assert np.array_equal(bp, f["pet2d_bp32"]) and np.array_equal(s, f["pet2d_sens32"])
    g3, sc3 = cases.c2_small()
    planes = int(f["pet3d_planes"])
    n3 = sc3.n_radial * sc3.n_angles * planes
    q32 = O.SiddonCase(g3, sc3.n_radial, sc3.n_angles, sc3.xy_table(), sc3.z[: 2 * planes], atten=f["pet3d_mu"],
                       atten_mode=1)
    assert np.array_equal(q32.fp(f["pet3d_img"], n3, 1), f["pet3d_fp32"])


def test_zero_attenuation_changes_nothing():
    f = _load()
    grid = cases.tiny_grid()
    x = f["list_x"]
    n = x.size // 6
    z0 = np.zeros(2, np.float32)
    c = O.SiddonCase(grid, 1, 1, x, z0, mode=O.MODE_LIST, atten=np.zeros(grid.N, np.float32), atten_mode=1)
    plain = O.SiddonCase(grid, 1, 1, x, z0, mode=O.MODE_LIST)
    assert np.all(c.atten_factors(n) == 1.0)
    assert np.array_equal(c.fp(f["list_img"], n, 1), plain.fp(f["list_img"], n, 1))
    assert np.array_equal(c.bp(f["list_yin"], n, nthreads=1), plain.bp(f["list_yin"], n, nthreads=1))


def _chords_2d(xy, grid):
    """In-plane chord of each [xs, ys, xd, yd] segment through the image box, float64 slab clipping."""
    t = xy.reshape(-1, 4).astype(np.float64)
    s, d = t[:, :2], t[:, 2:]
    lo = np.array([grid.bx, grid.by]); hi = lo + np.array([grid.Nx * grid.dx, grid.Ny * grid.dy])
    t0 = np.zeros(len(t)); t1 = np.ones(len(t))
    for k in range(2):
        dd = d[:, k] - s[:, k]
        with np.errstate(divide="ignore", invalid="ignore"):
            ta = (lo[k] - s[:, k]) / dd
            tb = (hi[k] - s[:, k]) / dd
        par = dd == 0
        inside = (s[:, k] >= lo[k]) & (s[:, k] <= hi[k])
        ta = np.where(par, np.where(inside, -np.inf, np.inf), ta)
        tb = np.where(par, np.where(inside, np.inf, -np.inf), tb)
        t0 = np.maximum(t0, np.minimum(ta, tb))
        t1 = np.minimum(t1, np.maximum(ta, tb))
    return np.maximum(t1 - t0, 0.0) * np.linalg.norm(d - s, axis=1)


def test_uniform_attenuation_closed_form():
    g2, sc = cases.orth_pet2d()
    c = 0.0123
    per = sc.n_radial * sc.n_angles
    case = O.SiddonCase(g2, sc.n_radial, sc.n_angles, sc.xy_table(), sc.z[:2], atten=np.full(g2.N, c, np.float32),
                        atten_mode=1)
    f = case.atten_factors(per).astype(np.float64)
    chord = _chords_2d(sc.xy_table(), g2)
    hit = chord > 1e-3
    assert hit.sum() > 0.9 * per
    np.testing.assert_allclose(f[hit], np.exp(-c * chord[hit]), rtol=2e-5)
    assert np.all(f[~hit] == 1.0)


def test_backprojection_prepass_is_unclipped():
    """The reference's BP attenuation walks the whole grid traversal: adjoint to its FP for LORs with
    both endpoints outside the image, not for LORs that start or end inside."""
    f = _load()
    grid, c64 = _list_case(f, np.float64)
    x = f["list_x"].reshape(-1, 6).astype(np.float64)
    n = len(x)
    lo = np.array([grid.bx, grid.by, grid.bz]); hi = lo + np.array([grid.Nx * grid.dx, grid.Ny * grid.dy, grid.Nz * grid.dz])
    inside = lambda p: np.all((p > lo) & (p < hi), axis=1)  # noqa: E731
    ends_in = inside(x[:, :3]) | inside(x[:, 3:])
    rng = np.random.default_rng(5)
    img = rng.random(grid.N)
    mism = []
    for i in range(n):
        e = np.zeros(n); e[i] = 1.0
        a = c64.fp(img, n, 1)[i]
        b = float(img @ c64.bp(e, n, nthreads=1))
        mism.append(abs(a - b) > 1e-12 * max(abs(a), 1e-300))
    mism = np.array(mism)
    assert not np.any(mism[~ends_in])
    assert np.any(mism[ends_in])
