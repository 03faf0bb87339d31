"""Several rays per LOR together with time of flight and with an attenuation image, in the CPU oracle.

Fixtures: tests/golden/make_combos.py (the reference's own CPU code, oracle/_ref, single-threaded).
Bar: forward projections bit-identical in float and double, single-threaded backprojections and
sensitivities bit-identical. Invariants: the TOF bins of a multi-ray LOR add up to its non-TOF
projection when the bins cover the whole LOR (1e-5, the bins' weights are normalised per voxel); the
float64 multi-ray TOF forward and backprojection are adjoint (1e-12). With an attenuation image they
are not, in the reference either: its forward projection sums mu over the walk's clipped lengths and
its backprojection's pre-pass over unclipped ones (projector_functions.h:1754-1775, :1938-1939), which
differ for LORs that start inside the image, as these list-mode LORs do.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests import cases

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "siddon_combos.npz")


def _f():
    return np.load(GOLD, allow_pickle=False)


def _list_kw(f, name):
    tof = dict(tof_bins=5, tof_sigma=np.float32(12.0), tof_centers=f["list_tofc"], tof_epps=1e-5)
    atn = dict(atten=f["list_mu"], atten_mode=1)
    return {"tof": tof, "atn": atn, "tofatn": dict(tof, **atn)}[name]


@pytest.mark.parametrize("name", ["tof", "atn", "tofatn"])
def test_listmode_matches_reference_code(name):
    f = _f()
    grid = cases.tiny_grid()
    n = f["list_x"].size // 6
    kw = dict(mode=O.MODE_LIST, rays=(3, 2), ray_pitch=(0.9, 1.1), **_list_kw(f, name))
    z0 = np.zeros(2, np.float32)
    c32 = O.SiddonCase(grid, 1, 1, f["list_x"], z0, **kw)
    assert np.array_equal(c32.fp(f["list_img"], n, 1), f[f"list_{name}_fp32"])
    c64 = O.SiddonCase(grid, 1, 1, f["list_x"], z0, dtype=np.float64, **kw)
    assert np.array_equal(c64.fp(f["list_img"].astype(np.float64), n, 1), f[f"list_{name}_fp64"])
    bp, s = c32.bp(f[f"list_{name}_yin"], n, sens=True, nthreads=1)
    assert np.array_equal(bp, f[f"list_{name}_bp32"]) and np.array_equal(s, f[f"list_{name}_sens32"])


def test_sinogram_tof_attenuation_rays_match_reference_code():
    f = _f()
    g2, sc = cases.orth_pet2d()
    planes = int(f["pet2d_planes"])
    n2 = sc.n_radial * sc.n_angles * planes
    p32 = O.SiddonCase(g2, sc.n_radial, sc.n_angles, sc.xy_table(), sc.z[: 2 * planes], rays=(3, 1),
                       ray_pitch=(2.5, 0.0), tof_bins=7, tof_sigma=np.float32(15.0), tof_centers=f["pet2d_tofc"],
                       tof_epps=1e-5, atten=f["pet2d_mu"], atten_mode=1)
    assert np.array_equal(p32.fp(f["pet2d_img"], n2, 1), f["pet2d_fp32"])
    bp, s = p32.bp(f["pet2d_yin"], n2, sens=True, nthreads=1)
    assert np.array_equal(bp, f["pet2d_bp32"]) and np.array_equal(s, f["pet2d_sens32"])


def test_tof_bins_sum_to_the_multiray_projection():
    grid = cases.tiny_grid()
    x = cases.list_rays(grid, 200, seed=3)
    n = x.size // 6
    img = np.random.default_rng(4).random(grid.N) + 0.1
    z0 = np.zeros(2)
    # 41 bins 5 mm apart, sigma 3 mm: every voxel of the ~60 mm LORs lies well inside the bins' span
    kw = dict(mode=O.MODE_LIST, rays=(2, 2), ray_pitch=(1.0, 0.8), dtype=np.float64)
    tof = O.SiddonCase(grid, 1, 1, x, z0, tof_bins=41, tof_sigma=3.0, tof_centers=np.arange(-100.0, 101.0, 5.0),
                       tof_epps=1e-12, **kw)
    plain = O.SiddonCase(grid, 1, 1, x, z0, **kw)
    yt = tof.fp(img, n).reshape(41, n).sum(0)
    yp = plain.fp(img, n)
    assert np.max(np.abs(yt - yp)) <= 1e-5 * np.max(np.abs(yp))


def test_multiray_tof_adjoint_float64():
    grid = cases.tiny_grid()
    x = cases.list_rays(grid, 150, seed=9)
    n = x.size // 6
    rng = np.random.default_rng(10)
    c = O.SiddonCase(grid, 1, 1, x, np.zeros(2), mode=O.MODE_LIST, rays=(3, 2), ray_pitch=(0.9, 1.1),
                     dtype=np.float64, tof_bins=5, tof_sigma=12.0, tof_centers=np.linspace(-30.0, 30.0, 5),
                     tof_epps=1e-12)
    xv = rng.random(grid.N)
    yv = rng.random(5 * n)
    lhs = float(c.fp(xv, n) @ yv)
    rhs = float(xv @ c.bp(yv, n))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
