"""GPU time-of-flight projections against the reference's own CPU outputs (tests/golden/siddon_tof.npz,
written by tests/golden/make_tof.py from the reference headers) and the oracle.

Bars: forward projections within 1e-6 relative per measurement of the reference (the Gaussian's
exponential is taken in double, as there, and rounded to float; a last-bit difference of the two
libraries' double exp can move a weight by one float ulp), backprojections and sensitivity within 1e-6
relative L2 (float atomics against the reference's sequential sums), fused OSEM subset iterations with
TOF (and a randoms term added to every bin) within 1e-5 of the oracle run.
"""
import os

import numpy as np
import pytest

from omega_amd import _lib
from omega_amd import geometry as G
from omega_amd.recon import OSEMRunner
from oracle import oracle as O
from tests import cases
from tests.hip_helpers import bp, fp, make_geom, rel_l2

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
EPPS = 1e-5


def _load():
    return np.load(os.path.join(GOLD, "siddon_tof.npz"), allow_pickle=False)


def _tof_geom(grid, x, z, size_x, size_y, bins, sigma, centers, mode=_lib.MODE_PET_SINO, **kw):
    return make_geom(grid, x, z, size_x, size_y, mode=mode, tof=(bins, float(sigma), centers, EPPS), **kw)


def _fp_close(got, ref):
    rel = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-30)
    print(f"  bit-identical {np.mean(got == ref):.4f}, max rel {rel.max():.2e}", flush=True)
    assert np.all(np.abs(got - ref) <= 1e-6 * np.abs(ref))


def test_listmode_against_reference(cuda):
    import torch
    f = _load()
    grid = cases.tiny_grid()
    x = f["list_x"]
    n = x.size // 6
    nb = int(f["list_bins"])
    print(f"[tof list-mode] n={n} bins={nb}", flush=True)
    geom = _tof_geom(grid, x, None, 1, 1, nb, f["list_sigma"], f["list_centers"], mode=_lib.MODE_LIST6)
    img = torch.from_numpy(f["list_img"]).to(cuda)
    _fp_close(fp(geom, 0, n, img).cpu().numpy(), f["list_fp32"])
    y = torch.from_numpy(f["list_yin"]).to(cuda)
    rhs, sens = bp(geom, 0, n, y, grid.N, sens=True)
    assert rel_l2(rhs.cpu().numpy(), f["list_bp32"]) < 1e-6
    assert rel_l2(sens.cpu().numpy(), f["list_sens32"]) < 1e-6
    geom = _tof_geom(grid, x, None, 1, 1, nb, f["list_sigma"], f["list_centers"], mode=_lib.MODE_LIST6,
                     norm=f["list_norm"], atten=f["list_fac"], atten_mode=2, scat=f["list_scat"], global_factor=0.9)
    _fp_close(fp(geom, 0, n, img).cpu().numpy(), f["list_all_fp32"])
    assert rel_l2(bp(geom, 0, n, y, grid.N).cpu().numpy(), f["list_all_bp32"]) < 1e-6


def test_sinograms_against_reference(cuda):
    import torch
    f = _load()
    g2, sc = cases.orth_pet2d()
    planes = int(f["pet2d_planes"])
    nb = int(f["pet2d_bins"])
    n = sc.n_radial * sc.n_angles * planes
    print(f"[tof pet2d] n={n} bins={nb}", flush=True)
    geom = _tof_geom(g2, sc.xy_table(), sc.z[: 2 * planes], sc.n_radial, sc.n_angles, nb, f["pet2d_sigma"],
                     f["pet2d_centers"])
    _fp_close(fp(geom, 0, n, torch.from_numpy(f["pet2d_img"]).to(cuda)).cpu().numpy(), f["pet2d_fp32"])
    rhs, sens = bp(geom, 0, n, torch.from_numpy(f["pet2d_yin"]).to(cuda), g2.N, sens=True)
    assert rel_l2(rhs.cpu().numpy(), f["pet2d_bp32"]) < 1e-6
    assert rel_l2(sens.cpu().numpy(), f["pet2d_sens32"]) < 1e-6
    g3, sc3 = cases.c2_small()
    planes3 = int(f["pet3d_planes"])
    nb3 = int(f["pet3d_bins"])
    n3 = sc3.n_radial * sc3.n_angles * planes3
    print(f"[tof pet3d] n={n3} bins={nb3}", flush=True)
    geom = _tof_geom(g3, sc3.xy_table(), sc3.z[: 2 * planes3], sc3.n_radial, sc3.n_angles, nb3, f["pet3d_sigma"],
                     f["pet3d_centers"])
    _fp_close(fp(geom, 0, n3, torch.from_numpy(f["pet3d_img"]).to(cuda)).cpu().numpy(), f["pet3d_fp32"])


def _tof_projector(subsets=3):
    from omega_amd.projector import projectorClass
    grid, sc = cases.orth_pet2d()
    p = projectorClass()
    p.Nx, p.Ny, p.Nz = grid.Nx, grid.Ny, grid.Nz
    p.FOVa_x, p.FOVa_y, p.axial_fov = grid.Nx * grid.dx, grid.Ny * grid.dy, grid.Nz * grid.dz
    p.x, p.z = sc.xy_table(), sc.z
    p.Ndist, p.Nang = sc.n_radial, sc.n_angles
    p.subsets, p.subsetType = subsets, 8
    p.TOF_bins, p.TOF_bins_used, p.TOF_width, p.TOF_FWHM = 7, 7, 100e-12, 300e-12
    p.addProjector()
    p.initProj()
    return grid, sc, p


@pytest.mark.parametrize("randoms", [False, True])
def test_osem_with_tof(cuda, randoms):
    import torch
    grid, sc, p = _tof_projector()
    nb = p.nBins
    rng = np.random.default_rng(21)
    xt = torch.from_numpy((rng.random(grid.N) * 2 + 0.2).astype(np.float32)).to(cuda)
    yk = [p.forwardProject(xt, k) for k in range(p.subsets)]
    y = torch.cat(yk).cpu().numpy()
    y = rng.poisson(y * 50.0).astype(np.float32)
    nlor = int(p.nTotMeas[-1])
    add = (rng.random(nlor) * 0.5 + 0.05).astype(np.float32) if randoms else None
    print(f"[tof osem] bins={nb} LORs={nlor} randoms={randoms}", flush=True)
    run = OSEMRunner(p, torch.from_numpy(y).to(cuda), additive=None if add is None else torch.from_numpy(add).to(cuda))
    x = torch.ones(grid.N, device=cuda)
    n_sub = 5
    done = 0
    for it in range(2):
        for k in range(p.subsets):
            if done < n_sub:
                run.subset_step(x, k)
                done += 1
        run.sens_done = True
    torch.cuda.synchronize()
    nmeas = np.diff(p.nMeas).astype(np.int64)
    tof = dict(tof_bins=nb, tof_sigma=np.float32(p.sigma_x), tof_centers=p.TOFCenter, tof_epps=float(p.epps))
    xo = O.osem(p.grid, p.nRowsD, p.nColsD, p.x, p.z, nmeas, y, 2, np.ones(grid.N, np.float32), epps=1e-5,
                subsets_to_run=n_sub, additive=add, tof=tof)
    # the same data with its bins summed, reconstructed without TOF
    per = p.nRowsD * p.nColsD
    y_sum = np.concatenate([c.reshape(nb, -1).sum(0) for c in np.split(y, nb * np.cumsum(nmeas * per)[:-1])])
    plain = O.osem(p.grid, p.nRowsD, p.nColsD, p.x, p.z, nmeas, y_sum.astype(np.float32), 2,
                   np.ones(grid.N, np.float32), epps=1e-5, subsets_to_run=n_sub, additive=add)
    e = rel_l2(x.cpu().numpy(), xo)
    print(f"  rel-L2 vs oracle {e:.2e}; TOF vs summed bins {rel_l2(xo, plain):.2e}", flush=True)
    assert e < 1e-5
    assert rel_l2(xo, plain) > 1e-3


def test_projector_class_tof(cuda):
    import torch
    grid, sc, p = _tof_projector()
    assert p.nBins == 7
    np.testing.assert_allclose(p.TOFCenter, G.tof_centers(7, 100e-12))
    geom = _tof_geom(p.grid, p.x, p.z, sc.n_radial, sc.n_angles, 7, p.sigma_x, p.TOFCenter)
    img = torch.rand(grid.N, device=cuda) + 0.1
    for k in range(p.subsets):
        y = p.forwardProject(img, k)
        n = y.numel() // 7
        assert np.array_equal(y.cpu().numpy(), fp(geom, int(p.nMeas[k]), n, img).cpu().numpy())
        b = p.backwardProject(y, k)
        assert rel_l2(b.cpu().numpy(), bp(geom, int(p.nMeas[k]), n, y, grid.N).cpu().numpy()) < 1e-6
