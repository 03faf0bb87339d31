"""GPU Siddon projections with several rays per LOR together with time of flight and with an attenuation
image, against the reference's own CPU outputs (tests/golden/siddon_combos.npz, written by
tests/golden/make_combos.py from the reference headers) and the oracle.

Bars, as for each feature alone: forward projections within 1e-6 relative per measurement (the TOF
Gaussian's double exponential and the attenuation factor's expf against glibc's can move a value by a
float ulp; the reference multiplies the attenuation factor in after norm and the global factor);
backprojections and sensitivities within 1e-6 relative L2 (float atomics against sequential sums) --
with an attenuation image only on sinograms, whose LORs end outside the image, since the reference's
backprojection integrates mu over the unclipped walk; fused OSEM subset iterations with TOF, an
attenuation image and three rays per LOR within 1e-5 of the oracle run.
"""
import os

import numpy as np
import pytest

from omega_amd import _lib
from omega_amd.recon import OSEMRunner
from oracle import oracle as O
from tests import cases
from tests.hip_helpers import bp, fp, make_geom, rel_l2

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "siddon_combos.npz")
EPPS = 1e-5


def _fp_close(got, ref):
    rel = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-30)
    print(f"  bit-identical {np.mean(got == ref):.4f}, max rel {rel.max():.2e}", flush=True)
    assert np.all(np.abs(got - ref) <= 1e-6 * np.abs(ref))


@pytest.mark.parametrize("name", ["tof", "atn", "tofatn"])
def test_listmode_against_reference(cuda, name):
    import torch
    f = np.load(GOLD, allow_pickle=False)
    grid = cases.tiny_grid()
    x = f["list_x"]
    n = x.size // 6
    kw = {}
    if "tof" in name:
        kw["tof"] = (5, 12.0, f["list_tofc"], EPPS)
    if "atn" in name:
        kw.update(atten=f["list_mu"], atten_mode=1)
    print(f"[combos list-mode {name}] n={n} rays=3x2", flush=True)
    geom = make_geom(grid, x, None, 1, 1, mode=_lib.MODE_LIST6, rays=(3, 2, 0.9, 1.1), **kw)
    _fp_close(fp(geom, 0, n, torch.from_numpy(f["list_img"]).to(cuda)).cpu().numpy(), f[f"list_{name}_fp32"])
    if name == "tof":
        rhs, sens = bp(geom, 0, n, torch.from_numpy(f["list_tof_yin"]).to(cuda), grid.N, sens=True)
        assert rel_l2(rhs.cpu().numpy(), f["list_tof_bp32"]) < 1e-6
        assert rel_l2(sens.cpu().numpy(), f["list_tof_sens32"]) < 1e-6


@pytest.mark.parametrize("generic", [False, True])
def test_sinogram_tof_atten_rays_against_reference(cuda, generic):
    import torch
    f = np.load(GOLD, allow_pickle=False)
    g2, sc = cases.orth_pet2d()
    planes = int(f["pet2d_planes"])
    n = sc.n_radial * sc.n_angles * planes
    print(f"[combos pet2d] n={n} rays=3 bins=7 attenuation image generic={generic}", flush=True)
    geom = make_geom(g2, sc.xy_table(), sc.z[: 2 * planes], sc.n_radial, sc.n_angles, rays=(3, 1, 2.5, 0.0),
                     tof=(7, 15.0, f["pet2d_tofc"], EPPS), atten=f["pet2d_mu"], atten_mode=1, generic=generic)
    _fp_close(fp(geom, 0, n, torch.from_numpy(f["pet2d_img"]).to(cuda)).cpu().numpy(), f["pet2d_fp32"])
    rhs, sens = bp(geom, 0, n, torch.from_numpy(f["pet2d_yin"]).to(cuda), g2.N, sens=True)
    assert rel_l2(rhs.cpu().numpy(), f["pet2d_bp32"]) < 1e-6
    assert rel_l2(sens.cpu().numpy(), f["pet2d_sens32"]) < 1e-6


def test_osem_with_tof_atten_rays(cuda):
    import torch
    from omega_amd.projector import projectorClass
    grid, sc = cases.orth_pet2d()
    p = projectorClass()
    p.Nx, p.Ny, p.Nz = grid.Nx, grid.Ny, grid.Nz
    p.FOVa_x, p.FOVa_y, p.axial_fov = grid.Nx * grid.dx, grid.Ny * grid.dy, grid.Nz * grid.dz
    p.x, p.z = sc.xy_table(), sc.z
    p.Ndist, p.Nang = sc.n_radial, sc.n_angles
    p.subsets, p.subsetType = 3, 8
    p.n_rays_transaxial, p.cr_pz, p.cr_p = 3, 2.5, 2.5
    p.TOF_bins, p.TOF_bins_used, p.TOF_width, p.TOF_FWHM = 7, 7, 100e-12, 300e-12
    rng = np.random.default_rng(23)
    mu = (0.01 * rng.random(grid.N)).astype(np.float32)
    p.attenuation_correction = True
    p.vaimennus = mu.reshape(grid.Nz, grid.Ny, grid.Nx).transpose(2, 1, 0)
    p.addProjector()
    p.initProj()
    xt = torch.from_numpy((rng.random(grid.N) * 2 + 0.2).astype(np.float32)).to(cuda)
    y = torch.cat([p.forwardProject(xt, k) for k in range(p.subsets)]).cpu().numpy()
    y = rng.poisson(y * 50.0).astype(np.float32)
    nb = p.nBins
    print(f"[combos osem] rays=3 bins={nb} attenuation image, values={y.size}", flush=True)
    run = OSEMRunner(p, torch.from_numpy(y).to(cuda))
    x = torch.ones(grid.N, device=cuda)
    n_sub, done = 5, 0
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
                subsets_to_run=n_sub, rays=(3, 1), ray_pitch=(2.5, 2.5), tof=tof, atten=p._attenuation()[0],
                atten_mode=1)
    e = rel_l2(x.cpu().numpy(), xo)
    print(f"  rel-L2 vs oracle {e:.2e}", flush=True)
    assert e < 1e-5
