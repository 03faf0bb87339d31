"""GPU parity of the interpolation projector (projector type 4) against the oracle's restatement.

Bars: forward projections bit-identical to the oracle (same float operations: exact FP32 trilinear sampling,
the same step arithmetic) for list-mode rays, PET sinograms with corrections and CT panels, with fixed
interpolation lengths and the Joseph variant; emission backprojections within 1e-5 relative L2 (float
atomics); the voxel-driven CT backprojection within 1e-6 (no atomics; reported as max |diff|); with an
attenuation image the per-measurement factors come from the device's expf against the host's exp, so the
forward projection is held to 1e-6 relative per measurement. OSEM with projector 4 on emission data and the
transmission OSEM on CT data are within 1e-5 of the oracle's step sequence.
"""
import numpy as np
import pytest

from omega_amd import _lib
from omega_amd.projector import projectorClass
from omega_amd.recon import OSEMRunner
from oracle import oracle as O
from tests import cases
from tests.hip_helpers import bp, fp, make_geom, rel_l2
from tests.test_oracle_interp import _outside_rays

pytestmark = pytest.mark.gpu


def _rand(n, seed, lo=0.05, scale=1.0):
    return (np.random.default_rng(seed).random(n) * scale + lo).astype(np.float32)


def _geom4(grid, x, z, sx, sy, dl, mode, kerroin=0.0, **kw):
    return make_geom(grid, x, z, sx, sy, mode=mode, projector_type=4, interp=(np.float32(dl), kerroin), **kw)


@pytest.mark.parametrize("dl", [0.4, 1.7, -1.0])
def test_listmode_fp_bitexact_bp(cuda, dl):
    import torch
    grid = cases.tiny_grid()
    rays = np.concatenate([_outside_rays(grid, 300, seed=1), cases.list_rays(grid, 200, seed=2).reshape(-1, 6)])
    rays = rays[np.linalg.norm(rays[:, 3:] - rays[:, :3], axis=1) > 0]   # a zero-length ray has no samples
    x = np.ascontiguousarray(rays.ravel())
    n = len(rays)
    print(f"[interp list] dL={dl} n={n}", flush=True)
    case = O.InterpCase(grid, 1, 1, x, np.zeros(2, np.float32), dl, mode=O.MODE_LIST)
    geom = _geom4(grid, x, None, 1, 1, dl, _lib.MODE_LIST6)
    img = _rand(grid.N, 3)
    y = fp(geom, 0, n, torch.from_numpy(img).to(cuda)).cpu().numpy()
    yo = case.fp(img, n)
    assert np.array_equal(y, yo), f"max |diff| {np.abs(y - yo).max()}"
    yin = _rand(n, 4)
    rhs, sens = bp(geom, 0, n, torch.from_numpy(yin).to(cuda), grid.N, sens=True)
    ro, so = case.bp(yin, n, sens=True)
    assert rel_l2(rhs.cpu().numpy(), ro) < 1e-5 and rel_l2(sens.cpu().numpy(), so) < 1e-5


@pytest.mark.parametrize("atten_mode", [0, 1, 2])
def test_pet_sinogram_with_corrections(cuda, atten_mode):
    import torch
    grid, sc = cases.c2_small()
    x, z = sc.xy_table(), sc.z
    n = sc.n_radial * sc.n_angles * (z.size // 2)
    norm = _rand(n, 5, lo=0.5)
    kw = dict(norm=norm, global_factor=1.3)
    ok = dict(norm=norm, global_factor=1.3)
    if atten_mode == 1:
        mu = _rand(grid.N, 6, lo=0.0, scale=0.02)
        kw.update(atten=mu, atten_mode=1)
        fac = O.InterpCase(grid, sc.n_radial, sc.n_angles, x, z, 0.8).atten_factors(mu, n)
        ok.update(atten=fac, atten_image=True)
    elif atten_mode == 2:
        fac = _rand(n, 7, lo=0.3, scale=0.7)
        kw.update(atten=fac, atten_mode=2)
        ok.update(atten=fac)
    print(f"[interp pet] atten_mode={atten_mode} n={n}", flush=True)
    case = O.InterpCase(grid, sc.n_radial, sc.n_angles, x, z, 0.8, **ok)
    geom = _geom4(grid, x, z, sc.n_radial, sc.n_angles, 0.8, _lib.MODE_PET_SINO, **kw)
    img = _rand(grid.N, 8)
    y = fp(geom, 0, n, torch.from_numpy(img).to(cuda)).cpu().numpy()
    yo = case.fp(img, n)
    if atten_mode == 1:
        assert np.all(np.abs(y - yo) <= 1e-6 * np.abs(yo) + 1e-12)
    else:
        assert np.array_equal(y, yo)
    yin = _rand(n, 9)
    rhs, sens = bp(geom, 0, n, torch.from_numpy(yin).to(cuda), grid.N, sens=True)
    ro, so = case.bp(yin, n, sens=True)
    assert rel_l2(rhs.cpu().numpy(), ro) < 1e-5 and rel_l2(sens.cpu().numpy(), so) < 1e-5


def test_ct_fp_bitexact_voxel_bp(cuda):
    import torch
    grid, x, uv, rows, cols, pitch = cases.ct_small()
    nv = x.size // 6
    n = rows * cols * nv
    kerroin = grid.dx * grid.dy * grid.dz / (pitch * pitch * 200.0)
    case = O.InterpCase(grid, rows, cols, x, uv, 0.5, mode=O.MODE_CT, dPitchZ=pitch, kerroin=kerroin, dPitchX=pitch,
                        dPitchY=pitch)
    geom = _geom4(grid, x, uv, rows, cols, 0.5, _lib.MODE_CT_FLAT, dPitchZ=pitch, dPitchX=pitch, dPitchY=pitch,
                  kerroin=kerroin)
    img = _rand(grid.N, 10)
    y = fp(geom, 0, n, torch.from_numpy(img).to(cuda)).cpu().numpy()
    assert np.array_equal(y, case.fp(img, n))
    yin = _rand(n, 11)
    rhs, sens = bp(geom, 0, n, torch.from_numpy(yin).to(cuda), grid.N, sens=True)
    ro, so = case.bp(yin, n, sens=True)
    e, es = rel_l2(rhs.cpu().numpy(), ro), rel_l2(sens.cpu().numpy(), so)
    print(f"[interp ct bp] rel-L2 {e:.2e} (max |diff| {np.abs(rhs.cpu().numpy() - ro).max():.2e}), sens {es:.2e}",
          flush=True)
    assert e < 1e-6 and es < 1e-6


def _pet_projector(subsets=4):
    grid, sc = cases.c2_small()
    p = projectorClass()
    p.Nx, p.Ny, p.Nz = grid.Nx, grid.Ny, grid.Nz
    p.FOVa_x, p.FOVa_y, p.axial_fov = grid.Nx * grid.dx, grid.Ny * grid.dy, grid.Nz * grid.dz
    p.x, p.z = sc.xy_table(), sc.z
    p.Ndist, p.Nang = sc.n_radial, sc.n_angles
    p.projector_type = 4
    p.dL = 0.5
    p.subsets, p.subsetType = subsets, 8
    p.addProjector()
    p.initProj()
    return p, p.grid, sc     # the projector's own grid (its FOV arithmetic rounds differently from the case's)


def test_projector_class_osem(cuda):
    """OSEM on emission data with projector 4: two iterations of 4 subsets through projectorClass /
    OSEMRunner against the oracle's FP -> ratio -> BP (+ sensitivity) -> EM update."""
    import torch
    p, grid, sc = _pet_projector()
    assert abs(p._interp_scalars()[0] - 0.5 * grid.dx) < 1e-6
    xt = torch.from_numpy(_rand(grid.N, 12, lo=0.2, scale=2.0)).to(cuda)
    y = torch.cat([p.forwardProject(xt, k) for k in range(p.subsets)])
    y = torch.poisson(y * 20.0).float()
    run = OSEMRunner(p, y)
    xg = run.run(2).cpu().numpy()
    ym = y.cpu().numpy()
    x = np.ones(grid.N, np.float32)
    S = [None] * p.subsets
    dl = float(p._interp_scalars()[0])
    for it in range(2):
        for k in range(p.subsets):
            first, nk = int(p.nMeas[k]), int(p.nMeasSubset[k, 0])
            case = O.InterpCase(grid, sc.n_radial, sc.n_angles, p.x, p.z, dl, plane_offset=first)
            yk = ym[int(p.nTotMeas[k]):int(p.nTotMeas[k]) + nk]
            r = O.ratio(yk, case.fp(x, nk), 1e-5)
            rhs, s = case.bp(r, nk, sens=True)
            if S[k] is None:
                S[k] = O.sens_clamp(s, 1e-5)
            x = O.osem_update(x, S[k], rhs, 1e-5)
    e = rel_l2(xg, x)
    print(f"[interp osem] rel-L2 vs oracle {e:.2e}", flush=True)
    assert e < 1e-5


def test_ct_transmission_osem(cuda):
    import torch
    grid, x, uv, rows, cols, pitch = cases.ct_small()
    p = projectorClass()
    p.CT = True
    p.projector_type = 4
    p.Nx, p.Ny, p.Nz = grid.Nx, grid.Ny, grid.Nz
    p.FOVa_x, p.FOVa_y, p.axial_fov = 32.0, 32.0, 24.0
    p.x, p.z = x, uv
    p.nRowsD, p.nColsD = rows, cols
    p.dPitchX = p.dPitchY = pitch
    p.sourceToDetector = 200.0
    p.subsets = 3
    p.flat = 1500.0
    p.addProjector()
    p.initProj()
    grid = p.grid
    from omega_amd import phantoms
    mu = (phantoms.attenuation_ellipsoids(grid, seed=3, n_ell=5) + 0.002).astype(np.float32)
    ax = torch.cat([p.forwardProject(torch.as_tensor(mu, device=cuda), k) for k in range(p.subsets)])
    g = torch.Generator(device=cuda)
    g.manual_seed(2)
    y = torch.poisson(1500.0 * torch.exp(-ax.clamp(min=0)), generator=g).float()
    xg = OSEMRunner(p, y).run(2, torch.full((grid.N,), 0.01, device=cuda)).cpu().numpy()
    dl, ker = p._interp_scalars()
    cs = [O.InterpCase(grid, rows, cols, p.x, p.z, dl, mode=O.MODE_CT, dPitchZ=pitch, plane_offset=int(p.nMeas[k]),
                       kerroin=ker, dPitchX=pitch, dPitchY=pitch) for k in range(p.subsets)]
    ym = y.cpu().numpy()
    parts = [ym[int(p.nTotMeas[k]):int(p.nTotMeas[k + 1])] for k in range(p.subsets)]
    xo = O.ct_osem(cs, parts, 2, np.full(grid.N, 0.01, np.float32), 1500.0, epps=1e-5)
    e = rel_l2(xg, xo)
    print(f"[interp ct osem] rel-L2 vs oracle {e:.2e}", flush=True)
    assert e < 1e-5


@pytest.mark.parametrize("atten_mode", [0, 1])
def test_direct_plane_sinogram_bp(cuda, atten_mode):
    """Direct-plane sinograms take the slice-in-LDS forward and backprojection (whole planes per call): the
    forward projection bit-identical to the oracle's, the backprojection within 1e-5, also through the generic
    kernels (omg_geom_desc.reserved[0]).
    Without attenuation the first and last samples lie on the image faces, so samples at x = Nx of the last
    row wrap into the next slice's first voxel, as in the reference; with an attenuation image the samples
    start a tenth of a step in."""
    import torch
    grid, sc = cases.orth_pet2d()
    x, z = sc.xy_table(), sc.z
    per = sc.n_radial * sc.n_angles
    n = per * (z.size // 2)
    kw, ok = {}, {}
    if atten_mode == 1:
        mu = _rand(grid.N, 16, lo=0.0, scale=0.02)
        kw.update(atten=mu, atten_mode=1)
        ok.update(atten=O.InterpCase(grid, sc.n_radial, sc.n_angles, x, z, 1.1).atten_factors(mu, n),
                  atten_image=True)
    case = O.InterpCase(grid, sc.n_radial, sc.n_angles, x, z, 1.1, **ok)
    yin = _rand(n, 17)
    ro, so = case.bp(yin, n, sens=True)
    for generic in (True, False):
        geom = _geom4(grid, x, z, sc.n_radial, sc.n_angles, 1.1, _lib.MODE_PET_SINO, generic=generic, **kw)
        img = _rand(grid.N, 18)
        y = fp(geom, 0, n, torch.from_numpy(img).to(cuda)).cpu().numpy()
        yo = case.fp(img, n)
        if atten_mode == 0:
            assert np.array_equal(y, yo), f"generic={generic}: FP differs from the oracle"
        else:   # the factors exp(jelppi): the device's expf against the host's
            assert np.all(np.abs(y - yo) <= 1e-6 * np.abs(yo) + 1e-12), f"generic={generic}"
        rhs, sens = bp(geom, 0, n, torch.from_numpy(yin).to(cuda), grid.N, sens=True)
        err = (rel_l2(rhs.cpu().numpy(), ro), rel_l2(sens.cpu().numpy(), so))
        print(f"[interp direct planes] atten_mode={atten_mode} generic={generic} n={n} rel-L2 {err}", flush=True)
        assert err[0] < 1e-5 and err[1] < 1e-5
        # two planes from a plane offset: the planes' share of the full backprojection's slices
        part = bp(geom, 3, 2 * per, torch.from_numpy(yin[3 * per:5 * per]).to(cuda), grid.N).cpu().numpy()
        assert rel_l2(part, case_part(grid, sc, x, z, ok, yin, 3, 2 * per)) < 1e-5


def case_part(grid, sc, x, z, ok, yin, plane, m):
    ok = dict(ok)
    if "atten" in ok:
        ok["atten"] = ok["atten"][plane * sc.n_radial * sc.n_angles:][:m]
    c = O.InterpCase(grid, sc.n_radial, sc.n_angles, x, z, 1.1, plane_offset=plane, **ok)
    return c.bp(yin[plane * sc.n_radial * sc.n_angles:][:m], m)
