"""GPU attenuation correction against the CPU oracle (itself bit-identical to the reference's own code).

The library turns an attenuation image into one factor per measurement, exp(-sum len * mu) over the
LOR's Siddon walk, once at geometry creation, and multiplies it into the LOR weight after 1/L, where
the reference's backprojection applies it. Bars:
* per-measurement factors (no exponential involved): forward projections bit-identical to the oracle,
  backprojections within 1e-6 relative L2 (atomic order);
* attenuation image: forward projections within 1e-6 relative per measurement (HIP's expf against
  glibc's, and the reference's forward projection multiplies the factor after norm and the global
  factor, a rounding-order difference); backprojections within 1e-6 relative L2 on sinograms whose
  LORs end outside the image (the reference's backprojection integrates mu over the unclipped walk,
  which differs only for LORs that end inside the image);
* fused OSEM subset iterations with the attenuation image within 1e-5 of the oracle run (the OSEM bar),
  through the direct-plane and the one-lane-per-LOR kernels, and the orthogonal projector with
  factors within its own bars (forward bit-identical, backprojection 1e-5).
"""
import numpy as np
import pytest

from omega_amd import _lib, workloads
from omega_amd.projector import Geometry
from omega_amd.recon import OSEMRunner
from oracle import oracle as O
from tests import cases
from tests.hip_helpers import bp, fp, make_geom, rel_l2

pytestmark = pytest.mark.gpu

FP_RTOL = 1e-6
BP_TOL = 1e-6


def _rand(n, seed, lo=0.05, scale=1.0):
    return (np.random.default_rng(seed).random(n) * scale + lo).astype(np.float32)


def _fp_close(got, ref):
    assert np.all(np.abs(got - ref) <= FP_RTOL * np.abs(ref)), f"max rel {np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1e-30)):.2e}"


def test_listmode_factors_bitexact(cuda):
    import torch
    grid = cases.tiny_grid()
    x = cases.list_rays(grid, 400, seed=3)
    n = x.size // 6
    fac = _rand(n, 1, lo=0.2, scale=0.8)
    print(f"[atten list-mode factors] n={n}", flush=True)
    case = O.SiddonCase(grid, 1, 1, x, np.zeros(2, np.float32), mode=O.MODE_LIST, atten=fac, atten_mode=2)
    geom = make_geom(grid, x, None, 1, 1, mode=_lib.MODE_LIST6, atten=fac, atten_mode=2)
    img = _rand(grid.N, 2)
    y = fp(geom, 0, n, torch.from_numpy(img).to(cuda)).cpu().numpy()
    assert np.array_equal(y, case.fp(img, n))
    yin = _rand(n, 3)
    rhs, sens = bp(geom, 0, n, torch.from_numpy(yin).to(cuda), grid.N, sens=True)
    rhs_o, sens_o = case.bp(yin, n, sens=True)
    assert rel_l2(rhs.cpu().numpy(), rhs_o) < BP_TOL
    assert rel_l2(sens.cpu().numpy(), sens_o) < BP_TOL


def test_listmode_all_weight_terms(cuda):
    """Norm, attenuation, multiplicative scatter correction and global factor together (axis-parallel
    LORs take a given per-LOR factor after norm, as the reference does): per-LOR factors bit-identical,
    an attenuation image within the forward bar."""
    import torch
    grid = cases.tiny_grid()
    x = cases.list_rays(grid, 400, seed=3)
    n = x.size // 6
    norm, scat, fac = _rand(n, 20, lo=0.5), _rand(n, 21, lo=0.7, scale=0.6), _rand(n, 22, lo=0.2, scale=0.8)
    mu = _rand(grid.N, 23, lo=0.0, scale=0.05)
    img = _rand(grid.N, 24)
    yin = _rand(n, 25)
    print(f"[atten list-mode all terms] n={n}", flush=True)
    for atn, mode in ((fac, 2), (mu, 1)):
        case = O.SiddonCase(grid, 1, 1, x, np.zeros(2, np.float32), mode=O.MODE_LIST, atten=atn, atten_mode=mode,
                            norm=norm, scat=scat, global_factor=0.9)
        geom = make_geom(grid, x, None, 1, 1, mode=_lib.MODE_LIST6, atten=atn, atten_mode=mode, norm=norm, scat=scat,
                         global_factor=0.9)
        y = fp(geom, 0, n, torch.from_numpy(img).to(cuda)).cpu().numpy()
        if mode == 2:
            assert np.array_equal(y, case.fp(img, n))
        else:
            _fp_close(y, case.fp(img, n))
        rhs = bp(geom, 0, n, torch.from_numpy(yin).to(cuda), grid.N)
        if mode == 2:
            assert rel_l2(rhs.cpu().numpy(), case.bp(yin, n)) < BP_TOL


def test_listmode_image_forward(cuda):
    """Every set-up branch, LORs that start or end inside the image and axis-parallel rows (whose
    factor integrates the whole image row): forward projections against the oracle."""
    import torch
    grid = cases.tiny_grid()
    x = cases.list_rays(grid, 400, seed=3)
    n = x.size // 6
    mu = _rand(grid.N, 4, lo=0.0, scale=0.05)
    print(f"[atten list-mode image] n={n}", flush=True)
    case = O.SiddonCase(grid, 1, 1, x, np.zeros(2, np.float32), mode=O.MODE_LIST, atten=mu, atten_mode=1)
    geom = make_geom(grid, x, None, 1, 1, mode=_lib.MODE_LIST6, atten=mu, atten_mode=1)
    img = _rand(grid.N, 5)
    y = fp(geom, 0, n, torch.from_numpy(img).to(cuda)).cpu().numpy()
    y_ref = case.fp(img, n)
    _fp_close(y, y_ref)
    plain = O.SiddonCase(grid, 1, 1, x, np.zeros(2, np.float32), mode=O.MODE_LIST).fp(img, n)
    assert rel_l2(plain, y_ref) > 1e-2        # the correction is there


@pytest.mark.parametrize("generic", [False, True])
def test_pet2d_image_fp_bp(cuda, generic):
    """2-D plane sinograms (angles 0 and 90 degrees included) through the direct-plane LDS kernels and
    through the one-lane-per-LOR kernels."""
    import torch
    grid, sc = cases.orth_pet2d()
    n = sc.n_radial * sc.n_angles * sc.n_planes
    mu = _rand(grid.N, 6, lo=0.0, scale=0.02)
    print(f"[atten pet2d] generic={generic} n={n}", flush=True)
    case = O.SiddonCase(grid, sc.n_radial, sc.n_angles, sc.xy_table(), sc.z, atten=mu, atten_mode=1)
    geom = make_geom(grid, sc.xy_table(), sc.z, sc.n_radial, sc.n_angles, atten=mu, atten_mode=1, generic=generic)
    img = _rand(grid.N, 7)
    _fp_close(fp(geom, 0, n, torch.from_numpy(img).to(cuda)).cpu().numpy(), case.fp(img, n))
    yin = _rand(n, 8)
    rhs, sens = bp(geom, 0, n, torch.from_numpy(yin).to(cuda), grid.N, sens=True)
    rhs_o, sens_o = case.bp(yin, n, sens=True)
    assert rel_l2(rhs.cpu().numpy(), rhs_o) < BP_TOL
    assert rel_l2(sens.cpu().numpy(), sens_o) < BP_TOL


def test_pet3d_image_fp_bp(cuda):
    import torch
    grid, sc = cases.c2_small()
    n = sc.n_radial * sc.n_angles * sc.n_planes
    mu = _rand(grid.N, 9, lo=0.0, scale=0.02)
    print(f"[atten pet3d] n={n}", flush=True)
    case = O.SiddonCase(grid, sc.n_radial, sc.n_angles, sc.xy_table(), sc.z, atten=mu, atten_mode=1)
    geom = make_geom(grid, sc.xy_table(), sc.z, sc.n_radial, sc.n_angles, atten=mu, atten_mode=1)
    img = _rand(grid.N, 10)
    _fp_close(fp(geom, 0, n, torch.from_numpy(img).to(cuda)).cpu().numpy(), case.fp(img, n))
    yin = _rand(n, 11)
    rhs = bp(geom, 0, n, torch.from_numpy(yin).to(cuda), grid.N)
    assert rel_l2(rhs.cpu().numpy(), case.bp(yin, n)) < BP_TOL


@pytest.fixture(scope="module")
def pet2d_atn(cuda):
    p = workloads.pet2d_projector(subsets=8)
    p.attenuation_correction = True
    g = p.grid
    # water-like cylinder (0.0096 / mm) in a 0.001 / mm background
    xc = (np.arange(g.Nx) + 0.5) * g.dx + g.bx
    yc = (np.arange(g.Ny) + 0.5) * g.dy + g.by
    r2 = xc[:, None] ** 2 + yc[None, :] ** 2
    mu2 = np.where(r2 < 120.0 ** 2, 0.0096, 0.001).astype(np.float32)
    p.vaimennus = np.repeat(mu2[:, :, None], g.Nz, axis=2)
    p.initProj()
    ph, y = workloads.pet2d_data(p, seed=3)
    return p, y


def _gpu_run(p, y, n_sub, generic=False):
    import torch
    run = OSEMRunner(p, y, beta=0.0, gamma=2.0)
    if generic:
        d, keep = p._descriptor()
        d.reserved[0] = 1
        gm = Geometry(d, keep)
        run.g = gm.handle
        run._keep = gm
    x = torch.ones(int(p.N[0]), device="cuda")
    done = 0
    for it in range(4):
        for k in range(p.subsets):
            if done < n_sub:
                run.subset_step(x, k)
                done += 1
        run.sens_done = True
    torch.cuda.synchronize()
    return x.cpu().numpy()


@pytest.mark.parametrize("generic", [False, True])
def test_osem_with_attenuation_image(pet2d_atn, generic):
    p, y = pet2d_atn
    n_sub = 10
    mu = p._attenuation()[0]
    nmeas = np.diff(p.nMeas).astype(np.int64)
    print(f"[atten osem] generic={generic} subset-iterations={n_sub}", flush=True)
    xg = _gpu_run(p, y, n_sub, generic=generic)
    xo = O.osem(p.grid, p.nRowsD, p.nColsD, p.x, p.z, nmeas, y.cpu().numpy(), 4, np.ones(int(p.N[0]), np.float32),
                epps=1e-5, subsets_to_run=n_sub, atten=mu, atten_mode=1)
    e = rel_l2(xg, xo)
    plain = O.osem(p.grid, p.nRowsD, p.nColsD, p.x, p.z, nmeas, y.cpu().numpy(), 4, np.ones(int(p.N[0]), np.float32),
                   epps=1e-5, subsets_to_run=n_sub)
    print(f"  rel-L2 vs oracle {e:.2e}; without attenuation {rel_l2(plain, xo):.2e}", flush=True)
    assert e < 1e-5
    assert rel_l2(plain, xo) > 1e-3


@pytest.mark.parametrize("mode", [1, 2])
def test_orth_with_factors(cuda, mode):
    """Orthogonal projector with attenuation, normalisation and a global factor, in the reference kernel's
    order: factors made from an image (mode 1) first, given per-LOR factors (mode 2) after the global
    factor (projectorType123.cl:975-988); FP bit-identical to the oracle with given factors, within FP_RTOL
    with factors from the image (exp on the device and on the host)."""
    import torch
    grid, sc = cases.orth_pet2d()
    n = sc.n_radial * sc.n_angles * sc.n_planes
    norm = _rand(n, 11, lo=0.5, scale=1.0)
    if mode == 2:
        atn = fac = _rand(n, 12, lo=0.3, scale=0.7)
    else:
        atn = _rand(grid.N, 12, lo=0.0, scale=0.02)
        fac = O.SiddonCase(grid, sc.n_radial, sc.n_angles, sc.xy_table(), sc.z, atten=atn,
                           atten_mode=1).atten_factors(n)
    print(f"[atten orth] mode={mode} n={n}", flush=True)
    case = O.OrthCase(grid, sc.n_radial, sc.n_angles, sc.xy_table(), sc.z, 3.0, flags=3, atten=fac, norm=norm,
                      global_factor=1.37, atten_lor=mode == 2)
    for generic in (False, True):
        geom = make_geom(grid, sc.xy_table(), sc.z, sc.n_radial, sc.n_angles, projector_type=2, orth_width=3.0,
                         orth_flags=3, atten=atn, atten_mode=mode, norm=norm, global_factor=1.37, generic=generic)
        img = _rand(grid.N, 13)
        got = fp(geom, 0, n, torch.from_numpy(img).to(cuda)).cpu().numpy()
        if mode == 2:
            assert np.array_equal(got, case.fp(img, n))
        else:   # factors from the image: the device's expf against the host's exp
            _fp_close(got, case.fp(img, n))
        yin = _rand(n, 14)
        rhs, sens = bp(geom, 0, n, torch.from_numpy(yin).to(cuda), grid.N, sens=True)
        rhs_o, sens_o = case.bp(yin, n, sens=True)
        assert rel_l2(rhs.cpu().numpy(), rhs_o) < 1e-5
        assert rel_l2(sens.cpu().numpy(), sens_o) < 1e-5


def test_projector_class_attenuation_options(cuda):
    """projectorClass: an (Nx, Ny, Nz) image in 1/cm (attIncm), and per-measurement attenuation factors
    with a multiplicative correction (additionalCorrection / corrVector) given in sinogram order, give
    the projections of the library called with the prepared tables."""
    import torch
    from omega_amd.projector import projectorClass
    grid, sc = cases.orth_pet2d()
    per = sc.n_radial * sc.n_angles
    mu = _rand(grid.N, 15, lo=0.0, scale=0.2)
    img = _rand(grid.N, 16)

    def make(**kw):
        p = projectorClass()
        p.Nx, p.Ny, p.Nz = grid.Nx, grid.Ny, grid.Nz
        p.FOVa_x, p.FOVa_y, p.axial_fov = grid.Nx * grid.dx, grid.Ny * grid.dy, grid.Nz * grid.dz
        p.x, p.z = sc.xy_table(), sc.z
        p.Ndist, p.Nang = sc.n_radial, sc.n_angles
        p.subsets, p.subsetType = 3, 8
        p.attenuation_correction = True
        for k, v in kw.items():
            setattr(p, k, v)
        p.addProjector()
        p.initProj()
        return p

    p = make(vaimennus=mu.reshape((grid.Nx, grid.Ny, grid.Nz), order="F") * 10.0, attIncm=True)
    geom = make_geom(p.grid, p.x, p.z, sc.n_radial, sc.n_angles, atten=p._attenuation()[0], atten_mode=1)
    np.testing.assert_allclose(p._attenuation()[0], mu, rtol=1e-6)
    it = torch.from_numpy(img).to(cuda)
    for k in range(p.subsets):
        y = p.forwardProject(it, k)
        assert np.array_equal(y.cpu().numpy(), fp(geom, int(p.nMeas[k]), y.numel(), it).cpu().numpy())
    fac_sino = _rand(per * sc.n_planes, 17, lo=0.3, scale=0.7)
    corr = _rand(per * sc.n_planes, 18, lo=0.7, scale=0.6)
    p = make(vaimennus=fac_sino, CTAttenuation=False, additionalCorrection=True, corrVector=corr)
    geom = make_geom(p.grid, p.x, p.z, sc.n_radial, sc.n_angles, atten=p.subset_order(fac_sino), atten_mode=2,
                     scat=p.subset_order(corr))
    for k in range(p.subsets):
        y = p.forwardProject(it, k)
        assert np.array_equal(y.cpu().numpy(), fp(geom, int(p.nMeas[k]), y.numel(), it).cpu().numpy())
