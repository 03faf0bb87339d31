"""The kernels at the measured configurations' own sizes, against the oracle and size-independent invariants.

* 3-D ring PET (32 rings of 448 crystals, span 1 up to ring difference 31: 224 x 224 x 1024 sinogram;
  128 x 128 x 63 image): backprojection with the sensitivity image of 24 ring-pair planes (direct and
  oblique) against the oracle (1e-5 relative L2, float atomics), its error against the FP64 re-run stated
  and no larger than the FP32 oracle's own (the walk's float lengths are ~1e-5 from FP64); 20 seeded adjoint pairs over the whole
  sinogram; central differences (h = 1e-3) of -L(x) + R(x) (2e7 counts, gamma 2, beta 0.1) at 64 voxels
  against the autograd gradient.
* The 24-frame dynamic study on that geometry (16 subsets): one batched subset step of all frames against
  the single-frame fused path, frame by frame (1e-5), and frame 0 against the oracle's
  FP -> ratio -> BP -> update (1e-5).
* Cone-beam CT with the branchless distance-driven projector (256^3 volume of 0.5 mm voxels, 512 x 384
  panel of 0.8 mm pixels, SOD 700 mm, SDD 1000 mm): forward and backprojection of a sample of the 360
  views (both slab orientations) against the oracle (1e-6: same float operations); over all 360 views,
  the uniform-volume projections against Siddon chord lengths, and the FP/BP inner-product ratio.
"""
import numpy as np
import pytest

from omega_amd import _lib, autograd as AG, geometry as G, phantoms, workloads
from omega_amd.projector import projectorClass
from oracle import oracle as O
from tests.hip_helpers import ptr, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pet3d(cuda):
    p = workloads.pet3d_projector(subsets=1)
    p.initProj()
    return p


def test_pet3d_backprojection_planes_match_oracle(pet3d, cuda):
    import torch
    p = pet3d
    per = int(p.nRowsD) * int(p.nColsD)
    rng = np.random.default_rng(3)
    for first in (0, 517, 1016):            # direct planes, mid ring differences, the largest ring difference
        n = 8 * per
        w = (rng.random(n) + 0.1).astype(np.float32)
        rhs = torch.zeros(p.grid.N, device=cuda)
        sens = torch.zeros(p.grid.N, device=cuda)
        _lib.call("omg_backward_project", p.geometry.handle, first, n, ptr(torch.from_numpy(w).to(cuda)), ptr(rhs),
                  ptr(sens), None)
        c32 = O.SiddonCase(p.grid, int(p.nRowsD), int(p.nColsD), p.x, p.z, plane_offset=first)
        bo, so = c32.bp(w, n, sens=True)
        c64 = O.SiddonCase(p.grid, int(p.nRowsD), int(p.nColsD), p.x, p.z, plane_offset=first, dtype=np.float64)
        b64 = c64.bp(w.astype(np.float64), n)
        e, es, e64 = rel_l2(rhs.cpu().numpy(), bo), rel_l2(sens.cpu().numpy(), so), rel_l2(rhs.cpu().numpy(), b64)
        ref64 = rel_l2(bo, b64)
        print(f"[c2 bp] planes {first}..{first + 7}: rel-L2 vs oracle {e:.2e} (sens {es:.2e}); vs FP64 {e64:.2e} "
              f"(the FP32 oracle's own FP64 error {ref64:.2e})", flush=True)
        assert e < 1e-5 and es < 1e-5
        # FP32 walk lengths: the GPU is as far from the FP64 re-run as the reference's FP32 algorithm is
        assert e64 <= 1.2 * ref64 + 1e-6


def test_pet3d_adjoint_20_pairs(pet3d, cuda):
    import torch
    p = pet3d
    worst = 0.0
    for seed in range(20):
        g = torch.Generator(device=cuda)
        g.manual_seed(seed)
        x = torch.rand(p.grid.N, device=cuda, generator=g)
        y = torch.rand(int(p.nMeasSubset[0, 0]), device=cuda, generator=g)
        lhs = float((p.forwardProject(x, 0).double() * y.double()).sum())
        rhs = float((x.double() * p.backwardProject(y, 0).double()).sum())
        worst = max(worst, abs(lhs - rhs) / abs(lhs))
    print(f"[c2 adjoint] worst relative mismatch over 20 pairs {worst:.2e}", flush=True)
    assert worst < 1e-6


def test_pet3d_finite_differences_64_voxels(pet3d, cuda):
    import torch
    p = pet3d
    ph = phantoms.ellipsoids(p.grid, seed=0) + 0.2
    meas, _ = workloads.poisson_data(p, ph, 2e7, seed=1)
    x = torch.as_tensor(ph, device=cuda).clone().requires_grad_(True)
    AG.penalised_nll(x, p, meas, gamma=2.0, beta=0.1).backward()
    grad = x.grad.double().cpu().numpy()
    rng = np.random.default_rng(7)
    h = 1e-3
    errs = []
    for j in rng.choice(p.grid.N, 64, replace=False):
        xp, xm = x.detach().clone(), x.detach().clone()
        xp[j] += h
        xm[j] -= h
        fd = (float(AG.penalised_nll(xp, p, meas)) - float(AG.penalised_nll(xm, p, meas))) / (2 * h)
        errs.append(abs(fd - grad[j]) / max(1.0, abs(grad[j])))
    print(f"[c2 fd] 64 voxels: max relative error {max(errs):.2e}, median {np.median(errs):.2e}", flush=True)
    assert max(errs) < 2e-3


def test_dynamic_24_frames_subset_step(cuda):
    import torch
    from omega_amd.dynamic import FramesOSEM
    p = workloads.pet3d_projector(subsets=16)
    p.initProj()
    F = 24
    frames = phantoms.dynamic_ellipsoids(p.grid, seed=0, durations=phantoms.FRAME_DURATIONS[:F])
    counts = workloads.frame_counts(F, seed=0)
    y = workloads.dynamic_data(p, frames, counts, seed=0, device=cuda)
    print(f"[c3] {F} frames x {y.shape[1]} measurements, counts {counts.min():.3g}..{counts.max():.3g}", flush=True)
    k = 5
    first, n, off = int(p.nMeas[k]), int(p.nMeasSubset[k, 0]), int(p.nTotMeas[k])
    run = FramesOSEM(p, y)
    run.sensitivity()
    run.step(k, 256, 8)
    got = run.x.cpu().numpy()
    S_batched = run.S[k].cpu().numpy()
    run.close()
    N = p.grid.N
    worst = 0.0
    for f in range(F):
        x = torch.ones(N, device=cuda)
        S = torch.zeros(N, device=cuda)
        yf = y[f, off:off + n].contiguous()
        _lib.call("omg_osem_subset_iteration", p.geometry.handle, first, n, ptr(x), ptr(yf), ptr(S), 1, 1e-5, 0.0,
                  2.0, 1, None)
        e = rel_l2(got[f], x.cpu().numpy())
        worst = max(worst, e)
        assert e <= 1e-5, f"frame {f}: {e:.2e}"
    print(f"[c3] batched subset step vs single-frame path: worst frame {worst:.2e}", flush=True)
    # frame 0 against the oracle
    case = O.SiddonCase(p.grid, int(p.nRowsD), int(p.nColsD), p.x, p.z, plane_offset=first)
    x0 = np.ones(N, np.float32)
    r = O.ratio(y[0, off:off + n].cpu().numpy(), case.fp(x0, n), 1e-5)
    rhs, sens = case.bp(r, n, sens=True)
    sens = O.sens_clamp(sens, 1e-5)
    xo = O.osem_update(x0, sens, rhs, 1e-5)
    e0, es = rel_l2(got[0], xo), rel_l2(S_batched, sens)
    print(f"[c3] frame 0 vs oracle {e0:.2e}, sensitivity {es:.2e}", flush=True)
    assert e0 <= 1e-5 and es <= 1e-5
    del y
    torch.cuda.empty_cache()


CT_VIEWS = (0, 44, 90, 135, 224, 301)     # |s.x| <= |s.y| and the other slab orientation


def _ct_projector(views):
    x, uv = G.cbct_tables(360, sod=700.0, sdd=1000.0, pitch_u=0.8)
    x6, uv2 = x.reshape(-1, 6), uv.reshape(-1, 2)
    idx = np.arange(360) if views is None else np.asarray(views)
    p = projectorClass()
    p.CT = True
    p.projector_type = 5
    p.Nx = p.Ny = p.Nz = 256
    p.FOVa_x = p.FOVa_y = p.axial_fov = 128.0
    p.x = np.ascontiguousarray(x6[idx].reshape(-1))
    p.z = np.ascontiguousarray(uv2[idx].reshape(-1))
    p.nRowsD, p.nColsD = 512, 384
    p.dPitchX = p.dPitchY = 0.8
    p.subsets = 1
    p.addProjector()
    p.initProj()
    return p


def test_cbct_bdd_sample_views_match_oracle(cuda):
    import torch
    p = _ct_projector(CT_VIEWS)
    case = O.BDDCase(p.grid, p.nRowsD, p.nColsD, p.x, p.z, p.dPitchX, p.dPitchY)
    mu = phantoms.attenuation_ellipsoids(p.grid, seed=0, n_ell=12)
    y = p.forwardProject(torch.as_tensor(mu, device=cuda), 0).cpu().numpy()
    yo = case.fp(mu)
    print(f"[c1 fp] {len(CT_VIEWS)} views: rel-L2 {rel_l2(y, yo):.2e} max|diff| {np.abs(y - yo).max():.2e}",
          flush=True)
    assert rel_l2(y, yo) < 1e-6
    w = (np.random.default_rng(5).random(case.M) + 0.05).astype(np.float32)
    sens = torch.zeros(p.grid.N, device=cuda)
    b = p.backwardProject(torch.as_tensor(w, device=cuda), 0, sens=sens).cpu().numpy()
    bo, so = case.bp(w, sens=True)
    eb, es = rel_l2(b, bo), rel_l2(sens.cpu().numpy(), so)
    print(f"[c1 bp] rel-L2 {eb:.2e}, sensitivity {es:.2e}", flush=True)
    assert eb < 1e-6 and es < 1e-6


def test_cbct_bdd_full_invariants(cuda):
    """All 360 views: a uniform volume projects to within 2% of the Siddon chord lengths on interior rays,
    and <A x, y> / <x, A^T y> is within 1% of 1 (BDD's FP and BP are not exact transposes)."""
    import torch
    p = _ct_projector(None)
    N = p.grid.N
    one = torch.ones(N, device=cuda)
    yb = p.forwardProject(one, 0)
    ps = projectorClass()
    ps.CT = True
    ps.projector_type = 1
    ps.Nx = ps.Ny = ps.Nz = 256
    ps.FOVa_x = ps.FOVa_y = ps.axial_fov = 128.0
    ps.x, ps.z = p.x.copy(), p.z.copy()
    ps.nRowsD, ps.nColsD = 512, 384
    ps.dPitchX = ps.dPitchY = 0.8
    ps.addProjector()
    ys = ps.forwardProject(one, 0)
    m = ys > 0.5 * ys.max()
    e = float(torch.linalg.vector_norm((yb - ys)[m]) / torch.linalg.vector_norm(ys[m]))
    print(f"[c1] uniform-volume projections vs chord lengths on {int(m.sum())} interior rays: {e:.2e}", flush=True)
    assert e < 0.02
    del ys, m
    g = torch.Generator(device=cuda)
    g.manual_seed(1)
    x = torch.rand(N, device=cuda, generator=g)
    yv = torch.rand(yb.numel(), device=cuda, generator=g)
    lhs = float((p.forwardProject(x, 0).double() * yv.double()).sum())
    rhs = float((x.double() * p.backwardProject(yv, 0).double()).sum())
    print(f"[c1] <Ax,y>/<x,A^T y> = {lhs / rhs:.6f}", flush=True)
    assert abs(lhs / rhs - 1.0) < 0.01
