"""Differentiable projector: autograd backward = the backprojection kernel; adjoint identity;
RDP value; central finite differences of -L(x) + R(x)."""
import numpy as np
import pytest

from omega_amd import autograd as AG
from omega_amd import phantoms, workloads
from tests.test_oracle_math import rdp_value as rdp_value_numpy

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pet3d_small(cuda):
    p = workloads.pet3d_projector(subsets=1, n_rings=8, crystals=96, max_ring_diff=7, radius=150.0,
                                  image=(48, 48, 15), fov_xy=200.0)
    p.initProj()
    return p


def test_backward_is_backprojection(pet3d_small):
    import torch
    p = pet3d_small
    g = torch.Generator(device="cuda")
    g.manual_seed(0)
    x = torch.rand(int(p.N[0]), device="cuda", generator=g).requires_grad_(True)
    y = AG.forward_project(x, p)
    v = torch.rand(y.numel(), device="cuda", generator=g)
    (y * v).sum().backward()
    ref = p.backwardProject(v, 0)
    # float atomics: equal up to the order of the additions
    assert float(torch.linalg.vector_norm(x.grad - ref) / torch.linalg.vector_norm(ref)) <= 1e-6


def test_adjoint_identity_20_pairs(pet3d_small):
    import torch
    p = pet3d_small
    worst = 0.0
    for seed in range(20):
        g = torch.Generator(device="cuda")
        g.manual_seed(seed)
        x = torch.rand(int(p.N[0]), device="cuda", generator=g)
        y = torch.rand(int(p.nMeasSubset[0, 0]), device="cuda", generator=g)
        lhs = float((p.forwardProject(x, 0).double() * y.double()).sum())
        rhs = float((x.double() * p.backwardProject(y, 0).double()).sum())
        worst = max(worst, abs(lhs - rhs) / abs(lhs))
    print(f"[adjoint] worst relative mismatch over 20 pairs {worst:.2e}")
    assert worst < 1e-5


def test_rdp_value_matches_float64(cuda):
    import torch
    shape = (11, 9, 7)
    u = (np.random.default_rng(3).random(int(np.prod(shape))) + 0.1).astype(np.float32)
    v = AG.rdp(torch.from_numpy(u).cuda(), shape, 2.0, 0.3)
    # numpy R counts every ordered interior pair (= 2 x unordered) and every ghost pair once
    full = rdp_value_numpy(u.astype(np.float64), shape, 2.0)
    Nx, Ny, Nz = shape
    w = u.astype(np.float64).reshape(Nz, Ny, Nx)
    ghost = 0.0
    for ax, n in ((2, Nx), (1, Ny), (0, Nz)):
        for idx in (0, n - 1):
            a = np.take(w, idx, axis=ax)
            ghost += np.sum(a * a / (a + 2.0 * a))
    # beta crosses the C ABI as float32
    expect = float(np.float32(0.3)) * ((full - ghost) / 2.0 + ghost)
    assert abs(float(v) - expect) < 1e-10 * abs(expect)


def test_finite_differences_penalised_nll(pet3d_small):
    import torch
    p = pet3d_small
    ph = phantoms.ellipsoids(p.grid, seed=5) + 0.2
    meas, _ = workloads.poisson_data(p, ph, 2e5, seed=6)
    x = torch.as_tensor(ph, device="cuda").clone().requires_grad_(True)
    loss = AG.penalised_nll(x, p, meas, gamma=2.0, beta=0.1)
    loss.backward()
    grad = x.grad.double().cpu().numpy()
    rng = np.random.default_rng(7)
    h = 1e-3
    errs = []
    for j in rng.choice(int(p.N[0]), 32, replace=False):
        xp = x.detach().clone()
        xm = x.detach().clone()
        xp[j] += h
        xm[j] -= h
        fd = (float(AG.penalised_nll(xp, p, meas)) - float(AG.penalised_nll(xm, p, meas))) / (2 * h)
        errs.append(abs(fd - grad[j]) / max(1.0, abs(grad[j])))
    print(f"[fd] max relative error over 32 voxels {max(errs):.2e}")
    assert max(errs) < 2e-3
