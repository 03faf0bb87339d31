"""The oracle's OSL-OSEM update against what the mathematics fixes.

The reference's one-step-late update (subiterStep.h:219-231, algorithms.h:24-27) is
x <- x / (S + beta grad R(x)) * A^T (y / (A x + eps)), positivity x < eps -> eps. Its fixed points with
x > eps are exactly the stationary points of the penalised Poisson log-likelihood

    F(x) = sum_i [y_i log((A x)_i + eps) - (A x)_i] - beta R(x),

i.e. A^T (y / (A x + eps)) - A^T 1 - beta grad R(x) = 0, because S = A^T 1. These tests run the oracle's
update (oracle_osl_update_f32 and its float64 twin) to convergence on a small dense Siddon system
whose every voxel is seen and whose matrix has full column rank, and check:

* the stationarity (KKT) residual of F, with A^T and S taken from the explicit matrix (numpy) and
  grad R from torch autograd of an independently written RDP value (eps -> 0);
* the converged image against scipy.optimize's L-BFGS-B minimiser of -F with x >= 0;
* with a large eps (0.05), the residual with the ratio's eps and the eps-guarded RDP gradient of the
  reference (pinned by its own golden vectors), so that misplacing eps (x / (S + dU) * (rhs + eps),
  for instance) leaves a residual of size eps and fails.

Plausible mistakes each of these catches: S - dU (the prior's sign), a missing beta, rhs + eps, the
prior gradient of the wrong neighbourhood or boundary, (S + dU) in the numerator.
"""
import numpy as np
import pytest

from omega_amd import geometry as G
from oracle import oracle as O

SHAPE = (6, 5, 3)
GAMMA = 2.0


def _system(dtype=np.float64):
    grid = G.Grid(SHAPE[0], SHAPE[1], SHAPE[2], 2.0, 2.5, 3.0, -6.0, -6.25, -4.5)
    rng = np.random.default_rng(21)
    # lines through uniform points of the box in uniform directions, endpoints 40 mm away
    lo = np.array([grid.bx, grid.by, grid.bz])
    ext = np.array([grid.Nx * grid.dx, grid.Ny * grid.dy, grid.Nz * grid.dz])
    p = lo + rng.random((1500, 3)) * ext
    u = rng.normal(size=(1500, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    rays = np.concatenate([p - 40 * u, p + 40 * u], axis=1).astype(np.float32)
    c64 = O.SiddonCase(grid, 1, 1, rays.reshape(-1).astype(np.float64), np.zeros(2), mode=O.MODE_LIST,
                       dtype=np.float64)
    n = rays.shape[0]
    A = np.stack([c64.fp(np.eye(grid.N)[j], n) for j in range(grid.N)], axis=1)
    keep = A.sum(1) > 0
    rays = np.ascontiguousarray(rays[keep].reshape(-1))
    A = A[keep]
    case = O.SiddonCase(grid, 1, 1, rays.astype(dtype), np.zeros(2), mode=O.MODE_LIST, dtype=dtype)
    rng = np.random.default_rng(5)
    xt = rng.random(grid.N) + 0.5
    y = rng.poisson(A @ xt * 40.0).astype(np.float64)
    return grid, case, A, y


def _rdp_value_torch(u, shape, gamma):
    """R(u) = sum over unordered 6-neighbour pairs inside the volume of phi(u_j, u_k) plus, for every voxel on
    the border, phi(u_j, 0) per missing neighbour; phi(a, b) = (a - b)^2 / (a + b + gamma |a - b|)."""
    import torch
    Nx, Ny, Nz = shape
    v = u.reshape(Nz, Ny, Nx)

    def phi(a, b):
        d = a - b
        return d * d / (a + b + gamma * d.abs())

    tot = phi(v[:, :, 1:], v[:, :, :-1]).sum() + phi(v[:, 1:, :], v[:, :-1, :]).sum() + \
        phi(v[1:, :, :], v[:-1, :, :]).sum()
    z = torch.zeros((), dtype=u.dtype)
    for face in (v[:, :, 0], v[:, :, -1], v[:, 0, :], v[:, -1, :], v[0, :, :], v[-1, :, :]):
        tot = tot + phi(face, z).sum()
    return tot


def _grad_R(x):
    import torch
    t = torch.tensor(x, dtype=torch.float64, requires_grad=True)
    _rdp_value_torch(t, SHAPE, GAMMA).backward()
    return t.grad.numpy()


def _osl(case, y, beta, eps, n_iter, dtype):
    """The oracle's OSL-OSEM with one subset: FP, ratio, BP, RDP gradient, update."""
    n = y.size
    f64 = np.dtype(dtype) == np.float64
    yd = y.astype(dtype)
    S = case.bp(np.ones(n, dtype), n)
    S = O._sens_clamp64(S, eps) if f64 else O.sens_clamp(S, eps)
    x = np.ones(case.grid.N, dtype)
    for _ in range(n_iter):   # one thread: thousands of tiny calls
        yhat = case.fp(x, n, nthreads=1)
        r = O._ratio64(yd, yhat, eps) if f64 else O.ratio(yd, yhat, eps)
        rhs = case.bp(r, n, nthreads=1)
        dU = O.rdp_gradient(x, SHAPE, GAMMA, eps, beta, dtype=dtype, nthreads=1)
        x = O._update64(x, S, dU, rhs, eps, True) if f64 else O.osl_update(x, S, dU, rhs, eps)
    return x


@pytest.fixture(scope="module")
def system():
    return _system()


@pytest.mark.parametrize("beta", [0.1, 0.5])
def test_osl_fixed_point_is_stationary(system, beta):
    grid, case, A, y = system
    eps = 1e-10
    x = _osl(case, y, beta, eps, 4000, np.float64)
    S = A.sum(0)
    g = A.T @ (y / (A @ x)) - S - beta * _grad_R(x)
    assert x.min() > 1e-3                       # an interior point: every voxel's residual must vanish
    assert np.max(np.abs(g) / S) < 1e-9
    # the same image is not stationary without the prior, nor with its sign flipped
    g0 = A.T @ (y / (A @ x)) - S
    g_flip = g0 + beta * _grad_R(x)
    assert np.max(np.abs(g0) / S) > 1e-3 and np.max(np.abs(g_flip) / S) > 1e-3


def test_osl_matches_scipy_minimiser(system):
    import torch
    from scipy.optimize import minimize
    grid, case, A, y = system
    beta = 0.5
    x_osl = _osl(case, y, beta, 1e-10, 4000, np.float64)

    def negF(v):
        t = torch.tensor(v, dtype=torch.float64, requires_grad=True)
        ax = torch.from_numpy(A) @ t
        f = -(torch.from_numpy(y) * torch.log(ax) - ax).sum() + beta * _rdp_value_torch(t, SHAPE, GAMMA)
        f.backward()
        return f.item(), t.grad.numpy()

    res = minimize(negF, np.ones(grid.N), jac=True, method="L-BFGS-B", bounds=[(1e-8, None)] * grid.N,
                   options={"maxiter": 20000, "ftol": 1e-15, "gtol": 1e-10})
    e = np.linalg.norm(res.x - x_osl) / np.linalg.norm(x_osl)
    assert e < 1e-5, f"OSL fixed point vs L-BFGS-B minimiser: rel L2 {e:.2e} ({res.message})"


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_osl_eps_placement(dtype):
    """With eps = 0.05 the fixed point satisfies A^T (y / (A x + eps)) = S + beta grad_eps R(x), grad_eps the
    reference's guarded RDP gradient; x / (S + dU) * (rhs + eps) would leave a residual of eps."""
    grid, case, A, y = _system(dtype)
    beta, eps = 0.5, 0.05
    x = _osl(case, y, beta, eps, 4000, dtype).astype(np.float64)
    S = A.sum(0)
    dU = O.rdp_gradient(x, SHAPE, GAMMA, eps, beta, dtype=np.float64)
    g = A.T @ (y / (A @ x + eps)) - S - dU
    inner = x > 2 * eps
    assert inner.mean() > 0.9
    tol = 1e-9 if dtype == np.float64 else 2e-5
    assert np.max(np.abs(g[inner]) / S[inner]) < tol
    # the misplaced-eps update's fixed point is off by eps / S relative, far above the tolerance
    assert eps / S.max() > 50 * tol
