"""The oracle's transmission (CT) OSEM / OSL-OSEM against what the mathematics fixes.

The reference's CT update (subiterStep.h:117-118,223-224 with computeForwardProject.h:80-88) is
x <- x / (S + dU) * flat A^T exp(-A x), S = A^T y. Its fixed points with x > eps are the stationary points
of the transmission log-likelihood L(x) = sum_i -y_i (A x)_i - flat exp(-(A x)_i) minus beta R(x):
A^T (flat exp(-A x)) - A^T y - beta grad R(x) = 0. Checked on a small cone-beam Siddon system with an
explicit matrix:

* noise-free counts y = flat exp(-A mu): one iteration from mu returns mu (float64, to rounding);
* noisy counts, OSL to convergence: the stationarity residual with grad R from torch autograd of an
  independently written RDP value vanishes, and the image equals scipy's L-BFGS-B minimiser;
* subsets: with noise-free counts every subset's update keeps mu.

A missing flat factor, S taken from A^T 1, exp(+A x) or the prior's sign flipped each fail one of these.
"""
import numpy as np
import pytest

from omega_amd import geometry as G
from oracle import oracle as O
from tests.test_oracle_osl import _grad_R, _rdp_value_torch, SHAPE, GAMMA

FLAT = 1000.0


def _ct_system(n_views=16, subsets=1):
    grid = G.Grid(SHAPE[0], SHAPE[1], SHAPE[2], 2.0, 2.5, 3.0, -6.0, -6.25, -4.5)
    x, uv = G.cbct_tables(n_views, sod=40.0, sdd=70.0, pitch_u=1.2)
    rows, cols = 12, 8
    idx, nmeas = G.subset_index_type8(n_views, subsets)
    x6 = x.reshape(-1, 6)[idx].reshape(-1)
    uv2 = uv.reshape(-1, 2)[idx].reshape(-1)
    cases, starts = [], np.concatenate([[0], np.cumsum(nmeas)])
    for k in range(subsets):
        cases.append(O.SiddonCase(grid, rows, cols, x6.astype(np.float64), uv2.astype(np.float64), mode=O.MODE_CT,
                                  dPitchZ=1.2, plane_offset=int(starts[k]), n_proj=n_views, dtype=np.float64))
    per = rows * cols
    A = []
    for k in range(subsets):
        n = int(nmeas[k]) * per
        A.append(np.stack([cases[k].fp(np.eye(grid.N)[j], n) for j in range(grid.N)], axis=1))
    return grid, cases, A


@pytest.fixture(scope="module")
def system():
    return _ct_system()


def test_noise_free_fixed_point():
    grid, cases, A = _ct_system(subsets=4)
    mu = np.random.default_rng(1).random(grid.N) * 0.03 + 0.005
    parts = [FLAT * np.exp(-Ak @ mu) for Ak in A]
    x = O.ct_osem(cases, parts, 1, mu, FLAT, epps=1e-12)
    assert np.max(np.abs(x - mu) / mu) < 1e-11
    # and from another start the image moves (the check is not vacuous)
    x1 = O.ct_osem(cases, parts, 1, np.full(grid.N, 0.02), FLAT, epps=1e-12, subsets_to_run=1)
    assert np.max(np.abs(x1 - 0.02)) > 1e-4


@pytest.mark.parametrize("beta", [0.0, 20.0])
def test_osl_fixed_point_is_stationary(system, beta):
    grid, cases, A = system
    A = A[0]
    assert np.linalg.matrix_rank(A) == grid.N and (A.sum(0) > 0).all()
    rng = np.random.default_rng(2)
    mu = rng.random(grid.N) * 0.03 + 0.005
    y = rng.poisson(FLAT * np.exp(-A @ mu)).astype(np.float64)
    x = O.ct_osem(cases, [y], 3000, np.full(grid.N, 0.01), FLAT, epps=1e-12, beta=beta, shape=SHAPE, nthreads=1)
    assert x.min() > 1e-4
    g = A.T @ (FLAT * np.exp(-A @ x)) - A.T @ y
    if beta:
        g = g - beta * _grad_R(x)
    scale = A.T @ y
    assert np.max(np.abs(g) / scale) < 1e-6
    if beta:
        g0 = A.T @ (FLAT * np.exp(-A @ x)) - A.T @ y
        assert np.max(np.abs(g0) / scale) > 1e-4


def test_osl_matches_scipy_minimiser(system):
    import torch
    from scipy.optimize import minimize
    grid, cases, A = system
    A = A[0]
    rng = np.random.default_rng(3)
    mu = rng.random(grid.N) * 0.03 + 0.005
    y = rng.poisson(FLAT * np.exp(-A @ mu)).astype(np.float64)
    beta = 20.0
    x_osl = O.ct_osem(cases, [y], 3000, np.full(grid.N, 0.01), FLAT, epps=1e-12, beta=beta, shape=SHAPE, nthreads=1)
    At = torch.from_numpy(A)
    yt = torch.from_numpy(y)

    def negF(v):
        t = torch.tensor(v, dtype=torch.float64, requires_grad=True)
        ax = At @ t
        f = (yt * ax + FLAT * torch.exp(-ax)).sum() + beta * _rdp_value_torch(t, SHAPE, GAMMA)
        f.backward()
        return f.item(), t.grad.numpy()

    res = minimize(negF, np.full(grid.N, 0.01), jac=True, method="L-BFGS-B", bounds=[(1e-8, None)] * grid.N,
                   options={"maxiter": 20000, "ftol": 1e-16, "gtol": 1e-12})
    e = np.linalg.norm(res.x - x_osl) / np.linalg.norm(x_osl)
    assert e < 1e-5, f"CT OSL fixed point vs L-BFGS-B: rel L2 {e:.2e} ({res.message})"
