"""Pins the CPU restatement of the built-in reconstruction loop (oracle/reconstruction.py).

Each algorithm is tied to something it does not restate itself:
* RDP gradient: the zero-boundary form equals the C oracle bit for bit (RDP.h order); the clamped form
  agrees with a float64 evaluation of the pair terms with edge-replicated neighbours;
* OSEM / OSL-OSEM: bit-identical to oracle.osem (itself pinned against the reference build and the
  OSL fixed point) on the same subsets;
* CGLS: equals oracle.cgls from x0 = 0 and reaches numpy's least-squares solution of a full-rank system;
* PDHG without subsets reaches the least-squares solution; with subsets its fixed point is the one the
  reference's uCP initialisation (a copy of x0, functions.hpp:2373-2375) produces:
  A^T (A x - y) = -x0 -- a mistake in the dual step, the accumulator or the extrapolation moves it;
* PKMA with RDP reaches the KKT point of max L(x) - R(x), x >= epps, and BSREM approaches the same point;
  MBSREM without a prior reaches scipy's maximiser of L over x >= epps.
The projector is a dense matrix built column by column from the oracle's Siddon forward projection of a
tiny list-mode problem.
"""
import numpy as np
import pytest

from omega_amd import geometry as G
from oracle import oracle as O
from oracle import reconstruction as R


class Dense:
    """A subset's projector as an explicit matrix (float32 products), with the oracle's interface."""

    def __init__(self, A):
        self.A = np.asarray(A, np.float32)

    def fp(self, img):
        return (self.A.astype(np.float64) @ np.asarray(img, np.float64)).astype(np.float32)

    def bp(self, y, sens=False):
        b = (self.A.T.astype(np.float64) @ np.asarray(y, np.float64)).astype(np.float32)
        if sens:
            return b, (self.A.T.astype(np.float64) @ np.ones(self.A.shape[0])).astype(np.float32)
        return b


class OracleCase:
    def __init__(self, case, n):
        self.c, self.n = case, n

    def fp(self, img):
        return self.c.fp(np.asarray(img, np.float32), self.n)

    def bp(self, y, sens=False):
        return self.c.bp(np.asarray(y, np.float32), self.n, sens=sens)


def _system(seed=0, n_lor=90, nx=5, ny=5):
    grid = G.Grid(nx, ny, 1, 2.0, 2.0, 2.0, -nx, -ny, -1.0)
    rng = np.random.default_rng(seed)
    a = rng.uniform(0, 2 * np.pi, n_lor)
    off = rng.uniform(-nx * 0.9, nx * 0.9, n_lor)
    c, s = np.cos(a), np.sin(a)
    p0 = np.stack([off * -s - 30 * c, off * c - 30 * s], 1)
    p1 = np.stack([off * -s + 30 * c, off * c + 30 * s], 1)
    rays = np.concatenate([p0[:, :1], p0[:, 1:], np.zeros((n_lor, 1)), p1[:, :1], p1[:, 1:], np.zeros((n_lor, 1))],
                          1).astype(np.float32).ravel()
    case = O.SiddonCase(grid, 1, 1, rays, np.zeros(2, np.float32), mode=O.MODE_LIST)
    N = grid.N
    A = np.stack([case.fp(np.eye(N, dtype=np.float32)[j], n_lor) for j in range(N)], 1)
    return grid, case, A


def test_rdp_zero_boundary_bitexact_with_c_oracle():
    shape = (9, 8, 7)
    u = (np.random.default_rng(1).random(int(np.prod(shape))) + 0.1).astype(np.float32)
    u[::13] = 0.0
    a = R.rdp_gradient(u, shape, 2.0, 1e-5, 0.3, clamp=False)
    b = O.rdp_gradient(u, shape, 2.0, 1e-5, 0.3)
    assert np.array_equal(a, b)


def test_rdp_clamp_boundary_float64():
    shape = (6, 5, 4)
    u = (np.random.default_rng(2).random(int(np.prod(shape))) + 0.2).astype(np.float32)
    g = R.rdp_gradient(u, shape, 1.5, 1e-5, 0.2, clamp=True)
    v = u.reshape(shape[::-1]).astype(np.float64)
    pad = np.pad(v, 1, mode="edge")
    c = pad[1:-1, 1:-1, 1:-1]
    acc = np.zeros_like(c)
    for k in (pad[1:-1, 1:-1, 2:], pad[1:-1, 1:-1, :-2], pad[1:-1, 2:, 1:-1], pad[1:-1, :-2, 1:-1],
              pad[2:, 1:-1, 1:-1], pad[:-2, 1:-1, 1:-1]):
        d = c - k
        den = c + k + 1.5 * np.abs(d)
        acc += d * (1.5 * np.abs(d) + c + 3 * k + 1e-10) / (den * den + 1e-5)
    ref = 0.2 * acc.ravel()
    assert np.max(np.abs(g - ref)) <= 1e-6 * np.max(np.abs(ref))


@pytest.mark.parametrize("beta", [0.0, 0.05])
def test_osem_and_osl_equal_oracle_osem(beta):
    grid = G.Grid.from_fov(16, 16, 4, 40.0, 40.0, 10.0)
    sc = G.pet_2d_planes(grid, radius=60.0, n_radial=16, n_angles=18, radial_pitch=2.5)
    idx, counts = G.subset_index_type8(sc.n_planes, 2)
    z = sc.z.reshape(-1, 2)[idx].ravel()
    per = 16 * 18
    rng = np.random.default_rng(3)
    y = rng.poisson(4.0, per * sc.n_planes).astype(np.float32)
    starts = np.concatenate([[0], np.cumsum(counts)])
    cases = [OracleCase(O.SiddonCase(grid, 16, 18, sc.xy_table(), z, plane_offset=int(starts[k])),
                        int(counts[k]) * per) for k in range(2)]
    parts = [y[starts[k] * per: starts[k + 1] * per] for k in range(2)]
    x0 = np.ones(grid.N, np.float32)
    p = R.Params(OSEM=beta == 0, OSL_OSEM=beta != 0, RDP=beta != 0, beta=beta, useImages=False, Niter=2,
                 shape=(16, 16, 4), enforcePositivity=True)
    a = R.omega_main(p, parts, cases, x0)["output"][-1]
    b = O.osem(grid, 16, 18, sc.xy_table(), z, counts, y, 2, x0, beta=beta, gamma=2.0)
    assert np.array_equal(a, b)


def test_cgls_matches_oracle_and_least_squares():
    grid, case, A = _system(0)
    rng = np.random.default_rng(4)
    xt = rng.random(grid.N).astype(np.float32) + 0.5
    y = (A.astype(np.float64) @ xt + rng.normal(0, 0.01, A.shape[0])).astype(np.float32)
    n = A.shape[0]
    # float32 round-off differences (gamma kept in float here, in double there) grow once the residual
    # nears the noise floor, so the step-by-step comparison is made over the first iterations
    p = R.Params(CGLS=True, CT=True, Niter=3, shape=(grid.Nx, grid.Ny, 1))
    a = R.omega_main(p, [y], [OracleCase(case, n)], np.zeros(grid.N, np.float32))["output"][-1]
    b = O.cgls(case, y, 3)
    assert np.linalg.norm(a - b) <= 1e-5 * np.linalg.norm(b)
    p.Niter = 60
    a = R.omega_main(p, [y], [OracleCase(case, n)], np.zeros(grid.N, np.float32))["output"][-1]
    ls = np.linalg.lstsq(A.astype(np.float64), y.astype(np.float64), rcond=None)[0]
    assert np.linalg.norm(a - ls) <= 1e-3 * np.linalg.norm(ls)


def test_pdhg_fixed_points():
    grid, case, A = _system(1)
    rng = np.random.default_rng(5)
    xt = rng.random(grid.N).astype(np.float32) + 0.5
    y = (A.astype(np.float64) @ xt).astype(np.float32)
    L = float(np.linalg.norm(A.astype(np.float64), 2) ** 2)
    x0 = np.full(grid.N, 0.3, np.float32)
    # no subsets: the least-squares solution
    p = R.Params(PDHG=True, CT=True, Niter=4000, tauCP=0.9 / L, sigmaCP=1.0, thetaCP=np.ones(4000),
                 shape=(grid.Nx, grid.Ny, 1))
    a = R.omega_main(p, [y], [Dense(A)], x0)["output"][-1]
    ls = np.linalg.lstsq(A.astype(np.float64), y.astype(np.float64), rcond=None)[0]
    assert np.linalg.norm(a - ls) <= 2e-3 * np.linalg.norm(ls)
    # three subsets: A^T (A x - y) = -x0 (uCP starts as x0)
    rows = [np.arange(k, A.shape[0], 3) for k in range(3)]
    it = 3000
    p = R.Params(PDHG=True, CT=True, Niter=it, tauCP=0.3 / L, sigmaCP=1.0, thetaCP=np.ones(3 * it),
                 shape=(grid.Nx, grid.Ny, 1))
    a = R.omega_main(p, [y[r] for r in rows], [Dense(A[r]) for r in rows], x0)["output"][-1]
    g = A.T.astype(np.float64) @ (A.astype(np.float64) @ a - y) + x0
    assert np.linalg.norm(g) <= 2e-3 * np.linalg.norm(x0)
    shifted = np.linalg.solve(A.T.astype(np.float64) @ A, A.T.astype(np.float64) @ y - x0)
    assert np.linalg.norm(a - shifted) <= 5e-3 * np.linalg.norm(shifted)


def _poisson(seed):
    grid, case, A = _system(seed, n_lor=120)
    rng = np.random.default_rng(seed + 10)
    xt = (rng.random(grid.N) * 3 + 1).astype(np.float32)
    y = rng.poisson(A.astype(np.float64) @ xt).astype(np.float32)
    return grid, A, y


def _kkt(A, y, x, beta, shape, epps=1e-5):
    """Projected gradient of L - R at x (float64): zero at the constrained maximum."""
    yh = A.astype(np.float64) @ x + epps
    gL = A.T.astype(np.float64) @ (y / yh - 1.0)
    gR = R.rdp_gradient(x.astype(np.float32), shape, 2.0, epps, beta, clamp=True).astype(np.float64)
    g = gL - gR
    free = x > 10 * epps
    return np.abs(g[free]).max() / np.abs(gL).max(), (g[~free] <= 1e-3 * np.abs(gL).max()).all()


def _pkma(A, y, shape, beta, it=1000):
    lam = (0.5 / (1.0 + np.arange(it) / 200.0)).astype(np.float32)
    p = R.Params(PKMA=True, lambdaN=lam, precondTypeImage=(False, True) + (False,) * 5,
                 alpha_PKMA=np.ones(it, np.float32), Niter=it, RDP=True, beta=beta, shape=shape,
                 enforcePositivity=True)
    return R.omega_main(p, [y], [Dense(A)], np.ones(A.shape[1], np.float32))["output"][-1]


def test_pkma_reaches_the_penalised_maximum():
    grid, A, y = _poisson(2)
    shape = (grid.Nx, grid.Ny, 1)
    x = _pkma(A, y, shape, 0.5)
    rel, bound_ok = _kkt(A, y, x, 0.5, shape)
    assert rel < 1e-5 and bound_ok, (rel, bound_ok)


@pytest.mark.parametrize("seed", [2, 3])
def test_bsrem_approaches_the_same_point(seed):
    """BSREM's relaxed EM step and end-of-iteration MAP step converge (as lambda -> 0) to the point PKMA
    reaches: the distance falls by more than 2.5x from 2000 to 8000 iterations and ends below 3e-3."""
    grid, A, y = _poisson(seed)
    shape = (grid.Nx, grid.Ny, 1)
    xp = _pkma(A, y, shape, 0.5)
    D = float(A.sum(0).max())
    dist = []
    for it in (2000, 8000):
        lam = (0.5 / (1.0 + np.arange(it) / 200.0) / D).astype(np.float32)
        p = R.Params(BSREM=True, lambdaN=lam, Niter=it, RDP=True, beta=0.5, shape=shape, enforcePositivity=True)
        x = R.omega_main(p, [y], [Dense(A)], np.ones(grid.N, np.float32))["output"][-1]
        dist.append(np.linalg.norm(x - xp) / np.linalg.norm(xp))
    assert dist[1] < 3e-3 and dist[1] * 2.5 < dist[0], dist


def test_mbsrem_reaches_the_bounded_ml_point():
    """MBSREM without a prior (EM preconditioner, upper bound U far away) ends at scipy's maximiser of
    the Poisson likelihood over x >= epps."""
    from scipy.optimize import minimize
    grid, A, y = _poisson(4)
    A64 = A.astype(np.float64)

    def negL(x):
        yh = A64 @ x + 1e-5
        return -(y @ np.log(yh) - yh.sum()), -(A64.T @ (y / yh - 1.0))

    ref = minimize(negL, np.ones(grid.N), jac=True, method="L-BFGS-B", bounds=[(1e-5, None)] * grid.N,
                   options=dict(maxiter=20000, ftol=1e-15, gtol=1e-12)).x
    it = 3000
    lam = (2.0 / (1.0 + np.arange(it) / 300.0)).astype(np.float32)
    p = R.Params(MBSREM=True, Niter=it, lambdaN=lam, U=1e4, precondTypeImage=(False, True) + (False,) * 5,
                 shape=(grid.Nx, grid.Ny, 1), enforcePositivity=True)
    x = R.omega_main(p, [y], [Dense(A)], np.ones(grid.N, np.float32))["output"][-1].astype(np.float64)
    assert np.linalg.norm(x - ref) <= 1e-3 * np.linalg.norm(ref)
