"""CGLS and PDHG recurrences of the oracle against closed forms on small dense problems.

The system matrix of a tiny Siddon geometry is assembled explicitly from one-hot forward
projections; CGLS must reach numpy.linalg.lstsq's least-squares solution (a library routine),
and PDHG with positivity must reach the non-negative least-squares solution (scipy nnls)."""
import numpy as np

from oracle import oracle as O
from tests import cases


def dense_system():
    grid = cases.tiny_grid()
    x = cases.list_rays(grid, 600, seed=11)
    n = x.size // 6
    c = O.SiddonCase(grid, 1, 1, x.astype(np.float64), np.zeros(2), mode=O.MODE_LIST, dtype=np.float64)
    A = np.stack([c.fp(np.eye(grid.N)[j], n) for j in range(grid.N)], axis=1)
    return c, A, n


def test_cgls_reaches_least_squares():
    c, A, n = dense_system()
    rng = np.random.default_rng(0)
    xt = rng.random(A.shape[1])
    y = A @ xt + 0.01 * rng.standard_normal(n)
    f = O.cgls(c, y, 1000, nthreads=1)   # rank-deficient A (rank 149 of 280 columns): CGLS converges in ~rank steps, slowly
    ls = np.linalg.lstsq(A, y, rcond=None)[0]
    # compare in the data space (A may have a null space on voxels no ray crosses)
    assert np.linalg.norm(A @ f - A @ ls) <= 1e-6 * np.linalg.norm(A @ ls)


def test_pdhg_reaches_nonnegative_least_squares():
    from scipy.optimize import nnls
    c, A, n = dense_system()
    rng = np.random.default_rng(1)
    y = A @ rng.random(A.shape[1])
    L = np.linalg.norm(A, 2) ** 2
    f = O.pdhg([c], [y], 3000, tau=1.0 / L, sigma=1.0, theta=1.0, epps=0.0, nthreads=1)
    ref = nnls(A, y, maxiter=5000)[0]
    assert np.linalg.norm(A @ f - A @ ref) <= 2e-3 * np.linalg.norm(A @ ref)
