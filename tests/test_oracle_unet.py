"""Pins the U-Net oracle: convolution against scipy.ndimage.correlate, the convolution adjoints
against finite differences, and the network gradient against central differences (float64)."""
import numpy as np
from scipy import ndimage

from oracle import unet_oracle as U


def test_conv3d_matches_scipy_correlate():
    rng = np.random.default_rng(0)
    X = rng.normal(size=(5, 6, 7, 3))
    W = rng.normal(size=(27, 2, 3))
    b = rng.normal(size=2)
    Y = U.conv3d(X, W, b)
    for co in range(2):
        ref = np.full(X.shape[:3], b[co])
        for ci in range(3):
            k = W[:, co, ci].reshape(3, 3, 3)       # k[dz+1, dy+1, dx+1]
            ref += ndimage.correlate(X[..., ci], k, mode="constant", cval=0.0)
        assert np.allclose(Y[..., co], ref, atol=1e-12)


def test_conv_adjoints_by_finite_differences():
    rng = np.random.default_rng(1)
    X = rng.normal(size=(4, 4, 6, 2))
    W = rng.normal(size=(27, 3, 2))
    R = rng.normal(size=(4, 4, 6, 3))
    f = lambda X_, W_: float((U.conv3d(X_, W_) * R).sum())
    dX = U.conv3d_dx(R, W)
    dW = U.conv3d_dw(R, X)
    h = 1e-6
    for idx in [(0, 0, 0, 0), (3, 2, 5, 1), (1, 3, 0, 1), (2, 1, 3, 0)]:
        E = np.zeros_like(X); E[idx] = h
        assert abs((f(X + E, W) - f(X - E, W)) / (2 * h) - dX[idx]) < 1e-6
    for idx in [(0, 0, 0), (13, 2, 1), (26, 1, 0), (5, 0, 1)]:
        E = np.zeros_like(W); E[idx] = h
        assert abs((f(X, W + E) - f(X, W - E)) / (2 * h) - dW[idx]) < 1e-6


def test_pool_up_adjoints():
    rng = np.random.default_rng(2)
    X = rng.normal(size=(4, 6, 8, 2))
    Yc = rng.normal(size=(2, 3, 4, 2))
    assert np.isclose((U.pool(X) * Yc).sum(), (X * U.pool_bwd(Yc)).sum())
    assert np.isclose((U.up(Yc) * X).sum(), (Yc * U.up_bwd(X)).sum())


def test_unet_gradient_central_differences():
    rng = np.random.default_rng(3)
    P = {}
    for name, cins, cout in U.LAYERS:
        P[name + ".w"] = rng.normal(0, np.sqrt(2.0 / (27 * sum(cins))), size=(27, cout, sum(cins)))
        P[name + ".b"] = rng.normal(0, 0.05, size=cout)
    P["head.w"] = rng.normal(0, 0.3, size=16)
    P["head.b"] = np.array([0.1])
    x0 = np.zeros((4, 4, 4, 16)); x0[..., 0] = rng.random((4, 4, 4))
    T = rng.normal(size=(4, 4, 4))
    loss = lambda P_: float(((U.unet_forward(x0, P_)[0] - T) ** 2).sum())
    head, A = U.unet_forward(x0, P)
    G = U.unet_backward(2 * (head - T), P, A)
    h = 1e-6
    for key, idx in [("e1a.w", (13, 3, 0)), ("e2b.w", (4, 7, 20)), ("btb.w", (20, 10, 50)), ("d2a.w", (13, 5, 70)),
                     ("d1a.w", (0, 2, 40)), ("d1b.b", (4,)), ("head.w", (3,)), ("bta.b", (17,))]:
        Pp = {k: v.copy() for k, v in P.items()}; Pm = {k: v.copy() for k, v in P.items()}
        Pp[key][idx] += h; Pm[key][idx] -= h
        fd = (loss(Pp) - loss(Pm)) / (2 * h)
        assert abs(fd - G[key][idx]) <= 1e-5 * max(1.0, abs(fd)), (key, fd, G[key][idx])


def test_bf16_round_matches_torch():
    import torch
    x = np.random.default_rng(4).normal(size=1000).astype(np.float32) * 100
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(U.bf16_round(x), ref)
