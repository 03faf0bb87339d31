"""CPU ORACLE (test infrastructure only) for the 3-D U-Net kernels: float64 numpy restatement.

The U-Net has no counterpart in the reference; this oracle restates the textbook operations the
library's kernels implement, in the same layouts (activations [D, H, W, C] channels-last, weights
[27, cout, cin] with tap = (dz * 3 + dy) * 3 + dx, zero padding 1):

* conv3d(X, W, b)[v, co] = b[co] + sum_{tap, ci} W[tap, co, ci] X[v + off(tap), ci]
* conv3d_dx(dY, W)[v, ci] = sum_{tap, co} W[tap, co, ci] dY[v - off(tap), co]
* conv3d_dw(dY, X)[tap, co, ci] = sum_v dY[v, co] X[v + off(tap), ci]
* 2x2x2 average pooling, nearest x2 upsampling, ReLU, the 1x1 head and the U-Net built from them.

Pinning: conv3d against scipy.ndimage.correlate (an independent implementation) per channel pair,
conv3d_dx / conv3d_dw against finite differences of <conv3d(X, W), R>, and the network gradient
against central differences (tests/test_oracle_unet.py). Only tests/ import this module.
"""
from __future__ import annotations

import numpy as np

TAPS = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def bf16_round(a):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).reshape(np.shape(a))


def _shift(X, dz, dy, dx):
    """S[z, y, x] = X[z + dz, y + dy, x + dx] with zeros outside."""
    D, H, W = X.shape[:3]
    P = np.pad(X, ((1, 1), (1, 1), (1, 1)) + ((0, 0),) * (X.ndim - 3))
    return P[1 + dz:1 + dz + D, 1 + dy:1 + dy + H, 1 + dx:1 + dx + W]


def conv3d(X, W, b=None):
    X = np.asarray(X, np.float64)
    W = np.asarray(W, np.float64)
    out = np.zeros(X.shape[:3] + (W.shape[1],))
    for t, (dz, dy, dx) in enumerate(TAPS):
        out += _shift(X, dz, dy, dx) @ W[t].T
    if b is not None:
        out += np.asarray(b, np.float64)
    return out


def conv3d_dx(dY, W):
    dY = np.asarray(dY, np.float64)
    W = np.asarray(W, np.float64)
    out = np.zeros(dY.shape[:3] + (W.shape[2],))
    for t, (dz, dy, dx) in enumerate(TAPS):
        out += _shift(dY, -dz, -dy, -dx) @ W[t]
    return out


def conv3d_dw(dY, X):
    dY = np.asarray(dY, np.float64)
    X = np.asarray(X, np.float64)
    co, ci = dY.shape[3], X.shape[3]
    out = np.zeros((27, co, ci))
    for t, (dz, dy, dx) in enumerate(TAPS):
        out[t] = dY.reshape(-1, co).T @ _shift(X, dz, dy, dx).reshape(-1, ci)
    return out


def pool(X):
    D, H, W, Cc = X.shape
    return X.reshape(D // 2, 2, H // 2, 2, W // 2, 2, Cc).mean(axis=(1, 3, 5))


def pool_bwd(dY):
    return np.repeat(np.repeat(np.repeat(dY, 2, 0), 2, 1), 2, 2) / 8.0


def up(X):
    return np.repeat(np.repeat(np.repeat(X, 2, 0), 2, 1), 2, 2)


def up_bwd(dY):
    D, H, W, Cc = dY.shape
    return dY.reshape(D // 2, 2, H // 2, 2, W // 2, 2, Cc).sum(axis=(1, 3, 5))


LAYERS = [
    ("e1a", (16,), 16), ("e1b", (16,), 16), ("e2a", (16,), 32), ("e2b", (32,), 32), ("bta", (32,), 64),
    ("btb", (64,), 64), ("d2a", (64, 32), 32), ("d2b", (32,), 32), ("d1a", (32, 16), 16), ("d1b", (16,), 16),
]


def _ident(v):
    return v


def bf16_store(v):
    """What storing a float64 result as BF16 does (through float32, as the kernels round)."""
    return bf16_round(np.asarray(v, np.float32)).astype(np.float64)


def unet_forward(x0, P, rnd=_ident):
    """x0: [D, H, W, 16] embedded input; P: parameters (float arrays). Returns (head output before the
    residual, activations). ``rnd`` is applied wherever the library stores an activation in BF16
    (``bf16_store`` emulates that; the default keeps float64 throughout)."""
    A = {"x0": np.asarray(x0, np.float64)}
    r = lambda v: rnd(np.maximum(v, 0.0))
    A["a1"] = r(conv3d(A["x0"], P["e1a.w"], P["e1a.b"]))
    A["s1"] = r(conv3d(A["a1"], P["e1b.w"], P["e1b.b"]))
    A["p1"] = rnd(pool(A["s1"]))
    A["a2"] = r(conv3d(A["p1"], P["e2a.w"], P["e2a.b"]))
    A["s2"] = r(conv3d(A["a2"], P["e2b.w"], P["e2b.b"]))
    A["p2"] = rnd(pool(A["s2"]))
    A["a3"] = r(conv3d(A["p2"], P["bta.w"], P["bta.b"]))
    A["b3"] = r(conv3d(A["a3"], P["btb.w"], P["btb.b"]))
    A["u2"] = up(A["b3"])
    A["a4"] = r(conv3d(np.concatenate([A["u2"], A["s2"]], -1), P["d2a.w"], P["d2a.b"]))
    A["d2"] = r(conv3d(A["a4"], P["d2b.w"], P["d2b.b"]))
    A["u1"] = up(A["d2"])
    A["a5"] = r(conv3d(np.concatenate([A["u1"], A["s1"]], -1), P["d1a.w"], P["d1a.b"]))
    A["d1"] = r(conv3d(A["a5"], P["d1b.w"], P["d1b.b"]))
    head = A["d1"] @ np.asarray(P["head.w"], np.float64) + float(P["head.b"][0])
    return head, A


def unet_backward(gout, P, A, rnd=_ident):
    """Parameter gradients for d(loss)/d(head) = gout [D, H, W]; ``rnd`` as in :func:`unet_forward`, at the
    points where the library stores a gradient in BF16."""
    G = {}
    g = np.asarray(gout, np.float64)
    G["head.w"] = np.einsum("zyxc,zyx->c", A["d1"], g)
    G["head.b"] = np.array([g.sum()])
    g_d1 = rnd(g[..., None] * np.asarray(P["head.w"], np.float64)) * (A["d1"] > 0)

    def layer(name, gy, xin):
        G[name + ".w"] = conv3d_dw(gy, xin)
        G[name + ".b"] = gy.reshape(-1, gy.shape[-1]).sum(0)
        return rnd(conv3d_dx(gy, P[name + ".w"]))

    g_a5 = layer("d1b", g_d1, A["a5"]) * (A["a5"] > 0)
    g_cat1 = layer("d1a", g_a5, np.concatenate([A["u1"], A["s1"]], -1))
    g_d2 = rnd(up_bwd(g_cat1[..., :32])) * (A["d2"] > 0)
    g_a4 = layer("d2b", g_d2, A["a4"]) * (A["a4"] > 0)
    g_cat2 = layer("d2a", g_a4, np.concatenate([A["u2"], A["s2"]], -1))
    g_b3 = rnd(up_bwd(g_cat2[..., :64])) * (A["b3"] > 0)
    g_a3 = layer("btb", g_b3, A["a3"]) * (A["a3"] > 0)
    g_p2 = layer("bta", g_a3, A["p2"])
    g_s2 = rnd(rnd(pool_bwd(g_p2)) + g_cat2[..., 64:]) * (A["s2"] > 0)
    g_a2 = layer("e2b", g_s2, A["a2"]) * (A["a2"] > 0)
    g_p1 = layer("e2a", g_a2, A["p1"])
    g_s1 = rnd(rnd(pool_bwd(g_p1)) + g_cat1[..., 32:]) * (A["s1"] > 0)
    g_a1 = layer("e1b", g_s1, A["a1"]) * (A["a1"] > 0)
    layer("e1a", g_a1, A["x0"])
    return G
