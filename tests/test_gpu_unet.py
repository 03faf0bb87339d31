"""U-Net kernels (BF16 MFMA, FP32 accumulation) against the float64 oracle.

Bar: a convolution's BF16 output within one BF16 rounding (2^-8 relative) of the oracle on BF16 inputs,
plus the recursive-summation bound n * 2^-24 * sum|x w| of the FP32 accumulation over its n = 27 * Cin
exact BF16 products (it matters only near cancellation; a dropped tap or a wrong neighbour is off by
O(1)); dX likewise; dW (FP32 output) within 1e-4 relative to its largest entry; the
whole network's output and every parameter gradient within 1e-2 relative L2 of the oracle run with
BF16 storage emulated at the same points (float64 arithmetic in between); against the pure float64
network (no BF16 storage) the output agrees to 1e-2 and the gradients to 0.25, the accumulated
effect of BF16 activations through ten layers; Adam training lowers the loss.
"""
import ctypes as C

import numpy as np
import pytest

from omega_amd import _lib
from omega_amd import unet as UN
from oracle import unet_oracle as U
from tests.hip_helpers import rel_l2

pytestmark = pytest.mark.gpu


def _p(t):
    return C.c_void_p(0 if t is None else t.data_ptr())


def _bf(torch, a, cuda):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(cuda).to(torch.bfloat16)


SHAPES = [((16,), 16), ((16,), 32), ((32,), 64), ((64, 32), 32), ((32, 16), 16), ((64,), 64)]


@pytest.mark.parametrize("dhw", [(4, 6, 10), (3, 5, 32)])
@pytest.mark.parametrize("cins,cout", SHAPES)
def test_conv_forward_dx_dw(cuda, cins, cout, dhw):
    """(4, 6, 10): 240 voxels, a partial last wave, rows not a multiple of 16; (3, 5, 32): rows of 32
    (the weight gradient's row-wise neighbour path)."""
    import torch
    rng = np.random.default_rng(sum(cins) + cout)
    D, H, W = dhw
    cin = sum(cins)
    print(f"[unet conv] cins={cins} cout={cout}", flush=True)
    Xs = [U.bf16_round(rng.normal(size=(D, H, W, c))) for c in cins]
    Wt = U.bf16_round(rng.normal(0, 0.2, size=(27, cout, cin)))
    b = rng.normal(0, 0.1, size=cout).astype(np.float32)
    X = np.concatenate(Xs, -1)
    ref = np.maximum(U.conv3d(X, Wt, b), 0.0)
    xt = [_bf(torch, x, cuda) for x in Xs]
    wt = _bf(torch, Wt, cuda)
    bt = torch.from_numpy(b).to(cuda)
    y = torch.empty((D * H * W, cout), dtype=torch.bfloat16, device=cuda)
    _lib.call("omg_conv3d_forward", _p(xt[0]), cins[0], _p(xt[1] if len(xt) > 1 else None),
              cins[1] if len(cins) > 1 else 0, _p(wt), _p(bt), _p(y), cout, D, H, W, 1, None)
    got = y.float().cpu().numpy().reshape(D, H, W, cout)
    acc = 27 * cin * 2.0 ** -24 * U.conv3d(np.abs(X), np.abs(Wt), np.abs(b))
    assert np.all(np.abs(got - ref) <= 2.0 ** -8 * np.abs(ref) + acc), np.abs(got - ref).max()
    # dX = conv of dY with the flipped transposed weights
    dY = U.bf16_round(rng.normal(size=(D, H, W, cout)))
    wf = torch.empty((27, cin, cout), dtype=torch.bfloat16, device=cuda)
    _lib.call("omg_conv3d_flipT", _p(wt), _p(wf), cout, cin, None)
    dyt = _bf(torch, dY, cuda)
    dx = torch.empty((D * H * W, cin), dtype=torch.bfloat16, device=cuda)
    _lib.call("omg_conv3d_forward", _p(dyt), cout, None, 0, _p(wf), None, _p(dx), cin, D, H, W, 0, None)
    ref_dx = U.conv3d_dx(dY, Wt)
    got_dx = dx.float().cpu().numpy().reshape(D, H, W, cin)
    acc_dx = 27 * cout * 2.0 ** -24 * U.conv3d_dx(np.abs(dY), np.abs(Wt))
    assert np.all(np.abs(got_dx - ref_dx) <= 2.0 ** -8 * np.abs(ref_dx) + acc_dx), np.abs(got_dx - ref_dx).max()
    # dW
    dw = torch.zeros((27, cout, cin), dtype=torch.float32, device=cuda)
    _lib.call("omg_conv3d_wgrad", _p(dyt), cout, _p(xt[0]), cins[0], _p(xt[1] if len(xt) > 1 else None),
              cins[1] if len(cins) > 1 else 0, _p(dw), D, H, W, None)
    ref_dw = U.conv3d_dw(dY, X)
    assert np.abs(dw.cpu().numpy() - ref_dw).max() <= 1e-4 * np.abs(ref_dw).max()


@pytest.mark.parametrize("dhw,cins,cout", [
    ((32, 128, 64), (16,), 16),      # 8-wave blocks of 16 x 8 x 4 (the full-resolution layers)
    ((32, 128, 64), (32, 16), 16),   # the same with two inputs, 48 channels
    ((16, 64, 64), (16,), 32),       # 4-wave blocks of 16 x 4 x 4
    ((16, 64, 64), (64, 32), 32),    # 96 input channels
    ((16, 32, 32), (64,), 64),       # one-row blocks (small volumes)
    ((13, 30, 40), (32,), 32),       # partial one-row blocks at every edge
    ((34, 126, 70), (16,), 16),      # partial 8-wave blocks at every edge
])
def test_conv_forward_tile_shapes(cuda, dhw, cins, cout):
    """The LDS-tiled convolution on volumes that select each of its block shapes, against the float64
    oracle with the bound of test_conv_forward_dx_dw."""
    import torch
    rng = np.random.default_rng(7 + sum(cins) + cout)
    D, H, W = dhw
    cin = sum(cins)
    print(f"[unet tile] {dhw} cins={cins} cout={cout}", flush=True)
    Xs = [U.bf16_round(rng.normal(size=(D, H, W, c))) for c in cins]
    Wt = U.bf16_round(rng.normal(0, 0.2, size=(27, cout, cin)))
    b = rng.normal(0, 0.1, size=cout).astype(np.float32)
    X = np.concatenate(Xs, -1)
    ref = np.maximum(U.conv3d(X, Wt, b), 0.0)
    xt = [_bf(torch, x, cuda) for x in Xs]
    wt = _bf(torch, Wt, cuda)            # held: the kernel reads them after the call returns
    bt = torch.from_numpy(b).to(cuda)
    y = torch.empty((D * H * W, cout), dtype=torch.bfloat16, device=cuda)
    _lib.call("omg_conv3d_forward", _p(xt[0]), cins[0], _p(xt[1] if len(xt) > 1 else None),
              cins[1] if len(cins) > 1 else 0, _p(wt), _p(bt), _p(y), cout, D, H, W, 1, None)
    got = y.float().cpu().numpy().reshape(D, H, W, cout)
    acc = 27 * cin * 2.0 ** -24 * U.conv3d(np.abs(X), np.abs(Wt), np.abs(b))
    assert np.all(np.abs(got - ref) <= 2.0 ** -8 * np.abs(ref) + acc), np.abs(got - ref).max()


def test_network_forward_and_gradients(cuda):
    import torch
    nx = ny = 8
    nz = 7                     # padded to 8 planes
    net = UN.UNet3D(nx, ny, nz, seed=3, device=cuda)
    rng = np.random.default_rng(11)
    # non-trivial head and biases so every path carries gradient
    for k in net.p:
        if k.endswith(".b") or k == "head.w":
            net.p[k].copy_(torch.from_numpy(rng.normal(0, 0.1, size=tuple(net.p[k].shape)).astype(np.float32)))
    net._sync_bf16()
    img = torch.from_numpy(rng.random(nx * ny * nz).astype(np.float32)).to(cuda)
    scale = 2.0
    out = net.forward(img, scale).cpu().numpy()
    target = torch.from_numpy(rng.random(nx * ny * nz).astype(np.float32)).to(cuda)
    mask = net.mask()
    net.zero_grad()
    net.loss(target, mask, scale)
    net.backward(net.act["gout"])
    # oracle on the GPU's BF16 weights and embedded input
    P = {}
    for name, cins, cout, _ in UN.LAYERS:
        P[name + ".w"] = net.wb[name].float().cpu().numpy().astype(np.float64)
        P[name + ".b"] = net.p[name + ".b"].cpu().numpy().astype(np.float64)
    P["head.w"] = net.p["head.w"].cpu().numpy().astype(np.float64)
    P["head.b"] = net.p["head.b"].cpu().numpy().astype(np.float64)
    D = net.D[0]
    x0 = net.act["x0"].float().cpu().numpy().reshape(D, ny, nx, 16)
    base = np.zeros(D * ny * nx); base[: nx * ny * nz] = img.cpu().numpy() * scale
    tgt = np.zeros(D * ny * nx); tgt[: nx * ny * nz] = target.cpu().numpy() * scale
    m = mask.cpu().numpy()
    # the loss and its gradient against the GPU's own output
    assert abs(float(net.act["loss"]) - np.sum(m * (out - tgt) ** 2) / m.sum()) <= 1e-6 * float(net.act["loss"])
    assert np.allclose(net.act["gout"].cpu().numpy(), 2.0 * (out - tgt) * m / m.sum(), rtol=1e-5, atol=1e-9)
    for rnd, tol, tol_out in ((U.bf16_store, 1e-2, 1e-3), (U._ident, 0.25, 1e-2)):
        head, A = U.unet_forward(x0, P, rnd)
        ref_out = head.reshape(-1) + base
        assert rel_l2(out, ref_out) <= tol_out
        gout = 2.0 * (ref_out - tgt) * m / m.sum()
        G = U.unet_backward(gout.reshape(D, ny, nx), P, A, rnd)
        for k in G:
            err = rel_l2(net.g[k].cpu().numpy(), G[k])
            print(f"  grad {k} ({'bf16 storage' if rnd is U.bf16_store else 'float64'}): rel L2 {err:.2e}", flush=True)
            assert err <= tol, k


def test_masked_mse_many_blocks(cuda):
    """300000 voxels: more than one pass of the capped grid-stride loop; the FP64 loss and the FP32
    gradient against numpy."""
    import torch
    rng = np.random.default_rng(5)
    V = 300_000
    out = rng.random(V).astype(np.float32)
    tgt = rng.random(V).astype(np.float32)
    m = (rng.random(V) < 0.7).astype(np.float32)
    g = torch.empty(V, dtype=torch.float32, device=cuda)
    loss = torch.zeros(1, dtype=torch.float64, device=cuda)
    tt = [torch.from_numpy(a).to(cuda) for a in (out, tgt, m)]
    inv_n = np.float32(1.0 / m.sum())
    _lib.call("omg_mse", _p(tt[0]), _p(tt[1]), _p(tt[2]), _p(g), _p(loss), V, float(inv_n), None)
    e = ((out - tgt) * m).astype(np.float64)
    assert abs(float(loss) - np.sum(e * e) * float(inv_n)) <= 1e-9 * float(loss)
    assert np.allclose(g.cpu().numpy(), 2.0 * e * float(inv_n), rtol=1e-6, atol=0)


def test_training_lowers_loss(cuda):
    import torch
    nx = ny = 16
    nz = 8
    net = UN.UNet3D(nx, ny, nz, seed=5, device=cuda)
    rng = np.random.default_rng(12)
    img = torch.from_numpy(rng.random(nx * ny * nz).astype(np.float32)).to(cuda)
    # target: a smoothed version of the input (a map the network can learn)
    a = img.reshape(nz, ny, nx)
    tgt = ((a + a.roll(1, 0) + a.roll(1, 1) + a.roll(1, 2)) / 4.0).reshape(-1).contiguous()
    mask = net.mask()
    losses = []
    for it in range(60):
        net.forward(img, 1.0)
        net.zero_grad()
        losses.append(float(net.loss(tgt, mask, 1.0)))
        net.backward(net.act["gout"])
        net.adam_step(lr=2e-3)
    print(f"  loss {losses[0]:.4e} -> {losses[-1]:.4e}", flush=True)
    assert losses[-1] < 0.5 * losses[0]
