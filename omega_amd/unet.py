"""3-D U-Net warm start for OSEM-RDP, on the library's BF16-MFMA kernels.

Network (3 levels, 16/32/64 channels, 3x3x3 convolutions, ReLU, 2x2x2 average pooling, nearest
upsampling, skip concatenation, a 1x1 FP32 output layer and a residual connection to the input):

    x0  = embed(image * scale)                       [V0, 16]   (channel 0 = image, 15 zero channels)
    a1  = relu(conv e1a(x0));  s1 = relu(conv e1b(a1))            16
    a2  = relu(conv e2a(pool s1)); s2 = relu(conv e2b(a2))        32
    a3  = relu(conv bta(pool s2)); b3 = relu(conv btb(a3))        64
    a4  = relu(conv d2a([up b3 | s2])); d2 = relu(conv d2b(a4))   32
    a5  = relu(conv d1a([up d2 | s1])); d1 = relu(conv d1b(a5))   16
    out = image * scale + head(d1)                   [V0] FP32

Activations and convolution weights are BF16 with FP32 accumulation (omg_conv3d_forward); the
backward pass is explicit (omg_conv3d_wgrad, the same convolution kernel with flipped transposed
weights for dX, and the elementwise kernels); Adam keeps FP32 master weights (omg_adam). The volume's
z extent is zero-padded to a multiple of 4 for the two pooling levels.

torch is used only to allocate device buffers.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib

# name, input channels per input tensor, output channels, level
LAYERS = [
    ("e1a", (16,), 16, 0), ("e1b", (16,), 16, 0),
    ("e2a", (16,), 32, 1), ("e2b", (32,), 32, 1),
    ("bta", (32,), 64, 2), ("btb", (64,), 64, 2),
    ("d2a", (64, 32), 32, 1), ("d2b", (32,), 32, 1),
    ("d1a", (32, 16), 16, 0), ("d1b", (16,), 16, 0),
]


def _p(t):
    return C.c_void_p(0 if t is None else t.data_ptr())


def padded_depth(nz: int) -> int:
    return int(math.ceil(nz / 4.0) * 4)


def init_params(seed: int = 0):
    """He-initialised FP32 parameters as numpy arrays: conv weights [27, cout, cin], biases, head."""
    rng = np.random.default_rng(seed)
    params = {}
    for name, cins, cout, _ in LAYERS:
        cin = sum(cins)
        fan_in = 27 * (1 if name == "e1a" else cin)   # the embedded input has one live channel
        w = rng.normal(0.0, math.sqrt(2.0 / fan_in), size=(27, cout, cin)).astype(np.float32)
        params[name + ".w"] = w
        params[name + ".b"] = np.zeros(cout, np.float32)
    params["head.w"] = (rng.normal(0.0, 1e-3, size=16)).astype(np.float32)
    params["head.b"] = np.zeros(1, np.float32)
    return params


class UNet3D:
    """Device-resident network for images of ``nx x ny x nz`` voxels (nx, ny multiples of 4)."""

    def __init__(self, nx: int, ny: int, nz: int, params=None, seed: int = 0, device="cuda"):
        import torch
        self.torch = torch
        if nx % 4 or ny % 4:
            raise ValueError("the image's x and y extents must be multiples of 4")
        self.nx, self.ny, self.nz = nx, ny, nz
        self.D = [padded_depth(nz) >> l for l in range(3)]
        self.H = [ny >> l for l in range(3)]
        self.W = [nx >> l for l in range(3)]
        self.V = [self.D[l] * self.H[l] * self.W[l] for l in range(3)]
        self.dev = device
        params = init_params(seed) if params is None else params
        f32 = dict(dtype=torch.float32, device=device)
        self.p = {k: torch.as_tensor(np.ascontiguousarray(v), **f32) for k, v in params.items()}
        self.g = {k: torch.zeros_like(v) for k, v in self.p.items()}
        self.m = {k: torch.zeros_like(v) for k, v in self.p.items()}
        self.v = {k: torch.zeros_like(v) for k, v in self.p.items()}
        self.wb = {}    # BF16 weights [27, cout, cin]
        self.wf = {}    # BF16 flipped-transposed weights [27, cin, cout]
        for name, cins, cout, _ in LAYERS:
            self.wb[name] = torch.empty((27, cout, sum(cins)), dtype=torch.bfloat16, device=device)
            self.wf[name] = torch.empty((27, sum(cins), cout), dtype=torch.bfloat16, device=device)
        self.step_count = 0
        self._sync_bf16()
        self.act = {}
        self._alloc()

    # ---- buffers ------------------------------------------------------------------------------------
    def _buf(self, level, ch):
        return self.torch.empty((self.V[level], ch), dtype=self.torch.bfloat16, device=self.dev)

    def _alloc(self):
        t = self.torch
        a = self.act
        a["x0"] = self._buf(0, 16)
        for name, lvl, ch in (("a1", 0, 16), ("s1", 0, 16), ("p1", 1, 16), ("a2", 1, 32), ("s2", 1, 32),
                              ("p2", 2, 32), ("a3", 2, 64), ("b3", 2, 64), ("u2", 1, 64), ("a4", 1, 32),
                              ("d2", 1, 32), ("u1", 0, 32), ("a5", 0, 16), ("d1", 0, 16)):
            a[name] = self._buf(lvl, ch)
        a["base"] = t.zeros(self.V[0], dtype=t.float32, device=self.dev)
        a["out"] = t.zeros(self.V[0], dtype=t.float32, device=self.dev)
        a["gout"] = t.zeros(self.V[0], dtype=t.float32, device=self.dev)
        # gradient buffers (reused across levels)
        a["g16_0"] = self._buf(0, 16)
        a["g16_0b"] = self._buf(0, 16)
        a["g48_0"] = self._buf(0, 48)
        a["g32_1"] = self._buf(1, 32)
        a["g32_1b"] = self._buf(1, 32)
        a["g96_1"] = self._buf(1, 96)
        a["g16_1"] = self._buf(1, 16)
        a["g64_2"] = self._buf(2, 64)
        a["g64_2b"] = self._buf(2, 64)
        a["g32_2"] = self._buf(2, 32)
        a["loss"] = t.zeros(1, dtype=t.float64, device=self.dev)

    def _s(self):
        return C.c_void_p(self.torch.cuda.current_stream().cuda_stream)

    def _sync_bf16(self):
        for name, cins, cout, _ in LAYERS:
            w = self.p[name + ".w"]
            _lib.call("omg_to_bf16", _p(w), _p(self.wb[name]), w.numel(), self._s())
            _lib.call("omg_conv3d_flipT", _p(self.wb[name]), _p(self.wf[name]), cout, sum(cins), self._s())

    # ---- layers -------------------------------------------------------------------------------------
    def _conv(self, name, X1, X2, Y, relu=True):
        spec = {n: (c, o, l) for n, c, o, l in LAYERS}[name]
        cins, cout, l = spec
        c2 = cins[1] if len(cins) > 1 else 0
        _lib.call("omg_conv3d_forward", _p(X1), cins[0], _p(X2), c2, _p(self.wb[name]), _p(self.p[name + ".b"]),
                  _p(Y), cout, self.D[l], self.H[l], self.W[l], int(relu), self._s())

    def _conv_dx(self, name, dY, dX):
        cins, cout, l = {n: (c, o, l) for n, c, o, l in LAYERS}[name]
        _lib.call("omg_conv3d_forward", _p(dY), cout, None, 0, _p(self.wf[name]), None, _p(dX), sum(cins),
                  self.D[l], self.H[l], self.W[l], 0, self._s())

    def _wgrad(self, name, dY, X1, X2=None):
        cins, cout, l = {n: (c, o, l) for n, c, o, l in LAYERS}[name]
        c2 = cins[1] if len(cins) > 1 else 0
        _lib.call("omg_conv3d_wgrad", _p(dY), cout, _p(X1), cins[0], _p(X2), c2, _p(self.g[name + ".w"]),
                  self.D[l], self.H[l], self.W[l], self._s())
        _lib.call("omg_bias_grad", _p(dY), _p(self.g[name + ".b"]), self.V[l], cout, self._s())

    def _relu_bwd(self, Y, dY):
        _lib.call("omg_relu_backward", _p(Y), _p(dY), dY.numel(), self._s())

    def _pool(self, X, Y, l, ch):
        _lib.call("omg_pool_forward", _p(X), _p(Y), self.D[l], self.H[l], self.W[l], ch, self._s())

    def _up(self, X, Y, l, ch):
        _lib.call("omg_upsample_forward", _p(X), _p(Y), self.D[l], self.H[l], self.W[l], ch, self._s())

    # ---- forward / backward -------------------------------------------------------------------------
    def forward(self, image, scale: float):
        """image: FP32 device tensor of nz*ny*nx voxels. Returns the FP32 output [V0] (padded planes included);
        the prediction for the image is out[:nz*ny*nx] / scale."""
        a = self.act
        t = self.torch
        img = image.reshape(-1).contiguous()
        _lib.call("omg_embed", _p(img), _p(a["x0"]), self.D[0], self.H[0], self.W[0], self.nz, float(scale), self._s())
        a["base"].zero_()
        a["base"][: img.numel()] = img * scale
        self._conv("e1a", a["x0"], None, a["a1"])
        self._conv("e1b", a["a1"], None, a["s1"])
        self._pool(a["s1"], a["p1"], 0, 16)
        self._conv("e2a", a["p1"], None, a["a2"])
        self._conv("e2b", a["a2"], None, a["s2"])
        self._pool(a["s2"], a["p2"], 1, 32)
        self._conv("bta", a["p2"], None, a["a3"])
        self._conv("btb", a["a3"], None, a["b3"])
        self._up(a["b3"], a["u2"], 1, 64)
        self._conv("d2a", a["u2"], a["s2"], a["a4"])
        self._conv("d2b", a["a4"], None, a["d2"])
        self._up(a["d2"], a["u1"], 0, 32)
        self._conv("d1a", a["u1"], a["s1"], a["a5"])
        self._conv("d1b", a["a5"], None, a["d1"])
        _lib.call("omg_head_forward", _p(a["d1"]), _p(self.p["head.w"]), _p(self.p["head.b"]), _p(a["base"]),
                  _p(a["out"]), self.V[0], 16, self._s())
        return a["out"]

    def zero_grad(self):
        for g in self.g.values():
            g.zero_()

    def backward(self, gout):
        """Accumulates parameter gradients for d(loss)/d(out) = gout (FP32 [V0])."""
        a = self.act
        V0 = self.V[0]
        g_d1 = a["g16_0"]
        _lib.call("omg_head_backward", _p(a["d1"]), _p(self.p["head.w"]), _p(gout), _p(g_d1), _p(self.g["head.w"]),
                  _p(self.g["head.b"]), V0, 16, self._s())
        self._relu_bwd(a["d1"], g_d1)
        self._wgrad("d1b", g_d1, a["a5"])
        g_a5 = a["g16_0b"]
        self._conv_dx("d1b", g_d1, g_a5)
        self._relu_bwd(a["a5"], g_a5)
        self._wgrad("d1a", g_a5, a["u1"], a["s1"])
        g_cat1 = a["g48_0"]
        self._conv_dx("d1a", g_a5, g_cat1)
        g_d2 = a["g32_1"]
        _lib.call("omg_upsample_backward", _p(g_cat1), 48, 0, _p(g_d2), self.D[1], self.H[1], self.W[1], 32, self._s())
        self._relu_bwd(a["d2"], g_d2)
        self._wgrad("d2b", g_d2, a["a4"])
        g_a4 = a["g32_1b"]
        self._conv_dx("d2b", g_d2, g_a4)
        self._relu_bwd(a["a4"], g_a4)
        self._wgrad("d2a", g_a4, a["u2"], a["s2"])
        g_cat2 = a["g96_1"]
        self._conv_dx("d2a", g_a4, g_cat2)
        g_b3 = a["g64_2"]
        _lib.call("omg_upsample_backward", _p(g_cat2), 96, 0, _p(g_b3), self.D[2], self.H[2], self.W[2], 64, self._s())
        self._relu_bwd(a["b3"], g_b3)
        self._wgrad("btb", g_b3, a["a3"])
        g_a3 = a["g64_2b"]
        self._conv_dx("btb", g_b3, g_a3)
        self._relu_bwd(a["a3"], g_a3)
        self._wgrad("bta", g_a3, a["p2"])
        g_p2 = a["g32_2"]
        self._conv_dx("bta", g_a3, g_p2)
        g_s2 = a["g32_1"]   # g_d2 is consumed
        _lib.call("omg_pool_backward", _p(g_p2), _p(g_s2), self.D[1], self.H[1], self.W[1], 32, 0, self._s())
        _lib.call("omg_take_channels", _p(g_cat2), 96, 64, _p(g_s2), self.V[1], 32, 1, self._s())
        self._relu_bwd(a["s2"], g_s2)
        self._wgrad("e2b", g_s2, a["a2"])
        g_a2 = a["g32_1b"]
        self._conv_dx("e2b", g_s2, g_a2)
        self._relu_bwd(a["a2"], g_a2)
        self._wgrad("e2a", g_a2, a["p1"])
        g_p1 = a["g16_1"]
        self._conv_dx("e2a", g_a2, g_p1)
        g_s1 = a["g16_0"]   # g_d1 is consumed
        _lib.call("omg_pool_backward", _p(g_p1), _p(g_s1), self.D[0], self.H[0], self.W[0], 16, 0, self._s())
        _lib.call("omg_take_channels", _p(g_cat1), 48, 32, _p(g_s1), V0, 16, 1, self._s())
        self._relu_bwd(a["s1"], g_s1)
        self._wgrad("e1b", g_s1, a["a1"])
        g_a1 = a["g16_0b"]
        self._conv_dx("e1b", g_s1, g_a1)
        self._relu_bwd(a["a1"], g_a1)
        self._wgrad("e1a", g_a1, a["x0"])

    def loss(self, target, mask, scale: float):
        """Masked MSE between out and target*scale over the image's voxels; fills act['gout']."""
        a = self.act
        t = self.torch
        tgt = t.zeros(self.V[0], dtype=t.float32, device=self.dev)
        tgt[: target.numel()] = target.reshape(-1) * scale
        a["loss"].zero_()
        n_valid = float(mask.sum())
        _lib.call("omg_mse", _p(a["out"]), _p(tgt), _p(mask), _p(a["gout"]), _p(a["loss"]), self.V[0], 1.0 / n_valid,
                  self._s())
        return a["loss"]

    def adam_step(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.step_count += 1
        for k, p in self.p.items():
            _lib.call("omg_adam", _p(p), _p(self.g[k]), _p(self.m[k]), _p(self.v[k]), None, p.numel(), lr, b1, b2,
                      eps, self.step_count, self._s())
        self._sync_bf16()

    def mask(self):
        t = self.torch
        m = t.zeros(self.V[0], dtype=t.float32, device=self.dev)
        m[: self.nx * self.ny * self.nz] = 1.0
        return m

    def predict(self, image, scale: float):
        out = self.forward(image, scale)
        n = self.nx * self.ny * self.nz
        return (out[:n] / scale).clamp_min(0.0)

    def state(self):
        return {k: v.detach().cpu().numpy() for k, v in self.p.items()}
