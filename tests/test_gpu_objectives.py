"""The objective kernels against their closed forms in float64 (numpy).

omg_poisson_loglik and omg_frames_loglik evaluate L = sum_i y_i log(yhat_i) - yhat_i (terms with y_i = 0
are -yhat_i); the learned warm start's yardstick (the plain run's log-likelihood) stands on them. Both read
float32 inputs and accumulate in double, so they agree with numpy's float64 sum of the same terms to
summation-order rounding (1e-12 relative here). Sizes: the 2-D plane PET sinogram (128 x 160 x 63 =
1,290,240 measurements) and a ragged length that is not a multiple of any block size.
"""
import numpy as np
import pytest

from omega_amd import _lib
from tests.hip_helpers import ptr

pytestmark = pytest.mark.gpu


def _ll64(y, yhat):
    y = y.astype(np.float64)
    h = yhat.astype(np.float64)
    return float(np.sum(np.where(y != 0, y * np.log(h), 0.0) - h))


def _data(n, seed):
    rng = np.random.default_rng(seed)
    y = rng.poisson(2.0, n).astype(np.float32)          # about 13% zeros
    yhat = (rng.random(n) * 5 + 0.05).astype(np.float32)
    return y, yhat


@pytest.mark.parametrize("n", [1_290_240, 100_003, 1, 0])
def test_poisson_loglik_closed_form(cuda, n):
    import torch
    y, yhat = _data(n, 1)
    out = torch.full((1,), 123.0, dtype=torch.float64, device=cuda)
    yt, ht = torch.from_numpy(y).to(cuda), torch.from_numpy(yhat).to(cuda)
    _lib.call("omg_poisson_loglik", ptr(yt), ptr(ht), n, ptr(out), None)
    ref = _ll64(y, yhat)
    got = float(out.item())
    print(f"[loglik] n={n} gpu {got!r} float64 {ref!r}", flush=True)
    assert abs(got - ref) <= 1e-12 * max(1.0, abs(ref))


def test_poisson_loglik_terms(cuda):
    """Single-term values: y log(yhat) - yhat, y = 0 gives -yhat, and the log is natural."""
    import torch
    y = np.array([0.0, 1.0, 3.0, 7.0], np.float32)
    h = np.array([2.5, np.e, 1.0, 0.5], np.float32)
    want = [-2.5, 1.0 - float(np.float32(np.e)), -1.0, 7.0 * np.log(0.5) - 0.5]
    for i in range(4):
        out = torch.zeros(1, dtype=torch.float64, device=cuda)
        yt, ht = torch.from_numpy(y[i:i + 1]).to(cuda), torch.from_numpy(h[i:i + 1]).to(cuda)
        _lib.call("omg_poisson_loglik", ptr(yt), ptr(ht), 1, ptr(out), None)
        assert abs(float(out.item()) - want[i]) < 1e-6


@pytest.mark.parametrize("n,frames", [(1_290_240, 3), (40_001, 24)])
def test_frames_loglik_closed_form(cuda, n, frames):
    import torch
    ys, hs = zip(*[_data(n, 10 + f) for f in range(frames)])
    Y = torch.from_numpy(np.stack(ys)).to(cuda)
    H = torch.from_numpy(np.stack(hs)).to(cuda)
    out = torch.zeros(frames, dtype=torch.float64, device=cuda)
    _lib.call("omg_frames_loglik", ptr(Y), ptr(H), n, frames, ptr(out), None)
    got = out.cpu().numpy()
    for f in range(frames):
        ref = _ll64(ys[f], hs[f])
        assert abs(got[f] - ref) <= 1e-12 * abs(ref), f"frame {f}: {got[f]!r} vs {ref!r}"
