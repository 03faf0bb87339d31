"""Learned warm start, end to end on a reduced 3-D PET scanner: pairs made by the method itself,
U-Net training on the library's kernels, held-out evaluation against the plain OSEM-RDP run.

Bar: the training loss falls; on held-out seeds, OSEM-RDP continued from the network's image reaches
the plain run's OSEM-RDP objective in at most as many iterations on average as the same continuation
from the short-OSEM image without the network (printed: what the network adds at this tiny scale);
the batched log-likelihood matches the single-frame forward projection + omg_poisson_loglik.
"""
import numpy as np
import pytest

from omega_amd import _lib, warmstart as WS, workloads
from omega_amd.unet import UNet3D
from tests.hip_helpers import ptr

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scanner(cuda):
    p = workloads.pet3d_projector(subsets=4, n_rings=8, crystals=96, max_ring_diff=7, radius=150.0,
                                  image=(40, 40, 15), fov_xy=160.0)
    p.initProj()
    return p


def test_batched_loglik_matches_single(scanner, cuda):
    import torch
    p = scanner
    y, X = WS.batch_data(p, [11, 12, 13, 14], counts=2e5)
    ll = WS.loglik(p, y, X)
    for j in range(4):
        x = X[j].contiguous()
        yhat = torch.cat([p.forwardProject(x, k) for k in range(p.subsets)])
        out = torch.zeros(1, dtype=torch.float64, device=cuda)
        yj = y[j].contiguous()
        _lib.call("omg_poisson_loglik", ptr(yj), ptr(yhat), yj.numel(), ptr(out), None)
        assert abs(float(out) - ll[j]) <= 1e-6 * abs(ll[j])


def test_pairs_training_and_gain(scanner, cuda):
    p = scanner
    seeds = list(range(100, 164))
    # beta scaled to this small scanner's sensitivity (OSL needs S + beta grad R > 0)
    xin, xt = WS.make_pairs(p, seeds, counts=5e5, n_short=2, n_long=10, beta=0.01)
    assert xin.shape[0] >= 60 and xin.shape == xt.shape and xin.shape[1] == p.grid.N
    net = UNet3D(40, 40, 15, seed=1, device=cuda)
    hist = WS.train(net, xin, xt, epochs=10, lr=1e-3, log=print)
    assert hist[-1] < hist[0]
    res = WS.evaluate(p, net, list(range(5000, 5008)), counts=5e5, n_short=2, n_long=10, beta=0.01, log=print)
    gain = np.array(res["ll_warm_start"]) - np.array(res["ll_input"])
    it_w = np.array(res["iters_warm"], float)
    it_n = np.array(res["iters_no_net"], float)
    it_w[it_w < 0] = 11
    it_n[it_n < 0] = 11
    print(f"  objective of the network image minus that of its input: {gain}", flush=True)
    print(f"  iterations to the plain 10-iteration objective: with the network {res['iters_warm']}, "
          f"without {res['iters_no_net']}", flush=True)
    print(f"  ms per image: plain {res['ms_plain_per_image']:.2f}, warm start {res['ms_warm_per_image']:.2f}",
          flush=True)
    assert it_w.mean() <= it_n.mean()
