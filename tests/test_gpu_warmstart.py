"""Learned warm start, end to end on a reduced 3-D PET scanner: pairs made by the method itself,
U-Net training on the library's kernels, held-out evaluation against the plain OSEM-RDP run.

Settings of the full-size workload: beta 0.1, angle subsets (type 4), and the support of voxels with at
least 30% of the peak sensitivity, used by every run (plain, short, continuations), so plain OSEM-RDP
on that support stays the yardstick. Bar: the training loss falls; on held-out seeds, OSEM-RDP continued from the network's image
ends its first iteration at a strictly higher OSEM-RDP objective than the same continuation from the
network's input (the 2-iteration OSEM image) on average, and reaches the plain run's objective in
strictly fewer iterations on average (a network that degrades its input fails both); the objective of
the network's image itself against its input's is printed (it is not asserted: the EM image fits the
data more closely and scores higher on some seeds, although the network's image is the better start); the batched log-likelihood matches the
single-frame forward projection + omg_poisson_loglik and a float64 sum on the host.
"""
import numpy as np
import pytest

from omega_amd import _lib, warmstart as WS, workloads
from omega_amd.unet import UNet3D
from tests.hip_helpers import ptr

pytestmark = pytest.mark.gpu


BETA = 0.1
SUPPORT = 0.3


@pytest.fixture(scope="module")
def scanner(cuda):
    p = workloads.pet3d_projector(subsets=4, n_rings=8, crystals=96, max_ring_diff=7, radius=150.0,
                                  image=(40, 40, 15), fov_xy=160.0, subset_type=4)
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
        yd, hd = yj.double().cpu().numpy(), yhat.double().cpu().numpy()
        host = float(np.sum(np.where(yd != 0, yd * np.log(np.where(yd != 0, hd, 1.0)), 0.0) - hd))
        assert abs(host - ll[j]) <= 1e-6 * abs(host)


def test_pairs_training_and_gain(scanner, cuda):
    import torch
    p = scanner
    support = WS.support_mask(p, SUPPORT)
    assert 0.5 * p.grid.N < int(support.sum()) < p.grid.N
    seeds = list(range(100, 228))
    xin, xt = WS.make_pairs(p, seeds, counts=5e5, n_short=2, n_long=20, beta=BETA, support=support)
    assert xin.shape[0] == len(seeds) and xin.shape == xt.shape and xin.shape[1] == p.grid.N
    assert bool(torch.isfinite(xt).all())
    net = UNet3D(40, 40, 15, seed=1, device=cuda)
    hist = WS.train(net, xin, xt, epochs=10, lr=1e-3, log=print, support=support)
    assert hist[-1] < hist[0]
    res = WS.evaluate(p, net, list(range(5000, 5016)), counts=5e5, n_short=2, n_long=20, beta=BETA, log=print,
                      support=support)
    gain = np.array(res["ll_warm_start"]) - np.array(res["ll_input"])
    first = np.array(res["ll_first_warm"]) - np.array(res["ll_first_no_net"])
    it_w = np.array(res["iters_warm"], float)
    it_n = np.array(res["iters_no_net"], float)
    it_w[it_w < 0] = 21
    it_n[it_n < 0] = 21
    print(f"  objective of the network image minus that of its input: {np.round(gain, 1)}", flush=True)
    print(f"  objective after one continuation iteration, from the network image minus from the input: "
          f"{np.round(first, 1)}", flush=True)
    print(f"  iterations to the plain 20-iteration objective: with the network {res['iters_warm']}, "
          f"without {res['iters_no_net']}", flush=True)
    print(f"  ms per image: plain {res['ms_plain_per_image']:.2f}, warm start {res['ms_warm_per_image']:.2f}",
          flush=True)
    assert first.mean() > 0, "one iteration from the network's image ends lower than one from its input"
    assert it_w.mean() < it_n.mean(), "the network does not shorten the continuation"
