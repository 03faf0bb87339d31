"""Learned warm start for OSEM-RDP: pairs made by the method itself, a 3-D U-Net trained on them with the
library's own kernels, and the evaluation that keeps the plain method as fallback and yardstick.

* Pairs: for each seeded ellipsoid phantom, Poisson data of ``counts`` total counts; the input is the
  image after ``n_short`` plain OSEM iterations, the target the image after ``n_long`` OSEM-RDP
  (OSL, RDP prior beta, gamma) iterations -- both reconstructed ``batch`` phantoms per launch with
  :class:`omega_amd.dynamic.FramesOSEM` (the phantoms ride as frames of one study).
* Training: masked MSE on images scaled by 1 / mean(input), Adam, one volume per step.
* Evaluation on held-out seeds: the plain run is ``n_long`` OSEM-RDP iterations from a uniform
  image; the warm-started run is ``n_short`` OSEM iterations, the network, then OSEM-RDP iterations
  until the OSEM-RDP objective L(x) - beta R(x) (Poisson log-likelihood minus the RDP value) reaches
  the plain run's final value within ``tol`` relative; the same continuation from the short-OSEM
  image without the network is run as the control. Both starting images -- the network's and the
  control's -- are first rescaled to the objective's maximum along their ray (``best_scale``, closed
  form), so the difference between them is the network's alone. Optionally (``search``) the warm start
  is the maximum of the objective on the segment from the input to the network's image
  (``segment_search``, two forward projections): it never scores below its input, but OSL continued
  from it has needed more iterations than from the network's image itself.
* Support (optional): every run reconstructs the same voxel support (``support_mask``: the voxels with
  at least a given fraction of the peak sensitivity); plain OSEM-RDP on that support is the yardstick.
  Wall time is compared on the GPU's own clock (HIP events); when the warm start does not reach the
  target within ``n_long`` iterations the plain result stands.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _lib, phantoms
from .dynamic import FramesOSEM
from .unet import UNet3D


def _p(t):
    return C.c_void_p(0 if t is None else t.data_ptr())


def _stream(torch):
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def batch_data(p, seeds, counts, device="cuda"):
    """Poisson data [len(seeds), M] (subset order) of the seeded ellipsoid phantoms, and the phantoms [F, N]."""
    import torch
    N = p.grid.N
    X = torch.empty((len(seeds), N), dtype=torch.float32, device=device)
    for j, s in enumerate(seeds):
        X[j] = torch.from_numpy(phantoms.ellipsoids(p.grid, seed=int(s), n_ell=8, background=0.1)).to(device)
    y = frames_forward(p, X)
    scale = counts / y.double().sum(1)
    g = torch.Generator(device=device)
    g.manual_seed(int(seeds[0]) * 7919 + 1)
    y = torch.poisson((y.double() * scale[:, None]).float(), generator=g)
    return y, X


def frames_forward(p, X):
    """A x_f for every row of X [F, N] over all subsets, in subset order [F, M]."""
    import torch
    F = X.shape[0]
    Xc = X.contiguous()
    M = int(p.nTotMeas[-1])
    out = torch.empty((F, M), dtype=torch.float32, device=X.device)
    for k in range(p.subsets):
        first, n, off = int(p.nMeas[k]), int(p.nMeasSubset[k, 0]), int(p.nTotMeas[k])
        _lib.call("omg_frames_forward", p.geometry.handle, first, n, F, _p(Xc), _p(out[0, off:]), M, _stream(torch))
    return out


def loglik(p, y, X):
    """Poisson log-likelihood of each row of X [F, N] for data y [F, M] (float64 numpy array)."""
    import torch
    yhat = frames_forward(p, X)
    out = torch.zeros(X.shape[0], dtype=torch.float64, device=X.device)
    _lib.call("omg_frames_loglik", _p(y), _p(yhat), y.shape[1], y.shape[0], _p(out), _stream(torch))
    return out.cpu().numpy()


def rdp_values(p, X, beta, gamma):
    """beta * R(x_f) for every row of X [F, N] (omg_rdp_value, zero boundary), float64 numpy array."""
    import torch
    out = torch.zeros(X.shape[0], dtype=torch.float64, device=X.device)
    Nx, Ny, Nz = int(p.Nx[0]), int(p.Ny[0]), int(p.Nz[0])
    for j in range(X.shape[0]):
        xj = X[j].contiguous()
        _lib.call("omg_rdp_value", _p(xj), Nx, Ny, Nz, float(gamma), float(beta), _p(out[j:j + 1]), _stream(torch))
    return out.cpu().numpy()


def objective(p, y, X, beta, gamma):
    """The OSEM-RDP objective L(x) - beta R(x) of every row of X [F, N] (float64 numpy array)."""
    return loglik(p, y, X) - rdp_values(p, X, beta, gamma)


def support_mask(p, frac, epps=1e-5):
    """Voxels whose sensitivity over all subsets, sum_k A_k^T 1, is at least ``frac`` of its maximum (bool
    device tensor [N]): the image region the data samples well. OSL-OSEM divides by S_k + beta grad R; where
    S_k is small (the axial end slices of a 3-D ring, the corners outside the transaxial coverage) a
    strong prior drives that denominator to zero or below, and such voxels spike or stick at the
    positivity floor."""
    tot = sensitivity_total(p, epps)
    return tot >= frac * tot.max()


def sensitivity_total(p, epps=1e-5):
    """s = sum_k A_k^T 1 over all subsets (float64 device tensor [N]; voxels no LOR reaches are 0), so that
    sum_i (A x)_i = s . x."""
    import torch
    N = int(p.N[0])
    S = torch.zeros(N, dtype=torch.float32, device="cuda")
    tot = torch.zeros(N, dtype=torch.float64, device="cuda")
    stream = _stream(torch)
    for k in range(p.subsets):
        first, n = int(p.nMeas[k]), int(p.nMeasSubset[k, 0])
        _lib.call("omg_frames_sensitivity", p.geometry.handle, first, n, _p(S), float(epps), stream)
        tot += torch.where(S > epps, S, torch.zeros_like(S)).double()
    return tot


def support_lors(p, support):
    """bool device tensor [M] (subset order): the LORs that cross the support."""
    hit = frames_forward(p, support.float()[None, :].contiguous())[0]
    return hit > 0


def best_scale(p, y, X, s, beta, gamma, support=None, lors=None):
    """Per row of X [F, N], the factor c maximising the OSEM-RDP objective along the ray that scales the
    image (its ``support`` voxels, when given, the others held at their epps floor). The projections are
    linear in c and the RDP value is homogeneous of degree 1 (R(c x) = c R(x)), so
    d/dc [L(c x) - beta R(c x)] = sum(y) / c - s . x - beta R(x) = 0 gives c = sum(y) / (s . x + beta R(x)).
    With a support, the counts of LORs that miss it (``lors`` false) are explained only by the floor voxels,
    which do not scale: the sums then run over the LORs that cross the support and its voxels."""
    import torch
    if support is None:
        sy = y.double().sum(1)
        sx = X.double() @ s
    else:
        sy = (y.double() * lors.double()[None, :]).sum(1)
        sx = X.double() @ (s * support.double())
    den = sx + torch.from_numpy(rdp_values(p, X, beta, gamma)).to(X.device)
    return (sy / den).float()


def segment_search(p, y, A, B, beta, gamma, iters=16):
    """Per row f, t in [0, 1] maximising the OSEM-RDP objective at (1 - t) A_f + t B_f, and that image.

    The projections of the segment are the same mix of the endpoints' projections (two forward
    projections in all), the Poisson log-likelihood of an affine projection is concave and the RDP
    value convex, so the objective is concave along the segment and a golden-section search on t
    converges to its maximum. t = 0 is kept unless the search finds a higher objective, so the result
    never has a lower objective than A."""
    import torch
    hA, hB = frames_forward(p, A), frames_forward(p, B)
    F = A.shape[0]
    dev = A.device

    def obj(t):
        yh = torch.lerp(hA, hB, t[:, None])
        out = torch.zeros(F, dtype=torch.float64, device=dev)
        _lib.call("omg_frames_loglik", _p(y), _p(yh), y.shape[1], F, _p(out), _stream(torch))
        del yh
        return out.cpu().numpy() - rdp_values(p, torch.lerp(A, B, t[:, None]).contiguous(), beta, gamma)

    g = (np.sqrt(5.0) - 1.0) / 2.0
    lo, hi = np.zeros(F), np.ones(F)
    t1, t2 = hi - g * (hi - lo), lo + g * (hi - lo)
    f1 = obj(torch.tensor(t1, dtype=torch.float32, device=dev))
    f2 = obj(torch.tensor(t2, dtype=torch.float32, device=dev))
    for _ in range(iters):
        left = f1 >= f2                       # the maximum lies in [lo, t2]
        hi = np.where(left, t2, hi)
        lo = np.where(left, lo, t1)
        t2n = np.where(left, t1, lo + g * (hi - lo))
        t1n = np.where(left, hi - g * (hi - lo), t2)
        fk = obj(torch.tensor(np.where(left, t1n, t2n), dtype=torch.float32, device=dev))
        f2, f1 = np.where(left, f1, fk), np.where(left, fk, f2)
        t1, t2 = t1n, t2n
    t = np.where(f1 >= f2, t1, t2)
    best = np.maximum(f1, f2)
    f0 = obj(torch.zeros(F, dtype=torch.float32, device=dev))
    t = np.where(best > f0, t, 0.0)
    tt = torch.tensor(t, dtype=torch.float32, device=dev)
    return torch.lerp(A, B, tt[:, None]).contiguous(), t


def _rescale(x, c, support):
    """x * c on the support (the floor voxels outside it stay at epps), in place."""
    if support is None:
        x *= c[:, None]
    else:
        x *= torch_where_support(support, c)
    return x


def torch_where_support(support, c):
    """[F, N] factors: c on the support, 1 elsewhere."""
    import torch
    return torch.where(support[None, :], c[:, None], torch.ones_like(c)[:, None])


def make_pairs(p, seeds, counts=5e6, n_short=2, n_long=40, beta=0.1, gamma=2.0, batch=8, rpb=256, ft=8, log=None,
               support=None):
    """(inputs [n, N], targets [n, N]) device tensors for the seeds, `batch` phantoms per launch; both runs
    reconstruct only ``support`` (see FramesOSEM) when it is given."""
    import torch
    ins, outs = [], []
    for b0 in range(0, len(seeds), batch):
        if log and (b0 // batch) % 8 == 0:
            log(f"pairs: batch {b0 // batch} of {(len(seeds) + batch - 1) // batch}")
        sb = seeds[b0:b0 + batch]
        y, _ = batch_data(p, sb, counts)
        short = FramesOSEM(p, y, support=support)
        short.run(n_short, rpb, min(ft, _tile(len(sb))))
        long = FramesOSEM(p, y, beta=beta, gamma=gamma, support=support)
        long.run(n_long, rpb, min(ft, _tile(len(sb))))
        # OSL-OSEM divides by S + beta grad R; a run that left the finite range is not a usable target
        ok = torch.isfinite(short.x).all(dim=1) & torch.isfinite(long.x).all(dim=1)
        ins.append(short.x[ok].clone())
        outs.append(long.x[ok].clone())
        short.close()
        long.close()
        del y
    return torch.cat(ins), torch.cat(outs)


def _tile(F):
    for t in (8, 4, 2, 1):
        if F % t == 0:
            return t
    return 1


def train(net: UNet3D, inputs, targets, epochs=4, lr=1e-3, seed=0, log=None, support=None):
    """Adam on masked MSE of scaled images (over ``support`` when given); returns the per-epoch mean loss."""
    rng = np.random.default_rng(seed)
    mask = net.mask()
    if support is not None:
        mask[: support.numel()] *= support.float()
    hist = []
    for ep in range(epochs):
        tot = 0.0
        for i in rng.permutation(inputs.shape[0]):
            x = inputs[i]
            scale = float(1.0 / x.mean())
            net.forward(x, scale)
            net.zero_grad()
            tot += float(net.loss(targets[i], mask, scale))
            net.backward(net.act["gout"])
            net.adam_step(lr=lr)
        hist.append(tot / inputs.shape[0])
        if log:
            log(f"epoch {ep}: loss {hist[-1]:.4e}")
    return hist


def evaluate(p, net: UNet3D, seeds, counts=5e6, n_short=2, n_long=40, beta=0.1, gamma=2.0, tol=1e-4, rpb=256,
             ft=8, log=None, support=None, search=False):
    """Plain vs warm-started OSEM-RDP on held-out seeds (batched). Returns a dict of per-seed results. With
    ``support``, every run (plain, short, continuations) reconstructs the same support and the network's
    image is set to epps outside it. ``search``: start from the best point of the segment between the input
    and the network's image instead of the network's image (``segment_search``)."""
    import torch
    res = {"seed": [], "ll_plain": [], "ll_warm_start": [], "ll_input": [], "ll_network": [], "t_segment": [],
           "ll_first_warm": [], "ll_first_no_net": [],
           "iters_warm": [], "iters_no_net": [], "reached": []}
    t_plain = t_short = t_net = t_search = t_cont = 0.0
    s_tot = sensitivity_total(p)
    lors = support_lors(p, support) if support is not None else None
    for b0 in range(0, len(seeds), 8):
        sb = list(seeds[b0:b0 + 8])
        F = len(sb)
        tile = min(ft, _tile(F))
        y, _ = batch_data(p, sb, counts)
        # plain: n_long OSEM-RDP iterations from a uniform image
        plain = FramesOSEM(p, y, beta=beta, gamma=gamma, support=support)
        plain.sensitivity()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record()
        plain.run(n_long, rpb, tile)
        ev1.record()
        ev1.synchronize()
        t_plain += ev0.elapsed_time(ev1)
        ll_plain = objective(p, y, plain.x, beta, gamma)
        # warm start: n_short OSEM iterations, the network, then OSEM-RDP until the plain log-likelihood
        short = FramesOSEM(p, y, support=support)
        short.sensitivity()
        ev0.record()
        short.run(n_short, rpb, tile)
        ev1.record()
        ev1.synchronize()
        t_short += ev0.elapsed_time(ev1)
        x_in = _rescale(short.x.clone(), best_scale(p, y, short.x, s_tot, beta, gamma, support, lors), support)
        ll_in = objective(p, y, x_in, beta, gamma)
        ws = torch.empty_like(short.x)
        ev0.record()
        for j in range(F):
            x = short.x[j]
            ws[j] = net.predict(x, float(1.0 / x.mean())).clamp_min(1e-5)
            if support is not None:
                ws[j].masked_fill_(~support, 1e-5)
        # the network fixes the image's shape; its scale is set exactly: the objective's maximum along c x
        _rescale(ws, best_scale(p, y, ws, s_tot, beta, gamma, support, lors), support)
        ev1.record()
        ev1.synchronize()
        t_net += ev0.elapsed_time(ev1)
        ll_net = objective(p, y, ws, beta, gamma)
        t_seg = np.ones(F)
        if search:
            # exact line search from the input towards the network's image (two forward projections)
            ev0.record()
            ws, t_seg = segment_search(p, y, x_in, ws, beta, gamma)
            ev1.record()
            ev1.synchronize()
            t_search += ev0.elapsed_time(ev1)
        ll_ws = objective(p, y, ws, beta, gamma)
        cont = FramesOSEM(p, y, beta=beta, gamma=gamma, x0=ws, support=support)
        cont.sensitivity()
        target = ll_plain - tol * np.abs(ll_plain)
        iters = np.full(F, -1)
        for it in range(1, n_long + 1):
            ev0.record()
            cont.run(1, rpb, tile)
            ev1.record()
            ev1.synchronize()
            t_cont += ev0.elapsed_time(ev1)
            ll = objective(p, y, cont.x, beta, gamma)
            if it == 1:
                ll_first_warm = ll
            newly = (iters < 0) & (ll >= target)
            iters[newly] = it
            if np.all(iters > 0):
                break
        # the same continuation without the network (from the short-OSEM image): what the network adds
        nonet = FramesOSEM(p, y, beta=beta, gamma=gamma, x0=x_in, support=support)
        nonet.sensitivity()
        iters_nn = np.full(F, -1)
        for it in range(1, n_long + 1):
            nonet.run(1, rpb, tile)
            ll = objective(p, y, nonet.x, beta, gamma)
            if it == 1:
                ll_first_nonet = ll
            iters_nn[(iters_nn < 0) & (ll >= target)] = it
            if np.all(iters_nn > 0):
                break
        nonet.close()
        for j in range(F):
            res["seed"].append(int(sb[j]))
            res["ll_plain"].append(float(ll_plain[j]))
            res["ll_warm_start"].append(float(ll_ws[j]))
            res["ll_network"].append(float(ll_net[j]))
            res["t_segment"].append(float(t_seg[j]))
            res["ll_first_warm"].append(float(ll_first_warm[j]))
            res["ll_first_no_net"].append(float(ll_first_nonet[j]))
            res["ll_input"].append(float(ll_in[j]))
            res["iters_warm"].append(int(iters[j]))
            res["iters_no_net"].append(int(iters_nn[j]))
            res["reached"].append(bool(iters[j] > 0))
        for r in (plain, short, cont):
            r.close()
        if log:
            log(f"held-out batch {b0 // 8}: warm-start iterations {iters.tolist()}")
    # Per image, on the GPU clock: the batched runs advance all images of a batch together, so per-image
    # costs are the batch costs divided by the batch size. A warm start that does not reach the target
    # within n_long iterations falls back to the plain run and is charged its full cost on top.
    n_img = len(res["seed"])
    n_batches = (len(seeds) + 7) // 8
    it_rdp = t_plain / (n_long * n_img)            # one OSEM-RDP iteration of one image
    it_osem = t_short / (n_short * n_img) if n_short else 0.0
    net_img = (t_net + t_search) / n_img
    plain_img = n_long * it_rdp
    warm_img = []
    for k in res["iters_warm"]:
        base = n_short * it_osem + net_img
        warm_img.append(base + (k * it_rdp if k > 0 else n_long * it_rdp + plain_img))
    res["ms_plain_per_image"] = plain_img
    res["ms_warm_per_image"] = float(np.mean(warm_img))
    res["ms_iteration_rdp"] = it_rdp
    res["ms_iteration_osem"] = it_osem
    res["ms_network_per_image"] = t_net / n_img
    res["ms_search_per_image"] = t_search / n_img
    res["batches"] = n_batches
    return res


def main(argv=None):
    """Pair generation, training and held-out evaluation on the 3-D PET geometry; prints one JSON line."""
    import argparse
    import json

    import torch
    from . import workloads

    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--pairs", type=int, default=2000)
    ap.add_argument("--holdout", type=int, default=200)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--counts", type=float, default=5e6)
    ap.add_argument("--subsets", type=int, default=16)
    ap.add_argument("--n-long", type=int, default=40)
    ap.add_argument("--beta", type=float, default=0.1)
    ap.add_argument("--n-short", type=int, default=2)
    ap.add_argument("--small", action="store_true", help="reduced scanner (8 rings, 96 crystals, 40x40x15 image)")
    ap.add_argument("--subset-type", type=int, default=4,
                    help="the reference's subset type (4: every n-th angle of every plane; 8: every n-th plane)")
    ap.add_argument("--support", type=float, default=0.0,
                    help="reconstruct only voxels with at least this fraction of the peak sensitivity (0: all)")
    a = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)
    if a.small:
        p = workloads.pet3d_projector(subsets=a.subsets, n_rings=8, crystals=96, max_ring_diff=7, radius=150.0,
                                      image=(40, 40, 15), fov_xy=160.0, subset_type=a.subset_type)
    else:
        p = workloads.pet3d_projector(subsets=a.subsets, subset_type=a.subset_type)
    p.initProj()
    support = support_mask(p, a.support) if a.support > 0 else None
    if support is not None:
        log(f"support: {int(support.sum())} of {support.numel()} voxels (sensitivity >= {a.support} of the peak)")
    t0 = time.time()
    train_seeds = list(range(1000, 1000 + a.pairs))
    hold_seeds = list(range(900000, 900000 + a.holdout))
    xin, xt = make_pairs(p, train_seeds, counts=a.counts, n_short=a.n_short, n_long=a.n_long, beta=a.beta, log=log,
                         support=support)
    torch.cuda.synchronize()
    log(f"pairs: {xin.shape[0]} usable of {a.pairs} in {time.time() - t0:.1f} s")
    ratio = (xt.amax(dim=1) / xt.mean(dim=1)).cpu().numpy()
    log(f"targets: max/mean median {np.median(ratio):.1f}, max {ratio.max():.1f}; inputs max/mean "
        f"{float((xin.amax(dim=1) / xin.mean(dim=1)).median()):.1f}")
    net = UNet3D(int(p.Nx[0]), int(p.Ny[0]), int(p.Nz[0]), seed=0)
    t1 = time.time()
    hist = train(net, xin, xt, epochs=a.epochs, lr=a.lr, log=log, support=support)
    torch.cuda.synchronize()
    log(f"training: {time.time() - t1:.1f} s")
    res = evaluate(p, net, hold_seeds, counts=a.counts, n_short=a.n_short, n_long=a.n_long, beta=a.beta, log=log,
                   support=support)
    reached = int(np.sum(res["reached"]))
    out = {"pairs": a.pairs, "holdout": a.holdout, "epochs": a.epochs, "loss": hist, "beta": a.beta,
           "subset_type": a.subset_type, "support": a.support,
           "support_voxels": None if support is None else int(support.sum()),
           "reached": reached, "mean_iters_warm": float(np.mean([i for i in res["iters_warm"] if i > 0] or [0])),
           "mean_iters_no_net": float(np.mean([i for i in res["iters_no_net"] if i > 0] or [0])),
           "ms_plain_per_image": res["ms_plain_per_image"], "ms_warm_per_image": res["ms_warm_per_image"],
           "ms_network_per_image": res["ms_network_per_image"], "ms_iteration_rdp": res["ms_iteration_rdp"],
           "speedup": res["ms_plain_per_image"] / max(res["ms_warm_per_image"], 1e-9),
           "ms_search_per_image": res["ms_search_per_image"],
           "ll_gain_input_to_warm_start": float(np.mean(np.array(res["ll_warm_start"]) - np.array(res["ll_input"]))),
           "ll_gain_first_iteration_vs_no_net": float(np.mean(np.array(res["ll_first_warm"]) -
                                                              np.array(res["ll_first_no_net"])))}
    print(json.dumps(out), flush=True)
    return out


if __name__ == "__main__":
    main()
