"""CPU restatement of the reference's built-in reconstruction loop for the algorithms omegaMain provides.

Test infrastructure only (imported by tests/): the checker of libomega_matrixfree*.so. It follows
reconstructionAF (source/cpp/reconstructionAF.h:1053-1508) step by step, with one oracle projector case
per subset for the forward and backward projections and the elementwise steps in float32 numpy written
in the reference's operation order:

  forward step        computeForwardStep, source/cpp/computeForwardProject.h:5-409
  image update        computeOSEstimates, source/cpp/subiterStep.h:5-377; algorithms.h:12-134,149-215,
                      243-269,306-400; PoissonUpdate / PDHGUpdate, source/opencl/auxKernels.cl:1997-2059
  end of iteration    computeOSEstimatesIter, source/cpp/iterStep.h:7-64
  initialisation      CGLS / PDHG, source/cpp/functions.hpp:2315-2385; D, reconstructionAF.h:687-822
  RDP gradient        RDPKernel, source/opencl/auxKernels.cl:664-873 (buffer: zero outside; image:
                      clamp to edge)

Pinning (tests/test_oracle_reconstruction.py): OSEM / OSL-OSEM equal oracle.osem on the same inputs;
CGLS equals oracle.cgls from x0 = 0 and reaches numpy's least-squares solution; non-subset PDHG reaches
scipy's bound-constrained least-squares solution; BSREM and PKMA with RDP reach a point where the KKT
conditions of the penalised Poisson likelihood hold (BSREM approaches that point); MBSREM without a prior reaches the bounded Poisson
maximum likelihood; the zero-boundary RDP gradient equals the C oracle's bit for bit and the clamped one a
float64 evaluation.
"""
from __future__ import annotations

import numpy as np

f32 = np.float32


def _rdp_term(c, k, gamma, epps):
    du = (c - k).astype(f32)
    ad = np.abs(du)
    den = ((c + k).astype(f32) + (f32(gamma) * ad).astype(f32)).astype(f32)
    num = ((((f32(gamma) * ad).astype(f32) + c).astype(f32) + (f32(3.0) * k).astype(f32)).astype(f32)
           + f32(epps) * f32(epps)).astype(f32)
    with np.errstate(divide="ignore", invalid="ignore"):
        return ((du * num).astype(f32) / ((den * den).astype(f32) + f32(epps))).astype(f32)


def rdp_gradient(u, shape, gamma, epps, beta, clamp):
    """beta * dR/du of the relative difference prior, 6-neighbourhood, float32 in the kernels' order:
    lane 0 sums the +x, +y, +z pairs, lane 1 the -x, -y, -z pairs, NaN lanes -> 0, beta * (lane0 + lane1).
    clamp: a missing neighbour is the voxel itself (image read), else zero (buffer read)."""
    Nx, Ny, Nz = shape
    v = np.asarray(u, f32).reshape(Nz, Ny, Nx)
    pad = np.pad(v, 1, mode="edge" if clamp else "constant")
    c = pad[1:-1, 1:-1, 1:-1]
    xp, xm = pad[1:-1, 1:-1, 2:], pad[1:-1, 1:-1, :-2]
    yp, ym = pad[1:-1, 2:, 1:-1], pad[1:-1, :-2, 1:-1]
    zp, zm = pad[2:, 1:-1, 1:-1], pad[:-2, 1:-1, 1:-1]
    l0 = (_rdp_term(c, xp, gamma, epps) + _rdp_term(c, yp, gamma, epps)).astype(f32)
    l0 = (l0 + _rdp_term(c, zp, gamma, epps)).astype(f32)
    l1 = (_rdp_term(c, xm, gamma, epps) + _rdp_term(c, ym, gamma, epps)).astype(f32)
    l1 = (l1 + _rdp_term(c, zm, gamma, epps)).astype(f32)
    l0[np.isnan(l0)] = 0
    l1[np.isnan(l1)] = 0
    return (f32(beta) * (l0 + l1).astype(f32)).astype(f32).reshape(-1)


def _finite(r, epps):
    r = r.astype(f32)
    r[~np.isfinite(r)] = f32(epps)
    return r


class Params:
    """The parameter block's values the loop reads (names as in the block). Unset algorithm switches are
    False; shape = (Nx, Ny, Nz)."""
    OSEM = OSL_OSEM = BSREM = MBSREM = PKMA = PDHG = CGLS = False
    CT = False
    RDP = False
    useImages = True
    beta = 0.0
    RDP_gamma = 2.0
    epps = 1e-5
    flat = 0.0
    U = 1e6
    Niter = 1
    enforcePositivity = False
    save_iter = False
    saveNiter = ()
    storeFP = False
    storeResidual = False
    lambdaN = None
    alpha_PKMA = None
    tauCP = 0.0
    sigmaCP = 1.0
    thetaCP = None
    precondTypeImage = (False,) * 7
    regEveryIter = 1
    shape = None

    def __init__(self, **kw):
        for k, v in kw.items():
            setattr(self, k, v)


def omega_main(p, sino_parts, cases, x0, add_parts=None, n_lor=None):
    """Runs the configured algorithm; returns dict(output=[saved images], fp=[forward projections per
    sub-iteration], residual=np.ndarray). sino_parts[k]: float32 measurements of subset k (TOF: bins
    blocks); cases[k]: object with fp(img) -> measurements and bp(y, sens=False) -> image (or
    (image, sens)); add_parts[k]: randoms, one per LOR of subset k, or None; n_lor[k]: LORs of subset k
    (defaults to the measurement count)."""
    S_n = len(cases)
    N = int(np.prod(p.shape))
    epps = f32(p.epps)
    x = np.asarray(x0, f32).copy()
    out, fps = [], []
    resid = np.zeros(max(1, p.Niter * S_n), f32)
    ys = [np.asarray(s, f32) for s in sino_parts]
    adds = add_parts or [None] * S_n
    n_lor = n_lor or [y.size for y in ys]
    prec = list(p.precondTypeImage)
    # D = (sum_k A_k^T 1) / subsets, D == 0 -> 1
    D = None
    if prec[0] or prec[1]:
        D = np.zeros(N, f32)
        for k in range(S_n):
            D = (D + cases[k].bp(np.ones(ys[k].size, f32))).astype(f32)
        if S_n > 1:
            D = (D / f32(S_n)).astype(f32)
        D[D == 0] = f32(1)
    S = [None] * S_n
    dU = np.zeros(N, f32)
    pcp = [np.zeros(y.size, f32) for y in ys]
    ucp = x.copy() if p.PDHG else None
    r_cg = f_cg = None
    gamma_cg = f32(0)
    ee = 0
    counter_max = S_n * p.Niter - 1

    def _shift(v, y, k):
        # every emission forward step starts with input += epps, then input += randoms
        # (computeForwardProject.h:20-21,66-69)
        if p.CT:
            return v
        v = (v + epps).astype(f32)
        if adds[k] is not None:
            v = (v + np.tile(np.asarray(adds[k], f32), y.size // n_lor[k])).astype(f32)
        return v

    def reg(counter, cmax):
        return p.regEveryIter <= 1 or counter == 0 or counter == cmax or counter % p.regEveryIter == 0

    def prior(grad, fresh):
        g = np.zeros(N, f32) if fresh else grad
        if p.RDP:
            g = (g + rdp_gradient(x, p.shape, p.RDP_gamma, p.epps, p.beta, p.useImages)).astype(f32)
        return g

    def precond(rhs, im, mirror):
        if D is None:
            return rhs
        if prec[0]:
            return (rhs / D).astype(f32)
        u = im.copy()
        if mirror:
            m = u >= f32(p.U) / f32(2)
            u[m] = (f32(p.U) - u[m]).astype(f32)
        return (rhs * (u / D).astype(f32)).astype(f32)

    for it in range(p.Niter):
        for k in range(S_n):
            kk = it * S_n + k
            y = ys[k]
            if p.CGLS and it == 0:
                r_cg = y.copy()
                f_cg = x.copy()
                x = cases[k].bp(y).astype(f32)
                gamma_cg = f32(np.dot(x.astype(np.float64), x.astype(np.float64)))
            v = cases[k].fp(x).astype(f32)
            if p.storeFP:
                fps.append(v.copy())
            sens_now = False
            if p.OSEM or p.OSL_OSEM or p.BSREM or p.MBSREM or p.PKMA:
                kind = 0 if (p.OSEM or p.OSL_OSEM) else (2 if p.PKMA else 1)
                if p.CT:
                    if (p.OSEM or p.OSL_OSEM) and it == 0:
                        s = cases[k].bp(y).astype(f32)
                        s[s < epps] = epps
                        S[k] = s
                    a = np.where(v < 0, f32(0), v).astype(f32)
                    e = np.exp(-a.astype(np.float64)).astype(f32)
                    fl = f32(p.flat)
                    r = e if kind == 0 else ((e * fl).astype(f32) - y if kind == 1 else y - (e * fl).astype(f32))
                else:
                    if (p.OSEM or p.OSL_OSEM) and it == 0:
                        sens_now = True
                    a = (v + epps).astype(f32)
                    if adds[k] is not None:
                        a = (a + np.tile(np.asarray(adds[k], f32), y.size // n_lor[k])).astype(f32)
                    with np.errstate(divide="ignore", invalid="ignore"):
                        q = (y / a).astype(f32)
                    r = q if kind == 0 else ((q - f32(1)) if kind == 1 else (f32(1) - q))
                bp_in = _finite(r, epps)
            elif p.PDHG:
                v = _shift(v, y, k)
                if p.storeResidual:
                    res = (v - y).astype(f32)
                    nr = f32(np.sqrt(np.dot(res.astype(np.float64), res.astype(np.float64))))
                    val = f32(nr * nr * f32(0.5))
                    np_ = f32(np.sqrt(np.dot(pcp[k].astype(np.float64), pcp[k].astype(np.float64))))
                    py = f32(np.dot(pcp[k].astype(np.float64), y.astype(np.float64)))
                    resid[kk] = f32(float(val) + (float(f32(np_ * np_)) * 0.5 + float(py)))
                sig = f32(p.sigmaCP)
                res = (v - y).astype(f32)
                pn = ((pcp[k] + (sig * res).astype(f32)).astype(f32) / (f32(1) + sig)).astype(f32)
                o = (pn - pcp[k]).astype(f32) if S_n > 1 else pn
                pcp[k] = pn
                bp_in = _finite(o, epps)
            else:   # CGLS
                v = _shift(v, y, k)
                nq = f32(np.sqrt(np.dot(v.astype(np.float64), v.astype(np.float64))))
                alpha_cg = f32(gamma_cg / f32(nq * nq))
                r_cg = (r_cg - (alpha_cg * v).astype(f32)).astype(f32)
                bp_in = _finite(r_cg.copy(), epps)
            if sens_now:
                rhs, s = cases[k].bp(bp_in, sens=True)
                s = np.asarray(s, f32).copy()
                s[s < epps] = epps
                S[k] = s
            else:
                rhs = cases[k].bp(bp_in)
            rhs = np.asarray(rhs, f32)
            pos_now = p.enforcePositivity and S_n > 1
            do_reg = reg(kk, counter_max)
            if p.OSEM or p.OSL_OSEM:
                if p.OSL_OSEM and do_reg:
                    dU = prior(dU, True)
                if p.CT:
                    num = (f32(p.flat) * rhs).astype(f32)
                    den = S[k] if p.OSEM else (S[k] + dU).astype(f32)
                    x = ((x / den).astype(f32) * num).astype(f32)
                elif p.OSEM:
                    x = ((x / S[k]).astype(f32) * (rhs + epps).astype(f32)).astype(f32)
                else:
                    x = ((x / (S[k] + dU).astype(f32)).astype(f32) * rhs).astype(f32)
                if pos_now:
                    x[x < epps] = epps
            elif p.BSREM:
                lam = f32(p.lambdaN[it])
                v2 = (x + ((lam * rhs).astype(f32) * x).astype(f32)).astype(f32)
                if p.enforcePositivity:
                    v2 = np.maximum(epps, v2)
                x = v2
                if pos_now:
                    x[x < epps] = epps
            elif p.MBSREM or p.PKMA:
                if do_reg:
                    rhs = prior(rhs, False)
                rhs = precond(rhs, x, p.MBSREM)
                lam = f32(p.lambdaN[it])
                if p.MBSREM:
                    v2 = (x + (lam * rhs).astype(f32)).astype(f32)
                    if p.enforcePositivity:
                        v2 = np.maximum(epps, v2)
                    U = f32(p.U)
                    v2 = np.where(v2 >= U, U - epps, v2).astype(f32)
                else:
                    al = f32(p.alpha_PKMA[kk])
                    v2 = (x - (lam * rhs).astype(f32)).astype(f32)
                    if p.enforcePositivity:
                        v2 = np.maximum(epps, v2)
                    v2 = (((f32(1) - al) * x).astype(f32) + (al * v2).astype(f32)).astype(f32)
                x = v2
                if pos_now:
                    x[x < epps] = epps
            elif p.PDHG:
                tau = f32(p.tauCP)
                if S_n > 1:
                    ucp = (f32(1) * rhs + f32(1) * ucp).astype(f32)
                    rhs = (f32(1) * ucp + (f32(S_n) * f32(p.thetaCP[kk])) * rhs).astype(f32)
                rhs = precond(rhs, x, False)
                if S_n > 1:
                    v2 = (x - (tau * rhs).astype(f32)).astype(f32)
                    if p.enforcePositivity:
                        v2 = np.maximum(epps, v2)
                    x = v2
                else:
                    un = (ucp - (tau * rhs).astype(f32)).astype(f32)
                    if p.enforcePositivity:
                        un = np.maximum(epps, un)
                    th = f32(p.thetaCP[kk])
                    x = (un + (th * (un - ucp).astype(f32)).astype(f32)).astype(f32)
                    ucp = un
            else:   # CGLS
                gamma_new = f32(np.dot(rhs.astype(np.float64), rhs.astype(np.float64)))
                beta = f32(gamma_new / gamma_cg)
                f_cg = ((alpha_cg * x).astype(f32) + f32(1) * f_cg).astype(f32)
                if it == p.Niter - 1:
                    x = f_cg.copy()
                else:
                    x = (f32(1) * rhs + (beta * x).astype(f32)).astype(f32)
                gamma_cg = gamma_new
        # end of the iteration
        if p.BSREM:
            if reg(it, p.Niter - 1):
                dU = prior(dU, True)
            lam = f32(p.lambdaN[it])
            v2 = (x - ((lam * x).astype(f32) * dU).astype(f32)).astype(f32)
            v2[v2 < epps] = epps
            x = v2
        mid = len(p.saveNiter)
        if p.save_iter or (mid > 0 and (it == p.Niter - 1 or (ee < mid and p.saveNiter[ee] == it))):
            if p.save_iter and it == 0:
                out.append(np.asarray(x0, f32).copy())
            out.append(x.copy())
            ee += 1
        if p.enforcePositivity and S_n == 1 and not p.PDHG:
            x[x < epps] = epps
    if not (p.save_iter or len(p.saveNiter) > 0):
        out.append(x.copy())
    return {"output": out, "fp": fps, "residual": resid}
