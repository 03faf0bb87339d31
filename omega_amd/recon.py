"""Device-resident OSEM / OSL-OSEM(RDP) drivers on the HIP library.

Restates the subset loop of the reference's built-in reconstruction
(source/cpp/reconstructionAF.h:1053-1531) for OSEM and OSL-OSEM with the RDP prior:

    for iter: for subset k:
        yhat = A_k x;  r = y_k / ((yhat + epps) + a_k), NaN/Inf -> epps  (computeForwardProject.h:20-21,66-69,94,400-401)
                       a_k: optional additive randoms / scatter term, 0 without it
        rhs = A_k^T r;  in iteration 0 also S_k = A_k^T 1, S_k < epps -> epps
                                                                         (reconstructionAF.h:1156-1264; functions.hpp:2996)
        OSEM:     x = x / S_k * (rhs + epps)                             (subiterStep.h:111-121)
        OSL-OSEM: dU = beta * grad RDP(x);  x = x / (S_k + dU) * rhs    (priors.h:447-451; subiterStep.h:219-231)
        positivity x < epps -> epps                                      (reconstructionAF.h:1461-1465)

For transmission (CT) data the same loop follows the reference's CT branch (reconstructionAF.h:322-327,
1224-1260; computeForwardProject.h:76-90; subiterStep.h:111-121,219-231): S_k = A_k^T y_k in iteration 0,
r = exp(-max(A_k x, 0)), rhs = A_k^T r, x = x / S_k * (flat * rhs) or x / (S_k + dU) * (flat * rhs), with
``flat`` the blank-scan value (omg_ct_osem_subset_iteration).

A subset iteration is one C-ABI call (omg_osem_subset_iteration): the fused FP -> ratio -> BP
projector kernel and one update kernel (plus the RDP gradient kernel for OSL-OSEM). Sensitivity
images are kept per subset (compute_norm_matrix = 2, reconstructionAF.h:313-317) and frozen after
the first iteration (:1508-1509). The RDP boundary of the fused path is the CPU reference's
(zero outside the volume).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


def _p(t):
    return C.c_void_p(0 if t is None else t.data_ptr())


class OSEMRunner:
    """OSEM / OSL-OSEM on a :class:`omega_amd.projectorClass` with subset-ordered data on the device."""

    def __init__(self, proj, meas, epps: float = 1e-5, beta: float = 0.0, gamma: float = 2.0,
                 enforce_positivity: bool = True, rdp_boundary: int = 0, stream=None, additive=None,
                 flat: float = 0.0):
        """``additive``: optional randoms (+ scatter) estimate, a float32 device tensor in the layout of
        ``meas`` (subset order), added to the forward projection in the ratio (randoms_correction,
        computeForwardProject.h:66-69). With TOF (``proj.nBins`` > 1) each subset's measurements are
        nBins blocks of the subset's LORs, and ``additive`` holds one value per LOR, added to every bin
        (af::tile, computeForwardProject.h:66-67). ``flat``: the blank-scan value of transmission (CT) data
        (projectorClass.flat; required for CT, whose ``meas`` are the measured counts)."""
        import torch
        self.torch = torch
        self.proj = proj
        if not proj.projectorInitialized:
            proj.initProj()
        self.g = proj.geometry.handle
        self.N = int(proj.N[0])
        self.Nxyz = (int(proj.Nx[0]), int(proj.Ny[0]), int(proj.Nz[0]))
        self.meas = meas.reshape(-1).contiguous()
        if self.meas.dtype != torch.float32 or not self.meas.is_cuda:
            raise ValueError("measurements must be a float32 device tensor")
        self.nb = int(getattr(proj, "nBins", 1))
        if self.meas.numel() != int(proj.nTotMeas[-1]) * self.nb:
            raise ValueError(f"measurement vector has {self.meas.numel()} elements, expected "
                             f"{int(proj.nTotMeas[-1]) * self.nb}")
        self.epps = float(epps)
        self.beta = float(beta)
        self.gamma = float(gamma)
        self.pos = int(bool(enforce_positivity))
        self.additive = None
        if additive is not None:
            self.additive = additive.reshape(-1).contiguous()
            if self.additive.dtype != torch.float32 or not self.additive.is_cuda or \
                    self.additive.numel() != int(proj.nTotMeas[-1]):
                raise ValueError("additive must be a float32 device tensor with one value per LOR")
        self.ct = bool(proj.CT)
        self.flat = float(flat if flat else getattr(proj, "flat", 0.0))
        if self.ct:
            if not self.flat > 0.0:
                raise ValueError("OSEM of CT data needs the blank-scan value flat > 0")
            if self.additive is not None:
                raise ValueError("OSEM of CT data is provided without a randoms / scatter term")
        self.rdp_boundary = int(rdp_boundary)
        dev = self.meas.device
        self.S = torch.zeros((proj.subsets, self.N), dtype=torch.float32, device=dev)
        self.rhs = torch.zeros(self.N, dtype=torch.float32, device=dev)
        self.sens_done = False
        self.stream = stream

    def _s(self):
        s = self.stream if self.stream is not None else self.torch.cuda.current_stream()
        return C.c_void_p(s.cuda_stream)

    def subset_step(self, x, k: int) -> None:
        """One subset iteration on image ``x`` (updated in place): omg_osem_subset_iteration."""
        proj = self.proj
        first = int(proj.nMeas[k])
        n = int(proj.nMeasSubset[k, 0])
        off = int(proj.nTotMeas[k])
        yk = self.meas[off * self.nb:(off + n) * self.nb]
        if self.ct:
            _lib.call("omg_ct_osem_subset_iteration", self.g, first, n, _p(x), _p(yk), _p(self.S[k]),
                      int(not self.sens_done), self.flat, self.epps, self.beta, self.gamma, self.pos, self._s())
            return
        ak = None if self.additive is None else self.additive[off:off + n]
        _lib.call("omg_osem_subset_iteration_ex", self.g, first, n, _p(x), _p(yk), _p(ak), _p(self.S[k]),
                  int(not self.sens_done), self.epps, self.beta, self.gamma, self.pos, self._s())

    def subset_step_unfused(self, x, k: int) -> None:
        """The same iteration through the separate entry points (data step, clamp, RDP, update)."""
        if self.ct:
            raise ValueError("the unfused iteration is for emission data")
        proj = self.proj
        first = int(proj.nMeas[k])
        n = int(proj.nMeasSubset[k, 0])
        off = int(proj.nTotMeas[k])
        st = self._s()
        sens = None if self.sens_done else self.S[k]
        if sens is not None:
            sens.zero_()
        yk = self.meas[off * self.nb:(off + n) * self.nb]
        ak = None if self.additive is None else self.additive[off:off + n]
        self.rhs.zero_()
        _lib.call("omg_osem_data_step_ex", self.g, first, n, _p(x), _p(yk), _p(ak), _p(self.rhs), _p(sens),
                  self.epps, st)
        if sens is not None:
            _lib.call("omg_sens_clamp", _p(sens), self.N, self.epps, st)
        if self.beta == 0.0:
            _lib.call("omg_osem_update", _p(x), _p(self.S[k]), _p(self.rhs), self.N, self.epps, self.pos, 1, st)
        else:
            dU = self.torch.zeros_like(x)
            _lib.call("omg_rdp_gradient", _p(x), _p(dU), *self.Nxyz, self.gamma, self.epps, self.beta,
                      self.rdp_boundary, st)
            _lib.call("omg_osl_update", _p(x), _p(self.S[k]), _p(dU), _p(self.rhs), self.N, self.epps,
                      self.pos, 1, st)

    def run(self, n_iter: int, x0=None):
        """Reconstruct from ``x0`` (default: proj.x0) for ``n_iter`` full iterations; returns the image."""
        torch = self.torch
        if x0 is None:
            x0 = torch.as_tensor(self.proj.x0, device=self.meas.device)
        x = x0.reshape(-1).to(torch.float32).clone().contiguous()
        self.rhs.zero_()
        self.sens_done = False
        for it in range(n_iter):
            for k in range(self.proj.subsets):
                self.subset_step(x, k)
            self.sens_done = True
        return x


def poisson_loglik(y, yhat, epps=0.0) -> float:
    """L(x) = sum_i y_i log(yhat_i) - yhat_i (FP64), the objective OSEM ascends."""
    y = np.asarray(y, dtype=np.float64)
    yh = np.asarray(yhat, dtype=np.float64) + epps
    m = yh > 0
    return float(np.sum(y[m] * np.log(yh[m])) - np.sum(yh))


def _dot(a, b, out) -> float:
    """<a, b> in double on the device (omg_dot); returns a Python float (synchronises)."""
    _lib.call("omg_dot", _p(a), _p(b), a.numel(), _p(out), None)
    return float(out.item())


class CGLSRunner:
    """CGLS without subsets, restating the reference's recurrences
    (init functions.hpp:2315-2340; step computeForwardProject.h:197-206; update algorithms.h:243-269):

        r = y; f = x0; p = A^T r; gamma = <p, p>
        per iteration: q = A p; alpha = gamma / <q, q>; r = r - alpha q; s = A^T r;
                       gamma' = <s, s>; f = f + alpha p; p = s + (gamma'/gamma) p; gamma = gamma'
    """

    def __init__(self, proj, meas):
        import torch
        self.torch = torch
        self.proj = proj
        if not proj.projectorInitialized:
            proj.initProj()
        if proj.subsets != 1:
            raise ValueError("CGLS is defined without subsets (set subsets = 1)")
        self.meas = meas.reshape(-1).contiguous()
        self.acc = torch.zeros(1, dtype=torch.float64, device=self.meas.device)

    def run(self, n_iter: int, x0=None):
        torch = self.torch
        p_ = self.proj
        N = int(p_.N[0])
        f = torch.zeros(N, device=self.meas.device) if x0 is None else x0.reshape(-1).float().clone()
        r = self.meas.clone()
        if x0 is not None and float(f.abs().sum()) != 0.0:
            ax = p_.forwardProject(f, 0)
            _lib.call("omg_axpby", _p(r), _p(ax), r.numel(), -1.0, 1.0, None)
        pv = p_.backwardProject(r, 0)
        gamma = _dot(pv, pv, self.acc)
        for _ in range(n_iter):
            q = p_.forwardProject(pv, 0)
            alpha = gamma / _dot(q, q, self.acc)
            _lib.call("omg_axpby", _p(r), _p(q), r.numel(), -alpha, 1.0, None)
            s = p_.backwardProject(r, 0)
            gamma_new = _dot(s, s, self.acc)
            beta = gamma_new / gamma
            _lib.call("omg_axpby", _p(f), _p(pv), N, alpha, 1.0, None)
            _lib.call("omg_axpby", _p(pv), _p(s), N, 1.0, beta, None)
            gamma = gamma_new
        return f


def power_method(proj, iterations: int = 10, seed: int = 0, subset: int = 0) -> float:
    """Step size tau = 1 / L with L = subsets * largest eigenvalue of A_k^T A_k (power iteration),
    as util/powermethod.py:9-168 computes it for PDHG."""
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    x = torch.randn(int(proj.N[0]), device="cuda", generator=g).abs()
    x = x / torch.linalg.vector_norm(x)
    L = 1.0
    for _ in range(iterations):
        x2 = proj.backwardProject(proj.forwardProject(x, subset), subset)
        L = float(torch.dot(x, x2) / torch.dot(x, x)) * proj.subsets
        x = x2 / torch.linalg.vector_norm(x2)
    return 1.0 / L


class PDHGRunner:
    """PDHG with subsets for the least-squares data term, in the operator-API form of the reference
    (source/Python/CBCT_custom_algorithm_example_PyTorchCUDA.py:235-264; dual step
    computeForwardProject.h:227-316; image step PDHGUpdate auxKernels.cl:2034-2059):

        p_k' = (p_k + sigma (A_k f - y_k)) / (1 + sigma);  dg = A_k^T (p_k' - p_k);  p_k = p_k'
        g1 += dg;  g = g1 + theta * subsets * dg;  f = max(epps, f - tau g)
    """

    def __init__(self, proj, meas, tau: float, sigma: float = 1.0, theta: float = 1.0, epps: float = 1e-5,
                 positivity: bool = True):
        import torch
        self.torch = torch
        self.proj = proj
        if not proj.projectorInitialized:
            proj.initProj()
        self.meas = meas.reshape(-1).contiguous()
        self.tau, self.sigma, self.theta, self.epps = float(tau), float(sigma), float(theta), float(epps)
        self.pos = int(bool(positivity))
        self.p = [torch.zeros(int(proj.nMeasSubset[k, 0]), device=self.meas.device) for k in range(proj.subsets)]

    def run(self, n_iter: int, x0=None):
        torch = self.torch
        pr = self.proj
        N = int(pr.N[0])
        f = torch.zeros(N, device=self.meas.device) if x0 is None else x0.reshape(-1).float().clone()
        g1 = torch.zeros(N, device=self.meas.device)
        for _ in range(n_iter):
            for k in range(pr.subsets):
                off, n = int(pr.nTotMeas[k]), int(pr.nMeasSubset[k, 0])
                yk = self.meas[off:off + n]
                ax = pr.forwardProject(f, k)
                pl = self.p[k].clone()
                _lib.call("omg_pdhg_dual_l2", _p(pl), _p(ax), _p(yk), n, self.sigma, None)
                diff = pl - self.p[k]
                dg = pr.backwardProject(diff, k)
                self.p[k] = pl
                g1 += dg
                gk = g1 + (self.theta * pr.subsets) * dg
                _lib.call("omg_pdhg_update", _p(f), _p(gk), None, N, self.epps, 1.0, self.tau, self.pos, 1, None)
        return f
