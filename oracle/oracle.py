"""ctypes bindings for the CPU oracle (liboracle.so) and the reference checker (_ref/libomega_ref.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the cpu_baseline
leg of bench.py. The product package never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libomega_ref.so")

_lib = None
_ref = None

MODE_PET, MODE_CT, MODE_LIST = 0, 1, 2

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


def build(force: bool = False) -> None:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", HERE, "liboracle.so"] + (["-B"] if force else []))


def build_ref(force: bool = False) -> bool:
    """Compile the reference checker if the reference sources are present; returns availability."""
    src = "/root/reference/source/cpp/projector_functions.h"
    if os.path.exists(src) and (force or not os.path.exists(REF_PATH)):
        subprocess.check_call(["make", "-s", "-C", HERE, "ref"] + (["-B"] if force else []))
    return os.path.exists(REF_PATH)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        for T, suf, p in ((C.c_float, "f32", _f32p), (C.c_double, "f64", _f64p)):
            geo = [C.c_void_p, p, p, C.c_uint32, C.c_uint32, C.c_int32, C.c_int64, T, C.c_int32, T, C.c_void_p,
                   C.c_void_p, C.c_int32, C.c_void_p, C.c_uint32, T, C.c_void_p, T, C.c_uint32, C.c_uint32, T, T,
                   C.c_int64]
            f = getattr(_lib, "oracle_siddon_fp_" + suf)
            f.argtypes = geo + [p, p, p, p, C.c_int]
            f.restype = C.c_int
            f = getattr(_lib, "oracle_siddon_bp_" + suf)
            f.argtypes = geo + [p, p, p, p, C.c_void_p, C.c_int]
            f.restype = C.c_int
            f = getattr(_lib, "oracle_siddon_trace_" + suf)
            f.argtypes = geo + [p, p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]
            f.restype = C.c_int64
            f = getattr(_lib, "oracle_siddon_atten_" + suf)
            f.argtypes = geo + [p, p, p, C.c_int]
            f.restype = C.c_int
            f = getattr(_lib, "oracle_rdp_gradient_" + suf)
            f.argtypes = [p, p, C.c_int, C.c_int, C.c_int, T, T, T, C.c_int]
            f.restype = C.c_int
        bgeo = [C.c_void_p, _f32p, _f32p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float, C.c_float, _f32p, _f32p]
        _lib.oracle_bdd_fp_f32.argtypes = bgeo + [_f32p, _f32p, C.c_int]
        _lib.oracle_bdd_bp_f32.argtypes = bgeo + [_f32p, _f32p, C.c_void_p, C.c_int]
        _lib.oracle_bdd_geom5_f32.argtypes = bgeo + [_f32p]
        ogeo = [C.c_void_p, _f32p, _f32p, C.c_uint32, C.c_uint32, C.c_int32, C.c_int64, C.c_float, C.c_void_p,
                C.c_void_p, C.c_void_p, C.c_float, C.c_int32, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, _f32p,
                _f32p]
        _lib.oracle_orth_fp_f32.argtypes = ogeo + [_f32p, _f32p, C.c_int]
        _lib.oracle_orth_bp_f32.argtypes = ogeo + [_f32p, _f32p, C.c_void_p, C.c_int]
        tofa = [C.c_uint32, C.c_float, _f32p, C.c_float]
        _lib.oracle_orth_tof_fp_f32.argtypes = ogeo + tofa + [_f32p, _f32p, C.c_int]
        _lib.oracle_orth_tof_bp_f32.argtypes = ogeo + tofa + [_f32p, _f32p, C.c_void_p, C.c_int]
        _lib.oracle_orth_trace_f32.argtypes = ogeo + [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]
        _lib.oracle_orth_trace_f32.restype = C.c_int64
        igeo = [C.c_void_p, _f32p, _f32p, C.c_uint32, C.c_uint32, C.c_int32, C.c_int64, C.c_float, C.c_int32,
                C.c_float, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, _f32p, _f32p, C.c_float, C.c_float,
                C.c_float, C.c_float, C.c_void_p, C.c_int64, _f32p, _f32p]
        _lib.oracle_interp_fp_f32.argtypes = igeo + [_f32p, _f32p, C.c_int]
        _lib.oracle_interp_tof_fp_f32.argtypes = igeo + [C.c_uint32, C.c_float, _f32p, _f32p, _f32p, C.c_int]
        _lib.oracle_interp_tof_bp_f32.argtypes = igeo + [C.c_uint32, C.c_float, _f32p, _f32p, _f32p, C.c_void_p,
                                                         C.c_int]
        _lib.oracle_interp_bp_f32.argtypes = igeo + [_f32p, _f32p, C.c_void_p, C.c_int]
        _lib.oracle_interp_atten_f32.argtypes = igeo + [_f32p, _f32p, C.c_int]
        _lib.oracle_ratio_f32.argtypes = [_f32p, _f32p, C.c_int64, C.c_float]
        _lib.oracle_sens_clamp_f32.argtypes = [_f32p, C.c_int64, C.c_float]
        _lib.oracle_osem_update_f32.argtypes = [_f32p, _f32p, _f32p, C.c_int64, C.c_float, C.c_int]
        _lib.oracle_osl_update_f32.argtypes = [_f32p, _f32p, _f32p, _f32p, C.c_int64, C.c_float, C.c_int]
    return _lib


def ref():
    """The reference checker, or None when it was not built (e.g. no reference sources)."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_PATH):
            return None
        _ref = C.CDLL(REF_PATH)
        for T, suf, p in ((C.c_float, "f32", _f32p), (C.c_double, "f64", _f64p)):
            f = getattr(_ref, "ref_siddon_" + suf)
            f.argtypes = [C.c_void_p, p, p, C.c_uint32, C.c_uint32, C.c_int32, C.c_int64, T, C.c_int32, T,
                          C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_uint32, T, C.c_void_p, T,
                          C.c_uint32, C.c_uint32, T, T, C.c_uint32, C.c_int64, p, p, p, p, C.c_void_p, C.c_int,
                          C.c_int]
            f.restype = C.c_int
        _ref.ref_rdp_f32.argtypes = [_f32p, _f32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float]
    return _ref


class SiddonCase:
    """Everything a Siddon call needs besides the data vectors."""

    def __init__(self, grid, size_x, size_y, x, z, mode=MODE_PET, plane_offset=0, dPitchZ=0.0, ct_pitch=0,
                 global_factor=1.0, norm=None, n_proj=None, dtype=np.float32, atten=None, atten_mode=0, scat=None,
                 tof_bins=0, tof_sigma=0.0, tof_centers=None, tof_epps=1e-5, rays=(1, 1), ray_pitch=(0.0, 0.0)):
        """atten: attenuation image mu (atten_mode 1, N values, 1/mm, x fastest) or one factor per
        measurement (atten_mode 2, norm's order); scat: per-measurement multiplicative correction;
        tof_bins > 1: time of flight with standard deviation tof_sigma (mm) and bin centres tof_centers
        (mm); measurement vectors then hold tof_bins blocks of nmeas values; rays = (transaxial, axial)
        rays per LOR spread over ray_pitch (mm)."""
        self.dtype = np.dtype(dtype)
        self.n3 = np.ascontiguousarray(grid.n3())
        self.d3 = grid.d3(self.dtype)
        self.b3 = grid.b3(self.dtype)
        self.grid = grid
        self.size_x, self.size_y = int(size_x), int(size_y)
        self.x = np.ascontiguousarray(x, dtype=self.dtype)
        self.z = np.ascontiguousarray(z, dtype=self.dtype)
        self.mode = mode
        self.plane_offset = int(plane_offset)
        self.dPitchZ = dPitchZ
        self.ct_pitch = ct_pitch
        self.global_factor = global_factor
        self.norm = None if norm is None else np.ascontiguousarray(norm, dtype=self.dtype)
        self.n_proj = n_proj if n_proj is not None else max(1, self.z.size // 2)
        self.atten = None if atten is None else np.ascontiguousarray(atten, dtype=self.dtype)
        self.atten_mode = int(atten_mode) if atten is not None else 0
        self.scat = None if scat is None else np.ascontiguousarray(scat, dtype=self.dtype)
        self.tof_bins = int(tof_bins) if tof_bins and tof_bins > 1 else 1
        self.tof_sigma = tof_sigma
        self.tof_centers = None if tof_centers is None else np.ascontiguousarray(tof_centers, dtype=self.dtype)
        self.tof_epps = tof_epps
        self.rays = (int(rays[0]), int(rays[1]))
        self.ray_pitch = (float(ray_pitch[0]), float(ray_pitch[1]))

    @property
    def suf(self):
        return "f32" if self.dtype == np.float32 else "f64"

    def _geo(self, nmeas):
        normp = None if self.norm is None else self.norm.ctypes.data
        atnp = None if self.atten is None else self.atten.ctypes.data
        scp = None if self.scat is None else self.scat.ctypes.data
        tcp = None if self.tof_centers is None else self.tof_centers.ctypes.data
        return [self.n3.ctypes.data, self.d3, self.b3, self.size_x, self.size_y, self.mode, self.plane_offset,
                self.dPitchZ, self.ct_pitch, self.global_factor, normp, atnp, self.atten_mode, scp,
                self.tof_bins, self.tof_sigma, tcp, self.tof_epps, self.rays[0], self.rays[1], self.ray_pitch[0],
                self.ray_pitch[1], nmeas]

    # ---- oracle -------------------------------------------------------------------------
    def fp(self, img, nmeas, nthreads=0):
        y = np.zeros(nmeas * self.tof_bins, dtype=self.dtype)
        getattr(lib(), "oracle_siddon_fp_" + self.suf)(*self._geo(nmeas), self.x, self.z,
                                                       np.ascontiguousarray(img, dtype=self.dtype), y, nthreads)
        return y

    def bp(self, yin, nmeas, sens=False, nthreads=0):
        n = self.grid.N
        rhs = np.zeros(n, dtype=self.dtype)
        s = np.zeros(n, dtype=self.dtype) if sens else None
        getattr(lib(), "oracle_siddon_bp_" + self.suf)(*self._geo(nmeas), self.x, self.z,
                                                       np.ascontiguousarray(yin, dtype=self.dtype), rhs,
                                                       None if s is None else s.ctypes.data, nthreads)
        return (rhs, s) if sens else rhs

    def atten_factors(self, nmeas, nthreads=0):
        """exp(jelppi) per measurement for the attenuation image of this case (forward-projection
        semantics: clipped walk lengths; axis-parallel LORs over their whole row); 1 for skipped LORs."""
        assert self.atten_mode == 1
        out = np.ones(nmeas, dtype=self.dtype)
        getattr(lib(), "oracle_siddon_atten_" + self.suf)(*self._geo(nmeas), self.x, self.z, out, nthreads)
        return out

    def trace(self, nmeas):
        """Per-LOR voxel index lists (traversal order), lengths and LOR weights."""
        f = getattr(lib(), "oracle_siddon_trace_" + self.suf)
        counts = np.zeros(nmeas, dtype=np.int64)
        w = np.zeros(nmeas, dtype=self.dtype)
        total = f(*self._geo(nmeas), self.x, self.z, counts.ctypes.data, 0, None, None, w.ctypes.data)
        idx = np.zeros(max(total, 1), dtype=np.uint32)
        ln = np.zeros(max(total, 1), dtype=self.dtype)
        f(*self._geo(nmeas), self.x, self.z, None, total, idx.ctypes.data, ln.ctypes.data, None)
        return counts, idx[:total], ln[:total], w

    # ---- reference checker ---------------------------------------------------------------
    def ref_fp(self, img, nmeas, nthreads=0):
        r = ref()
        y = np.zeros(nmeas * self.tof_bins, dtype=self.dtype)
        g = self._geo(nmeas)
        getattr(r, "ref_siddon_" + self.suf)(*g[:-1], self.n_proj, nmeas, self.x, self.z,
                                             np.ascontiguousarray(img, dtype=self.dtype).copy(), y, None, 1, nthreads)
        return y

    def ref_bp(self, yin, nmeas, sens=False, nthreads=0):
        r = ref()
        n = self.grid.N
        rhs = np.zeros(n, dtype=self.dtype)
        s = np.zeros(n, dtype=self.dtype) if sens else None
        g = self._geo(nmeas)
        getattr(r, "ref_siddon_" + self.suf)(*g[:-1], self.n_proj, nmeas, self.x, self.z,
                                             np.ascontiguousarray(yin, dtype=self.dtype).copy(), rhs,
                                             None if s is None else s.ctypes.data, 2, nthreads)
        return (rhs, s) if sens else rhs


def rdp_gradient(u, shape, gamma, epps, beta, grad=None, dtype=np.float32, nthreads=0):
    Nx, Ny, Nz = shape
    suf = "f32" if np.dtype(dtype) == np.float32 else "f64"
    u = np.ascontiguousarray(u, dtype=dtype)
    g = np.zeros_like(u) if grad is None else np.ascontiguousarray(grad, dtype=dtype).copy()
    getattr(lib(), "oracle_rdp_gradient_" + suf)(g, u, Nx, Ny, Nz, gamma, epps, beta, nthreads)
    return g


def ref_rdp_gradient(u, shape, gamma, epps, beta, grad=None):
    Nx, Ny, Nz = shape
    u = np.ascontiguousarray(u, dtype=np.float32)
    g = np.zeros_like(u) if grad is None else np.ascontiguousarray(grad, dtype=np.float32).copy()
    ref().ref_rdp_f32(g, u, Nx, Ny, Nz, gamma, epps, beta)
    return g


def ratio(y, yhat, epps):
    r = np.ascontiguousarray(yhat, dtype=np.float32).copy()
    lib().oracle_ratio_f32(np.ascontiguousarray(y, dtype=np.float32), r, r.size, epps)
    return r


def ratio_additive(y, yhat, add, epps):
    """The ratio with the randoms / scatter term, in float32 and the reference's order
    (computeForwardProject.h:20-21, 66-69, 94): a = yhat + epps; a = a + add; r = y / a; NaN/Inf -> epps."""
    f = np.float32
    a = (np.asarray(yhat, f) + f(epps)).astype(f)
    a = (a + np.asarray(add, f)).astype(f)
    with np.errstate(divide="ignore", invalid="ignore"):
        r = (np.asarray(y, f) / a).astype(f)
    r[~np.isfinite(r)] = f(epps)
    return r


def sens_clamp(s, epps):
    s = np.ascontiguousarray(s, dtype=np.float32).copy()
    lib().oracle_sens_clamp_f32(s, s.size, epps)
    return s


def osem_update(x, S, rhs, epps, positivity=True):
    x = np.ascontiguousarray(x, dtype=np.float32).copy()
    lib().oracle_osem_update_f32(x, np.ascontiguousarray(S, dtype=np.float32),
                                 np.ascontiguousarray(rhs, dtype=np.float32), x.size, epps, int(positivity))
    return x


def osl_update(x, S, dU, rhs, epps, positivity=True):
    x = np.ascontiguousarray(x, dtype=np.float32).copy()
    lib().oracle_osl_update_f32(x, np.ascontiguousarray(S, dtype=np.float32), np.ascontiguousarray(dU, dtype=np.float32),
                                np.ascontiguousarray(rhs, dtype=np.float32), x.size, epps, int(positivity))
    return x


def _ratio64(y, yhat, epps, add=None):
    """oracle_ratio_f32's steps in float64: r = y / ((yhat + epps) + add), NaN/Inf -> epps."""
    with np.errstate(divide="ignore", invalid="ignore"):
        den = np.asarray(yhat, np.float64) + epps
        if add is not None:
            den = den + np.asarray(add, np.float64)
        r = np.asarray(y, np.float64) / den
    r[~np.isfinite(r)] = epps
    return r


def _sens_clamp64(s, epps):
    s = np.asarray(s, np.float64).copy()
    s[s < epps] = epps
    return s


def _update64(x, S, dU, rhs, epps, positivity):
    """oracle_osem_update_f32 (dU None) / oracle_osl_update_f32 in float64, same operation order."""
    v = x / S * (rhs + epps) if dU is None else x / (S + dU) * rhs
    if positivity:
        v = np.where(v < epps, epps, v)
    return v


def osem(grid, size_x, size_y, xy, z_subset_ordered, nmeas_planes, meas, n_iter, x0, epps=1e-5, beta=0.0,
         gamma=2.0, use_ref=False, nthreads=0, positivity=True, subsets_to_run=None, dtype=np.float32,
         additive=None, atten=None, atten_mode=0, tof=None, rays=(1, 1), ray_pitch=(0.0, 0.0)):
    """OSEM / OSL-OSEM(RDP) on the CPU in the reference's step order (see omega_amd/recon.py).

    ``meas`` is the full measurement vector in subset order. With ``use_ref`` the forward and
    backward projections come from the reference build (oracle/_ref), the elementwise steps from
    the oracle. ``subsets_to_run`` limits the run to the first that many subset-iterations.
    ``dtype=np.float64`` is the FP64 re-run: the same projector and steps in double precision (not
    with ``use_ref``). ``additive`` (optional, laid out like ``meas``) is the randoms / scatter term
    of the ratio. ``atten``: attenuation image (``atten_mode`` 1) or per-measurement factors laid out
    like ``meas`` (``atten_mode`` 2). ``tof``: optional dict of SiddonCase's tof_* arguments; each
    subset's measurements are then tof_bins blocks of its LORs and ``additive`` (one value per LOR)
    is added to every bin. ``rays`` / ``ray_pitch``: rays per LOR and their spread (SiddonCase). Returns the
    image (float32, or float64 for the FP64 re-run).
    """
    tof = dict(tof or {})
    nb = int(tof.get("tof_bins", 1)) if int(tof.get("tof_bins", 1)) > 1 else 1
    f64 = np.dtype(dtype) == np.float64
    assert not (f64 and use_ref), "the reference build is FP32 only"
    per = size_x * size_y
    starts = np.concatenate([[0], np.cumsum(nmeas_planes)]).astype(np.int64)
    x = np.ascontiguousarray(x0, dtype=dtype).copy()
    S = [None] * len(nmeas_planes)
    done = 0
    shape = (grid.Nx, grid.Ny, grid.Nz)
    for it in range(n_iter):
        for k in range(len(nmeas_planes)):
            if subsets_to_run is not None and done >= subsets_to_run:
                return x
            n = int(nmeas_planes[k]) * per
            atn = atten
            if atten is not None and atten_mode == 2:
                atn = atten[starts[k] * per: starts[k] * per + n]
            case = SiddonCase(grid, size_x, size_y, xy, z_subset_ordered, plane_offset=int(starts[k]), dtype=dtype,
                              atten=atn, atten_mode=atten_mode, rays=rays, ray_pitch=ray_pitch, **tof)
            yk = meas[nb * starts[k] * per: nb * (starts[k] * per + n)]
            yhat = case.ref_fp(x, n, nthreads) if use_ref else case.fp(x, n, nthreads)
            if additive is not None:
                ak = np.tile(additive[starts[k] * per: starts[k] * per + n], nb)
                r = _ratio64(yk, yhat, epps, ak) if f64 else ratio_additive(yk, yhat, ak, epps)
            else:
                r = _ratio64(yk, yhat, epps) if f64 else ratio(yk, yhat, epps)
            if it == 0:
                rhs, s = (case.ref_bp if use_ref else case.bp)(r, n, sens=True, nthreads=nthreads)
                S[k] = _sens_clamp64(s, epps) if f64 else sens_clamp(s, epps)
            else:
                rhs = (case.ref_bp if use_ref else case.bp)(r, n, nthreads=nthreads)
            dU = rdp_gradient(x, shape, gamma, epps, beta, dtype=dtype, nthreads=nthreads) if beta != 0.0 else None
            if f64:
                x = _update64(x, S[k], dU, rhs, epps, positivity)
            elif dU is not None:
                x = osl_update(x, S[k], dU, rhs, epps, positivity)
            else:
                x = osem_update(x, S[k], rhs, epps, positivity)
            done += 1
    return x


def ct_osem(cases, meas_parts, n_iter, x0, flat, epps=1e-5, beta=0.0, gamma=2.0, positivity=True, shape=None,
            subsets_to_run=None, nthreads=0):
    """Transmission (CT) OSEM / OSL-OSEM without randoms in the reference's step order, one oracle case
    (Siddon or BDD) per subset and the subset's measured counts in ``meas_parts``:
    S_k = A_k^T y_k in iteration 0 with S_k < epps -> epps (reconstructionAF.h:322-327,1224-1260);
    yhat = A_k x, yhat < 0 -> 0, r = exp(-yhat) (computeForwardProject.h:80-88); rhs = A_k^T r;
    x = x / S_k * (flat * rhs) (subiterStep.h:117-118) or x / (S_k + dU) * (flat * rhs) (:223-224),
    dU = beta * grad RDP(x) (``shape`` = (Nx, Ny, Nz)); positivity x < epps -> epps. The arithmetic is in
    the cases' dtype; the exponential is evaluated in double and rounded once, as the library does."""
    dt = np.dtype(getattr(cases[0], "dtype", np.float32))
    f = dt.type

    def fp(case, v, n):
        return case.fp(v, nthreads=nthreads) if isinstance(case, BDDCase) else case.fp(v, n, nthreads)

    def bp(case, v, n):
        return case.bp(v, nthreads=nthreads) if isinstance(case, BDDCase) else case.bp(v, n, nthreads=nthreads)

    x = np.ascontiguousarray(x0, dtype=dt).copy()
    S = [None] * len(cases)
    done = 0
    for it in range(n_iter):
        for k, case in enumerate(cases):
            if subsets_to_run is not None and done >= subsets_to_run:
                return x
            y = np.asarray(meas_parts[k], dtype=dt)
            n = y.size
            if S[k] is None:
                s = bp(case, y, n)
                S[k] = np.where(s < f(epps), f(epps), s).astype(dt)
            yhat = fp(case, x, n)
            yhat = np.where(yhat < 0, f(0), yhat).astype(dt)
            r = np.exp(-yhat.astype(np.float64)).astype(dt)   # exp in double, rounded once
            rhs = bp(case, r, n)
            num = (f(flat) * rhs).astype(dt)
            if beta != 0.0:
                dU = rdp_gradient(x, shape, gamma, epps, beta, dtype=dt, nthreads=nthreads)
                v = (x / (S[k] + dU) * num).astype(dt)
            else:
                v = (x / S[k] * num).astype(dt)
            if positivity:
                v = np.where(v < f(epps), f(epps), v).astype(dt)
            x = v
            done += 1
    return x


def cgls(case, meas, n_iter, nmeas=None, nthreads=0):
    """CGLS recurrences (functions.hpp:2315-2340, computeForwardProject.h:197-206, algorithms.h:243-269)
    with the oracle projector of ``case`` (float64 vectors)."""
    n = meas.size if nmeas is None else nmeas
    y = np.asarray(meas, dtype=case.dtype)
    f = np.zeros(case.grid.N, dtype=case.dtype)
    r = y.copy()
    p = case.bp(r, n, nthreads=nthreads)
    gamma = float(np.dot(p.astype(np.float64), p.astype(np.float64)))
    for _ in range(n_iter):
        q = case.fp(p, n, nthreads)
        alpha = gamma / float(np.dot(q.astype(np.float64), q.astype(np.float64)))
        r = r - alpha * q
        s = case.bp(r, n, nthreads=nthreads)
        gamma_new = float(np.dot(s.astype(np.float64), s.astype(np.float64)))
        f = f + alpha * p
        p = s + (gamma_new / gamma) * p
        gamma = gamma_new
    return f


def pdhg(cases, meas_parts, n_iter, tau, sigma=1.0, theta=1.0, epps=1e-5, nthreads=0):
    """PDHG with subsets, least-squares data term (CBCT_custom_algorithm_example_PyTorchCUDA.py:235-264,
    PDHGUpdate auxKernels.cl:2034-2059), one oracle case per subset."""
    N = cases[0].grid.N
    f = np.zeros(N, dtype=np.float64)
    g1 = np.zeros(N, dtype=np.float64)
    p = [np.zeros(m.size) for m in meas_parts]
    S = len(cases)
    for _ in range(n_iter):
        for k, c in enumerate(cases):
            n = meas_parts[k].size
            ax = c.fp(f.astype(c.dtype), n, nthreads).astype(np.float64)
            pl = (p[k] + sigma * (ax - meas_parts[k])) / (1.0 + sigma)
            dg = c.bp((pl - p[k]).astype(c.dtype), n, nthreads=nthreads).astype(np.float64)
            p[k] = pl
            g1 = g1 + dg
            g = g1 + (theta * S) * dg
            f = np.maximum(epps, f - tau * g)
    return f


class OrthCase:
    """Orthogonal distance-based projector (GPU-kernel semantics), float32.

    ``flags``: bit 0 transaxial row sweep (tube_width_xy > 0), bit 1 axial sweep (tube_width_z > 0);
    ``width`` is the kernel's tube width (tube_width_z); ``atten``, ``scat``: optional per-measurement
    attenuation factors and multiplicative corrections: the weight is atten * norm * scat * global_factor
    for factors made from an attenuation image, norm * scat * global_factor * atten for factors given per
    LOR (``atten_lor``; norm * scat * atten * global_factor on the axis-parallel path).
    ``volume``: (V, bmin, bmax, Vmax) selects the volume-of-intersection projector (projector type 3),
    ``width`` then being the tube radius."""

    def __init__(self, grid, size_x, size_y, x, z, width, flags=3, mode=MODE_PET, plane_offset=0,
                 global_factor=1.0, norm=None, atten=None, scat=None, volume=None, atten_lor=False):
        self.n3 = np.ascontiguousarray(grid.n3())
        self.d3 = grid.d3(np.float32)
        self.b3 = grid.b3(np.float32)
        self.grid = grid
        self.size_x, self.size_y = int(size_x), int(size_y)
        self.x = np.ascontiguousarray(x, dtype=np.float32)
        self.z = np.ascontiguousarray(z, dtype=np.float32)
        # flags bit 2: `atten` holds given per-LOR factors (-DATNM order), not factors made from an image
        self.width, self.flags = float(width), int(flags) | (4 if atten_lor else 0)
        self.mode, self.plane_offset, self.global_factor = mode, int(plane_offset), global_factor
        self.norm = None if norm is None else np.ascontiguousarray(norm, dtype=np.float32)
        self.atten = None if atten is None else np.ascontiguousarray(atten, dtype=np.float32)
        self.scat = None if scat is None else np.ascontiguousarray(scat, dtype=np.float32)
        self.vol = self.vol_s = None
        if volume is not None:
            self.vol = np.ascontiguousarray(volume[0], dtype=np.float32)
            self.vol_s = np.array(volume[1:4], dtype=np.float32)

    def _geo(self, nmeas):
        normp = None if self.norm is None else self.norm.ctypes.data
        atnp = None if self.atten is None else self.atten.ctypes.data
        scp = None if self.scat is None else self.scat.ctypes.data
        volp = None if self.vol is None else self.vol.ctypes.data
        vsp = None if self.vol is None else self.vol_s.ctypes.data
        nv = 0 if self.vol is None else self.vol.size
        return [self.n3.ctypes.data, self.d3, self.b3, self.size_x, self.size_y, self.mode, self.plane_offset,
                self.global_factor, normp, atnp, scp, self.width, self.flags, volp, nv, vsp, nmeas, self.x, self.z]

    def fp(self, img, nmeas, nthreads=0):
        y = np.zeros(nmeas, dtype=np.float32)
        lib().oracle_orth_fp_f32(*self._geo(nmeas), np.ascontiguousarray(img, dtype=np.float32), y, nthreads)
        return y

    def bp(self, yin, nmeas, sens=False, nthreads=0):
        rhs = np.zeros(self.grid.N, dtype=np.float32)
        s = np.zeros(self.grid.N, dtype=np.float32) if sens else None
        lib().oracle_orth_bp_f32(*self._geo(nmeas), np.ascontiguousarray(yin, dtype=np.float32), rhs,
                                 None if s is None else s.ctypes.data, nthreads)
        return (rhs, s) if sens else rhs

    def _tof(self, tof):
        bins, sigma, centres, epps = tof
        c = np.ascontiguousarray(centres, dtype=np.float32)
        assert c.size == bins
        return [int(bins), float(sigma), c, float(epps)]

    def tof_fp(self, img, nmeas, tof, nthreads=0):
        """Time-of-flight forward projection, ``tof`` = (bins, sigma, centres, epps): bins blocks of nmeas."""
        y = np.zeros(nmeas * int(tof[0]), dtype=np.float32)
        lib().oracle_orth_tof_fp_f32(*self._geo(nmeas), *self._tof(tof), np.ascontiguousarray(img, dtype=np.float32),
                                     y, nthreads)
        return y

    def tof_bp(self, yin, nmeas, tof, sens=False, nthreads=0):
        rhs = np.zeros(self.grid.N, dtype=np.float32)
        s = np.zeros(self.grid.N, dtype=np.float32) if sens else None
        lib().oracle_orth_tof_bp_f32(*self._geo(nmeas), *self._tof(tof), np.ascontiguousarray(yin, dtype=np.float32),
                                     rhs, None if s is None else s.ctypes.data, nthreads)
        return (rhs, s) if sens else rhs

    def trace(self, nmeas):
        """Per-LOR voxel lists in visiting order: (counts, indices, tube weights, LOR weights)."""
        f = lib().oracle_orth_trace_f32
        counts = np.zeros(nmeas, dtype=np.int64)
        lw = np.zeros(nmeas, dtype=np.float32)
        total = f(*self._geo(nmeas), counts.ctypes.data, 0, None, None, lw.ctypes.data)
        idx = np.zeros(max(total, 1), dtype=np.int64)
        w = np.zeros(max(total, 1), dtype=np.float32)
        f(*self._geo(nmeas), None, total, idx.ctypes.data, w.ctypes.data, None)
        return counts, idx[:total], w[:total], lw


class InterpCase:
    """Interpolation projector (projector type 4, GPU-kernel semantics with exact FP32 trilinear sampling) of
    the oracle, float32. ``dL``: interpolation length in mm (<= 0: the Joseph variant); ``kerroin``: the CT
    backprojection scale (dx dy dz) / (dPitchX dPitchY sourceToDetector); ``atten``: per-measurement
    attenuation factors, made from an attenuation image (``atten_image``) or given per LOR; CT
    backprojection is voxel-driven over the (size_x x size_y)-pixel panels of the call's projections.
    ``rays``: (transaxial, axial, pitch_xy, pitch_z) for several rays per LOR (PET); image attenuation factors
    are then per ray, ray r's at r * nmeas + lo (as ``atten_factors`` returns them)."""

    def __init__(self, grid, size_x, size_y, x, z, dL, mode=MODE_PET, plane_offset=0, dPitchZ=0.0, ct_pitch=0,
                 global_factor=1.0, norm=None, atten=None, atten_image=False, scat=None, kerroin=0.0, dPitchX=0.0,
                 dPitchY=0.0, rays=None):
        self.grid = grid
        self.n3 = np.ascontiguousarray(grid.n3())
        self.d3 = grid.d3(np.float32)
        self.b3 = grid.b3(np.float32)
        # box far corner b + N d in double, rounded once; 1 / (N d) (init.py:705-735, coordinates.py:224-226)
        self.bmax3 = np.array([float(b) + n * float(d) for b, n, d in zip(self.b3, self.n3, self.d3)], np.float32)
        self.scale3 = np.array([1.0 / (float(d) * n) for d, n in zip(self.d3, self.n3)], np.float32)
        self.size_x, self.size_y = int(size_x), int(size_y)
        self.x = np.ascontiguousarray(x, dtype=np.float32)
        self.z = np.ascontiguousarray(z, dtype=np.float32)
        self.dL, self.mode, self.plane_offset = float(dL), mode, int(plane_offset)
        self.dPitchZ, self.ct_pitch, self.global_factor = float(dPitchZ), int(ct_pitch), float(global_factor)
        self.norm = None if norm is None else np.ascontiguousarray(norm, dtype=np.float32)
        self.atten = None if atten is None else np.ascontiguousarray(atten, dtype=np.float32)
        self.atten_image = int(bool(atten_image))
        self.scat = None if scat is None else np.ascontiguousarray(scat, dtype=np.float32)
        self.kerroin, self.dPitchX, self.dPitchY = float(kerroin), float(dPitchX), float(dPitchY)
        self.rays = None if rays is None else np.asarray(rays, dtype=np.float32)
        self.n_rays = 1 if rays is None else int(rays[0]) * int(rays[1])

    def _geo(self, nmeas):
        p = lambda a: None if a is None else a.ctypes.data  # noqa: E731
        return [self.n3.ctypes.data, self.d3, self.b3, self.size_x, self.size_y, self.mode, self.plane_offset,
                self.dPitchZ, self.ct_pitch, self.global_factor, p(self.norm), p(self.atten), self.atten_image,
                p(self.scat), self.bmax3, self.scale3, self.dL, self.kerroin, self.dPitchX, self.dPitchY,
                p(self.rays), nmeas, self.x, self.z]

    def fp(self, img, nmeas, nthreads=0):
        y = np.zeros(nmeas, dtype=np.float32)
        lib().oracle_interp_fp_f32(*self._geo(nmeas), np.ascontiguousarray(img, dtype=np.float32), y, nthreads)
        return y

    def atten_factors(self, mu, nmeas, nthreads=0):
        """exp(-sum stepLen * mu(p)) over the forward projection's samples (1 where the ray misses)."""
        out = np.ones(nmeas * self.n_rays, dtype=np.float32)
        lib().oracle_interp_atten_f32(*self._geo(nmeas), np.ascontiguousarray(mu, dtype=np.float32), out, nthreads)
        return out

    def tof_fp(self, img, nmeas, tof, nthreads=0):
        """Time-of-flight forward projection (PET), ``tof`` = (bins, sigma, centres): bins blocks of nmeas."""
        bins, sigma, centres = tof[:3]
        y = np.zeros(nmeas * int(bins), dtype=np.float32)
        lib().oracle_interp_tof_fp_f32(*self._geo(nmeas), int(bins), float(sigma),
                                       np.ascontiguousarray(centres, dtype=np.float32),
                                       np.ascontiguousarray(img, dtype=np.float32), y, nthreads)
        return y

    def tof_bp(self, yin, nmeas, tof, sens=False, nthreads=0):
        bins, sigma, centres = tof[:3]
        rhs = np.zeros(self.grid.N, dtype=np.float32)
        s = np.zeros(self.grid.N, dtype=np.float32) if sens else None
        lib().oracle_interp_tof_bp_f32(*self._geo(nmeas), int(bins), float(sigma),
                                       np.ascontiguousarray(centres, dtype=np.float32),
                                       np.ascontiguousarray(yin, dtype=np.float32), rhs,
                                       None if s is None else s.ctypes.data, nthreads)
        return (rhs, s) if sens else rhs

    def bp(self, yin, nmeas, sens=False, nthreads=0):
        rhs = np.zeros(self.grid.N, dtype=np.float32)
        s = np.zeros(self.grid.N, dtype=np.float32) if sens else None
        lib().oracle_interp_bp_f32(*self._geo(nmeas), np.ascontiguousarray(yin, dtype=np.float32), rhs,
                                   None if s is None else s.ctypes.data, nthreads)
        return (rhs, s) if sens else rhs


class BDDCase:
    """Branchless distance-driven projector of the oracle on a CT flat-panel geometry."""

    def __init__(self, grid, nRows, nCols, xyz, uv, pitchX, pitchY):
        self.grid = grid
        self.n3 = np.ascontiguousarray(grid.n3())
        self.d3 = grid.d3(np.float32)
        self.b3 = grid.b3(np.float32)
        self.nRows, self.nCols = int(nRows), int(nCols)
        self.xyz = np.ascontiguousarray(xyz, dtype=np.float32)
        self.uv = np.ascontiguousarray(uv, dtype=np.float32)
        self.nProj = self.xyz.size // 6
        self.pitchX, self.pitchY = float(pitchX), float(pitchY)

    def _g(self):
        return [self.n3.ctypes.data, self.d3, self.b3, self.nRows, self.nCols, self.nProj, self.pitchX, self.pitchY,
                self.xyz, self.uv]

    @property
    def M(self):
        return self.nRows * self.nCols * self.nProj

    def fp(self, vol, nthreads=0):
        y = np.zeros(self.M, dtype=np.float32)
        lib().oracle_bdd_fp_f32(*self._g(), np.ascontiguousarray(vol, dtype=np.float32), y, nthreads)
        return y

    def bp(self, y, sens=False, nthreads=0):
        v = np.zeros(self.grid.N, dtype=np.float32)
        s = np.zeros(self.grid.N, dtype=np.float32) if sens else None
        lib().oracle_bdd_bp_f32(*self._g(), np.ascontiguousarray(y, dtype=np.float32), v,
                                None if s is None else s.ctypes.data, nthreads)
        return (v, s) if sens else v

    def geom5(self):
        out = np.zeros(16 * self.nProj, dtype=np.float32)
        lib().oracle_bdd_geom5_f32(*self._g(), out)
        return out
