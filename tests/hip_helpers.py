"""Helpers that drive libomega_hip.so through its C ABI from tests."""
from __future__ import annotations

import ctypes as C

import numpy as np

from omega_amd import _lib
from omega_amd.projector import Geometry


def make_geom(grid, x, z, size_x, size_y, mode=_lib.MODE_PET_SINO, n_proj=None, projector_type=1, dPitchZ=0.0,
              ct_pitch=0, global_factor=1.0, norm=None, orth_width=0.0, orth_flags=0, dPitchX=0.0, dPitchY=0.0,
              generic=False, x_fastest_only=False, atten=None, atten_mode=0, scat=None, tof=None, rays=None,
              volume=None, interp=None, tube_budget_mib=0):
    """Geometry handle through the C ABI; ``generic`` routes every call through the one-lane-per-LOR
    kernels (omg_geom_desc.reserved[0]); ``x_fastest_only`` keeps those kernels in the x-fastest image
    (reserved[3]); ``atten`` is an attenuation image (``atten_mode`` 1) or per-measurement factors (2);
    ``scat`` a per-measurement multiplicative correction; ``tof`` = (bins, sigma_mm, centres_mm, epps);
    ``rays`` = (transaxial, axial, pitch_xy, pitch_z); ``volume`` = (V, bmin, bmax, Vmax) for projector 3;
    ``interp`` = (dL mm, kerroin) for projector 4; ``tube_budget_mib``: projector 3's tube-list budget
    (reserved[4]; 0 default, < 0 off: every call traces the tubes)."""
    x = np.ascontiguousarray(x, dtype=np.float32).ravel()
    z = np.ascontiguousarray(z, dtype=np.float32).ravel() if z is not None else np.zeros(2, np.float32)
    d = _lib.GeomDesc()
    d.struct_size = C.sizeof(_lib.GeomDesc)
    d.projector_type = projector_type
    d.mode = mode
    d.Nx, d.Ny, d.Nz = grid.Nx, grid.Ny, grid.Nz
    d.dx, d.dy, d.dz = grid.dx, grid.dy, grid.dz
    d.bx, d.by, d.bz = grid.bx, grid.by, grid.bz
    d.size_x, d.size_y = size_x, size_y
    if n_proj is None:
        n_proj = (x.size // 6) if mode != _lib.MODE_PET_SINO else z.size // 2
    d.n_proj = n_proj
    d.ct_pitch = ct_pitch
    d.dPitchZ = dPitchZ
    d.global_factor = global_factor
    d.x, d.x_len = x.ctypes.data, x.size
    d.z, d.z_len = z.ctypes.data, z.size
    keep = [x, z]
    if norm is not None:
        norm = np.ascontiguousarray(norm, dtype=np.float32)
        keep.append(norm)
        d.norm, d.norm_len = norm.ctypes.data, norm.size
    if atten is not None:
        atten = np.ascontiguousarray(atten, dtype=np.float32).ravel()
        keep.append(atten)
        d.atten_mode = atten_mode
        d.atten, d.atten_len = atten.ctypes.data, atten.size
    if scat is not None:
        scat = np.ascontiguousarray(scat, dtype=np.float32).ravel()
        keep.append(scat)
        d.scat, d.scat_len = scat.ctypes.data, scat.size
    if rays is not None:
        d.rays_xy, d.rays_z, d.cr_xy, d.cr_z = int(rays[0]), int(rays[1]), float(rays[2]), float(rays[3])
    if tof is not None:
        bins, sigma, centres, epps = tof
        centres = np.ascontiguousarray(centres, dtype=np.float32)
        keep.append(centres)
        d.tof_bins, d.tof_sigma, d.tof_epps, d.tof_centers = int(bins), float(sigma), float(epps), centres.ctypes.data
    if volume is not None:
        V = np.ascontiguousarray(volume[0], dtype=np.float32)
        keep.append(V)
        d.vol_V, d.vol_len = V.ctypes.data, V.size
        d.vol_bmin, d.vol_bmax, d.vol_vmax = (float(v) for v in volume[1:4])
    if interp is not None:
        d.interp_dl, d.interp_kerroin = float(interp[0]), float(interp[1])
    d.orth_width = orth_width
    d.orth_flags = orth_flags
    d.dPitchX = dPitchX
    d.dPitchY = dPitchY
    if generic:
        d.reserved[0] = 1
    if x_fastest_only:
        d.reserved[3] = 1
    d.reserved[4] = int(tube_budget_mib)
    return Geometry(d, keep)


def ptr(t):
    return C.c_void_p(0 if t is None else t.data_ptr())


def fp(geom, first, n, img_t):
    """y = A img for measurements [first, first + n): n values per TOF bin."""
    import torch
    bins = max(1, int(geom.desc.tof_bins))
    y = torch.empty(n * bins, dtype=torch.float32, device=img_t.device)
    _lib.call("omg_forward_project", geom.handle, first, n, ptr(img_t), ptr(y), None)
    return y


def bp(geom, first, n, y_t, N, sens=False):
    import torch
    rhs = torch.zeros(N, dtype=torch.float32, device=y_t.device)
    s = torch.zeros(N, dtype=torch.float32, device=y_t.device) if sens else None
    _lib.call("omg_backward_project", geom.handle, first, n, ptr(y_t), ptr(rhs), ptr(s), None)
    return (rhs, s) if sens else rhs


def trace(geom, first, n, device):
    """Per-measurement voxel lists from the GPU traversal: (counts, idx, len)."""
    import torch
    counts = torch.zeros(n, dtype=torch.int32, device=device)
    _lib.call("omg_siddon_trace", geom.handle, first, n, ptr(counts), None, None, None, None)
    c = counts.to(torch.int64)
    offs = torch.zeros(n, dtype=torch.int64, device=device)
    if n > 1:
        offs[1:] = torch.cumsum(c, 0)[:-1]
    total = int(c.sum().item())
    idx = torch.zeros(max(total, 1), dtype=torch.int32, device=device)
    ln = torch.zeros(max(total, 1), dtype=torch.float32, device=device)
    _lib.call("omg_siddon_trace", geom.handle, first, n, ptr(counts), ptr(offs), ptr(idx), ptr(ln), None)
    return counts.cpu().numpy().astype(np.int64), idx[:total].cpu().numpy().view(np.uint32), ln[:total].cpu().numpy()


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))
