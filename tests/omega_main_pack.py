"""Fills omegaMain's parameter block as the reference front end does and calls the library.

``transfer`` walks every member of the block and copies the option of the same name, scalars as their
ctypes type and arrays as pointers into contiguous numpy arrays of the member's element type (empty
arrays when the option is unset), as transferData does (source/Python/omegatomo/reconstruction/
recomain.py:23-375; the options here carry the block's member names). ``from_projector`` takes the
geometry, subsets and tables of a configured projectorClass, the way recomain.py:427-700 hands the
prepared options over.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from omega_amd import omega_main as M

_NP = {C.c_float: np.float32, C.c_double: np.float64, C.c_uint32: np.uint32, C.c_int32: np.int32,
       C.c_uint16: np.uint16, C.c_int16: np.int16, C.c_uint8: np.uint8, C.c_int8: np.int8, C.c_int64: np.int64,
       C.c_uint64: np.uint64, C.c_bool: np.bool_}


class Options:
    """Option values by block member name; unset scalars are 0 / False, unset arrays empty. The defaults
    set here are the front end's (proj.py) for the members omegaMain reads."""

    def __init__(self, **kw):
        self.Niter = 1
        self.Nt = 1
        self.subsets = 1
        self.subsetType = 8
        self.epps = 1e-5
        self.global_factor = 1.0
        self.U = 1e6
        self.RDP_gamma = 2.0
        self.useImages = True
        self.CTAttenuation = True
        self.loadTOF = True
        self.useL2Ball = True
        self.useFDKWeights = True
        self.useTotLength = True
        self.n_rays_transaxial = 1
        self.n_rays_axial = 1
        self.TOF_bins = 1
        self.nLayers = 1
        self.regEveryIter = 1
        self.precondTypeImage = np.zeros(7, dtype=bool)
        self.precondTypeMeas = np.zeros(2, dtype=bool)
        for k, v in kw.items():
            setattr(self, k, v)


def transfer(o: Options):
    """(parameters, keep): the packed block and the arrays its pointers reference."""
    param = M.parameters()
    keep = []
    for name, ct in M.parameters._fields_:
        val = getattr(o, name, None)
        if hasattr(ct, "_type_") and isinstance(ct._type_, type) and issubclass(ct, C._Pointer):
            arr = np.ascontiguousarray(np.asarray(val if val is not None else [], dtype=_NP[ct._type_]).ravel())
            keep.append(arr)
            setattr(param, name, arr.ctypes.data_as(ct))
        else:
            setattr(param, name, ct(0 if val is None else val).value if ct is not C.c_bool else bool(val))
    return param, keep


def from_projector(p, **kw) -> Options:
    """Options for omegaMain from a configured (addProjector'ed) projectorClass: image grid, tables,
    subset boundaries (pituus = nMeas), measurement-level subset indices, sizes."""
    o = Options(**kw)
    per = int(p.nRowsD) * int(p.nColsD)
    o.Nx = np.array([p.grid.Nx], np.uint32)
    o.Ny = np.array([p.grid.Ny], np.uint32)
    o.Nz = np.array([p.grid.Nz], np.uint32)
    o.dx, o.dy, o.dz = (np.array([v], np.float32) for v in (p.grid.dx, p.grid.dy, p.grid.dz))
    o.bx, o.by, o.bz = (np.array([v], np.float32) for v in (p.grid.bx, p.grid.by, p.grid.bz))
    o.x, o.z = p.x, p.z
    o.sizeX, o.sizeZ = p.x.size, p.z.size
    o.nRowsD, o.nColsD = int(p.nRowsD), int(p.nColsD)
    o.Ndist, o.Nang = int(p.nRowsD), int(p.nColsD)
    o.nProjections = int(p.nProjections)
    o.subsets, o.subsetType = int(p.subsets), int(p.subsetType)
    o.pituus = np.asarray(p.nMeas, np.int64)
    o.projector_type = int(p.projector_type)
    o.CT = bool(p.CT)
    o.listmode = int(bool(p.listmode))
    o.PET = not p.CT and not p.listmode and (p.subsets == 1 or p.subsetType >= 8)
    if p.subsets > 1 and p.subsetType < 8 and not p.CT and not p.listmode:
        o.xy_index = (p.index % per).astype(np.uint32)
        o.z_index = (p.index // per).astype(np.uint16)
        o.sizeXYind = o.sizeZind = p.index.size
    o.dPitchX, o.dPitchY = float(p.dPitchX), float(p.dPitchY)
    o.flat = float(p.flat)
    return o


def run(o: Options, sino, n_images: int, fp_size: int = 0, resid_size: int = 1):
    """Calls omegaMain with the measurement type of ``sino``; returns (status, output, FP, residual)."""
    sino = np.ascontiguousarray(sino)
    o.measElem = sino.size
    o.mDim = sino.size // max(1, int(o.Nt))
    lib = M.library(sino.dtype)
    param, keep = transfer(o)
    N = int(np.asarray(o.Nx)[0]) * int(np.asarray(o.Ny)[0]) * int(np.asarray(o.Nz)[0])
    out = np.zeros(N * n_images, np.float32)
    fp = np.zeros(max(fp_size, 0), np.float32)
    resid = np.zeros(max(resid_size, 1), np.float32)
    elem = {np.dtype(np.float32): C.c_float, np.dtype(np.uint16): C.c_uint16, np.dtype(np.uint8): C.c_uint8}[sino.dtype]
    rc = lib.omegaMain(param, b"", sino.ctypes.data_as(C.POINTER(elem)), out.ctypes.data_as(C.POINTER(C.c_float)),
                       fp.ctypes.data_as(C.POINTER(C.c_float)), resid.ctypes.data_as(C.POINTER(C.c_float)))
    del keep
    return rc, out.reshape(n_images, N), fp, resid
