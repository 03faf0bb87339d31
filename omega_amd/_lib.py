"""ctypes binding of libomega_hip.so (the C ABI declared in include/omega_hip.h).

There is no fallback: if the library is missing or fails to load, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libomega_hip.so")

OMG_OK = 0
OMG_ERR_ARG = -1
OMG_ERR_HIP = -2
OMG_ERR_UNSUPPORTED = -3

MODE_PET_SINO, MODE_CT_FLAT, MODE_LIST6 = 0, 1, 2
PROJ_SIDDON, PROJ_ORTH, PROJ_VOLUME, PROJ_INTERP, PROJ_BDD = 1, 2, 3, 4, 5


class OmegaHipError(RuntimeError):
    """A libomega_hip.so entry point returned an error code."""

    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} failed with code {code}: {msg}")
        self.code = code


class GeomDesc(C.Structure):
    """Mirror of omg_geom_desc (include/omega_hip.h); field order and types must match."""
    _fields_ = [
        ("struct_size", C.c_uint32),
        ("projector_type", C.c_int32),
        ("mode", C.c_int32),
        ("Nx", C.c_uint32), ("Ny", C.c_uint32), ("Nz", C.c_uint32),
        ("dx", C.c_float), ("dy", C.c_float), ("dz", C.c_float),
        ("bx", C.c_float), ("by", C.c_float), ("bz", C.c_float),
        ("size_x", C.c_uint32),
        ("size_y", C.c_uint32),
        ("n_proj", C.c_uint32),
        ("ct_pitch", C.c_int32),
        ("dPitchZ", C.c_float),
        ("global_factor", C.c_float),
        ("x", C.c_void_p), ("x_len", C.c_int64),
        ("z", C.c_void_p), ("z_len", C.c_int64),
        ("norm", C.c_void_p), ("norm_len", C.c_int64),
        ("orth_width", C.c_float),
        ("orth_flags", C.c_int32),
        ("dPitchX", C.c_float),
        ("dPitchY", C.c_float),
        ("reserved", C.c_int32 * 8),
        ("meas_index", C.c_void_p), ("meas_index_len", C.c_int64),
        ("atten_mode", C.c_int32), ("atten", C.c_void_p), ("atten_len", C.c_int64),
        ("scat", C.c_void_p), ("scat_len", C.c_int64),
        ("tof_bins", C.c_uint32), ("tof_sigma", C.c_float), ("tof_epps", C.c_float), ("tof_centers", C.c_void_p),
        ("rays_xy", C.c_uint32), ("rays_z", C.c_uint32), ("cr_xy", C.c_float), ("cr_z", C.c_float),
        ("vol_V", C.c_void_p), ("vol_len", C.c_int64), ("vol_bmin", C.c_float), ("vol_bmax", C.c_float),
        ("vol_vmax", C.c_float),
        ("interp_dl", C.c_float), ("interp_kerroin", C.c_float),
    ]


_P = C.c_void_p
_I64 = C.c_int64
_F = C.c_float
_I32 = C.c_int32
_U32 = C.c_uint32

# name -> argtypes (all return int)
SIGNATURES = {
    "omg_geom_create": [C.POINTER(GeomDesc), C.POINTER(_P)],
    "omg_geom_destroy": [_P],
    "omg_forward_project": [_P, _I64, _I64, _P, _P, _P],
    "omg_backward_project": [_P, _I64, _I64, _P, _P, _P, _P],
    "omg_osem_data_step": [_P, _I64, _I64, _P, _P, _P, _P, _F, _P],
    "omg_siddon_trace": [_P, _I64, _I64, _P, _P, _P, _P, _P],
    "omg_osem_subset_iteration": [_P, _I64, _I64, _P, _P, _P, _I32, _F, _F, _F, _I32, _P],
    "omg_osem_data_step_ex": [_P, _I64, _I64, _P, _P, _P, _P, _P, _F, _P],
    "omg_osem_subset_iteration_ex": [_P, _I64, _I64, _P, _P, _P, _P, _I32, _F, _F, _F, _I32, _P],
    "omg_osem_subset_iteration_noalloc": [_P, _I64, _I64, _P, _P, _P, _P, _I32, _F, _F, _F, _I32, _P],
    "omg_ct_osem_subset_iteration": [_P, _I64, _I64, _P, _P, _P, _I32, _F, _F, _F, _F, _I32, _P],
    "omg_frames_sensitivity": [_P, _I64, _I64, _P, _F, _P],
    "omg_frames_forward": [_P, _I64, _I64, _I32, _P, _P, _I64, _P],
    "omg_frames_loglik": [_P, _P, _I64, _I32, _P, _P],
    "omg_frames_osem_step": [_P, _I64, _I64, _I32, _P, _P, _I64, _P, _P, _P, _F, _F, _F, _I32, _I32, _I32, _I32, _P],
    "omg_frames_graph_create": [_P, _I32, _P, _P, _P, _I32, _P, _P, _I64, _P, _P, _P, _F, _F, _F, _I32, _I32, _I32,
                                _I32, _P],
    "omg_graph_launch": [_P, _P],
    "omg_graph_destroy": [_P],
    "omg_conv3d_forward": [_P, _I32, _P, _I32, _P, _P, _P, _I32, _I32, _I32, _I32, _I32, _P],
    "omg_conv3d_wgrad": [_P, _I32, _P, _I32, _P, _I32, _P, _I32, _I32, _I32, _P],
    "omg_conv3d_flipT": [_P, _P, _I32, _I32, _P],
    "omg_bias_grad": [_P, _P, _I32, _I32, _P],
    "omg_relu_backward": [_P, _P, _I64, _P],
    "omg_pool_forward": [_P, _P, _I32, _I32, _I32, _I32, _P],
    "omg_pool_backward": [_P, _P, _I32, _I32, _I32, _I32, _I32, _P],
    "omg_upsample_forward": [_P, _P, _I32, _I32, _I32, _I32, _P],
    "omg_upsample_backward": [_P, _I32, _I32, _P, _I32, _I32, _I32, _I32, _P],
    "omg_take_channels": [_P, _I32, _I32, _P, _I64, _I32, _I32, _P],
    "omg_head_forward": [_P, _P, _P, _P, _P, _I32, _I32, _P],
    "omg_head_backward": [_P, _P, _P, _P, _P, _P, _I32, _I32, _P],
    "omg_mse": [_P, _P, _P, _P, _P, _I32, _F, _P],
    "omg_adam": [_P, _P, _P, _P, _P, _I32, _F, _F, _F, _F, _I32, _P],
    "omg_to_bf16": [_P, _P, _I64, _P],
    "omg_embed": [_P, _P, _I32, _I32, _I32, _I32, _F, _P],
    "omg_timing_enable": [_I32],
    "omg_timing_collect": [C.POINTER(C.c_double), C.POINTER(C.c_int64)],
    "omg_rdp_gradient": [_P, _P, _U32, _U32, _U32, _F, _F, _F, _I32, _P],
    "omg_ratio": [_P, _P, _I64, _F, _P],
    "omg_sens_clamp": [_P, _I64, _F, _P],
    "omg_osem_update": [_P, _P, _P, _I64, _F, _I32, _I32, _P],
    "omg_osl_update": [_P, _P, _P, _P, _I64, _F, _I32, _I32, _P],
    "omg_pdhg_update": [_P, _P, _P, _I64, _F, _F, _F, _I32, _I32, _P],
    "omg_pdhg_dual_l2": [_P, _P, _P, _I64, _F, _P],
    "omg_dot": [_P, _P, _I64, _P, _P],
    "omg_axpby": [_P, _P, _I64, _F, _F, _P],
    "omg_poisson_loglik": [_P, _P, _I64, _P, _P],
    "omg_rdp_value": [_P, _U32, _U32, _U32, _F, _F, _P, _P],
}

_lib = None


def load():
    """Load libomega_hip.so; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built: run `python -c 'import __graft_entry__ as g; g.build()'` "
                           "or `python omega_amd/build.py`")
    lib = C.CDLL(LIB_PATH)
    for name, argt in SIGNATURES.items():
        f = getattr(lib, name)
        f.argtypes = argt
        f.restype = C.c_int
    lib.omg_version.restype = C.c_char_p
    lib.omg_version.argtypes = []
    lib.omg_last_error.restype = C.c_char_p
    lib.omg_last_error.argtypes = []
    _lib = lib
    return lib


def call(name: str, *args) -> None:
    """Call an entry point and raise OmegaHipError on a non-zero return code."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != OMG_OK:
        raise OmegaHipError(name, rc, lib.omg_last_error().decode(errors="replace"))


def last_error() -> str:
    return load().omg_last_error().decode(errors="replace")


def version() -> str:
    return load().omg_version().decode()


def exported_symbols():
    return list(SIGNATURES) + ["omg_version", "omg_last_error"]
