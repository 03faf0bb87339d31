"""The C ABI on the CPU: the library loads, exports every entry point the header declares, and
rejects bad descriptors with OMG_ERR_ARG before touching the device."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from omega_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "omega_hip.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(omg_\w+)\s*\(", text, flags=re.M)))


def test_header_and_binding_agree():
    names = declared()
    assert len(names) >= 15
    assert sorted(set(_lib.exported_symbols())) == names


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in declared():
        assert hasattr(lib, name), f"{name} missing from libomega_hip.so"
    assert b"gfx950" in lib.omg_version()


def test_descriptor_layout_matches_header(tmp_path):
    """ctypes' omg_geom_desc layout equals the C compiler's (sizeof and every field offset)."""
    import subprocess
    fields = [f[0] for f in _lib.GeomDesc._fields_]
    src = tmp_path / "layout.c"
    body = "".join(f'printf("%s %zu\\n", "{f}", offsetof(omg_geom_desc, {f}));' for f in fields)
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "omega_hip.h"\n'
                   'int main(void){printf("sizeof %zu\\n", sizeof(omg_geom_desc));' + body + 'return 0;}')
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    out = dict(line.split() for line in subprocess.check_output([str(exe)]).decode().splitlines())
    assert int(out["sizeof"]) == C.sizeof(_lib.GeomDesc)
    for f in fields:
        assert int(out[f]) == getattr(_lib.GeomDesc, f).offset, f


def _desc(**kw):
    x = np.zeros(4 * 4 * 3, np.float32)
    z = np.zeros(2 * 2, np.float32)
    d = _lib.GeomDesc()
    d.struct_size = C.sizeof(_lib.GeomDesc)
    d.projector_type = 1
    d.mode = _lib.MODE_PET_SINO
    d.Nx = d.Ny = d.Nz = 4
    d.dx = d.dy = d.dz = 1.0
    d.size_x, d.size_y, d.n_proj = 4, 3, 2
    d.x, d.x_len = x.ctypes.data, x.size
    d.z, d.z_len = z.ctypes.data, z.size
    for k, v in kw.items():
        setattr(d, k, v)
    return d, (x, z)


@pytest.mark.parametrize("field,value,msg", [
    ("struct_size", 8, "size mismatch"),
    ("Nx", 0, "empty image"),
    ("dx", -1.0, "positive"),
    ("mode", 7, "unknown data mode"),
    ("projector_type", 6, "projector_type 6"),
    ("x_len", 3, "x table"),
    ("n_proj", 9, "z table"),
])
def test_create_rejects_bad_descriptors(field, value, msg):
    d, keep = _desc(**{field: value})
    h = C.c_void_p()
    lib = _lib.load()
    rc = lib.omg_geom_create(C.byref(d), C.byref(h))
    assert rc == _lib.OMG_ERR_ARG
    assert msg in _lib.last_error()


def test_create_rejects_bad_interp_settings():
    """Projector 4 needs a finite interpolation length and, on CT data, the backprojection scale and pitches."""
    lib = _lib.load()
    d, keep = _desc(projector_type=4, interp_dl=float("nan"))
    h = C.c_void_p()
    assert lib.omg_geom_create(C.byref(d), C.byref(h)) == _lib.OMG_ERR_ARG
    assert "interp_dl" in _lib.last_error()
    x = np.zeros(6 * 2, np.float32)
    uv = np.ones(2 * 2, np.float32)
    d, keep = _desc(projector_type=4, mode=_lib.MODE_CT_FLAT, interp_dl=0.5)
    d.x, d.x_len, d.z, d.z_len = x.ctypes.data, x.size, uv.ctypes.data, uv.size
    assert lib.omg_geom_create(C.byref(d), C.byref(h)) == _lib.OMG_ERR_ARG
    assert "interp_kerroin" in _lib.last_error()


def test_create_rejects_bad_subset_index():
    """A measurement index table must point into the sinogram (4 x 3 bins x 2 planes = 24 bins) and is
    for PET sinogram data only."""
    lib = _lib.load()
    h = C.c_void_p()
    bad = np.array([0, 5, 24], np.uint32)
    d, keep = _desc(meas_index=bad.ctypes.data, meas_index_len=bad.size)
    assert lib.omg_geom_create(C.byref(d), C.byref(h)) == _lib.OMG_ERR_ARG
    assert "meas_index[2] = 24" in _lib.last_error()
    ok = np.array([0, 5, 23], np.uint32)
    d, keep = _desc(meas_index=ok.ctypes.data, meas_index_len=ok.size, mode=_lib.MODE_LIST6, x_len=12)
    assert lib.omg_geom_create(C.byref(d), C.byref(h)) == _lib.OMG_ERR_ARG
    assert "PET sinogram" in _lib.last_error()


def test_create_rejects_bad_correction_tables():
    """Per-measurement tables (norm, attenuation factors) must cover the 4 x 3 x 2 = 24 measurements; an
    attenuation image must match the 4 x 4 x 4 grid; attenuation is for PET data; the mode is 0-2."""
    lib = _lib.load()
    h = C.c_void_p()
    short = np.ones(23, np.float32)
    img_bad = np.ones(63, np.float32)
    checks = [
        (dict(norm=short.ctypes.data, norm_len=short.size), "norm holds 23"),
        (dict(atten_mode=2, atten=short.ctypes.data, atten_len=short.size), "attenuation table holds 23"),
        (dict(atten_mode=1, atten=img_bad.ctypes.data, atten_len=img_bad.size), "image grid 64"),
        (dict(atten_mode=1), "without an atten table"),
        (dict(atten_mode=3, atten=short.ctypes.data, atten_len=short.size), "atten_mode 3"),
        (dict(atten_mode=2, atten=short.ctypes.data, atten_len=short.size, mode=_lib.MODE_CT_FLAT, x_len=12),
         "PET data"),
    ]
    for kw, msg in checks:
        d, keep = _desc(**kw)
        assert lib.omg_geom_create(C.byref(d), C.byref(h)) == _lib.OMG_ERR_ARG, msg
        assert msg in _lib.last_error(), (msg, _lib.last_error())


def test_create_rejects_bad_tof():
    """TOF needs the Siddon projector on PET data, at most 64 bins, a positive sigma and the centres."""
    lib = _lib.load()
    h = C.c_void_p()
    c = np.zeros(65, np.float32)
    checks = [
        (dict(tof_bins=65, tof_sigma=10.0, tof_centers=c.ctypes.data), "exceeds 64"),
        (dict(tof_bins=3, tof_sigma=0.0, tof_centers=c.ctypes.data), "tof_sigma"),
        (dict(tof_bins=3, tof_sigma=10.0), "without tof_centers"),
        (dict(tof_bins=3, tof_sigma=10.0, tof_centers=c.ctypes.data, projector_type=2), "projector_type 1"),
    ]
    for kw, msg in checks:
        d, keep = _desc(**kw)
        assert lib.omg_geom_create(C.byref(d), C.byref(h)) == _lib.OMG_ERR_ARG, msg
        assert msg in _lib.last_error(), (msg, _lib.last_error())


def test_create_rejects_bad_multiray():
    """Several rays per LOR: Siddon on PET data (with or without TOF and attenuation), at most 256."""
    lib = _lib.load()
    h = C.c_void_p()
    checks = [
        (dict(rays_xy=17, rays_z=16), "exceeds 256"),
        (dict(rays_xy=2, projector_type=2), "improved Siddon"),
        (dict(rays_xy=2, cr_xy=float("inf")), "must be finite"),
    ]
    for kw, msg in checks:
        d, keep = _desc(**kw)
        assert lib.omg_geom_create(C.byref(d), C.byref(h)) == _lib.OMG_ERR_ARG, msg
        assert msg in _lib.last_error(), (msg, _lib.last_error())


def test_bdd_requires_ct():
    d, keep = _desc(projector_type=5)
    h = C.c_void_p()
    assert _lib.load().omg_geom_create(C.byref(d), C.byref(h)) == _lib.OMG_ERR_ARG
    assert "CT" in _lib.last_error()


def test_elementwise_argument_checks():
    lib = _lib.load()
    assert lib.omg_ratio(None, None, 10, 1e-5, None) == _lib.OMG_ERR_ARG
    assert lib.omg_rdp_gradient(None, None, 4, 4, 4, 2.0, 1e-5, 0.1, 0, None) == _lib.OMG_ERR_ARG
    assert lib.omg_pdhg_update(C.c_void_p(8), C.c_void_p(8), None, 10, 1e-5, 1.0, 0.1, 1, 0, None) == _lib.OMG_ERR_ARG


def test_unet_size_checks():
    """Zero or negative sizes are refused before any launch (an empty grid is not a valid launch)."""
    lib = _lib.load()
    p = C.c_void_p(256)        # never dereferenced: each call fails its argument check first
    assert lib.omg_mse(p, p, None, p, p, 0, C.c_float(1.0), None) == _lib.OMG_ERR_ARG
    assert lib.omg_bias_grad(p, p, 0, 16, None) == _lib.OMG_ERR_ARG
    assert lib.omg_pool_forward(p, p, 0, 2, 2, 16, None) == _lib.OMG_ERR_ARG
    assert lib.omg_upsample_forward(p, p, 2, 2, 2, 0, None) == _lib.OMG_ERR_ARG
    assert lib.omg_head_forward(p, p, p, p, p, 0, 16, None) == _lib.OMG_ERR_ARG
    assert lib.omg_embed(p, p, 4, 4, 4, 0, C.c_float(1.0), None) == _lib.OMG_ERR_ARG
    assert lib.omg_conv3d_forward(p, 16, None, 0, p, None, p, 16, 0, 4, 4, 0, None) == _lib.OMG_ERR_ARG


def test_create_rejects_bad_volume_table():
    """Projector 3 needs its table and 0 <= bmin < bmax, Vmax > 0; it runs on PET data."""
    lib = _lib.load()
    h = C.c_void_p()
    V = np.ones(8, np.float32)
    ok = dict(projector_type=3, orth_width=1.0, orth_flags=3, vol_V=V.ctypes.data, vol_len=V.size, vol_bmin=0.1,
              vol_bmax=2.0, vol_vmax=4.0)
    checks = [
        (dict(ok, vol_V=None, vol_len=0), "needs the volume table"),
        (dict(ok, vol_bmin=2.0), "vol_bmin < vol_bmax"),
        (dict(ok, vol_vmax=0.0), "vol_vmax > 0"),
        (dict(ok, vol_bmin=float("nan")), "vol_bmin < vol_bmax"),
    ]
    for kw, msg in checks:
        d, keep = _desc(**kw)
        assert lib.omg_geom_create(C.byref(d), C.byref(h)) == _lib.OMG_ERR_ARG, msg
        assert msg in _lib.last_error(), (msg, _lib.last_error())
