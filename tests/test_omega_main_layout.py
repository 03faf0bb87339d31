"""omegaMain's binary interface, without a GPU.

* the parameter block of include/omega_main.h has the members, order, types and one-byte packing of the
  block the reference's Python front end packs (tests/golden/inputstruct_layout.json, written from
  proj.py's ``parameters`` by make_inputstruct_layout.py);
* the three libraries load and export ``omegaMain``;
* configurations the library does not provide are refused with a message and a nonzero status before
  any device work (so these calls run on a machine without a GPU).
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

from omega_amd import omega_main as M
from tests import omega_main_pack as P

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "inputstruct_layout.json")


def _ctype(name):
    """ctypes type of a fixture type name ('c_uint32', '*c_float', ...)."""
    return C.POINTER(getattr(C, name[1:])) if name.startswith("*") else getattr(C, name)


def test_block_matches_front_end_layout():
    ref = json.load(open(GOLDEN))
    assert ref["pack"] == 1
    mine = M.parameters._fields_
    assert len(mine) == len(ref["members"]) == 320
    for (n, t), (rn, rt) in zip(mine, ref["members"]):
        assert n == rn and t is _ctype(rt), f"member {n} ({t}) differs from the front end's {rn} ({rt})"

    class Front(C.Structure):   # the front end's block rebuilt from the fixture
        _pack_ = 1
        _fields_ = [(n, _ctype(t)) for n, t in ref["members"]]

    assert C.sizeof(Front) == C.sizeof(M.parameters) == 1346
    for n, _ in M.parameters._fields_:
        assert getattr(Front, n).offset == getattr(M.parameters, n).offset


@pytest.mark.parametrize("dtype", [np.float32, np.uint16, np.uint8])
def test_libraries_export_omegaMain(dtype):
    lib = M.library(dtype)
    assert hasattr(lib, "omegaMain")


def _refused(capfd, sino_dtype=np.float32, **kw):
    o = P.Options(OSEM=True, Niter=1, x=np.zeros(8, np.float32), sizeX=8, Nx=[4], Ny=[4], Nz=[1],
                  dx=[1.0], dy=[1.0], dz=[1.0], bx=[-2.0], by=[-2.0], bz=[-0.5], pituus=[0, 1],
                  x0=np.ones(16, np.float32), nRowsD=2, nColsD=1, nProjections=1)
    for k, v in kw.items():
        setattr(o, k, v)
    rc, *_ = P.run(o, np.ones(2, sino_dtype), 1)
    out = capfd.readouterr().out
    assert rc != 0
    assert "Reconstruction failed!" in out
    return out


@pytest.mark.parametrize("flag,text", [
    ("NLM", "NLM prior"), ("TV", "TV prior"), ("FISTA", "FISTA algorithm"), ("ACOSEM", "ACOSEM algorithm"),
    ("LSQR", "LSQR algorithm"), ("use_psf", "PSF"), ("SPECT", "SPECT"), ("largeDim", "largeDim"),
    ("useMaskFP", "masks"), ("RDPIncludeCorners", "corners"), ("stochasticSubsetSelection", "stochastic"),
])
def test_unsupported_switches_are_refused(capfd, flag, text):
    out = _refused(capfd, **{flag: True})
    assert text in out


def test_refusals_name_the_reason(capfd):
    assert "exactly one reconstruction algorithm" in _refused(capfd, OSEM=True, PKMA=True)
    assert "exactly one reconstruction algorithm" in _refused(capfd, OSEM=False)
    assert "dynamic data" in _refused(capfd, Nt=2)
    assert "multi-volume" in _refused(capfd, nMultiVolumes=1)
    assert "preconditioner type 5" in _refused(capfd, MBSREM=True, OSEM=False,
                                                precondTypeImage=[0, 0, 0, 0, 0, 1, 0])
    assert "measurement-based" in _refused(capfd, precondTypeMeas=[0, 1])
    assert "RDP prior is used with" in _refused(capfd, RDP=True)
    assert "CGLS runs without subsets" in _refused(capfd, OSEM=False, CGLS=True, subsets=2, pituus=[0, 1, 1])
    assert "adaptive PDHG" in _refused(capfd, OSEM=False, PDHG=True, PDAdaptiveType=1)
    assert "randoms correction of CT" in _refused(capfd, CT=True, randoms_correction=1)
