"""Python side of the built-in reconstruction entry point ``omegaMain`` (include/omega_main.h).

``parameters`` is the packed ctypes mirror of ``inputStruct``, built from the header itself so that the
C and Python views of the block cannot drift apart; it plays the role of the reference front end's
``projectorClass.parameters`` (source/Python/omegatomo/projector/proj.py:1959-2282), and a caller fills it
member by member as ``transferData`` does (source/Python/omegatomo/reconstruction/recomain.py:23-375).
``library(dtype)`` loads the variant for the measurement type, as recomain.py:652-700 picks
``*_matrixfree_lib`` / ``*_uint16_lib`` / ``*_uint8_lib``.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = os.path.join(HERE, "..", "include", "omega_main.h")
LIBS = {
    np.dtype(np.float32): os.path.join(HERE, "libomega_matrixfree.so"),
    np.dtype(np.uint16): os.path.join(HERE, "libomega_matrixfree_uint16.so"),
    np.dtype(np.uint8): os.path.join(HERE, "libomega_matrixfree_uint8.so"),
}

_CTYPES = {
    "uint8_t": C.c_uint8, "int8_t": C.c_int8, "uint16_t": C.c_uint16, "int16_t": C.c_int16,
    "uint32_t": C.c_uint32, "int32_t": C.c_int32, "uint64_t": C.c_uint64, "int64_t": C.c_int64,
    "float": C.c_float, "double": C.c_double, "bool": C.c_bool,
}


def header_fields(path: str = HEADER):
    """(name, ctypes type) of every member of inputStruct, in declaration order, parsed from the header."""
    text = open(path).read()
    body = re.search(r"typedef struct inputStruct \{(.*?)\} inputStruct;", text, re.S)
    if body is None:
        raise RuntimeError(f"inputStruct not found in {path}")
    src = re.sub(r"/\*.*?\*/", " ", body.group(1), flags=re.S)
    fields = []
    for stmt in src.split(";"):
        stmt = " ".join(stmt.split())
        if not stmt:
            continue
        m = re.match(r"(\w+)\s*(.*)", stmt)
        base, rest = m.group(1), m.group(2)
        if base not in _CTYPES:
            raise RuntimeError(f"unknown member type {base!r} in inputStruct")
        for decl in rest.split(","):
            decl = decl.strip()
            ptr = decl.startswith("*")   # `float* x` and `float *x, *z` both leave '*' on the declarator
            name = decl.lstrip("* ").strip()
            if not re.fullmatch(r"\w+", name):
                raise RuntimeError(f"cannot parse declarator {decl!r} in inputStruct")
            t = _CTYPES[base]
            fields.append((name, C.POINTER(t) if ptr else t))
    return fields


class parameters(C.Structure):
    """The packed parameter block of omegaMain (one byte alignment)."""
    _pack_ = 1
    _fields_ = header_fields()


_loaded = {}


def library(dtype=np.float32):
    """The omegaMain library for measurement data of ``dtype`` (float32, uint16 or uint8), loaded once."""
    dt = np.dtype(dtype)
    if dt not in LIBS:
        raise ValueError(f"omegaMain takes float32, uint16 or uint8 measurements, not {dt}")
    if dt not in _loaded:
        path = LIBS[dt]
        if not os.path.exists(path):
            raise RuntimeError(f"{os.path.basename(path)} is not built (run omega_amd/build.py)")
        lib = C.CDLL(path)
        elem = {np.dtype(np.float32): C.c_float, np.dtype(np.uint16): C.c_uint16, np.dtype(np.uint8): C.c_uint8}[dt]
        lib.omegaMain.argtypes = [parameters, C.c_char_p, C.POINTER(elem), C.POINTER(C.c_float),
                                  C.POINTER(C.c_float), C.POINTER(C.c_float)]
        lib.omegaMain.restype = C.c_int
        _loaded[dt] = lib
    return _loaded[dt]
