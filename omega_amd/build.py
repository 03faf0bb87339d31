"""Builds libomega_hip.so from omega_amd/csrc/*.hip with hipcc for gfx950 (in-tree), then the three
omegaMain libraries (omega_amd/csrc/main/omega_main.hip, one per measurement type) linked against it.

Each translation unit is compiled to an object in omega_amd/_build/ (in parallel), then linked.
"""
from __future__ import annotations

import glob
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libomega_hip.so")
INCLUDE = os.path.join(HERE, "..", "include")

# -ffp-contract=off + IEEE div/sqrt: Siddon index lists and forward projections are bit-identical
#   to the reference CPU build;
# -munsafe-fp-atomics: float atomicAdd lowers to global_atomic_add_f32 (no CAS loop);
# -packed-fp32-ops disabled: no v_pk_*_f32 emitted.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


# omegaMain variants: library name -> measurement element type (include/omega_main.h)
MAIN_SRC = os.path.join(CSRC, "main", "omega_main.hip")
MAIN_LIBS = {
    "libomega_matrixfree.so": "float",
    "libomega_matrixfree_uint16.so": "uint16_t",
    "libomega_matrixfree_uint8.so": "uint8_t",
}


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found; the HIP library cannot be built")


def sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")))


def _headers_mtime() -> float:
    hs = glob.glob(os.path.join(CSRC, "*.hpp")) + glob.glob(os.path.join(INCLUDE, "*.h"))
    return max((os.path.getmtime(h) for h in hs), default=0.0)


def _compile(src: str, incremental: bool, verbose: bool):
    obj = os.path.join(OBJ, os.path.basename(src)[:-4] + ".o")
    if incremental and os.path.exists(obj) and os.path.getmtime(obj) >= max(os.path.getmtime(src), _headers_mtime()):
        return obj, None
    cmd = [hipcc()] + HIPCC_FLAGS + ["-I", INCLUDE, "-c", src, "-o", obj + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        return obj, r.stdout + r.stderr
    os.replace(obj + ".tmp", obj)
    return obj, None


def _build_main(name: str, meas_type: str, verbose: bool):
    """Compile and link one omegaMain library; returns an error text or None."""
    out = os.path.join(HERE, name)
    obj = os.path.join(OBJ, name[:-3] + ".o")
    cmd = [hipcc()] + HIPCC_FLAGS + ["-I", INCLUDE, f"-DOMG_MAIN_MEAS_TYPE={meas_type}", "-c", MAIN_SRC, "-o", obj]
    link = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", obj, "-L", HERE, "-lomega_hip",
            "-Wl,-rpath,$ORIGIN", "-o", out + ".tmp"]
    for c in (cmd, link):
        if verbose:
            print(" ".join(c), flush=True)
        r = subprocess.run(c, capture_output=True, text=True)
        if r.returncode != 0:
            return r.stdout + r.stderr
    os.replace(out + ".tmp", out)
    return None


def build(verbose: bool = False, incremental: bool = False, jobs: int = 0) -> str:
    """Compile every HIP source and link omega_amd/libomega_hip.so. By default every source is
    recompiled; incremental=True reuses objects newer than their source and the headers."""
    os.makedirs(OBJ, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(lambda s: _compile(s, incremental, verbose), sources()))
    errors = [e for _, e in results if e]
    if errors:
        sys.stderr.write("\n".join(errors))
        raise RuntimeError("hipcc failed building libomega_hip.so")
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC"] + [o for o, _ in results] + ["-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("linking libomega_hip.so failed")
    os.replace(tmp, LIB)
    with ThreadPoolExecutor(max_workers=min(jobs, len(MAIN_LIBS))) as ex:
        errors = [e for e in ex.map(lambda kv: _build_main(kv[0], kv[1], verbose), MAIN_LIBS.items()) if e]
    if errors:
        sys.stderr.write("\n".join(errors))
        raise RuntimeError("hipcc failed building the omegaMain libraries")
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, incremental="-i" in sys.argv))
