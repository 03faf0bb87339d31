"""Writes the golden fixtures in tests/golden/ from the reference's own CPU code.

Runs only where the reference sources are present: it compiles oracle/_ref (the unmodified
source/cpp/projector_functions.h and RDP.h of villekf/OMEGA behind a C shim, oracle/ref_driver.cpp)
and records inputs and outputs of
  * improved Siddon FP (float and double) and BP (float, with sensitivity) on an asymmetric
    8x7x5 grid with list-mode LORs covering every set-up branch;
  * one-hot backprojections (y = e_i) of those LORs: the exact voxel set and length*weight of each LOR;
  * Siddon FP on two planes of the 2-D plane PET geometry (128 x 160 bins, 128x128x63 grid),
    on a 3-D ring-pair geometry and on a cone-beam flat-panel geometry (inputs rebuilt by
    tests/cases.py; a checksum of the tables is stored);
  * the RDP gradient (zero boundary) on a 16x12x10 image.

    python tests/golden/make_golden.py
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from omega_amd import geometry as G  # noqa: E402
from tests import cases  # noqa: E402


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return np.frombuffer(h.digest(), dtype=np.uint8)


def main():
    assert O.build_ref(force=True), "reference sources not found: fixtures need /root/reference"
    rng = np.random.default_rng(1234)

    # --- tiny grid, list mode -------------------------------------------------------------------
    grid = cases.tiny_grid()
    x = cases.list_rays(grid, 400, seed=1)
    n = x.size // 6
    img = rng.random(grid.N).astype(np.float32) + 0.05
    yin = rng.random(n).astype(np.float32)
    c32 = O.SiddonCase(grid, 1, 1, x, np.zeros(2, np.float32), mode=O.MODE_LIST)
    c64 = O.SiddonCase(grid, 1, 1, x, np.zeros(2, np.float32), mode=O.MODE_LIST, dtype=np.float64)
    fp32 = c32.ref_fp(img, n, 1)
    fp64 = c64.ref_fp(img.astype(np.float64), n, 1)
    bp32, s32 = c32.ref_bp(yin, n, sens=True, nthreads=1)
    # one-hot BP per LOR: (lor, voxel, value) triplets
    lor_ids, vox, val = [], [], []
    for i in range(n):
        e = np.zeros(n, np.float32)
        e[i] = 1.0
        b = c32.ref_bp(e, n, nthreads=1)
        nz = np.nonzero(b)[0]
        lor_ids.append(np.full(nz.size, i, np.int32))
        vox.append(nz.astype(np.int32))
        val.append(b[nz])
    np.savez_compressed(os.path.join(HERE, "siddon_tiny_listmode.npz"),
                        grid=np.array([grid.Nx, grid.Ny, grid.Nz, grid.dx, grid.dy, grid.dz, grid.bx, grid.by, grid.bz],
                                      dtype=np.float64),
                        x=x, img=img, yin=yin, fp32=fp32, fp64=fp64, bp32=bp32, sens32=s32,
                        onehot_lor=np.concatenate(lor_ids), onehot_vox=np.concatenate(vox),
                        onehot_val=np.concatenate(val))

    # --- 2-D plane PET: two planes -----------------------------------------------------------------
    grid, sc = cases.c0()
    xy = sc.xy_table()
    per = sc.n_radial * sc.n_angles
    img = np.random.default_rng(77).random(grid.N).astype(np.float32) + 0.05   # rebuilt from the seed by the tests
    planes = [0, 37]
    out32, out64 = [], []
    for p in planes:
        c = O.SiddonCase(grid, sc.n_radial, sc.n_angles, xy, sc.z, plane_offset=p)
        out32.append(c.ref_fp(img, per))
        c = O.SiddonCase(grid, sc.n_radial, sc.n_angles, xy, sc.z, plane_offset=p, dtype=np.float64)
        out64.append(c.ref_fp(img.astype(np.float64), per))
    np.savez_compressed(os.path.join(HERE, "siddon_pet2d_planes.npz"), planes=np.array(planes), img_seed=77,
                        fp32=np.stack(out32), fp64=np.stack(out64), tables_sha256=digest(xy, sc.z))

    # --- 3-D ring pairs ----------------------------------------------------------------------------
    grid, sc = cases.c2_small()
    xy = sc.xy_table()
    per = sc.n_radial * sc.n_angles
    img = rng.random(grid.N).astype(np.float32) + 0.05
    nplanes = 6
    c = O.SiddonCase(grid, sc.n_radial, sc.n_angles, xy, sc.z, plane_offset=10)
    fp = c.ref_fp(img, nplanes * per)
    np.savez_compressed(os.path.join(HERE, "siddon_pet3d_rings.npz"), first=10, nplanes=nplanes, img=img, fp32=fp,
                        tables_sha256=digest(xy, sc.z))

    # --- cone-beam flat panel ----------------------------------------------------------------------
    grid, xct, uv, rows, cols, pitch = cases.ct_small()
    nv = xct.size // 6
    img = rng.random(grid.N).astype(np.float32)
    c = O.SiddonCase(grid, rows, cols, xct, uv, mode=O.MODE_CT, dPitchZ=pitch, n_proj=nv)
    fp = c.ref_fp(img, nv * rows * cols)
    np.savez_compressed(os.path.join(HERE, "siddon_ct_flat.npz"), img=img, fp32=fp, tables_sha256=digest(xct, uv))

    # --- RDP gradient ------------------------------------------------------------------------------
    shape = (16, 12, 10)
    u = (rng.random(int(np.prod(shape))) * 2.0 + 0.01).astype(np.float32)
    u[::37] = 0.0   # zero voxels exercise the epsilon guards
    grad = O.ref_rdp_gradient(u, shape, 2.0, 1e-5, 0.1)
    np.savez_compressed(os.path.join(HERE, "rdp_zero_boundary.npz"), shape=np.array(shape), u=u, gamma=2.0,
                        epps=1e-5, beta=0.1, grad32=grad)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
