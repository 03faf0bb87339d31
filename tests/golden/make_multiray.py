"""Writes tests/golden/siddon_multiray.npz from the reference's own CPU code.

Runs only where the reference sources are present: oracle/_ref is the unmodified
source/cpp/projector_functions.h of villekf/OMEGA behind oracle/ref_driver.cpp, with paramStruct
nRays2D / nRays3D and dPitchXY / dPitchZ set. Records, single-threaded:
  * list-mode LORs on the asymmetric 8x7x5 grid (every Siddon set-up branch), 3 transaxial x 2 axial
    rays over 0.9 / 1.1 mm: FP in float and double, BP with sensitivity in float, and FP / BP with
    normalisation, per-LOR attenuation, multiplicative scatter and a global factor;
  * 2-D plane sinograms (32 radial x 40 angles, 3 planes of the 32x32x9 grid), 5 transaxial rays over
    2.5 mm: FP and BP with sensitivity;
  * the reduced 3-D ring-pair case, 4 planes, 2 x 3 rays over 2.0 / 4.0 mm: FP.

    python tests/golden/make_multiray.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from tests import cases  # noqa: E402


def main():
    assert O.build_ref(force=True), "reference sources not found: fixtures need /root/reference"
    rng = np.random.default_rng(93)
    out = {}
    grid = cases.tiny_grid()
    x = cases.list_rays(grid, 400, seed=5)
    n = x.size // 6
    img = (rng.random(grid.N) + 0.05).astype(np.float32)
    yin = rng.random(n).astype(np.float32)
    z0 = np.zeros(2, np.float32)
    kw = dict(mode=O.MODE_LIST, rays=(3, 2), ray_pitch=(0.9, 1.1))
    c32 = O.SiddonCase(grid, 1, 1, x, z0, **kw)
    c64 = O.SiddonCase(grid, 1, 1, x, z0, dtype=np.float64, **kw)
    out.update(list_x=x, list_img=img, list_yin=yin, list_fp32=c32.ref_fp(img, n, 1),
               list_fp64=c64.ref_fp(img.astype(np.float64), n, 1))
    out["list_bp32"], out["list_sens32"] = c32.ref_bp(yin, n, sens=True, nthreads=1)
    norm = (0.5 + rng.random(n)).astype(np.float32)
    fac = (0.2 + 0.8 * rng.random(n)).astype(np.float32)
    scat = (0.7 + 0.6 * rng.random(n)).astype(np.float32)
    a32 = O.SiddonCase(grid, 1, 1, x, z0, norm=norm, atten=fac, atten_mode=2, scat=scat, global_factor=0.9, **kw)
    out.update(list_norm=norm, list_fac=fac, list_scat=scat, list_all_fp32=a32.ref_fp(img, n, 1),
               list_all_bp32=a32.ref_bp(yin, n, nthreads=1))

    g2, sc = cases.orth_pet2d()
    planes = 3
    n2 = sc.n_radial * sc.n_angles * planes
    img2 = (rng.random(g2.N) + 0.05).astype(np.float32)
    y2 = rng.random(n2).astype(np.float32)
    p32 = O.SiddonCase(g2, sc.n_radial, sc.n_angles, sc.xy_table(), sc.z[: 2 * planes], rays=(5, 1),
                       ray_pitch=(2.5, 0.0))
    out.update(pet2d_img=img2, pet2d_yin=y2, pet2d_planes=np.int64(planes), pet2d_fp32=p32.ref_fp(img2, n2, 1))
    out["pet2d_bp32"], out["pet2d_sens32"] = p32.ref_bp(y2, n2, sens=True, nthreads=1)

    g3, sc3 = cases.c2_small()
    planes3 = 4
    n3 = sc3.n_radial * sc3.n_angles * planes3
    img3 = (rng.random(g3.N) + 0.05).astype(np.float32)
    q32 = O.SiddonCase(g3, sc3.n_radial, sc3.n_angles, sc3.xy_table(), sc3.z[: 2 * planes3], rays=(2, 3),
                       ray_pitch=(2.0, 4.0))
    out.update(pet3d_img=img3, pet3d_planes=np.int64(planes3), pet3d_fp32=q32.ref_fp(img3, n3, 1))
    np.savez_compressed(os.path.join(HERE, "siddon_multiray.npz"), **out)
    print({k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
