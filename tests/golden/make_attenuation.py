"""Writes tests/golden/siddon_attenuation.npz from the reference's own CPU code.

Runs only where the reference sources are present: oracle/_ref is the unmodified
source/cpp/projector_functions.h of villekf/OMEGA behind oracle/ref_driver.cpp, with
paramStruct::attenuationCorrection set and either CTAttenuation (an attenuation image, the sum
exp(-sum len * mu) along each LOR) or per-measurement factors. Records, single-threaded:
  * list-mode LORs on the asymmetric 8x7x5 grid (every Siddon set-up branch, LORs that start or end
    inside the image, axis-parallel rows): FP in float and double, BP with sensitivity in float, for
    an attenuation image and for per-measurement factors, and both together with a normalisation, a
    multiplicative scatter correction (paramStruct::scatterCoefs) and a global factor;
  * the 2-D plane PET case of the orthogonal tests (32 radial x 40 angles x 9 planes on a 32x32x9
    grid, angles 0 and 90 degrees included): FP and BP with an attenuation image;
  * the reduced 3-D ring-pair case: FP with an attenuation image on 16 planes.

    python tests/golden/make_attenuation.py
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
    rng = np.random.default_rng(77)
    out = {}

    grid = cases.tiny_grid()
    x = cases.list_rays(grid, 400, seed=3)
    n = x.size // 6
    mu = (rng.random(grid.N) * 0.05).astype(np.float32)
    img = (rng.random(grid.N) + 0.05).astype(np.float32)
    yin = rng.random(n).astype(np.float32)
    fac = (0.2 + 0.8 * rng.random(n)).astype(np.float32)
    z0 = np.zeros(2, np.float32)
    c32 = O.SiddonCase(grid, 1, 1, x, z0, mode=O.MODE_LIST, atten=mu, atten_mode=1)
    c64 = O.SiddonCase(grid, 1, 1, x, z0, mode=O.MODE_LIST, atten=mu.astype(np.float64), atten_mode=1,
                       dtype=np.float64)
    l32 = O.SiddonCase(grid, 1, 1, x, z0, mode=O.MODE_LIST, atten=fac, atten_mode=2)
    out.update(list_x=x, list_mu=mu, list_img=img, list_yin=yin, list_fac=fac,
               list_fp32=c32.ref_fp(img, n, 1), list_fp64=c64.ref_fp(img.astype(np.float64), n, 1))
    out["list_bp32"], out["list_sens32"] = c32.ref_bp(yin, n, sens=True, nthreads=1)
    out["list_lor_fp32"] = l32.ref_fp(img, n, 1)
    out["list_lor_bp32"], out["list_lor_sens32"] = l32.ref_bp(yin, n, sens=True, nthreads=1)
    # every per-LOR weight term at once: norm, per-LOR attenuation, multiplicative scatter, global factor
    norm = (0.5 + rng.random(n)).astype(np.float32)
    scat = (0.7 + 0.6 * rng.random(n)).astype(np.float32)
    a32 = O.SiddonCase(grid, 1, 1, x, z0, mode=O.MODE_LIST, atten=fac, atten_mode=2, norm=norm, scat=scat,
                       global_factor=0.9)
    i32 = O.SiddonCase(grid, 1, 1, x, z0, mode=O.MODE_LIST, atten=mu, atten_mode=1, norm=norm, scat=scat,
                       global_factor=0.9)
    out.update(list_norm=norm, list_scat=scat, list_all_fp32=a32.ref_fp(img, n, 1),
               list_img_all_fp32=i32.ref_fp(img, n, 1))
    out["list_all_bp32"] = a32.ref_bp(yin, n, nthreads=1)
    out["list_img_all_bp32"] = i32.ref_bp(yin, n, nthreads=1)

    g2, sc = cases.orth_pet2d()
    mu2 = (rng.random(g2.N) * 0.02).astype(np.float32)
    img2 = (rng.random(g2.N) + 0.05).astype(np.float32)
    n2 = sc.n_radial * sc.n_angles * sc.n_planes
    y2 = rng.random(n2).astype(np.float32)
    p32 = O.SiddonCase(g2, sc.n_radial, sc.n_angles, sc.xy_table(), sc.z, atten=mu2, atten_mode=1)
    out.update(pet2d_mu=mu2, pet2d_img=img2, pet2d_yin=y2, pet2d_fp32=p32.ref_fp(img2, n2, 1))
    out["pet2d_bp32"], out["pet2d_sens32"] = p32.ref_bp(y2, n2, sens=True, nthreads=1)

    g3, sc3 = cases.c2_small()
    planes = 16
    mu3 = (rng.random(g3.N) * 0.02).astype(np.float32)
    img3 = (rng.random(g3.N) + 0.05).astype(np.float32)
    n3 = sc3.n_radial * sc3.n_angles * planes
    q32 = O.SiddonCase(g3, sc3.n_radial, sc3.n_angles, sc3.xy_table(), sc3.z[: 2 * planes], atten=mu3, atten_mode=1)
    out.update(pet3d_mu=mu3, pet3d_img=img3, pet3d_planes=np.int64(planes), pet3d_fp32=q32.ref_fp(img3, n3, 1))

    np.savez_compressed(os.path.join(HERE, "siddon_attenuation.npz"), **out)
    print({k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
