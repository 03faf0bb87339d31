"""Writes tests/golden/siddon_tof.npz from the reference's own CPU code.

Runs only where the reference sources are present: oracle/_ref is the unmodified
source/cpp/projector_functions.h of villekf/OMEGA behind oracle/ref_driver.cpp, with paramStruct::TOF,
nBins, sigma_x, TOFCenters and epps set. Bin centres and sigma follow projectorClass
(omega_amd.geometry.tof_centers / tof_sigma). Records, single-threaded:
  * list-mode LORs on the asymmetric 8x7x5 grid (every Siddon set-up branch, LORs that start or end
    inside the image, axis-parallel rows), 5 bins: FP in float and double, BP with sensitivity in
    float; FP and BP with normalisation, per-LOR attenuation, multiplicative scatter and a global factor;
  * 2-D plane sinograms (32 radial x 40 angles, 3 planes of the 32x32x9 grid), 7 bins: FP and BP with
    sensitivity;
  * the reduced 3-D ring-pair case, 4 planes, 5 bins: FP.

    python tests/golden/make_tof.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from omega_amd import geometry as G  # noqa: E402
from tests import cases  # noqa: E402

EPPS = 1e-5


def main():
    assert O.build_ref(force=True), "reference sources not found: fixtures need /root/reference"
    rng = np.random.default_rng(91)
    out = {}

    grid = cases.tiny_grid()
    x = cases.list_rays(grid, 400, seed=4)
    n = x.size // 6
    nb = 5
    tc = G.tof_centers(nb, 20e-12)
    sig = np.float32(G.tof_sigma(40e-12))
    img = (rng.random(grid.N) + 0.05).astype(np.float32)
    yin = rng.random(n * nb).astype(np.float32)
    z0 = np.zeros(2, np.float32)
    kw = dict(mode=O.MODE_LIST, tof_bins=nb, tof_sigma=sig, tof_centers=tc, tof_epps=EPPS)
    c32 = O.SiddonCase(grid, 1, 1, x, z0, **kw)
    c64 = O.SiddonCase(grid, 1, 1, x, z0, dtype=np.float64, **kw)
    out.update(list_x=x, list_img=img, list_yin=yin, list_centers=tc, list_sigma=sig, list_bins=np.int64(nb),
               list_fp32=c32.ref_fp(img, n, 1), list_fp64=c64.ref_fp(img.astype(np.float64), n, 1))
    out["list_bp32"], out["list_sens32"] = c32.ref_bp(yin, n, sens=True, nthreads=1)
    norm = (0.5 + rng.random(n)).astype(np.float32)
    fac = (0.2 + 0.8 * rng.random(n)).astype(np.float32)
    scat = (0.7 + 0.6 * rng.random(n)).astype(np.float32)
    a32 = O.SiddonCase(grid, 1, 1, x, z0, norm=norm, atten=fac, atten_mode=2, scat=scat, global_factor=0.9, **kw)
    out.update(list_norm=norm, list_fac=fac, list_scat=scat, list_all_fp32=a32.ref_fp(img, n, 1),
               list_all_bp32=a32.ref_bp(yin, n, nthreads=1))

    g2, sc = cases.orth_pet2d()
    nb2, planes = 7, 3
    tc2 = G.tof_centers(nb2, 100e-12)
    sig2 = np.float32(G.tof_sigma(300e-12))
    n2 = sc.n_radial * sc.n_angles * planes
    img2 = (rng.random(g2.N) + 0.05).astype(np.float32)
    y2 = rng.random(n2 * nb2).astype(np.float32)
    p32 = O.SiddonCase(g2, sc.n_radial, sc.n_angles, sc.xy_table(), sc.z[: 2 * planes], tof_bins=nb2, tof_sigma=sig2,
                       tof_centers=tc2, tof_epps=EPPS)
    out.update(pet2d_img=img2, pet2d_yin=y2, pet2d_centers=tc2, pet2d_sigma=sig2, pet2d_bins=np.int64(nb2),
               pet2d_planes=np.int64(planes), pet2d_fp32=p32.ref_fp(img2, n2, 1))
    out["pet2d_bp32"], out["pet2d_sens32"] = p32.ref_bp(y2, n2, sens=True, nthreads=1)

    g3, sc3 = cases.c2_small()
    nb3, planes3 = 5, 4
    tc3 = G.tof_centers(nb3, 80e-12)
    sig3 = np.float32(G.tof_sigma(250e-12))
    n3 = sc3.n_radial * sc3.n_angles * planes3
    img3 = (rng.random(g3.N) + 0.05).astype(np.float32)
    q32 = O.SiddonCase(g3, sc3.n_radial, sc3.n_angles, sc3.xy_table(), sc3.z[: 2 * planes3], tof_bins=nb3,
                       tof_sigma=sig3, tof_centers=tc3, tof_epps=EPPS)
    out.update(pet3d_img=img3, pet3d_centers=tc3, pet3d_sigma=sig3, pet3d_bins=np.int64(nb3),
               pet3d_planes=np.int64(planes3), pet3d_fp32=q32.ref_fp(img3, n3, 1))

    np.savez_compressed(os.path.join(HERE, "siddon_tof.npz"), **out)
    print({k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
