"""Writes tests/golden/siddon_combos.npz from the reference's own CPU code: several rays per LOR
together with time of flight and with an attenuation image.

Runs only where the reference sources are present: oracle/_ref is the unmodified
source/cpp/projector_functions.h of villekf/OMEGA behind oracle/ref_driver.cpp, with paramStruct
nRays2D / nRays3D, dPitchXY / dPitchZ, TOF (nBins, sigma_x, TOFCenters) and CTAttenuation set. Records,
single-threaded:
  * list-mode LORs on the asymmetric 8x7x5 grid, 3 transaxial x 2 axial rays over 0.9 / 1.1 mm, with
    5 TOF bins; with an attenuation image; with both: FP in float and double, BP with sensitivity in
    float;
  * 2-D plane sinograms (32 radial x 40 angles, 3 planes of the 32x32x9 grid), 3 transaxial rays over
    2.5 mm with 7 TOF bins and an attenuation image: FP and BP with sensitivity in float.

    python tests/golden/make_combos.py
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
    rng = np.random.default_rng(211)
    out = {}
    grid = cases.tiny_grid()
    x = cases.list_rays(grid, 300, seed=7)
    n = x.size // 6
    img = (rng.random(grid.N) + 0.05).astype(np.float32)
    mu = (0.01 * rng.random(grid.N)).astype(np.float32)
    tofc = np.linspace(-30.0, 30.0, 5).astype(np.float32)
    out.update(list_x=x, list_img=img, list_mu=mu, list_tofc=tofc)
    z0 = np.zeros(2, np.float32)
    variants = {"tof": dict(tof_bins=5, tof_sigma=np.float32(12.0), tof_centers=tofc, tof_epps=1e-5),
                "atn": dict(atten=mu, atten_mode=1)}
    variants["tofatn"] = dict(variants["tof"], **variants["atn"])
    for name, kw in variants.items():
        nb = kw.get("tof_bins", 1)
        yin = rng.random(n * nb).astype(np.float32)
        c32 = O.SiddonCase(grid, 1, 1, x, z0, mode=O.MODE_LIST, rays=(3, 2), ray_pitch=(0.9, 1.1), **kw)
        c64 = O.SiddonCase(grid, 1, 1, x, z0, mode=O.MODE_LIST, rays=(3, 2), ray_pitch=(0.9, 1.1),
                           dtype=np.float64, **kw)
        out[f"list_{name}_yin"] = yin
        out[f"list_{name}_fp32"] = c32.ref_fp(img, n, 1)
        out[f"list_{name}_fp64"] = c64.ref_fp(img.astype(np.float64), n, 1)
        out[f"list_{name}_bp32"], out[f"list_{name}_sens32"] = c32.ref_bp(yin, n, sens=True, nthreads=1)

    g2, sc = cases.orth_pet2d()
    planes = 3
    n2 = sc.n_radial * sc.n_angles * planes
    img2 = (rng.random(g2.N) + 0.05).astype(np.float32)
    mu2 = (0.01 * rng.random(g2.N)).astype(np.float32)
    tofc2 = np.linspace(-45.0, 45.0, 7).astype(np.float32)
    y2 = rng.random(n2 * 7).astype(np.float32)
    p32 = O.SiddonCase(g2, sc.n_radial, sc.n_angles, sc.xy_table(), sc.z[: 2 * planes], rays=(3, 1),
                       ray_pitch=(2.5, 0.0), tof_bins=7, tof_sigma=np.float32(15.0), tof_centers=tofc2, tof_epps=1e-5,
                       atten=mu2, atten_mode=1)
    out.update(pet2d_img=img2, pet2d_mu=mu2, pet2d_tofc=tofc2, pet2d_yin=y2, pet2d_planes=np.int64(planes),
               pet2d_fp32=p32.ref_fp(img2, n2, 1))
    out["pet2d_bp32"], out["pet2d_sens32"] = p32.ref_bp(y2, n2, sens=True, nthreads=1)
    np.savez_compressed(os.path.join(HERE, "siddon_combos.npz"), **out)
    print({k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
