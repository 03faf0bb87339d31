"""Writes tests/golden/scanner_geometry.npz: OMEGA's own PET scanner coordinates from the reference's Python
(detectorCoordinates, sinogramCoordinates2D and sinogramCoordinates3D,
source/Python/omegatomo/projector/detcoord.py:326-896).

Nothing else pins these tables (the reference's tests hold none), so this script imports that module
(numpy only) from the reference sources and records, for several block-ring scanners: the crystal
coordinates, the sinogram table [xs, ys, xd, yd] per (radial, angle) bin and the axial table [zs, zd]
per sinogram (span 1 and span > 1, with the reference's segment table; with ring gaps between axial
blocks; where the reference raises, as its span > 1 path does for span >= 5 and its pseudo-detector and
angular-mashing paths always do, the error is recorded). Runs only where the reference sources are
present:

    python tests/golden/make_scanner_geometry.py
"""
import os
import sys
import types

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, "/root/reference/source/Python")
sys.path.insert(0, ROOT)

from omega_amd import geometry as G  # noqa: E402

RINGS = [  # blocks_per_ring, cryst_per_block, cr_p, diameter, Ndist, ndist_side
    (16, 8, 4.0, 550.0, 64, 1),
    (18, 6, 3.2, 400.0, 40, -1),
    (24, 13, 4.0, 820.0, 101, 1),
    (38, 8, 2.4, 665.0, 128, 1),
]
AXIAL = [  # rings, cr_pz, span, ring_difference, image Nz
    (8, 4.0, 3, 7, 15),
    (13, 4.0, 5, 12, 25),
    (8, 4.0, 1, 7, 15),
    (10, 3.3, 3, 6, 19),
    (16, 4.0, 7, 15, 31),
]
GAPS = [  # rings, cr_pz, span, ring_difference, image Nz, cryst_per_block_axial, ring gaps (mm)
    (12, 4.0, 1, 11, 23, 4, [1.5, 2.25]),
    (12, 4.0, 3, 11, 23, 4, [1.5, 2.25]),
    (15, 3.3, 1, 14, 29, 5, [0.8, 0.8]),
    (15, 3.3, 3, 10, 29, 5, [2.0, 0.5]),
]


def main():
    from omegatomo.projector import detcoord
    out = {}
    for n, (bpr, cpb, cr_p, dia, nd, side) in enumerate(RINGS):
        det = bpr * cpb
        o = types.SimpleNamespace(cr_p=cr_p, diameter=dia, cryst_per_block=cpb, blocks_per_ring=bpr, DOI=0.0,
                                  transaxial_multip=1, det_w_pseudo=det, det_per_ring=det, nLayers=1,
                                  flip_image=False, offangle=0.0, Nang=det // 2, Ndist=nd, ndist_side=side)
        x, y = detcoord.detectorCoordinates(o)
        xs, ys = detcoord.sinogramCoordinates2D(o, x, y)
        out[f"ring{n}_x"], out[f"ring{n}_y"] = np.asarray(x, np.float32), np.asarray(y, np.float32)
        out[f"ring{n}_table"] = np.column_stack((xs[:, 0], ys[:, 0], xs[:, 1], ys[:, 1])).astype(np.float32)
    for n, (rings, cr_pz, span, rd, nz) in enumerate(AXIAL):
        seg = G.omega_segment_table(rings, span, rd, nz)
        tot = rings * rings if span == 1 else int(seg.sum())
        o = types.SimpleNamespace(cr_pz=cr_pz, rings=rings, ringGaps=np.empty(0), nLayers=1, span=span,
                                  ring_difference=rd, segment_table=seg, TotSinos=tot,
                                  cryst_per_block_axial=rings, linear_multip=1)
        out[f"axial{n}_seg"] = seg
        try:
            out[f"axial{n}_z"] = np.asarray(detcoord.sinogramCoordinates3D(o), np.float64)
        except ValueError as e:   # the reference's span > 1 path fails for span >= 5 (detcoord.py:791)
            out[f"axial{n}_error"] = np.array(str(e))
    for n, (rings, cr_pz, span, rd, nz, cpba, gaps) in enumerate(GAPS):
        seg = G.omega_segment_table(rings, span, rd, nz)
        tot = rings * rings if span == 1 else int(seg.sum())
        o = types.SimpleNamespace(cr_pz=cr_pz, rings=rings, ringGaps=np.asarray(gaps), nLayers=1, span=span,
                                  ring_difference=rd, segment_table=seg, TotSinos=tot,
                                  cryst_per_block_axial=cpba, linear_multip=rings // cpba)
        out[f"gaps{n}_z"] = np.asarray(detcoord.sinogramCoordinates3D(o), np.float64)
    # pseudo detectors and angular mashing: the reference's Python raises on both (recorded; the tests pin
    # these paths by invariants against the tables above)
    for name, (bpr, cpb, mash) in {"pseudo": (16, 8, 1), "mashing": (16, 8, 2)}.items():
        det = bpr * cpb
        detp = bpr * (cpb + 1) if name == "pseudo" else det
        o = types.SimpleNamespace(cr_p=4.0, diameter=550.0, cryst_per_block=cpb, blocks_per_ring=bpr, DOI=0.0,
                                  transaxial_multip=1, det_w_pseudo=detp, det_per_ring=det, nLayers=1,
                                  flip_image=False, offangle=0.0, Nang=detp // 2 // mash, Ndist=64, ndist_side=1)
        try:
            x, y = detcoord.detectorCoordinates(o)
            detcoord.sinogramCoordinates2D(o, x, y)
            out[f"{name}_error"] = np.array("")
        except Exception as e:  # noqa: BLE001 -- the reference's own failure is the record
            out[f"{name}_error"] = np.array(f"{type(e).__name__}: {e}")
    path = os.path.join(HERE, "scanner_geometry.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}: " + ", ".join(f"{k} {v.shape}" for k, v in out.items()))


if __name__ == "__main__":
    main()
