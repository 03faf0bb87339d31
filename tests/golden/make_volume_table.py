"""Writes tests/golden/volume_table.npz: the look-up table of the reference's volume-of-intersection
projector (computeVoxelVolumes, source/Python/omegatomo/projector/coordinates.py:133-207).

Nothing else pins this table (the reference's tests hold none), so this script imports that module
(numpy + scipy) from the reference sources and records V, bmin, bmax and Vmax for a few tube radii,
voxel radius scales and voxel sizes. Runs only where the reference sources are present:

    python tests/golden/make_volume_table.py
"""
import importlib.util
import os
import types
import warnings

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/source/Python/omegatomo/projector/coordinates.py"
CASES = [  # tube_radius, voxel_radius, (dx, dy, dz)
    (2.0, 1.0, (2.5, 2.5, 2.5)),
    (1.2, 1.0, (1.0, 1.0, 2.0)),
    (3.5, 0.8, (2.0, 2.0, 2.0)),
]


def main():
    spec = importlib.util.spec_from_file_location("omega_ref_coordinates", REF)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    out = {}
    for n, (tr, vr, d) in enumerate(CASES):
        o = types.SimpleNamespace(projector_type=3, tube_radius=tr, voxel_radius=vr, dx=np.array([d[0]]),
                                  dy=np.array([d[1]]), dz=np.array([d[2]]))
        with warnings.catch_warnings(), np.errstate(all="ignore"):
            warnings.simplefilter("ignore")
            mod.computeVoxelVolumes(o)
        out[f"c{n}_V"] = np.asarray(o.V, np.float32)
        out[f"c{n}_scalars"] = np.array([o.bmin, o.bmax, o.Vmax, o.voxel_radius], np.float64)
    path = os.path.join(HERE, "volume_table.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}: " + ", ".join(f"{k} {v.shape}" for k, v in out.items()))


if __name__ == "__main__":
    main()
