"""Writes tests/golden/subset_indices.npz: the measurement index vectors of the reference's subset
types, from the reference's own Python (indexMaker, source/Python/omegatomo/projector/indices.py:20-254).

Nothing else pins these conventions (the reference's tests hold no vectors for them), so this script
imports that one file by path -- it needs numpy only -- and records, for small sinograms
(Ndist x Nang x NSinos) and several subset counts, the concatenated index vector and the
per-subset lengths of subset types 1, 2, 3 (seeded), 4 and 5 (index into the sinogram, radial bin
fastest, then angle, then plane), of type 8 (plane numbers), and, for several projection counts
(primes and composites), of the projection-level types 9 (seeded permutation) and 11 (prime-factor
order). Where the reference raises (subset
type 4 with Nang % subsets != 0 fails in its remainder branch, indices.py:58-66, with a numpy
broadcasting error), the case is recorded as an error. Runs only where the reference sources are
present:

    python tests/golden/make_subset_indices.py
"""
import importlib.util
import os
import types

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/source/Python/omegatomo/projector/indices.py"

CASES = [
    # (Ndist, Nang, NSinos, subsets)
    (8, 6, 5, 3),
    (8, 7, 4, 3),      # Nang % subsets != 0 (subset type 4's remainder branch)
    (6, 10, 3, 4),
    (5, 9, 7, 2),
]
TYPES = [1, 2, 3, 4, 5, 8]
# projection-level types 9 (seeded permutation) and 11 (prime-factor order): (nProjections, subsets)
PROJ_CASES = [(12, 3), (18, 4), (30, 7), (7, 2), (20, 6), (63, 8)]


def main():
    spec = importlib.util.spec_from_file_location("omega_ref_indices", REF)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    out = {}
    for (nd, na, ns, sub) in CASES:
        for st in TYPES:
            o = types.SimpleNamespace(subsets=sub, sampling=1, Ndist=nd, Nang=na, NSinos=ns, CT=False,
                                      subsetType=st, use_raw_data=0, listmode=0, Nt=1, seed=7,
                                      nProjections=ns, PET=True, SPECT=False)
            key = f"t{st}_{nd}x{na}x{ns}_s{sub}"
            try:
                mod.indexMaker(o)
            except ValueError as e:
                out[key + "_error"] = np.array(str(e))
                continue
            out[key + "_index"] = np.asarray(o.index).astype(np.int64).ravel()
            out[key + "_nmeas"] = np.asarray(o.nMeas).astype(np.int64).ravel()
    for (npj, sub) in PROJ_CASES:
        for st in (9, 11):
            o = types.SimpleNamespace(subsets=sub, sampling=1, Ndist=4, Nang=3, NSinos=npj, CT=False,
                                      subsetType=st, use_raw_data=0, listmode=0, Nt=1, seed=11,
                                      nProjections=npj, PET=True, SPECT=False)
            key = f"p{st}_{npj}_s{sub}"
            mod.indexMaker(o)
            out[key + "_index"] = np.asarray(o.index).astype(np.int64).ravel()
            out[key + "_nmeas"] = np.asarray(o.nMeas).astype(np.int64).ravel()
    path = os.path.join(HERE, "subset_indices.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}: {sum(k.endswith('_index') for k in out)} index sets, "
          f"{sum(k.endswith('_error') for k in out)} reference errors")


if __name__ == "__main__":
    main()
