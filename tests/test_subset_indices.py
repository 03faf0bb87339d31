"""Subset index vectors (subset types 1-5, 8, 9 and 11) against the reference's own indexMaker output
(tests/golden/subset_indices.npz, written by tests/golden/make_subset_indices.py), including the
cases where the reference raises; plus the partition property every measurement-level subset type
has (disjoint subsets covering every bin, type 3 less the one bin its last slice drops); the
projection-level types are permutations of the projections."""
import os

import numpy as np
import pytest

from omega_amd import geometry as G

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "subset_indices.npz"))
CASES = [(8, 6, 5, 3), (8, 7, 4, 3), (6, 10, 3, 4), (5, 9, 7, 2)]


@pytest.mark.parametrize("st", [1, 2, 3, 4, 5, 8])
@pytest.mark.parametrize("case", CASES)
def test_matches_reference(st, case):
    nd, na, ns, sub = case
    key = f"t{st}_{nd}x{na}x{ns}_s{sub}"
    if key + "_error" in GOLD.files:
        with pytest.raises(ValueError):
            G.subset_index_measurements(st, sub, nd, na, ns, seed=7)
        return
    if st == 8:
        idx, cnt = G.subset_index_type8(ns, sub)
    else:
        idx, cnt = G.subset_index_measurements(st, sub, nd, na, ns, seed=7)
    assert np.array_equal(idx, GOLD[key + "_index"])
    assert np.array_equal(cnt, GOLD[key + "_nmeas"])


@pytest.mark.parametrize("st", [1, 2, 3, 4, 5])
def test_partition(st):
    nd, na, ns, sub = 12, 8, 6, 4
    idx, cnt = G.subset_index_measurements(st, sub, nd, na, ns, seed=3)
    total = nd * na * ns
    assert cnt.sum() == idx.size
    u = np.unique(idx)
    assert u.size == idx.size                            # disjoint
    assert u.size == (total - 1 if st == 3 else total)   # covering
    assert u.min() >= 0 and u.max() < total


PROJ_CASES = [(12, 3), (18, 4), (30, 7), (7, 2), (20, 6), (63, 8)]


@pytest.mark.parametrize("st", [8, 9, 11])
@pytest.mark.parametrize("case", PROJ_CASES)
def test_projection_level_matches_reference(st, case):
    npj, sub = case
    idx, cnt = G.subset_index_projections(st, npj, sub, seed=11)
    assert np.array_equal(np.sort(idx), np.arange(npj))           # a permutation of the projections
    assert cnt.sum() == npj
    if st == 8:
        return   # type 8 is pinned on the sinogram cases above
    key = f"p{st}_{npj}_s{sub}"
    assert np.array_equal(idx, GOLD[key + "_index"])
    assert np.array_equal(cnt, GOLD[key + "_nmeas"])


def test_type_10_is_refused():
    with pytest.raises(ValueError):
        G.subset_index_projections(10, 12, 3)
