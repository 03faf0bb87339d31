"""The list-mode sensitivity image's LOR tables (geometry.listmode_sensitivity_tables) follow the reference's
-DSENS work-item layout: item (i, j, k) is the LOR from crystal i in ring k / rings to crystal j in ring
k % rings (projectorType123.cl:283-287, getDetectorCoordinatesRaw general_opencl_functions.h:1538-1543),
and its measurement index is i + j * det + k * det^2 (:287)."""
import numpy as np

from omega_amd import geometry as G


def test_layout_matches_work_items():
    rng = np.random.default_rng(3)
    det, rings = 7, 3
    c = rng.normal(size=(det, 2)).astype(np.float32) * 100
    zr = np.sort(rng.normal(size=rings)).astype(np.float32) * 10
    x, z = G.listmode_sensitivity_tables(c, zr)
    assert x.size == 4 * det * det and z.size == 2 * rings * rings
    for m in range(det * det * rings * rings):
        i, j, k = m % det, (m // det) % det, m // (det * det)
        b = m % (det * det)
        assert tuple(x[4 * b:4 * b + 4]) == (c[i, 0], c[i, 1], c[j, 0], c[j, 1])
        assert tuple(z[2 * k:2 * k + 2]) == (zr[k // rings], zr[k % rings])
