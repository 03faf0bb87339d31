"""The volume-of-intersection projector's look-up table (projector type 3) against the reference's own
computeVoxelVolumes output (tests/golden/volume_table.npz, written by tests/golden/make_volume_table.py),
and against what the geometry fixes independently of the closed form: the volume of a sphere of radius
r whose centre lies at distance b from the axis of a cylinder of radius R is the integral over the
sphere's z-slices of the area of the lens where the slice's disk meets the cylinder's circle.

Table quirks pinned here: the reference drops the two entries at b = R - 1e-3 and b = R (NaN and a
non-decreasing step of the closed form), so from there on V[k] holds the volume two steps beyond
bmin + k * 1e-3; its last entry is kept only by the appended 0 of its monotonicity filter and holds the
closed form's cancellation residue at bmax, not the volume there (which is zero)."""
import math
import os

import numpy as np
import pytest
from scipy.integrate import quad

from omega_amd import geometry as G

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "volume_table.npz"))
CASES = [(2.0, 1.0, (2.5, 2.5, 2.5)), (1.2, 1.0, (1.0, 1.0, 2.0)), (3.5, 0.8, (2.0, 2.0, 2.0))]


def _lens(R, rho, d):
    """Area of the intersection of circles of radii R and rho whose centres are d apart."""
    if d >= R + rho:
        return 0.0
    if d <= abs(R - rho):
        return math.pi * min(R, rho) ** 2
    a1 = math.acos((d * d + R * R - rho * rho) / (2 * d * R))
    a2 = math.acos((d * d + rho * rho - R * R) / (2 * d * rho))
    return R * R * a1 + rho * rho * a2 - 0.5 * math.sqrt((-d + R + rho) * (d + R - rho) * (d - R + rho) * (d + R + rho))


def _volume(R, r, b):
    return quad(lambda z: _lens(R, math.sqrt(max(r * r - z * z, 0.0)), b), -r, r, limit=200, epsabs=1e-12)[0]


@pytest.mark.parametrize("n", range(len(CASES)))
def test_matches_reference(n):
    tr, vr, d = CASES[n]
    V, bmin, bmax, vmax, r = G.voxel_volume_table(tr, vr, *d)
    assert np.array_equal(V, GOLD[f"c{n}_V"])
    assert np.array_equal(np.array([bmin, bmax, vmax, r]), GOLD[f"c{n}_scalars"])


@pytest.mark.parametrize("n", range(len(CASES)))
def test_closed_form_against_lens_integral(n):
    tr, vr, d = CASES[n]
    V, bmin, bmax, vmax, r, b = G.voxel_volume_table(tr, vr, *d, return_grid=True)
    assert r == pytest.approx(math.sqrt(2.0) * vr * max(d) / 2)
    assert vmax == pytest.approx(4.0 / 3.0 * math.pi * r ** 3)
    assert bmax == pytest.approx(tr + r)
    assert bmin == pytest.approx(max(0.0, tr - r), abs=1.5e-3)
    steps = np.round(b * 1e3).astype(np.int64)
    full = np.arange(steps[0], steps[-1] + 1)
    missing = np.setdiff1d(full, steps)
    k_r = int(round(tr * 1e3))
    assert np.array_equal(missing, [k_r - 1, k_r])
    for k in list(range(0, V.size - 1, 53)) + [V.size - 2]:
        assert abs(V[k] - _volume(tr, r, b[k])) <= 1e-5 * vmax, (k, b[k])
    if tr >= r:   # a voxel whose centre lies on the axis is inside the tube
        assert V[0] == pytest.approx(vmax, rel=1e-5)
    assert np.all(np.diff(V[:-1]) <= 0)
