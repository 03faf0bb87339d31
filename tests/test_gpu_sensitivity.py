"""GPU list-mode sensitivity image (projectorClass.computeSensitivityImage) against the oracle's
backprojection of ones over the same LORs given one by one as list-mode events: within 1e-5 relative L2
(float atomics), for improved Siddon, the orthogonal and the volume-of-intersection projectors, with and
without an attenuation image (the reference's -DSENS kernel applies it along every LOR,
ProjectorClass.h:3937-3953), on three rings and on one ring (every LOR in a transaxial plane: the
direct-plane kernels run), with the LORs split over several launches."""
import math

import numpy as np
import pytest

from omega_amd import geometry as G
from oracle import oracle as O
from tests.hip_helpers import rel_l2

pytestmark = pytest.mark.gpu

TUBE = 3.0


def _scanner(rings):
    det, R = 32, 60.0
    ang = 2 * math.pi * (np.arange(det) + 0.5) / det
    c = np.stack([R * np.cos(ang), R * np.sin(ang)], 1).astype(np.float32)
    zr = np.array([-6.0, 0.0, 6.0], np.float32) if rings == 3 else np.array([1.5], np.float32)
    return c, zr


def _projector(pt, mu=None):
    from omega_amd.projector import projectorClass
    p = projectorClass()
    p.Nx, p.Ny, p.Nz = 16, 16, 6
    p.FOVa_x, p.FOVa_y, p.axial_fov = 64.0, 64.0, 18.0
    p.listmode = 1
    p.x = np.zeros(6, np.float32)
    p.projector_type = pt
    if pt == 2:
        p.tube_width_xy = p.tube_width_z = TUBE
    if pt == 3:
        p.tube_radius = TUBE
    if mu is not None:
        p.attenuation_correction = True
        p.vaimennus = mu
    p.addProjector()
    p.initProj()
    return p


@pytest.mark.parametrize("rings", [3, 1])
@pytest.mark.parametrize("atten", [False, True])
@pytest.mark.parametrize("pt", [1, 2, 3])
def test_sensitivity_image_against_oracle(cuda, pt, atten, rings):
    c, zr = _scanner(rings)
    mu = None
    if atten:
        mu = (np.random.default_rng(7).random(16 * 16 * 6) * 0.02).astype(np.float32)
    p = _projector(pt, mu)
    print(f"[sensitivity] projector {pt} attenuation {atten} det={c.shape[0]} rings={zr.size}", flush=True)
    sens = p.computeSensitivityImage(c, zr, max_lors=2048).cpu().numpy()
    x, z = G.listmode_sensitivity_tables(c, zr)
    det, nr = c.shape[0], zr.size
    per = det * det
    lm = np.empty((per * nr * nr, 6), np.float32)
    xy = x.reshape(-1, 4)
    for k in range(nr * nr):
        lm[k * per:(k + 1) * per, 0:2] = xy[:, 0:2]
        lm[k * per:(k + 1) * per, 3:5] = xy[:, 2:4]
        lm[k * per:(k + 1) * per, 2] = z[2 * k]
        lm[k * per:(k + 1) * per, 5] = z[2 * k + 1]
    n = lm.shape[0]
    ones = np.ones(n, np.float32)
    z0 = np.zeros(2, np.float32)
    if pt == 1:
        ref = O.SiddonCase(p.grid, 1, 1, lm.ravel(), z0, mode=O.MODE_LIST, atten=mu,
                           atten_mode=1 if atten else 0).bp(ones, n)
    else:
        fac = None
        if atten:
            fac = O.SiddonCase(p.grid, 1, 1, lm.ravel(), z0, mode=O.MODE_LIST, atten=mu, atten_mode=1).atten_factors(n)
        vol = None if pt == 2 else (p.V, np.float32(p.bmin), np.float32(p.bmax), np.float32(p.Vmax))
        ref = O.OrthCase(p.grid, 1, 1, lm.ravel(), z0, TUBE, flags=3, mode=O.MODE_LIST, atten=fac,
                         volume=vol).bp(ones, n)
    assert ref.sum() > 0
    e = rel_l2(sens, ref)
    print(f"  rel-L2 vs oracle {e:.2e}", flush=True)
    assert e <= 1e-5
    if atten:   # the attenuation image changes the image by far more than the tolerance
        assert rel_l2(_projector(pt).computeSensitivityImage(c, zr).cpu().numpy(), ref) > 1e-3


def test_sensitivity_image_rejects_per_event_factors(cuda):
    from omega_amd.projector import projectorClass
    p = projectorClass()
    p.Nx, p.Ny, p.Nz = 16, 16, 6
    p.FOVa_x, p.FOVa_y, p.axial_fov = 64.0, 64.0, 18.0
    p.listmode = 1
    p.x = np.zeros(12, np.float32)
    p.attenuation_correction = True
    p.CTAttenuation = False
    p.vaimennus = np.ones(2, np.float32)
    p.addProjector()
    p.initProj()
    c, zr = _scanner(1)
    with pytest.raises(ValueError, match="per-event"):
        p.computeSensitivityImage(c, zr)
