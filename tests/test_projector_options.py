"""projectorClass refuses configurations the library does not provide, before any device work (CPU)."""
import numpy as np
import pytest

from omega_amd.projector import projectorClass
from tests import cases


def _pet2d(pt):
    grid, sc = cases.orth_pet2d()
    p = projectorClass()
    p.Nx, p.Ny, p.Nz = grid.Nx, grid.Ny, grid.Nz
    p.FOVa_x, p.FOVa_y, p.axial_fov = grid.Nx * grid.dx, grid.Ny * grid.dy, grid.Nz * grid.dz
    p.x, p.z = sc.xy_table(), sc.z
    p.Ndist, p.Nang = sc.n_radial, sc.n_angles
    p.projector_type = pt
    if pt == 2:
        p.tube_width_xy = p.tube_width_z = 3.0
    if pt == 3:
        p.tube_radius = 3.0
    return p


def test_tof_with_the_orthogonal_projector_is_refused():
    """TOF data (TOF_bins == TOF_bins_used > 1) with projector type 2 raises (as proj.py:1247-1248 does)
    instead of being projected as non-TOF data of the wrong size."""
    p = _pet2d(2)
    p.TOF_bins, p.TOF_bins_used, p.TOF_width, p.TOF_FWHM = 7, 7, 100e-12, 300e-12
    p.addProjector()
    assert p.nBins == 7
    with pytest.raises(ValueError, match="TOF"):
        p._descriptor()


def test_tof_with_the_volume_projector_reaches_the_descriptor():
    p = _pet2d(3)
    p.TOF_bins, p.TOF_bins_used, p.TOF_width, p.TOF_FWHM = 7, 7, 100e-12, 300e-12
    p.addProjector()
    assert p.nBins == 7
    d, _ = p._descriptor()
    assert d.tof_bins == 7 and d.projector_type == 3 and d.tof_sigma > 0


@pytest.mark.parametrize("pt", [1, 2, 3])
def test_summed_tof_bins_are_plain_data(pt):
    """TOF_bins_used != TOF_bins asks for the bins summed: a non-TOF reconstruction for every type."""
    p = _pet2d(pt)
    p.TOF_bins, p.TOF_bins_used = 7, 1
    p.addProjector()
    assert p.nBins == 1
    d, _ = p._descriptor()
    assert d.tof_bins == 0
