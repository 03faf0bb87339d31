"""GPU CGLS and PDHG drivers against the oracle's recurrences (same projector, CPU)."""
import numpy as np
import pytest

from omega_amd import geometry as G
from omega_amd import phantoms
from omega_amd.projector import projectorClass
from omega_amd.recon import CGLSRunner, PDHGRunner, power_method
from oracle import oracle as O
from tests import cases
from tests.hip_helpers import rel_l2

pytestmark = pytest.mark.gpu


def ct_projector(subsets=1):
    grid, x, uv, rows, cols, pitch = cases.ct_small()
    p = projectorClass()
    p.CT = True
    p.Nx, p.Ny, p.Nz = grid.Nx, grid.Ny, grid.Nz
    p.FOVa_x, p.FOVa_y, p.axial_fov = 32.0, 32.0, 24.0
    p.x, p.z = x, uv
    p.nRowsD, p.nColsD = rows, cols
    p.dPitchX = p.dPitchY = pitch
    p.subsets = subsets
    p.addProjector()
    p.initProj()
    return p, grid, x, uv, rows, cols, pitch


def test_cgls_ct_matches_oracle(cuda):
    import torch
    p, grid, x, uv, rows, cols, pitch = ct_projector()
    ph = phantoms.ellipsoids(grid, seed=2, background=0.0) * 0.02
    n = int(p.nMeasSubset[0, 0])
    y = p.forwardProject(torch.as_tensor(ph, device=cuda), 0)
    f = CGLSRunner(p, y).run(6).cpu().numpy()
    case = O.SiddonCase(grid, rows, cols, x, uv, mode=O.MODE_CT, dPitchZ=pitch, n_proj=x.size // 6)
    fo = O.cgls(case, y.cpu().numpy(), 6, n)
    e = rel_l2(f, fo)
    print(f"[cgls] rel-L2 vs oracle after 6 iterations {e:.2e}")
    assert e < 1e-5


def test_pdhg_subsets_matches_oracle(cuda):
    import torch
    p, grid, x, uv, rows, cols, pitch = ct_projector(subsets=4)
    ph = phantoms.ellipsoids(grid, seed=3, background=0.0) * 0.02
    y = torch.cat([p.forwardProject(torch.as_tensor(ph, device=cuda), k) for k in range(p.subsets)])
    tau = power_method(p, iterations=15)
    f = PDHGRunner(p, y, tau=tau).run(3).cpu().numpy()
    cs = [O.SiddonCase(grid, rows, cols, p.x, p.z, mode=O.MODE_CT, dPitchZ=pitch, plane_offset=int(p.nMeas[k]),
                       n_proj=x.size // 6) for k in range(p.subsets)]
    ym = y.cpu().numpy()
    parts = [ym[int(p.nTotMeas[k]):int(p.nTotMeas[k + 1])] for k in range(p.subsets)]
    fo = O.pdhg(cs, parts, 3, tau)
    e = rel_l2(f, fo)
    print(f"[pdhg] rel-L2 vs oracle after 3 iterations x 4 subsets {e:.2e}")
    assert e < 1e-5
