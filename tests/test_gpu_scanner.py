"""projectorClass on OMEGA's block-ring scanner model (tables from omega_amd.geometry.omega_*, pinned in
test_scanner_geometry.py): per-subset forward projections bit-identical to the oracle on the same
tables, backprojections within 1e-6 relative L2."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.hip_helpers import rel_l2
from tests.test_scanner_geometry import _scanner_projector

pytestmark = pytest.mark.gpu


def test_scanner_model_projections(cuda):
    import torch
    p = _scanner_projector(subsets=3)
    p.addProjector()
    p.initProj()
    rng = np.random.default_rng(6)
    img = (rng.random(p.grid.N) + 0.1).astype(np.float32)
    it = torch.from_numpy(img).to(cuda)
    per = int(p.nRowsD) * int(p.nColsD)
    print(f"[scanner model] sinograms {int(p.nProjections)} bins {per}", flush=True)
    for k in range(p.subsets):
        y = p.forwardProject(it, k)
        n = y.numel()
        case = O.SiddonCase(p.grid, int(p.nRowsD), int(p.nColsD), p.x, p.z, plane_offset=int(p.nMeas[k]))
        assert np.array_equal(y.cpu().numpy(), case.fp(img, n))
        yv = rng.random(n).astype(np.float32)
        b = p.backwardProject(torch.from_numpy(yv).to(cuda), k).cpu().numpy()
        assert rel_l2(b, case.bp(yv, n)) < 1e-6
