"""The measured workloads, built from their stated sizes with seeded synthetic data.

``dynamic``: the 3-D PET geometry of ``pet3d_projector`` with 24 frames whose images follow
seeded one-tissue-compartment curves (``phantoms.dynamic_ellipsoids``) and whose total counts are
lognormal in [1e5, 1e7] (``frame_counts``); 16 subsets x 3 iterations, all frames per launch.

``pet2d_osem``: 2-D plane PET OSEM with improved Siddon. Ring radius 400 mm, 63 direct sinogram
planes of 128 radial bins x 160 angles over 180 degrees (2.5 mm radial pitch); image 128x128x63
with 2.5 mm voxels; phantom of 8 seeded ellipsoids, activity U(0.5, 2.0) on a 0.1 background;
data Poisson(A x) with A x scaled to 5e6 total counts; 8 subsets (type 8) x 4 iterations.
"""
from __future__ import annotations

import numpy as np

from . import geometry as G
from . import phantoms
from .projector import projectorClass


def pet2d_projector(subsets: int = 8) -> projectorClass:
    grid = G.Grid.from_fov(128, 128, 63, 320.0, 320.0, 157.5)
    sc = G.pet_2d_planes(grid, radius=400.0, n_radial=128, n_angles=160, radial_pitch=2.5)
    p = projectorClass()
    p.Nx, p.Ny, p.Nz = 128, 128, 63
    p.FOVa_x, p.FOVa_y, p.axial_fov = 320.0, 320.0, 157.5
    p.projector_type = 1
    p.x = sc.xy_table()
    p.z = sc.z
    p.Ndist, p.Nang = sc.n_radial, sc.n_angles
    p.subsets = subsets
    p.subsetType = 8
    p.addProjector()
    return p


def pet2d_data(p: projectorClass, seed: int = 0, total_counts: float = 5e6, device="cuda"):
    """Phantom and Poisson data (subset order) for the 2-D plane PET workload.

    The noiseless projection is computed with the library's forward projector over all subsets.
    Returns (phantom numpy, measurements as a float32 device tensor in subset order).
    """
    import torch
    ph = phantoms.ellipsoids(p.grid, seed=seed, n_ell=8, background=0.1)
    x = torch.as_tensor(ph, device=device)
    parts = [p.forwardProject(x, k) for k in range(p.subsets)]
    ax = torch.cat(parts).double().cpu().numpy()
    scale = total_counts / ax.sum()
    rng = np.random.default_rng(seed + 1)
    y = rng.poisson(ax * scale).astype(np.float32)
    return ph, torch.as_tensor(y, device=device)


def pet3d_projector(subsets: int = 1, n_rings: int = 32, crystals: int = 448, ring_pitch: float = 4.0,
                    max_ring_diff: int = 31, radius: float = 400.0, image=(128, 128, 63),
                    fov_xy: float = 320.0, subset_type: int = 8) -> projectorClass:
    """3-D PET: ``n_rings`` rings at ``ring_pitch`` mm, ``crystals`` per ring, span-1 ring pairs up to
    ``max_ring_diff``; ``crystals/2`` angles x ``crystals/2`` radial bins; image ``image`` voxels over a
    ``fov_xy`` mm transaxial FOV and the rings' axial extent. Defaults: 224 x 224 x 1024 sinogram,
    128 x 128 x 63 image (2.5 mm transaxial, 128/63 mm axial voxels). ``subset_type``: the reference's
    subset types (8: every n-th sinogram plane; 4: every n-th angle of every plane; ...)."""
    sc = G.pet_3d_rings(radius=radius, n_rings=n_rings, ring_pitch=ring_pitch, crystals=crystals,
                        max_ring_diff=max_ring_diff)
    p = projectorClass()
    p.Nx, p.Ny, p.Nz = image
    p.FOVa_x, p.FOVa_y, p.axial_fov = fov_xy, fov_xy, n_rings * ring_pitch
    p.projector_type = 1
    p.x = sc.xy_table()
    p.z = sc.z
    p.Ndist, p.Nang = sc.n_radial, sc.n_angles
    p.subsets = subsets
    p.subsetType = subset_type
    p.addProjector()
    return p


def ct_cgls_projector(n=256, voxel=0.5, n_views=360, rows=512, cols=384, pixel=0.8, sod=700.0, sdd=1000.0):
    """CBCT with the branchless distance-driven projector: n^3 volume of `voxel` mm, `n_views` projections
    over 360 degrees on a rows x cols flat panel of `pixel` mm, source-origin `sod`, source-detector `sdd`."""
    x, uv = G.cbct_tables(n_views, sod=sod, sdd=sdd, pitch_u=pixel)
    p = projectorClass()
    p.CT = True
    p.projector_type = 5
    p.Nx = p.Ny = p.Nz = n
    p.FOVa_x = p.FOVa_y = p.axial_fov = n * voxel
    p.x, p.z = x, uv
    p.nRowsD, p.nColsD = rows, cols
    p.dPitchX = p.dPitchY = pixel
    p.subsets = 1
    p.addProjector()
    return p


def ct_cgls_data(p, seed=0, noise=0.01, device="cuda"):
    """Attenuation phantom (12 seeded ellipsoids, mu ~ U(0.01, 0.03)/mm) and its line integrals A mu plus
    Gaussian noise of sigma = noise * max."""
    import torch
    mu = phantoms.attenuation_ellipsoids(p.grid, seed=seed, n_ell=12)
    y = p.forwardProject(torch.as_tensor(mu, device=device), 0)
    g = torch.Generator(device=device)
    g.manual_seed(seed + 3)
    y = y + noise * float(y.max()) * torch.randn(y.shape, device=device, generator=g)
    return mu, y.contiguous()


def frame_counts(n_frames: int, seed: int, lo: float = 1e5, hi: float = 1e7):
    """Per-frame total counts, lognormal (log10 ~ N(6, 0.6)) clipped to [lo, hi]."""
    rng = np.random.default_rng(seed + 104729)
    return 10.0 ** np.clip(rng.normal(6.0, 0.6, n_frames), np.log10(lo), np.log10(hi))


def dynamic_data(p: projectorClass, frames, counts, seed: int, device="cuda"):
    """Frame-major Poisson data [F, M] in subset order: frame f is Poisson(A x_f scaled to counts[f]);
    ``frames`` is [N, F] (phantoms.dynamic_ellipsoids)."""
    import torch
    F = frames.shape[1]
    M = int(p.nTotMeas[-1])
    y = torch.empty((F, M), dtype=torch.float32, device=device)
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    for f in range(F):
        x = torch.as_tensor(np.ascontiguousarray(frames[:, f]), device=device)
        ax = torch.cat([p.forwardProject(x, k) for k in range(p.subsets)])
        scale = float(counts[f]) / float(ax.double().sum())
        y[f] = torch.poisson(ax * scale, generator=g)
        del ax
    return y


def poisson_data(p: projectorClass, phantom, total_counts: float, seed: int, device="cuda"):
    """Poisson data (subset order) of A x scaled to ``total_counts``, x = ``phantom``."""
    import torch
    x = torch.as_tensor(phantom, device=device)
    ax = torch.cat([p.forwardProject(x, k) for k in range(p.subsets)]).double()
    scale = total_counts / float(ax.sum())
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    y = torch.poisson(ax * scale, generator=g).float()
    return y, scale
