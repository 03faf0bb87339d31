"""Scanner geometry and coordinate tables in OMEGA's projector formats.

Coordinate tables follow the layouts the reference projector consumes
(source/cpp/projector_functions.h:357-498, source/opencl/general_opencl_functions.h:1601-1630):

* PET sinogram: ``x`` holds 4 floats per (radial, angle) bin ``[xs, ys, xd, yd]`` with the
  radial bin fastest; ``z`` holds 2 floats per sinogram plane ``[zs, zd]``.
* CT flat panel: ``x`` holds 6 floats per projection (source xyz, detector-centre xyz); ``z``
  holds 2 floats per projection, the transaxial panel direction scaled by the pixel pitch
  (``uV``, source/Python/omegatomo/projector/detcoord.py:163-185).
* List mode: 6 floats per LOR ``[xs, ys, zs, xd, yd, zd]``.

Voxel grid placement restates ``computePixelSize`` (source/Python/omegatomo/projector/coordinates.py:42-62):
boundaries are ``linspace(-FOV/2, FOV/2, N+1)`` in float32, ``d = xx[1] - xx[0]``, ``b = xx[0]``.

The scanner builders are this project's own (the reference reads measured geometry
from files); they generate rings, sinogram planes and flat-panel trajectories with the
sizes of the workloads measured here.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Grid:
    """Voxel grid: counts, voxel sizes and the lower-left corner of the FOV (mm)."""
    Nx: int
    Ny: int
    Nz: int
    dx: float
    dy: float
    dz: float
    bx: float
    by: float
    bz: float

    @property
    def N(self) -> int:
        return self.Nx * self.Ny * self.Nz

    @staticmethod
    def from_fov(Nx: int, Ny: int, Nz: int, FOVa_x: float, FOVa_y: float, axial_fov: float,
                 oOffsetX: float = 0.0, oOffsetY: float = 0.0, oOffsetZ: float = 0.0) -> "Grid":
        fov = np.array([FOVa_x, FOVa_y, axial_fov], dtype=np.float32)
        e = -fov / 2
        xx = np.linspace(e[0] + oOffsetX, -e[0] + oOffsetX, Nx + 1, dtype=np.float32)
        yy = np.linspace(e[1] + oOffsetY, -e[1] + oOffsetY, Ny + 1, dtype=np.float32)
        zz = np.linspace(e[2] + oOffsetZ, -e[2] + oOffsetZ, Nz + 1, dtype=np.float32)
        return Grid(Nx, Ny, Nz,
                    float(np.float32(xx[1] - xx[0])), float(np.float32(yy[1] - yy[0])), float(np.float32(zz[1] - zz[0])),
                    float(xx[0]), float(yy[0]), float(zz[0]))

    def n3(self) -> np.ndarray:
        return np.array([self.Nx, self.Ny, self.Nz], dtype=np.uint32)

    def d3(self, dtype=np.float32) -> np.ndarray:
        return np.array([self.dx, self.dy, self.dz], dtype=dtype)

    def b3(self, dtype=np.float32) -> np.ndarray:
        return np.array([self.bx, self.by, self.bz], dtype=dtype)

    def centers(self):
        """Voxel centre coordinates per axis (coordinates.py:124-126)."""
        xc = (np.float32(self.bx) + np.arange(self.Nx, dtype=np.float32) * np.float32(self.dx)
              + np.float32(self.dx) / np.float32(2)).astype(np.float32)
        yc = (np.float32(self.by) + np.arange(self.Ny, dtype=np.float32) * np.float32(self.dy)
              + np.float32(self.dy) / np.float32(2)).astype(np.float32)
        zc = (np.float32(self.bz) + np.arange(self.Nz, dtype=np.float32) * np.float32(self.dz)
              + np.float32(self.dz) / np.float32(2)).astype(np.float32)
        return xc, yc, zc


def ring_sinogram_xy(n_radial: int, n_angles: int, radial_pitch: float, radius: float,
                     angle_span: float = math.pi) -> np.ndarray:
    """Transaxial sinogram table ``[xs, ys, xd, yd]`` per (radial, angle), radial fastest.

    Bin (r, a) is the chord of the detector ring at signed distance
    ``s = (r - (n_radial - 1)/2) * radial_pitch`` from the centre, with normal at angle
    ``phi = a * angle_span / n_angles``. Endpoints lie on the ring. Computed in float64,
    stored as float32.
    """
    r = np.arange(n_radial, dtype=np.float64)
    a = np.arange(n_angles, dtype=np.float64)
    s = (r - (n_radial - 1) / 2.0) * radial_pitch
    phi = a * angle_span / n_angles
    S, PHI = np.meshgrid(s, phi, indexing="xy")       # shape (n_angles, n_radial): radial fastest
    t0 = np.sqrt(np.maximum(radius * radius - S * S, 0.0))
    c, sn = np.cos(PHI), np.sin(PHI)
    xs = S * c + t0 * sn
    ys = S * sn - t0 * c
    xd = S * c - t0 * sn
    yd = S * sn + t0 * c
    tab = np.stack([xs, ys, xd, yd], axis=-1)           # (n_angles, n_radial, 4)
    return np.ascontiguousarray(tab.reshape(-1).astype(np.float32))


def direct_planes_z(grid: Grid) -> np.ndarray:
    """One direct (in-plane) sinogram per voxel slice at the slice centre: ``[zs, zd]`` per plane."""
    zc = grid.centers()[2]
    return np.ascontiguousarray(np.stack([zc, zc], axis=-1).reshape(-1).astype(np.float32))


def ring_pair_planes_z(n_rings: int, ring_pitch: float, max_ring_diff: int) -> np.ndarray:
    """Span-1 plane table ``[zs, zd]`` for every ring pair with ``|r1 - r2| <= max_ring_diff``.

    Planes are ordered ring-difference major (0, +1, -1, +2, -2, ...) and by the first ring
    within a difference, the Michelogram ordering of span-1 segments.
    """
    zr = (np.arange(n_rings, dtype=np.float64) - (n_rings - 1) / 2.0) * ring_pitch
    planes = []
    diffs = [0]
    for d in range(1, max_ring_diff + 1):
        diffs += [d, -d]
    for d in diffs:
        for r1 in range(n_rings):
            r2 = r1 + d
            if 0 <= r2 < n_rings:
                planes.append((zr[r1], zr[r2]))
    return np.ascontiguousarray(np.asarray(planes, dtype=np.float64).reshape(-1).astype(np.float32))


def cbct_tables(n_views: int, sod: float, sdd: float, pitch_u: float, span: float = 2 * math.pi,
                start: float = 0.0):
    """Circular cone-beam trajectory: ``x`` (6 floats/view) and ``uV`` (2 floats/view).

    The source sits at distance ``sod`` from the rotation axis, the panel centre at
    ``sdd - sod`` on the far side; ``uV`` is the in-plane panel direction times ``pitch_u``.
    """
    th = start + np.arange(n_views, dtype=np.float64) * span / n_views
    c, s = np.cos(th), np.sin(th)
    src = np.stack([sod * c, sod * s, np.zeros_like(c)], axis=-1)
    det = np.stack([-(sdd - sod) * c, -(sdd - sod) * s, np.zeros_like(c)], axis=-1)
    x = np.concatenate([src, det], axis=-1).reshape(-1).astype(np.float32)
    uv = np.stack([-s * pitch_u, c * pitch_u], axis=-1).reshape(-1).astype(np.float32)
    return np.ascontiguousarray(x), np.ascontiguousarray(uv)


def subset_index_type8(n_proj: int, subsets: int):
    """Subset type 8: every ``subsets``-th plane/projection (indices.py:158-170).

    Returns (index, nMeas) where ``index`` lists planes in subset order and ``nMeas[k]`` the
    number of planes of subset ``k``.
    """
    if subsets <= 1:
        return np.arange(n_proj, dtype=np.int64), np.array([n_proj], dtype=np.int64)
    idx = [np.arange(k, n_proj, subsets, dtype=np.int64) for k in range(subsets)]
    return np.concatenate(idx), np.array([len(i) for i in idx], dtype=np.int64)


def _prime_factors(n: int):
    f, d = [], 2
    while n > 1:
        while n % d == 0:
            f.append(d)
            n //= d
        d += 1
    return f


def subset_index_projections(subset_type: int, n_proj: int, subsets: int, seed: int = -1):
    """Projection-level subsets (indices.py:158-240): a permutation of the planes / projections cut into
    ``subsets`` consecutive runs, the first n_proj % subsets of them one longer.

    * 8: every ``subsets``-th projection (interleaved, not cut into runs);
    * 9: a random permutation (numpy ``default_rng(seed).permutation``; unseeded when seed < 0);
    * 11: the prime-factor ordering: with n_proj = f1 f2 ... fN (prime factors in increasing order),
      projection r (digits d1 .. dN of r in that mixed radix, d1 fastest) goes to position
      d1 f2..fN + d2 f3..fN + ... + dN, which spreads consecutive positions over the whole range.

    Returns (index, nMeas): the projections in subset order and the number of projections per subset.
    """
    if subsets <= 1:
        return np.arange(n_proj, dtype=np.int64), np.array([n_proj], dtype=np.int64)
    if subset_type == 8:
        idx = [np.arange(k, n_proj, subsets, dtype=np.int64) for k in range(subsets)]
        return np.concatenate(idx), np.array([len(i) for i in idx], dtype=np.int64)
    if subset_type == 9:
        rng = np.random.default_rng() if seed < 0 else np.random.default_rng(seed)
        order = rng.permutation(n_proj).astype(np.int64)
    elif subset_type == 11:
        f = _prime_factors(n_proj)
        if not f:
            raise ValueError("subset type 11 needs at least two projections")
        weights = np.cumprod(np.asarray(f[1:][::-1], dtype=np.int64))[::-1]   # f(k+1) ... fN for k < N
        digits = np.zeros((len(f), n_proj), dtype=np.int64)
        r = np.arange(n_proj, dtype=np.int64)
        step = 1
        for k, fk in enumerate(f):
            digits[k] = (r // step) % fk
            step *= fk
        order = (digits[: len(weights)] * weights[:, None]).sum(0) + digits[-1]
    elif subset_type == 10:
        raise ValueError("Subset type 10 (golden angle) is not supported in the Python version of the reference either")
    else:
        raise ValueError(f"subset type {subset_type} is not a projection-level subset type (8, 9, 11)")
    base, extra = divmod(n_proj, subsets)
    sizes = np.array([base + (1 if k < extra else 0) for k in range(subsets)], dtype=np.int64)
    return order, sizes


def subset_index_measurements(subset_type: int, subsets: int, n_dist: int, n_ang: int, n_sinos: int,
                              seed: int = -1):
    """Measurement-level subsets of a sinogram (indexMaker, indices.py:51-123).

    Bins are numbered radial bin fastest, then angle, then plane (``r + a*n_dist + s*n_dist*n_ang``).
    Subset ``i`` of

    * type 1 takes every ``subsets``-th bin counted angle-fastest, starting at ``i``;
    * type 2 every ``subsets``-th bin in sinogram order, starting at ``i``;
    * type 3 the ``i``-th block of ``total // subsets`` bins of a permutation drawn with
      ``numpy.random.default_rng(seed)`` (seed < 0: unseeded); the last block ends one bin before the
      end of the permutation, as the reference's slice ``[port*i:-1]`` does;
    * type 4 every ``subsets``-th angle starting at ``i`` (whole radial rows), plane by plane; the
      reference fails with ``n_ang % subsets != 0`` (its remainder branch does not broadcast), so that
      case raises ``ValueError`` here;
    * type 5 every ``subsets``-th (radial bin, plane) row counted radial-fastest, each with all angles.

    Returns (index, counts): the subsets' bins concatenated in subset order (int64) and the number of
    bins per subset.
    """
    total = n_dist * n_ang * n_sinos
    parts = []
    if subset_type == 1:
        for i in range(subsets):
            k = np.arange(i, total, subsets, dtype=np.int64)
            a, rest = k % n_ang, k // n_ang
            parts.append(rest % n_dist + a * n_dist + (rest // n_dist) * (n_dist * n_ang))
    elif subset_type == 2:
        parts = [np.arange(i, total, subsets, dtype=np.int64) for i in range(subsets)]
    elif subset_type == 3:
        rng = np.random.default_rng(seed) if seed >= 0 else np.random.default_rng()
        perm = rng.permutation(total).astype(np.int64)
        port = total // subsets
        parts = [perm[port * i: port * (i + 1)] for i in range(subsets - 1)] + [perm[port * (subsets - 1): -1]]
    elif subset_type == 4:
        if n_ang % subsets:
            raise ValueError("Subset type 4 needs the number of angles to be divisible by the number of subsets "
                             "(the reference's indexMaker fails otherwise)")
        r = np.arange(n_dist, dtype=np.int64)
        s = np.arange(n_sinos, dtype=np.int64)
        for i in range(subsets):
            a = np.arange(i, n_ang, subsets, dtype=np.int64)
            parts.append((s[:, None, None] * (n_dist * n_ang) + a[None, :, None] * n_dist + r[None, None, :]).ravel())
    elif subset_type == 5:
        a = np.arange(n_ang, dtype=np.int64)
        for i in range(subsets):
            rows = np.arange(i, n_dist * n_sinos, subsets, dtype=np.int64)
            r, s = rows % n_dist, rows // n_dist
            parts.append((r[:, None] + a[None, :] * n_dist + s[:, None] * (n_dist * n_ang)).ravel())
    else:
        raise ValueError(f"subset type {subset_type} is not a measurement-level subset type (1-5)")
    return np.concatenate(parts), np.array([p.size for p in parts], dtype=np.int64)


SPEED_OF_LIGHT_MM_S = 2.99792458e11


def tof_sigma(fwhm_s: float) -> float:
    """TOF standard deviation along the LOR (mm) for a timing resolution FWHM in seconds
    (projectorClass, proj.py:693-694: the distance is c * t / 2)."""
    return (SPEED_OF_LIGHT_MM_S * fwhm_s / 2.0) / (2.0 * math.sqrt(2.0 * math.log(2.0)))


def tof_centers(bins: int, width_s: float, offset_s: float = 0.0) -> np.ndarray:
    """TOF bin centres along the LOR (mm), in the reference's bin order: the central bin, then
    alternately the next bin on the + and the - side (proj.py:695-703), times -c/2. That ordering needs
    an odd bin count (the reference's slice assignment fails for an even one)."""
    if bins < 1 or bins % 2 == 0:
        raise ValueError(f"TOF bin centres are generated for an odd number of bins (got {bins}); give TOFCenter instead")
    edges = np.linspace(-width_s * bins / 2, width_s * bins / 2, bins + 1, dtype=np.float32)
    mid = edges[:-1] + np.float32(width_s / 2.0)
    c = np.zeros(mid.shape, dtype=np.float32)
    h = mid.size // 2
    c[0] = mid[h]
    c[1::2] = mid[h + 1:]
    c[2::2] = mid[h - 1::-1]
    if offset_s > 0:
        c = c + np.float32(offset_s)
    return np.float32(-c * SPEED_OF_LIGHT_MM_S / 2.0)


@dataclass
class PETScanner:
    """A cylindrical PET scanner sampled as parallel-chord sinograms."""
    radius: float
    n_radial: int
    n_angles: int
    radial_pitch: float
    z: np.ndarray = field(repr=False)      # [zs, zd] per plane, float32

    @property
    def n_planes(self) -> int:
        return self.z.size // 2

    def xy_table(self) -> np.ndarray:
        return ring_sinogram_xy(self.n_radial, self.n_angles, self.radial_pitch, self.radius)


def pet_2d_planes(grid: Grid, radius: float = 400.0, n_radial: int = 128, n_angles: int = 160,
                  radial_pitch: float = 2.5) -> PETScanner:
    """2-D plane PET: one direct sinogram per voxel slice."""
    return PETScanner(radius, n_radial, n_angles, radial_pitch, direct_planes_z(grid))


def pet_3d_rings(radius: float = 400.0, n_rings: int = 32, ring_pitch: float = 4.0, crystals: int = 448,
                 max_ring_diff: int = 31) -> PETScanner:
    """3-D PET with span-1 ring-pair planes; ``crystals/2`` angles and radial bins at the
    arc sampling of the ring (half the crystal pitch)."""
    n_ang = crystals // 2
    pitch = math.pi * radius / crystals
    return PETScanner(radius, n_ang, n_ang, pitch, ring_pair_planes_z(n_rings, ring_pitch, max_ring_diff))


# ---- OMEGA's own PET scanner geometry (source/Python/omegatomo/projector/detcoord.py) ----------------
# The reference builds a block-ring scanner's crystal coordinates, the sinogram's LOR endpoints and the
# axial (ring-pair) coordinates procedurally; the tables here are float-for-float the same (checked
# against the reference's own output), so data binned by OMEGA can be projected directly.

def omega_crystal_ring(blocks_per_ring: int, cryst_per_block: int, cr_p: float, diameter: float,
                       transaxial_multip: int = 1, doi: float = 0.0, flip_image: bool = False,
                       offangle: float = 0.0, pseudo: bool = False):
    """Transaxial crystal centres (x, y), float32, of a ring of flat blocks (detectorCoordinates /
    computeCoordinates, detcoord.py:326-556, m-files/computeCoordinates.m): blocks at angles 90,
    90 - 360/B, ... degrees; crystals of pitch cr_p along each block; the gap between blocks chosen so that
    the ring spans the diameter (+ 2 DOI). Positions are stepped crystal by crystal, each stored as float32
    and the next one stepped from it. Without pseudo detectors the first quarter of the ring is stepped
    and the rest follows by the ring's mirror symmetries. With ``pseudo`` every block ends in a pseudo
    detector one pitch past its last crystal (det_w_pseudo = B (cryst_per_block + 1) positions), the next
    block stepping on from the last real crystal, and the whole ring is stepped block by block, as the
    reference's MATLAB version does (its Python translation steps one block past the end of the array and
    raises IndexError)."""
    B = int(blocks_per_ring) * int(transaxial_multip)
    c = int(cryst_per_block)
    cc = c + 1 if pseudo else c          # positions per block
    d = float(diameter) + float(doi) * 2
    deg = np.arange(90, -271, -360 / B)
    h = B // 2
    w_x = c * cr_p * np.cos(np.radians(deg[:h + 1]))
    w_y = c * cr_p * np.sin(np.radians(deg[:h]))
    gx = (d - np.sum(np.abs(w_x))) / np.sum(np.cos(np.radians(deg[:h + 1]))) / 2
    gy = (d - np.sum(np.abs(w_y))) / np.sum(np.abs(np.sin(np.radians(deg[:h])))) / 2
    cs = lambda k: np.cos(np.radians(deg[k]))  # noqa: E731  (scalar calls, as the reference makes them)
    sn = lambda k: np.sin(np.radians(deg[k]))  # noqa: E731
    x = np.zeros(B * cc, dtype=np.float32)
    y = np.zeros(B * cc, dtype=np.float32)

    def block_start(k, px, py):
        # first crystal of block k: half a pitch, the gaps on both sides of the block edge, half a pitch
        return (px - (cr_p * 0.5) * cs(k) - gx * cs(k - 1) - gx * cs(k) - (cr_p * 0.5) * cs(k - 1),
                py + (cr_p * 0.5) * sn(k) + gy * sn(k - 1) + gy * sn(k) + (cr_p * 0.5) * sn(k - 1))

    def step_blocks(blocks, n, px, py):
        # every position of the listed blocks; after a block's pseudo detector the walk steps back to its
        # last real crystal
        for k in blocks:
            for m in range(1, cc + 1):
                if k > 0 and m == 1:
                    x[n], y[n] = block_start(k, px, py)
                else:
                    x[n], y[n] = px - cr_p * cs(k), py + cr_p * sn(k)
                if pseudo and m == cc:
                    px, py = x[n] + cr_p * cs(k), y[n] - cr_p * sn(k)
                else:
                    px, py = x[n], y[n]
                n += 1
        return n, px, py

    quarter = -(-B // 4)
    n, px, py = step_blocks(range(quarter), 0, d / 2., -(c / 2. + 0.5) * cr_p)
    n4 = (B * cc) // 4
    if B % 4 == 0:
        k = quarter                      # the top block, stepped without the pseudo step-back
        x[n], y[n] = block_start(k, px, py)
        px, py = x[n], y[n]
        n += 1
        for _ in range(2, cc + 1):
            x[n], y[n] = px - cr_p * cs(k), py + cr_p * sn(k)
            px, py = x[n], y[n]
            n += 1
        if pseudo:
            px, py = x[n - 1] + cr_p * cs(k), y[n - 1] - cr_p * sn(k)
            step_blocks(range(quarter + 1, B), n, px, py)
        m = n4
    elif pseudo:
        step_blocks(range(quarter, B), n, px, py)
    else:
        m = n4 + c // 2
    if not pseudo:
        x[n:n + m] = -np.flip(x[:m])
        y[n:n + m] = np.flip(y[:m])
        x[n + m:] = np.flip(x[c:(B * c) // 2])
        y[n + m:] = -y[c:(B * c) // 2]
    if flip_image:
        x, y = np.flip(x), np.flip(y)
    if offangle != 0:
        x, y = np.roll(x, int(np.round(offangle))), np.roll(y, int(np.round(offangle)))
    return x, y


def omega_sinogram_xy(x, y, n_dist: int, n_ang: int, ndist_side: int = 1) -> np.ndarray:
    """Sinogram LOR endpoints [xs, ys, xd, yd] per (radial, angle) bin, radial bin fastest, from the
    crystal coordinates of one ring (sinogramCoordinates2D, detcoord.py:589-701): every crystal pair
    (a <= b) gets angle ((a + b + det/2 + 1) mod det) / 2 and signed radial offset |b - a - det/2|
    (negative where the pair wraps), endpoints swapped at the sinogram's corners; the pairs within
    n_dist radial bins are kept (ndist_side picks the extra bin of an even n_dist). Angular mashing
    (n_ang < det / 2, detcoord.py:616-620,688-695): the det / 2 angles are binned as without it and every m
    = det / 2 / n_ang consecutive angle columns are averaged (the reference's block_reduce(..., np.mean):
    a float32 sum in column order divided by m)."""
    det = len(x)
    mashing = 1
    if n_ang < det // 2:
        mashing = det // n_ang // 2
        if n_ang * mashing != det // 2:
            raise ValueError(f"Angular mashing needs Nang to divide det_w_pseudo / 2 ({det // 2}); got Nang = {n_ang}")
        n_ang = n_ang * mashing
    a, b = np.triu_indices(det)
    L = np.stack([a, b], 1).astype(np.int64)
    hi, lo = b.astype(np.int64), a.astype(np.int64)
    j = ((hi + lo + det // 2 + 1) % det) // 2
    i = np.abs(hi - lo - det // 2)
    i = np.where((lo < j) | (j + det // 2 < hi), -i, i)
    swap = (j * 2 < -i) | (i <= (j - det // 2) * 2)
    L[swap] = L[swap][:, ::-1]
    if n_dist % 2 == 0:
        keep = (i <= n_dist // 2 + min(0, ndist_side)) & (i >= -n_dist // 2 + max(0, ndist_side))
    else:
        keep = (i <= n_dist // 2) & (i >= -(n_dist // 2))
    j = j // (det // 2 // n_ang)
    i, j, L = i[keep], j[keep], L[keep]
    if i.min() < 0:
        i = i + np.abs(i.min())
    grids = [np.zeros((n_dist, n_ang), dtype=np.float32, order="F") for _ in range(4)]
    for g, v in zip(grids, (x[L[:, 0]], y[L[:, 0]], x[L[:, 1]], y[L[:, 1]])):
        np.add.at(g, (i, j), v)
    if mashing > 1:
        def mash(g):
            acc = g[:, 0::mashing].copy()
            for q in range(1, mashing):
                acc = (acc + g[:, q::mashing]).astype(np.float32)
            return (acc / np.float32(mashing)).astype(np.float32)
        grids = [mash(g) for g in grids]
    return np.ascontiguousarray(np.stack([g.ravel("F") for g in grids], 1).astype(np.float32))


def omega_segment_table(rings: int, span: int, ring_difference: int, image_nz: int) -> np.ndarray:
    """Sinograms per segment of a span-compressed ring-pair set (projectorClass, proj.py:668-670): the
    direct segment (2 rings - 1), then each oblique segment twice (+ and - ring difference). As in the
    reference the bound uses the image's Nz."""
    first = rings * 2 - 1
    rest = np.arange(first - (span + 1), max(image_nz - ring_difference * 2, rings - ring_difference), -span * 2)
    return np.insert(np.repeat(rest, 2), 0, first).astype(np.int64)


def omega_sinogram_z(rings: int, cr_pz: float, span: int, ring_difference: int, image_nz: int,
                     ring_gaps=None, cryst_per_block_axial: int = 0) -> np.ndarray:
    """Axial endpoints [zs, zd] per sinogram (sinogramCoordinates3D, detcoord.py:703-896, single layer).
    Span 1: every ring pair (a, b), a outer. Span > 1: ring z_a = (a + 1) cr_pz - (rings + 1)
    cr_pz / 2; sinogram s holds the mean (z_a, z_b) of the ring pairs binned into it: the direct segment
    holds |a - b| <= span / 2 at position a + b, oblique segment i the differences d in
    [k_i, k_i + span - 1] (k_i = ceil(span / 2) + span i), b - a = d in the segment's first copy and a - b = d
    in its second, at position (a + b) - k_i within the segment.
    ring_gaps (mm, one per boundary between axial blocks of cryst_per_block_axial rings) lengthen the axial
    extent by their sum and shift the rings as the reference does: with span > 1 block k's rings move by
    the cumulative gap before it (the ring pitch itself stretched by sum(gaps) / (rings + 1), :726-748);
    with span 1 the second ring of a pair moves by its block's cumulative gap and the first by
    n * gaps[n - 1], n counting the multiples of cryst_per_block_axial + 1 up to the pair's first ring
    (:872-880)."""
    gaps = np.asarray(ring_gaps if ring_gaps is not None else [], dtype=np.float64).ravel()
    has_gaps = gaps.size > 0 and gaps.sum() > 0
    if has_gaps and cryst_per_block_axial <= 0:
        raise ValueError("ring gaps need cryst_per_block_axial (rings per axial block)")
    maxz = float(rings + 1) * cr_pz + (np.sum(gaps) if has_gaps else 0.0)
    if span <= 1:
        t = np.repeat(np.arange(rings), rings).astype(np.float64)
        ring = np.arange(0, cr_pz * rings, cr_pz)
        z = np.stack([cr_pz * t * 1.0, np.tile(ring, rings)], 1)
        if has_gaps:
            per_ring = np.repeat(np.insert(np.cumsum(gaps), 0, 0), cryst_per_block_axial)[:rings]
            n = 0
            for a in range(rings):
                z[a * rings:(a + 1) * rings, 1] += per_ring
                if (a + 1) % (cryst_per_block_axial + 1) == 0:
                    n += 1
                if n > 0:
                    z[a * rings:(a + 1) * rings, 0] += n * gaps[n - 1]
        return z - (maxz / 2 - cr_pz)
    zr = np.linspace(cr_pz, maxz + cr_pz, rings + 2)[:rings]
    if has_gaps:
        cg = np.cumsum(gaps)
        for kk in range(1, int(rings / cryst_per_block_axial)):
            zr[cryst_per_block_axial * kk:cryst_per_block_axial * (kk + 1)] += cg[kk - 1]
    zr = zr - maxz / 2
    seg = omega_segment_table(rings, span, ring_difference, image_nz)
    off = np.cumsum(seg)
    tot = int(off[-1])
    z = np.zeros((tot, 2))
    cnt = np.zeros(tot)
    nz = rings * 2 - 1
    z[:nz + 1:2, 0] = zr
    z[:nz + 1:2, 1] = zr
    cnt[:nz + 1:2] = 1
    for e in range(1, span // 2 + 1):              # direct segment, |a - b| = e: both orders per position
        lo = np.arange(rings - e)
        p = 2 * lo + e
        z[p, 0] += zr[lo + e] + zr[lo]
        z[p, 1] += zr[lo] + zr[lo + e]
        cnt[p] += 2
    k = [int(np.ceil(span / 2)) + span * i for i in range((ring_difference - int(np.ceil(span / 2))) // span + 1)]
    for si in range(len(seg) // 2):
        for dd in range(k[si], k[si] + span):
            lo = np.arange(rings - dd)
            pos = lo * 2 + dd - k[si]
            z[off[2 * si] + pos, 0] += zr[lo + dd]     # first copy: pairs (a = b + d, b)
            z[off[2 * si] + pos, 1] += zr[lo]
            cnt[off[2 * si] + pos] += 1
            z[off[2 * si + 1] + pos, 0] += zr[lo]      # second copy: pairs (a, b = a + d)
            z[off[2 * si + 1] + pos, 1] += zr[lo + dd]
            cnt[off[2 * si + 1] + pos] += 1
    cnt[cnt == 0] = 1
    return z / cnt[:, None]


# ---- volume-of-intersection projector (projector type 3) ------------------------------------------------

def sphere_cylinder_volume(R: float, r: float, b):
    """Volume of the intersection of a sphere of radius r whose centre lies at distance b from the axis
    of an infinite cylinder of radius R (the closed form in complete elliptic integrals of the first and
    second kind and a third-kind integral evaluated by quadrature, as the reference uses it,
    coordinates.py:150-195). b: array; returns float32."""
    import warnings
    from scipy.integrate import quad
    from scipy.special import ellipe, ellipk
    with warnings.catch_warnings(), np.errstate(divide="ignore", invalid="ignore"):
        warnings.simplefilter("ignore")   # b = R divides by zero and quad warns there, as in the reference
        return _sphere_cylinder_volume(R, r, np.asarray(b), quad, ellipk, ellipe)


def _sphere_cylinder_volume(R, r, b, quad, ellipk, ellipe):
    A = np.maximum(r ** 2, (b + R) ** 2)
    B = np.minimum(r ** 2, (b + R) ** 2)
    C = (b - R) ** 2
    k = np.abs((B - C) / (A - C))
    alpha = (B - C) / C
    s = (b + R) * (b - R)
    third = np.array([quad(lambda x, al=al, kk=kk: 1. / ((1. + al * x ** 2) * np.sqrt(1. - x ** 2)
                                                          * np.sqrt(1. - kk * x ** 2)), 0, 1)[0]
                      for al, kk in zip(np.ravel(alpha), np.ravel(k))]).reshape(np.shape(alpha))
    K, E = ellipk(k), ellipe(k)
    step = np.zeros(b.shape[0], dtype=np.float32)
    step[R - b > 0] = 1
    step[R - b == 0] = 0.5
    V = np.zeros(b.shape[0], dtype=np.float32)
    sph = (4. * np.pi) / 3. * r ** 3
    out = r > b + R
    if np.any(out):
        a_, b_, c_, s_ = A[out], B[out], C[out], s[out]
        V[out] = sph * step[out] + 4. / (3. * np.sqrt(a_ - c_)) * (
            third[out] * (a_ ** 2 * s_) / c_ - K[out] * (a_ * s_ - ((a_ - b_) * (a_ - c_)) / 3.)
            - E[out] * (a_ - c_) * (s_ + (4. * a_ - 2. * b_ - 2. * c_) / 3.))
    inn = r < b + R
    if np.any(inn):
        a_, b_, c_, s_ = A[inn], B[inn], C[inn], s[inn]
        V[inn] = sph * step[inn] + ((4. / 3.) / (np.sqrt(a_ - c_))) * (
            third[inn] * ((b_ ** 2 * s_) / c_) + K[inn] * (s_ * (a_ - 2. * b_) + (a_ - b_) * ((3. * b_ - c_ - 2 * a_) / 3.))
            + E[inn] * (a_ - c_) * (-s_ + (2. * a_ - 4. * b_ + 2. * c_) / 3.))
    return V


def voxel_volume_table(tube_radius: float, voxel_radius: float, dx: float, dy: float, dz: float,
                       return_grid: bool = False):
    """Look-up table of the volume projector (computeVoxelVolumes, coordinates.py:133-207): a voxel is a
    sphere of radius sqrt(2) * voxel_radius * max(d) / 2; V[k] is its volume of intersection with the tube
    (radius tube_radius) at axis distance bmin + k * 1e-3 mm (the kernel indexes it by
    rint((d - bmin) * 1e3)); closer than bmin the voxel counts Vmax, beyond bmax nothing.
    Returns (V float32, bmin, bmax, Vmax, effective voxel radius), and with return_grid the distances
    the entries were evaluated at: the reference drops the entries where the closed form is NaN or not
    decreasing (at and next to b = tube_radius), so past that point V[k] holds the volume at a distance
    one or more 1e-3 steps beyond bmin + k * 1e-3, and the kernel's lookup reads it there as well."""
    # numpy float64 scalars where the reference has them (they set the dtype of the table's arithmetic)
    dp = np.max(np.array([dx, dy, dz], dtype=np.float64))
    vr = math.sqrt(2.) * voxel_radius * (dp / 2.)
    bmax = tube_radius + vr
    b = np.linspace(0, bmax, 10000, dtype=np.float32)
    b = b[tube_radius <= (b + vr)]
    b = np.unique(np.round(b * 10 ** 3) / 10 ** 3)
    V = sphere_cylinder_volume(tube_radius, vr, b)
    keep = np.append(np.diff(V), 0) <= 0
    b = b[keep]
    V = np.abs(V[keep]).astype(np.float32)
    vmax = (4 * np.pi) / 3 * vr ** 3
    if return_grid:
        return V, float(np.min(b)), float(bmax), float(vmax), vr, b
    return V, float(np.min(b)), float(bmax), float(vmax), vr


def listmode_sensitivity_tables(crystal_xy, ring_z):
    """LOR tables of the list-mode sensitivity image (the reference's -DSENS kernel: every ordered pair of
    crystals (i, j) in every ordered pair of rings (a, b), projectorType123.cl:283-300,426-432,
    getDetectorCoordinatesRaw general_opencl_functions.h:1527-1543) laid out as a sinogram: bin i + j * det
    holds (x_i, y_i, x_j, y_j) and plane a * rings + b holds (z_a, z_b), so measurement
    i + j * det + (a * rings + b) * det^2 is the kernel's work-item (i, j, a * rings + b).
    crystal_xy: (det, 2) crystal centres; ring_z: (rings,) ring positions. Returns (x table, z table)."""
    c = np.ascontiguousarray(crystal_xy, dtype=np.float32).reshape(-1, 2)
    zr = np.ascontiguousarray(ring_z, dtype=np.float32).ravel()
    det = c.shape[0]
    i = np.tile(np.arange(det), det)
    j = np.repeat(np.arange(det), det)
    x = np.stack([c[i, 0], c[i, 1], c[j, 0], c[j, 1]], axis=1).ravel()
    a = np.repeat(np.arange(zr.size), zr.size)
    b = np.tile(np.arange(zr.size), zr.size)
    z = np.stack([zr[a], zr[b]], axis=1).ravel()
    return np.ascontiguousarray(x), np.ascontiguousarray(z)
