"""The oracle's interpolation projector (projector type 4) against brute force in float64.

The forward projection samples the ray's line inside the box every dL (or, with dL <= 0, at every
crossing of a voxel-centre plane of the dominant axis: the Joseph variant) and sums stepLen times the
image's trilinear interpolant with clamp-to-edge. Brute force here: the sample positions from a float64
enumeration (steps from the entry, or the plane crossings themselves), and the interpolant as a sum over
all voxels of the edge-padded image times products of tent functions -- no floor/fraction arithmetic. The
emission backprojection deposits stepLen temp y into the voxel containing each sample (voxel 0 skipped,
the reference's quirk); the CT backprojection is checked voxel by voxel: the ray from each source through
the voxel centre meets the panel plane, where the panel is interpolated bilinearly (again with tents on the
edge-padded panel), weighted L^3 / |v|^2 kerroin. Attenuation factors: exp(-sum stepLen mu(p)).
"""
import numpy as np
import pytest

from omega_amd import geometry as G
from oracle import oracle as O
from tests import cases


def _grid():
    return cases.tiny_grid()


def _outside_rays(grid, n, seed):
    """Chords through uniform points of the box with both endpoints 30 mm outside it."""
    rng = np.random.default_rng(seed)
    lo = np.array([grid.bx, grid.by, grid.bz])
    ext = np.array([grid.Nx * grid.dx, grid.Ny * grid.dy, grid.Nz * grid.dz])
    p = lo + rng.random((n, 3)) * ext
    u = rng.normal(size=(n, 3))
    u[: n // 8, 2] = 0.0                     # some in-plane rays
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    return np.concatenate([p - 30 * u, p + 30 * u], axis=1).astype(np.float32)


def _box(grid):
    b = np.array([grid.bx, grid.by, grid.bz], np.float64)
    d = np.array([grid.dx, grid.dy, grid.dz], np.float64)
    n = np.array([grid.Nx, grid.Ny, grid.Nz])
    return b, d, n


def _tent_interp(img, grid, p):
    """Trilinear interpolant with clamp-to-edge at world points p (M, 3): edge-padded image, tent products."""
    b, d, n = _box(grid)
    v = np.pad(img.reshape(grid.Nz, grid.Ny, grid.Nx).astype(np.float64), 1, mode="edge")
    u = (p - b) / d - 0.5 + 1.0          # padded voxel coordinates of the points
    u = np.clip(u, 0.0, np.array(v.shape[::-1], np.float64) - 1.0)   # beyond the ring: the edge values
    out = np.zeros(len(p))
    for k in range(v.shape[0]):
        wz = np.maximum(0.0, 1.0 - np.abs(u[:, 2] - k))
        if not wz.any():
            continue
        for j in range(v.shape[1]):
            wy = np.maximum(0.0, 1.0 - np.abs(u[:, 1] - j)) * wz
            if not wy.any():
                continue
            for i in range(v.shape[2]):
                w = np.maximum(0.0, 1.0 - np.abs(u[:, 0] - i)) * wy
                out += w * v[k, j, i]
    return out


def _line_box(grid, s, v):
    b, d, n = _box(grid)
    with np.errstate(divide="ignore", invalid="ignore"):
        t1 = (b - s) / v
        t2 = (b + n * d - s) / v
    tmin = np.nanmax(np.fmin(t1, t2))
    tmax = np.nanmin(np.fmax(t1, t2))
    return tmin, tmax


def _samples64(grid, ray, dL):
    """World sample points and step length of one ray (float64), or None when the line misses the box."""
    s = ray[:3].astype(np.float64)
    v = ray[3:].astype(np.float64) - s
    t0, t1 = _line_box(grid, s, v)
    if not t0 < t1:
        return None
    L = np.linalg.norm(v)
    b, d, n = _box(grid)
    if dL > 0:
        step = dL / L
        m = int(np.floor((t1 - t0) / step)) + 1
        ts = t0 + step * np.arange(m)
        return s + np.outer(ts, v), dL, L
    # Joseph: the dominant axis has the most voxel planes crossed; sample where the line crosses its voxel-centre
    # planes, starting at the centre at or below the entry coordinate (behind the entry when the line runs up
    # that axis, ahead of it when it runs down), until the exit
    cr = np.abs(v) / d
    a = int(np.argmax(cr))
    entry = (s[a] + v[a] * t0 - b[a]) / d[a] - 0.5        # entry coordinate in voxel-centre units
    k = np.floor(entry) + np.sign(v[a]) * np.arange(0, n[a] + 3)
    ts = (b[a] + (k + 0.5) * d[a] - s[a]) / v[a]
    ts = ts[ts <= t1]
    step = d[a] / abs(v[a])
    return s + np.outer(ts, v), step * L, L


@pytest.mark.parametrize("dL", [0.4, 1.3, -1.0])
def test_forward_projection_brute_force(dL):
    grid = _grid()
    rays = _outside_rays(grid, 120, seed=3)
    img = (np.random.default_rng(4).random(grid.N) + 0.2).astype(np.float32)
    case = O.InterpCase(grid, 1, 1, rays.ravel(), np.zeros(2, np.float32), dL, mode=O.MODE_LIST)
    y = case.fp(img, len(rays))
    checked = 0
    for i, r in enumerate(rays):
        smp = _samples64(grid, r, dL)
        if smp is None:
            assert y[i] == 0.0
            continue
        pts, step_len, L = smp
        ref = step_len * _tent_interp(img, grid, pts).sum() / L
        assert abs(y[i] - ref) <= 2e-5 * abs(ref) + 1e-7, f"ray {i}: {y[i]} vs {ref}"
        checked += 1
    assert checked > 100


def test_constant_image_counts_steps():
    """A uniform image of 1 projects to nSteps * dL / L: the sample count from the clipped line."""
    grid = _grid()
    rays = _outside_rays(grid, 200, seed=5)
    case = O.InterpCase(grid, 1, 1, rays.ravel(), np.zeros(2, np.float32), 0.25, mode=O.MODE_LIST)
    y = case.fp(np.ones(grid.N, np.float32), len(rays)).astype(np.float64)
    for i, r in enumerate(rays):
        smp = _samples64(grid, r, 0.25)
        if smp is None:
            continue
        pts, step_len, L = smp
        assert abs(y[i] * L - len(pts) * 0.25) < 1e-4 * len(pts) * 0.25


def test_emission_backprojection_brute_force():
    grid = _grid()
    rays = _outside_rays(grid, 150, seed=6)
    n = len(rays)
    yin = (np.random.default_rng(7).random(n) + 0.5).astype(np.float32)
    case = O.InterpCase(grid, 1, 1, rays.ravel(), np.zeros(2, np.float32), 0.5, mode=O.MODE_LIST)
    rhs, sens = case.bp(yin, n, sens=True)
    b, d, nn = _box(grid)
    ref = np.zeros(grid.N)
    refs = np.zeros(grid.N)
    amb = np.zeros(grid.N)      # deposits of samples within 1e-4 voxel of a voxel face: either side is right
    for i, r in enumerate(rays):
        smp = _samples64(grid, r, 0.5)
        if smp is None:
            continue
        pts, step_len, L = smp
        u = (pts - b) / d
        ijk = np.maximum(np.floor(u).astype(np.int64), 0)
        ind = ijk[:, 0] + ijk[:, 1] * grid.Nx + ijk[:, 2] * grid.Nx * grid.Ny
        near = np.any(np.abs(u - np.round(u)) < 1e-4, axis=1)
        ok = (ind > 0) & (ind < grid.N)
        np.add.at(ref, ind[ok], yin[i] * step_len / L)
        np.add.at(refs, ind[ok], step_len / L)
        for q in np.nonzero(near)[0]:
            lo_ = np.maximum(np.floor(u[q] - 1e-4).astype(np.int64), 0)
            hi_ = np.maximum(np.floor(u[q] + 1e-4).astype(np.int64), 0)
            for c in {int(lo_[0] + lo_[1] * grid.Nx + lo_[2] * grid.Nx * grid.Ny),
                      int(hi_[0] + hi_[1] * grid.Nx + hi_[2] * grid.Nx * grid.Ny)}:
                if 0 < c < grid.N:
                    amb[c] += yin[i] * step_len / L
    tol = 1e-5 * np.abs(ref).max()
    assert np.all(np.abs(rhs - ref) <= amb + tol)
    assert np.all(np.abs(sens - refs) <= amb / yin.min() + 1e-5 * np.abs(refs).max())
    assert amb.sum() < 0.1 * ref.sum()   # mostly each ray's first sample, on the entry face


def _ct_case(n_views=6, rows=10, cols=8, pitch=1.2):
    grid = _grid()
    x, uv = G.cbct_tables(n_views, sod=30.0, sdd=55.0, pitch_u=pitch)
    sdd = 55.0
    kerroin = grid.dx * grid.dy * grid.dz / (pitch * pitch * sdd)
    case = O.InterpCase(grid, rows, cols, x, uv, 0.4, mode=O.MODE_CT, dPitchZ=pitch, kerroin=kerroin, dPitchX=pitch,
                        dPitchY=pitch)
    return grid, x, uv, rows, cols, pitch, kerroin, case


def test_ct_backprojection_brute_force():
    grid, x, uv, rows, cols, pitch, kerroin, case = _ct_case()
    nv = x.size // 6
    M = rows * cols * nv
    yin = (np.random.default_rng(8).random(M) + 0.1).astype(np.float32)
    out, sens = case.bp(yin, M, sens=True)
    b, d, n = _box(grid)
    zc = b[2] + (np.arange(grid.Nz) + 0.5) * d[2]
    yc = b[1] + (np.arange(grid.Ny) + 0.5) * d[1]
    xc = b[0] + (np.arange(grid.Nx) + 0.5) * d[0]
    Z, Y, X = np.meshgrid(zc, yc, xc, indexing="ij")
    vox = np.stack([X, Y, Z], -1).reshape(-1, 3)
    ref = np.zeros(grid.N)
    refs = np.zeros(grid.N)
    panels = yin.reshape(nv, cols, rows).astype(np.float64)
    for k in range(nv):
        s = x[6 * k:6 * k + 3].astype(np.float64)
        d1 = x[6 * k + 3:6 * k + 6].astype(np.float64)
        ux = np.array([uv[2 * k], uv[2 * k + 1], 0.0]) / pitch          # unit row direction
        uz = np.array([0.0, 0.0, 1.0])
        corner = d1 - ux * pitch * rows / 2 - uz * pitch * cols / 2
        nrm = np.cross(ux, uz)
        v = vox - s
        t = ((d1 - s) @ nrm) / (v @ nrm)
        p = s + v * t[:, None]
        px = ((p - corner) @ ux) / (pitch * rows)
        py = ((p - corner) @ uz) / (pitch * cols)
        inside = (px >= 0) & (px <= 1) & (py >= 0) & (py <= 1)
        # bilinear interpolant of the panel (rows fastest) with clamp to edge: tents on the padded panel
        pad = np.pad(panels[k], 1, mode="edge")
        ur = np.clip(px * rows - 0.5 + 1.0, 0.0, rows + 1.0)
        uc = np.clip(py * cols - 0.5 + 1.0, 0.0, cols + 1.0)
        val = np.zeros(len(vox))
        for c in range(pad.shape[0]):
            wc = np.maximum(0.0, 1.0 - np.abs(uc - c))
            for r_ in range(pad.shape[1]):
                val += wc * np.maximum(0.0, 1.0 - np.abs(ur - r_)) * pad[c, r_]
        L = np.linalg.norm(p - s, axis=1)
        w = L ** 3 / np.sum(v * v, axis=1) * kerroin
        use = inside & (val != 0)
        ref[use] += (val * w)[use]
        refs[use] += w[use]
    assert np.linalg.norm(out - ref) <= 1e-5 * np.linalg.norm(ref)
    assert np.linalg.norm(sens - refs) <= 1e-5 * np.linalg.norm(refs)


def test_attenuation_factors_brute_force():
    grid = _grid()
    rays = _outside_rays(grid, 60, seed=9)
    mu = (np.random.default_rng(10).random(grid.N) * 0.02).astype(np.float32)
    case = O.InterpCase(grid, 1, 1, rays.ravel(), np.zeros(2, np.float32), 0.5, mode=O.MODE_LIST)
    f = case.atten_factors(mu, len(rays))
    for i, r in enumerate(rays):
        smp = _samples64(grid, r, 0.5)
        if smp is None:
            assert f[i] == 1.0
            continue
        pts, step_len, L = smp
        ref = np.exp(-step_len * _tent_interp(mu, grid, pts).sum())
        assert abs(f[i] - ref) <= 2e-6 * ref
