"""Batched dynamic-study OSEM (all frames per launch, hipGraph replay, autotuned launch shape).

Bar: after a full iteration every frame equals the single-frame fused OSEM path
(omg_osem_subset_iteration, itself at oracle parity) within 1e-5 relative L2 -- both compute the same
per-frame sums, only the order of the float atomics differs; every (rays-per-block,
frames-per-tile) launch shape agrees to 1e-5; each of three graph replays, with eager work in
between, matches the eager iteration to 1e-5; one subset step of frame 0 matches the oracle's
FP -> ratio -> BP -> update to 1e-5.
"""
import ctypes as C

import numpy as np
import pytest

from omega_amd import _lib, phantoms, workloads
from omega_amd.dynamic import FramesOSEM, FramesTuner
from oracle import oracle as O
from tests.hip_helpers import ptr, rel_l2

pytestmark = pytest.mark.gpu

F = 8


@pytest.fixture(scope="module")
def study(cuda):
    import torch
    p = workloads.pet3d_projector(subsets=4, n_rings=8, crystals=96, max_ring_diff=7, radius=150.0,
                                  image=(40, 40, 15), fov_xy=160.0)
    p.initProj()
    frames = phantoms.dynamic_ellipsoids(p.grid, seed=5, durations=phantoms.FRAME_DURATIONS[::3])
    assert frames.shape[1] == F
    counts = workloads.frame_counts(F, seed=5, lo=2e4, hi=2e5)
    y = workloads.dynamic_data(p, frames, counts, seed=5, device=cuda)
    print(f"[frames study] N={p.grid.N} M={y.shape[1]} F={F}", flush=True)
    return p, y


def _single_frame_iteration(p, y, cuda):
    import torch
    N = p.grid.N
    out = np.zeros((F, N), np.float32)
    for f in range(F):
        x = torch.ones(N, dtype=torch.float32, device=cuda)
        S = torch.zeros((p.subsets, N), dtype=torch.float32, device=cuda)
        yf = y[f].contiguous()
        for k in range(p.subsets):
            first, n, off = int(p.nMeas[k]), int(p.nMeasSubset[k, 0]), int(p.nTotMeas[k])
            _lib.call("omg_osem_subset_iteration", p.geometry.handle, first, n, ptr(x), ptr(yf[off:off + n]),
                      ptr(S[k]), 1, 1e-5, 0.0, 2.0, 1, None)
        out[f] = x.cpu().numpy()
    return out


def test_frames_match_single_frame_path(study, cuda):
    p, y = study
    ref = _single_frame_iteration(p, y, cuda)
    for rpb, ft, gm in [(256, 8, 0), (64, 1, 1 << 30), (128, 4, 0), (64, 2, 1 << 30)]:
        run = FramesOSEM(p, y)
        run.iteration(rpb, ft, gm)
        got = run.x.cpu().numpy()
        for f in range(F):
            assert rel_l2(got[f], ref[f]) <= 1e-5, f"rpb={rpb} ft={ft} grid={gm} frame {f}"
        run.close()


def test_graph_replays_match_eager(study, cuda):
    import torch
    p, y = study
    eager = FramesOSEM(p, y)
    graph = FramesOSEM(p, y)
    for it in range(3):
        eager.iteration(128, 8)
        graph.replay(128, 8)
        # eager work of the program between replays
        tmp = torch.randn(1 << 20, device=cuda)
        float(tmp.sum())
        torch.cuda.synchronize()
        e, g = eager.x.cpu().numpy(), graph.x.cpu().numpy()
        assert rel_l2(g, e) <= 1e-5, f"replay {it}"
        assert float(graph.rhs.abs().max()) == 0.0
    eager.close()
    graph.close()


def test_step_matches_oracle_frame0(study, cuda):
    p, y = study
    run = FramesOSEM(p, y)
    run.step(0, 256, 8)
    k = 0
    first, n, off = int(p.nMeas[k]), int(p.nMeasSubset[k, 0]), int(p.nTotMeas[k])
    case = O.SiddonCase(p.grid, int(p.nRowsD), int(p.nColsD), p.x, p.z, plane_offset=first)
    x = np.ones(p.grid.N, np.float32)
    meas = y[0, off:off + n].cpu().numpy()
    r = case.fp(x, n)
    O.lib().oracle_ratio_f32(meas, r, n, 1e-5)
    rhs, sens = case.bp(r, n, sens=True)
    O.lib().oracle_sens_clamp_f32(sens, p.grid.N, 1e-5)
    O.lib().oracle_osem_update_f32(x, sens, rhs, p.grid.N, 1e-5, 1)
    assert rel_l2(run.S[0].cpu().numpy(), sens) <= 1e-5
    assert rel_l2(run.x[0].cpu().numpy(), x) <= 1e-5
    run.close()


def test_autotuner_picks_a_valid_shape(study, tmp_path):
    p, y = study
    run = FramesOSEM(p, y)
    tuner = FramesTuner(str(tmp_path / "tune.json"))
    best = tuner.tune(run, k=1, reps=2)
    assert best in tuner.last_timings and len(tuner.last_timings) == 24
    assert min(tuner.last_timings.values()) == tuner.last_timings[best]
    again = FramesTuner(str(tmp_path / "tune.json"))
    assert again.tune(run, k=1) == best
    assert float(run.x.min()) == 1.0 and float(run.x.max()) == 1.0   # tuning leaves the study untouched
    run.close()


def test_rejects_bad_launch_shapes(study):
    p, y = study
    run = FramesOSEM(p, y)
    run.sensitivity()
    with pytest.raises(_lib.OmegaHipError):
        run.step(0, 96, 8)
    with pytest.raises(_lib.OmegaHipError):
        run.step(0, 256, 3)
    run.close()


def test_osl_rdp_frames_match_single_frame(study, cuda):
    """OSL-RDP: one subset step to 1e-5; a full iteration to 5e-5 -- the OSL update divides by S + beta grad R,
    which amplifies the float-atomic ordering differences of the two paths where the prior gradient
    nearly cancels the sensitivity."""
    import torch
    p, y = study
    N = p.grid.N
    beta, gamma = 0.1, 2.0
    for n_steps, tol in ((1, 1e-5), (p.subsets, 5e-5)):
        run = FramesOSEM(p, y, beta=beta, gamma=gamma)
        for k in range(n_steps):
            run.step(k, 128, 4)
        got = run.x.cpu().numpy()
        run.close()
        for f in (0, 5):
            x = torch.ones(N, dtype=torch.float32, device=cuda)
            S = torch.zeros((p.subsets, N), dtype=torch.float32, device=cuda)
            yf = y[f].contiguous()
            for k in range(n_steps):
                first, n, off = int(p.nMeas[k]), int(p.nMeasSubset[k, 0]), int(p.nTotMeas[k])
                _lib.call("omg_osem_subset_iteration", p.geometry.handle, first, n, ptr(x), ptr(yf[off:off + n]),
                          ptr(S[k]), 1, 1e-5, beta, gamma, 1, None)
            assert rel_l2(got[f], x.cpu().numpy()) <= tol, f"frame {f} after {n_steps} subset steps"


@pytest.fixture(scope="module")
def study_angles(cuda):
    """The same study with angle subsets (type 4: every 4th angle of every plane, through the measurement
    index table): every subset covers every voxel."""
    p = workloads.pet3d_projector(subsets=4, n_rings=8, crystals=96, max_ring_diff=7, radius=150.0,
                                  image=(40, 40, 15), fov_xy=160.0, subset_type=4)
    p.initProj()
    assert p.index is not None and p.index.size == int(p.nTotMeas[-1])
    frames = phantoms.dynamic_ellipsoids(p.grid, seed=6, durations=phantoms.FRAME_DURATIONS[::6])
    counts = workloads.frame_counts(frames.shape[1], seed=6, lo=2e4, hi=2e5)
    y = workloads.dynamic_data(p, frames, counts, seed=6, device=cuda)
    print(f"[frames study, angle subsets] N={p.grid.N} M={y.shape[1]} F={y.shape[0]}", flush=True)
    return p, y


def test_angle_subsets_match_single_frame_and_oracle(study_angles, cuda):
    """Subset type 4: one OSEM iteration and one OSL-RDP subset step of every frame against the single-frame
    path, and frame 0's first step against the oracle on the subset's LORs as list-mode rays."""
    import torch
    p, y = study_angles
    Fa, N = y.shape[0], p.grid.N
    for beta, n_steps in ((0.0, p.subsets), (0.1, 1)):
        run = FramesOSEM(p, y, beta=beta)
        for k in range(n_steps):
            run.step(k, 128, 4 if Fa % 4 == 0 else 1)
        got = run.x.cpu().numpy()
        run.close()
        for f in range(Fa):
            x = torch.ones(N, dtype=torch.float32, device=cuda)
            S = torch.zeros((p.subsets, N), dtype=torch.float32, device=cuda)
            yf = y[f].contiguous()
            for k in range(n_steps):
                first, n, off = int(p.nMeas[k]), int(p.nMeasSubset[k, 0]), int(p.nTotMeas[k])
                _lib.call("omg_osem_subset_iteration", p.geometry.handle, first, n, ptr(x), ptr(yf[off:off + n]),
                          ptr(S[k]), 1, 1e-5, beta, 2.0, 1, None)
            assert rel_l2(got[f], x.cpu().numpy()) <= 1e-5, f"beta={beta} frame {f} after {n_steps} steps"
    run = FramesOSEM(p, y)
    run.step(0, 256, 1)
    per = int(p.nRowsD) * int(p.nColsD)
    bins = p.index[int(p.nMeas[0]):int(p.nMeas[1])]
    xy = p.x.reshape(-1, 4)[bins % per]
    zz = p.z.reshape(-1, 2)[bins // per]
    rays = np.ascontiguousarray(np.stack([xy[:, 0], xy[:, 1], zz[:, 0], xy[:, 2], xy[:, 3], zz[:, 1]], -1)
                                .astype(np.float32).ravel())
    n = rays.size // 6
    case = O.SiddonCase(p.grid, 1, 1, rays, np.zeros(2, np.float32), mode=O.MODE_LIST)
    x = np.ones(N, np.float32)
    r = case.fp(x, n)
    O.lib().oracle_ratio_f32(y[0, :n].cpu().numpy(), r, n, 1e-5)
    rhs, sens = case.bp(r, n, sens=True)
    O.lib().oracle_sens_clamp_f32(sens, N, 1e-5)
    O.lib().oracle_osem_update_f32(x, sens, rhs, N, 1e-5, 1)
    assert rel_l2(run.S[0].cpu().numpy(), sens) <= 1e-5
    assert rel_l2(run.x[0].cpu().numpy(), x) <= 1e-5
    # the voxels the data sees are seen by every angle subset, up to grazing edge voxels (type 8 plane
    # subsets leave whole end slices out of some subsets)
    seen = run.S > 1e-5
    assert float((seen.any(0) & ~seen.all(0)).float().mean()) < 0.01
    run.close()
