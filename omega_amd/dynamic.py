"""Batched dynamic-study OSEM: all frames of a study per launch, hipGraph-captured iterations, autotuned launch.

The reference reconstructs the frames (timesteps) of a dynamic study one after another inside every
subset (source/cpp/reconstructionAF.h:1089-1453; per-frame measurements :429-436, frame tiling of
``nMeas`` in source/Python/omegatomo/projector/indices.py:271-272). Here the frames share one
geometry handle and advance together:

* ``x`` is a frame-major image stack ``[F, N]`` and ``meas`` a frame-major measurement stack ``[F, M]``,
  each frame in the projector's subset order (the order of ``projectorClass.nTotMeas``), so every frame
  keeps the cache footprint of a single-frame reconstruction;
* one ``omg_frames_osem_step`` launch performs a subset step of every frame (OSEM, or OSL-OSEM with the
  RDP prior when ``beta != 0``); a persistent grid walks (frame tile, LOR block) work items, so
  rays-per-block and frames-per-tile are launch parameters;
* an iteration (all subset steps) is captured once with ``omg_frames_graph_create`` and replayed;
* :class:`FramesTuner` times every (rays-per-block, frames-per-tile) pair on the real data and keeps
  the fastest, per geometry and frame count.
"""
from __future__ import annotations

import ctypes as C
import json
import os

import numpy as np

from . import _lib

RAYS_PER_BLOCK = (64, 128, 256)
FRAMES_PER_TILE = (1, 2, 4, 8)
# persistent grid (one resident wave of workgroups walking the work items) or one workgroup per item
GRID_MODES = (0, 1 << 30)


def _p(t):
    return C.c_void_p(0 if t is None else t.data_ptr())


class FramesOSEM:
    """OSEM of ``F`` frames that share the geometry of ``proj`` (improved Siddon, emission data)."""

    def __init__(self, proj, meas, epps: float = 1e-5, enforce_positivity: bool = True, x0=None, beta: float = 0.0,
                 gamma: float = 2.0, support=None):
        """``support`` (optional bool device tensor [N]): the voxels that are reconstructed. The others have
        their sensitivity set to 1e30, so every OSEM / OSL update sends them below ``epps``, where the
        positivity clamp holds them; they then add epps-level terms to the projections and act as near-zero
        neighbours for the prior."""
        import torch
        self.torch = torch
        if not proj.projectorInitialized:
            proj.initProj()
        self.proj = proj
        self.g = proj.geometry.handle
        self.N = int(proj.N[0])
        if meas.dim() != 2 or meas.dtype != torch.float32 or not meas.is_cuda:
            raise ValueError("meas must be a float32 device tensor of shape [n_frames, n_measurements]")
        if meas.shape[1] != int(proj.nTotMeas[-1]):
            raise ValueError(f"meas has {meas.shape[1]} columns, the projector has {int(proj.nTotMeas[-1])} "
                             "measurements")
        self.meas = meas.contiguous()
        self.F = int(meas.shape[0])
        self.M = int(meas.shape[1])
        dev = meas.device
        self.x = torch.ones((self.F, self.N), dtype=torch.float32, device=dev) if x0 is None else \
            x0.reshape(self.F, self.N).to(torch.float32).contiguous().clone()
        self.S = torch.zeros((proj.subsets, self.N), dtype=torch.float32, device=dev)
        self.rhs = torch.zeros((self.F, self.N), dtype=torch.float32, device=dev)
        self.beta, self.gamma = float(beta), float(gamma)
        self.dU = torch.zeros((self.F, self.N), dtype=torch.float32, device=dev) if self.beta != 0.0 else None
        self.epps = float(epps)
        self.pos = int(bool(enforce_positivity))
        if support is not None:
            if support.dtype != torch.bool or support.numel() != self.N or support.device != dev:
                raise ValueError("support must be a bool tensor of N voxels on the measurements' device")
            if not self.pos:
                raise ValueError("a support mask relies on the positivity clamp")
        self.support = support
        self.sens_done = False
        self._graphs = {}

    # ---- subset bookkeeping -------------------------------------------------------------------------
    def _subset(self, k):
        p = self.proj
        return int(p.nMeas[k]), int(p.nMeasSubset[k, 0]), int(p.nTotMeas[k])

    def _stream(self):
        return C.c_void_p(self.torch.cuda.current_stream().cuda_stream)

    def sensitivity(self) -> None:
        """S_k = max(A_k^T 1, epps) for every subset (frame-independent, computed once)."""
        for k in range(self.proj.subsets):
            first, n, _ = self._subset(k)
            _lib.call("omg_frames_sensitivity", self.g, first, n, _p(self.S[k]), self.epps, self._stream())
        if self.support is not None:
            self.S[:, ~self.support] = 1e30
        self.sens_done = True

    def step(self, k: int, rays_per_block: int = 256, frames_per_tile: int = 8, grid_blocks: int = 0, x=None) -> None:
        """One subset step of every frame (eager launch)."""
        if not self.sens_done:
            self.sensitivity()
        first, n, off = self._subset(k)
        x = self.x if x is None else x
        _lib.call("omg_frames_osem_step", self.g, first, n, self.F, _p(x), _p(self.meas[0, off:]), self.M,
                  _p(self.S[k]), _p(self.rhs), _p(self.dU), self.beta, self.gamma, self.epps, self.pos,
                  int(rays_per_block), int(frames_per_tile), int(grid_blocks), self._stream())

    def iteration(self, rays_per_block: int = 256, frames_per_tile: int = 8, grid_blocks: int = 0) -> None:
        for k in range(self.proj.subsets):
            self.step(k, rays_per_block, frames_per_tile, grid_blocks)

    # ---- graph capture ------------------------------------------------------------------------------
    def capture(self, rays_per_block: int = 256, frames_per_tile: int = 8, grid_blocks: int = 0):
        """Capture one iteration (every subset step of every frame) as a hipGraph."""
        key = (int(rays_per_block), int(frames_per_tile), int(grid_blocks))
        if key in self._graphs:
            return self._graphs[key]
        if not self.sens_done:
            self.sensitivity()
        self.torch.cuda.synchronize()
        K = self.proj.subsets
        first = np.array([int(self.proj.nMeas[k]) for k in range(K)], dtype=np.int64)
        n = np.array([int(self.proj.nMeasSubset[k, 0]) for k in range(K)], dtype=np.int64)
        off = np.array([int(self.proj.nTotMeas[k]) for k in range(K)], dtype=np.int64)
        h = C.c_void_p()
        _lib.call("omg_frames_graph_create", self.g, K, first.ctypes.data, n.ctypes.data, off.ctypes.data, self.F,
                  _p(self.x), _p(self.meas), self.M, _p(self.S), _p(self.rhs), _p(self.dU), self.beta, self.gamma,
                  self.epps, self.pos, key[0], key[1], key[2], C.byref(h))
        self._graphs[key] = h
        return h

    def replay(self, rays_per_block: int = 256, frames_per_tile: int = 8, grid_blocks: int = 0) -> None:
        h = self.capture(rays_per_block, frames_per_tile, grid_blocks)
        _lib.call("omg_graph_launch", h, self._stream())

    def run(self, n_iter: int, rays_per_block: int = 256, frames_per_tile: int = 8, graph: bool = True,
            grid_blocks: int = 0) -> None:
        for _ in range(n_iter):
            if graph:
                self.replay(rays_per_block, frames_per_tile, grid_blocks)
            else:
                self.iteration(rays_per_block, frames_per_tile, grid_blocks)

    def close(self) -> None:
        for h in self._graphs.values():
            _lib.call("omg_graph_destroy", h)
        self._graphs.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FramesTuner:
    """Picks rays-per-block x frames-per-tile x grid mode for :class:`FramesOSEM` by timing them on the real
    data (grid mode: a persistent grid of one resident wave, or one workgroup per work item).

    Every candidate runs ``reps`` eager subset steps (subset ``k``) on a scratch copy of the image
    stack after one warm-up step; the median event-timed step wins. Results are cached per key
    (image size, measurements of the subset, frame count) in memory and, when ``path`` is given, in a
    JSON file.
    """

    def __init__(self, path: str | None = None):
        self.path = path
        self.cache = {}
        if path and os.path.exists(path):
            with open(path) as f:
                self.cache = json.load(f)
        self.last_timings = {}

    @staticmethod
    def key(run: FramesOSEM, k: int = 0) -> str:
        p = run.proj
        return f"N{run.N}-M{int(p.nMeasSubset[k, 0])}-F{run.F}-z{int(p.Nz[0])}"

    def tune(self, run: FramesOSEM, k: int = 0, reps: int = 3, candidates=None) -> tuple[int, int, int]:
        torch = run.torch
        key = self.key(run, k)
        if key in self.cache:
            return tuple(self.cache[key])
        if not run.sens_done:
            run.sensitivity()
        if candidates is None:
            candidates = [(r, f, gm) for r in RAYS_PER_BLOCK for f in FRAMES_PER_TILE for gm in GRID_MODES
                          if run.F % f == 0]
        xs = run.x.clone()
        timings = {}
        for rpb, ft, gm in candidates:
            times = []
            for rep in range(reps + 1):
                xs.copy_(run.x)
                a = torch.cuda.Event(enable_timing=True)
                b = torch.cuda.Event(enable_timing=True)
                a.record()
                run.step(k, rpb, ft, gm, x=xs)
                b.record()
                b.synchronize()
                if rep:
                    times.append(a.elapsed_time(b))
            timings[(rpb, ft, gm)] = float(np.median(times))
        run.rhs.zero_()
        best = min(timings, key=timings.get)
        self.last_timings = timings
        self.cache[key] = list(best)
        if self.path:
            os.makedirs(os.path.dirname(os.path.abspath(self.path)), exist_ok=True)
            with open(self.path, "w") as f:
                json.dump(self.cache, f, indent=1)
        return best
