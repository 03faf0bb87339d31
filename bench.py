"""Benchmark: OSEM subset-iterations/s on the 2-D plane PET Siddon workload, one MI355X.

Workload (one "step"): a full OSEM reconstruction of the 2-D plane PET data set -- 8 subsets
(type 8) x 4 iterations = 32 subset-iterations, each = fused Siddon FP -> ratio -> Siddon BP
(+ sensitivity image in iteration 0) and the EM update -- from x0 = 1, with the data already
resident in device memory. value = 32 * steps / wall-seconds of the timed region.

Also reported:
  roofline      for the dominant kernel (the fused FP/ratio/BP projector launch of iterations >= 1):
                algorithmic bytes per launch / its average duration, timed with HIP events on the
                launch stream inside the timed region (the library's omg_timing_* hooks).
                Algorithmic bytes of one such launch = 4 * sum(Np) + 8 * sum(Np, y != 0) + 8 * M_k (4 B gather
                of x per intersected voxel, 8 B read-modify-write of rhs per voxel of the LORs with counts --
                a zero ratio is not backprojected --, 4 B measurement + 4 B LOR table per LOR);
                sum(Np) = exact count of (LOR, voxel) intersections with length > 0.
  cpu_baseline  the reference CPU projector (oracle/_ref, the reference's own headers compiled
                with OpenMP) or, without it, the oracle restatement, on the host cores, on the
                first subset-iterations of the same reconstruction (bounded to ~15 s); its image
                is compared with the GPU image after the same number of subset-iterations.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "OSEM subset-iters/s; wall-time to plain 40-iter log-lik, learned vs plain"
HBM_PEAK_GBS = 8000.0
EXP64_INSTS = 19                                 # FP64 VALU instructions of exp(double) on gfx950
EXP64_PEAK = 78.6e12 / 2 / EXP64_INSTS          # exp(double) per second at the FP64 vector rate


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "siddon_planar_pet2d_traffic.txt"),
                    help="summary of the rocprofv3 --pmc HBM traffic of the dominant kernel, if present")
    ap.add_argument("--issue", default=os.path.join(ROOT, "profiles", "siddon_planar_pet2d_issue.txt"),
                    help="summary of the rocprofv3 --pmc VALU issue counters of the dominant kernel, if present")
    ap.add_argument("--workload", default="pet2d_osem",
                    choices=["pet2d_osem", "pet3d_frames", "ct_cgls_bdd", "pet2d_orth", "pet2d_volume", "pet3d_autograd",
                             "pet2d_tof", "pet2d_volume_tof", "pet2d_interp", "pet2d_interp_tof", "unet"],
                    help="pet2d_osem (default, the headline metric); the others print their own line")
    ap.add_argument("--frames", type=int, default=24, help="pet3d_frames: frames of the dynamic study")
    ap.add_argument("--voxel-bp", action="store_true",
                    help="pet2d_osem: opt into the voxel-driven planar backprojection (closed-form chords; "
                         "matches an FP64 Siddon re-run, not the FP32 walk's lengths) instead of the LDS-atomic one")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1).",
              file=sys.stderr)
        sys.exit(2)

    if args.workload != "pet2d_osem":
        {"pet3d_frames": bench_frames, "ct_cgls_bdd": bench_cgls, "pet2d_orth": bench_orth,
         "pet2d_volume": lambda a: bench_orth(a, volume=True),
         "pet2d_volume_tof": lambda a: bench_orth(a, volume=True, tof=True),
         "pet2d_interp": bench_interp, "pet2d_interp_tof": lambda a: bench_interp(a, tof=True),
         "pet3d_autograd": bench_autograd, "pet2d_tof": bench_tof, "unet": bench_unet}[args.workload](args)
        return

    import ctypes as C

    import numpy as np
    import torch

    from omega_amd import _lib, workloads
    from omega_amd.recon import OSEMRunner

    assert torch.cuda.is_available(), "bench.py needs the GPU"
    _lib.load()
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    p = workloads.pet2d_projector(subsets=8)
    if args.voxel_bp:
        p.voxel_backprojection = True
    p.initProj()
    ph, y = workloads.pet2d_data(p, seed=0)
    n_iter = 4
    sub_iters = p.subsets * n_iter

    # exact intersection counts per subset (GPU traversal; equal to the oracle's, see tests), all LORs and
    # the LORs with counts (only those are backprojected: a zero ratio adds nothing and is skipped)
    np_sum, np_nz = [], []
    for k in range(p.subsets):
        n = int(p.nMeasSubset[k, 0])
        counts = torch.zeros(n, dtype=torch.int32, device=dev)
        _lib.call("omg_siddon_trace", p.geometry.handle, int(p.nMeas[k]), n, C.c_void_p(counts.data_ptr()),
                  None, None, None, None)
        c64 = counts.to(torch.int64)
        off = int(p.nTotMeas[k])
        np_sum.append(int(c64.sum().item()))
        np_nz.append(int((c64 * (y[off:off + n] != 0)).sum().item()))
    M = [int(p.nMeasSubset[k, 0]) for k in range(p.subsets)]

    runner = OSEMRunner(p, y)
    x0 = torch.ones(int(p.N[0]), dtype=torch.float32, device=dev)

    def one_step(timed):
        """One full OSEM run; with ``timed`` the projector launches of iterations >= 1 are event-timed."""
        runner.sens_done = False
        x = x0.clone()
        for it in range(n_iter):
            if timed and it == 1:
                _lib.call("omg_timing_enable", 1)
            for k in range(p.subsets):
                runner.subset_step(x, k)
            runner.sens_done = True
        if timed:
            _lib.call("omg_timing_enable", 0)
        return x

    for _ in range(args.warmup):
        one_step(False)
    # calibration copy for HBM-traffic counters: a plain 256 MiB device-to-device copy whose
    # FETCH_SIZE / WRITE_SIZE give the counter scale (outside the timed region)
    cal_src = torch.ones(64 << 20, dtype=torch.float32, device=dev)
    cal_dst = torch.empty_like(cal_src)
    cal_dst.copy_(cal_src)
    del cal_src, cal_dst
    torch.cuda.synchronize()
    tot = C.c_double()
    cnt = C.c_int64()
    _lib.call("omg_timing_collect", C.byref(tot), C.byref(cnt))   # clear any earlier record
    t0 = time.perf_counter()
    for _ in range(args.steps):
        x_last = one_step(True)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    value = sub_iters * args.steps / dt
    _lib.call("omg_timing_collect", C.byref(tot), C.byref(cnt))

    # dominant kernel: the fused FP -> ratio -> BP projector launch(es) of iterations >= 1, timed with HIP
    # events on their stream: k_siddon_planar<640, true, true, false> (LDS-atomic BP, the default) or,
    # opted in, k_planar_fpq + k_planar_vbp (voxel-driven BP, ring sinograms)
    kernel_name = "k_planar_fpq<640,false> + k_planar_vbp<false,false>" if args.voxel_bp else \
        "k_siddon_planar<640,true,true,false>"
    avg_s = tot.value * 1e-3 / max(cnt.value, 1)
    bytes_per_launch = float(np.mean([4.0 * np_sum[k] + 8.0 * np_nz[k] + 8.0 * M[k] for k in range(p.subsets)]))
    achieved = bytes_per_launch / avg_s / 1e9
    traffic = None if args.voxel_bp else read_traffic(args.traffic)   # the summary covers the LDS kernel
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": kernel_name, "launches_timed": int(cnt.value),
                "avg_launch_us": round(avg_s * 1e6, 2), "algorithmic_bytes_per_launch": int(bytes_per_launch)}
    issue = None if args.voxel_bp else read_issue(args.issue)
    if issue:
        # the kernel's real limiter: its VALU issue rate, from the committed counter summary
        roofline["issue_bound"] = issue

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(p, y, x0, args.cpu_seconds, runner, n_iter)

    line = {
        "metric": METRIC, "value": round(value, 2), "unit": "subset-iters/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded ellipsoid phantom, Poisson data at 5e6 counts)",
        "config": {"workload": "pet2d_osem_siddon", "image": [128, 128, 63], "voxel_mm": 2.5,
                   "sinogram": [128, 160, 63], "ring_radius_mm": 400.0, "subsets": 8, "subset_type": 8,
                   "iterations": n_iter, "subset_iters_per_step": sub_iters, "counts": 5e6,
                   "projector": "improved Siddon", "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if cpu and cpu.get("value"):
        line["speedup_vs_cpu"] = round(value / cpu["value"], 1)
    print(json.dumps(line), flush=True)


def _progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _common(args, workload_line):
    line = {"n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32"}
    line.update(workload_line)
    print(json.dumps(line), flush=True)


def _trace_counts(p, k, dev):
    """Exact (LOR, voxel) visit counts of subset k (omg_siddon_trace)."""
    import ctypes as C
    import torch
    from omega_amd import _lib
    n = int(p.nMeasSubset[k, 0])
    counts = torch.zeros(n, dtype=torch.int32, device=dev)
    _lib.call("omg_siddon_trace", p.geometry.handle, int(p.nMeas[k]), n, C.c_void_p(counts.data_ptr()), None, None,
              None, None)
    return int(counts.to(torch.int64).sum().item())


def bench_frames(args):
    """Dynamic 3-D PET: F frames on the 3-D ring geometry, 16 subsets x 3 iterations, all frames per launch,
    hipGraph-replayed iterations, launch shape from the autotuner. value = frame-subset-iterations/s."""
    import numpy as np
    import torch
    from omega_amd import _lib, phantoms, workloads
    from omega_amd.dynamic import FramesOSEM, FramesTuner
    dev = torch.device("cuda:0")
    _lib.load()
    p = workloads.pet3d_projector(subsets=16)
    p.initProj()
    F = args.frames
    durations = (phantoms.FRAME_DURATIONS * ((F + 23) // 24))[:F]
    frames = phantoms.dynamic_ellipsoids(p.grid, seed=0, durations=durations)
    counts = workloads.frame_counts(F, seed=0)
    t0 = time.perf_counter()
    y = workloads.dynamic_data(p, frames, counts, seed=0, device=dev)
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    _progress(f"frames: data for {F} frames in {t_data:.1f} s")
    run = FramesOSEM(p, y)
    tuner = FramesTuner()
    t0 = time.perf_counter()
    rpb, ft, gm = tuner.tune(run, k=0, reps=2)
    t_tune = time.perf_counter() - t0
    _progress(f"frames: tuner picked rays_per_block={rpb} frames_per_tile={ft} grid={gm} in {t_tune:.1f} s")
    n_iter = 3
    for _ in range(args.warmup):
        run.x.fill_(1.0)
        run.run(1, rpb, ft, grid_blocks=gm)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    t0 = time.perf_counter()
    ev[0].record()
    for i in range(args.steps):
        run.x.fill_(1.0)
        for it in range(n_iter):
            run.run(1, rpb, ft, grid_blocks=gm)
            torch.cuda.synchronize()
            _progress(f"frames: step {i} iteration {it} done")
    ev[1].record()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    fsi = F * p.subsets * n_iter * args.steps / dt
    # one subset step timed alone (k_frames_osem + update) for the roofline of the dominant kernel
    xs = run.x.clone()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    run.step(0, rpb, ft, gm, x=xs)
    b.record()
    b.synchronize()
    step_s = a.elapsed_time(b) * 1e-3
    import ctypes as C
    M0 = int(p.nMeasSubset[0, 0])
    cnt = torch.zeros(M0, dtype=torch.int32, device=dev)
    _lib.call("omg_siddon_trace", p.geometry.handle, int(p.nMeas[0]), M0, C.c_void_p(cnt.data_ptr()), None, None,
              None, None)
    c64 = cnt.to(torch.int64)
    np0 = int(c64.sum().item())
    np_nz = int((c64[None, :] * (y[:, :M0] != 0)).sum().item())    # summed over frames
    bytes_step = 4.0 * np0 * F + 8.0 * np_nz + 8.0 * M0 * F
    # the same work as F serial single-frame fused iterations (the reference's frame loop), one subset
    from omega_amd.recon import OSEMRunner
    t_ser = 0.0
    for f in range(min(F, 4)):
        r1 = OSEMRunner(p, y[f].contiguous())
        x1 = torch.ones(p.grid.N, device=dev)
        r1.subset_step(x1, 0)
        r1.sens_done = True
        torch.cuda.synchronize()
        a.record()
        r1.subset_step(x1, 0)
        b.record()
        b.synchronize()
        t_ser += a.elapsed_time(b) * 1e-3
    serial_step_s = t_ser / min(F, 4) * F
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        _progress("frames: CPU baseline")
        case = O.SiddonCase(p.grid, p.nRowsD, p.nColsD, p.x, p.z, plane_offset=int(p.nMeas[0]))
        rate, per, reps = _cpu_subset_iteration(p, y[0, :M0].cpu().numpy(), case, M0, args.cpu_seconds)
        cpu = {"value": round(rate, 3), "unit": "frame-subset-iters/s", "cores": _threads(), "kind": "port",
               "sample": f"{reps} OSEM subset-iterations of frame 0, subset 0 ({per:.2f} s each) with the oracle's "
                         f"Siddon projector, one frame after another as the reference's frame loop does; OpenMP "
                         f"threads={_threads()}; {_cpu_model()}"}
    _common(args, {
        "metric": "dynamic PET OSEM frame-subset-iterations/s (all frames per launch, hipGraph, autotuned)",
        "value": round(fsi, 2), "unit": "frame-subset-iters/s", "ms_per_step": round(dt / args.steps * 1e3, 2),
        "data": "synthetic (ellipsoid phantom with one-tissue-compartment curves, lognormal frame counts)",
        "config": {"workload": "pet3d_frames_osem_siddon", "frames": F, "image": [128, 128, 63],
                   "sinogram": [224, 224, 1024], "rings": 32, "crystals": 448, "subsets": 16, "iterations": n_iter,
                   "rays_per_block": rpb, "frames_per_tile": ft, "grid": "persistent" if gm == 0 else "per-item",
                   "tuner_ms": tuner.last_timings and
                   {f"{k[0]}x{k[1]}{'p' if k[2] == 0 else 'i'}": round(v, 3) for k, v in tuner.last_timings.items()},
                   "data_gen_s": round(t_data, 1), "tune_s": round(t_tune, 1)},
        "roofline": {"bound": "hbm", "achieved": round(bytes_step / step_s / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(bytes_step / step_s / 1e9 / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": f"k_frames_osem<{ft},{rpb}> + k_frames_update", "avg_launch_us": round(step_s * 1e6, 1),
                     "algorithmic_bytes_per_launch": int(bytes_step)},
        "serial_frames_subset_step_ms": round(serial_step_s * 1e3, 2),
        "batched_subset_step_ms": round(step_s * 1e3, 2),
        "speedup_vs_serial_frames": round(serial_step_s / step_s, 2),
        "cpu_baseline": cpu,
    })


def bench_cgls(args):
    """CBCT CGLS with the branchless distance-driven projector: 256^3 volume, 360 views of 512 x 384.
    value = CGLS iterations/s (10 iterations per step)."""
    import torch
    from omega_amd import _lib, workloads
    from omega_amd.recon import CGLSRunner
    dev = torch.device("cuda:0")
    _lib.load()
    p = workloads.ct_cgls_projector()
    p.initProj()
    mu, y = workloads.ct_cgls_data(p, seed=0)
    _progress("cgls: data ready")
    runner = CGLSRunner(p, y)
    n_iter = 10
    for _ in range(args.warmup):
        runner.run(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        f = runner.run(n_iter)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    import numpy as np
    rel = float(np.linalg.norm(f.cpu().numpy() - mu) / np.linalg.norm(mu))
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    x = torch.rand(p.grid.N, device=dev)
    a.record()
    yy = p.forwardProject(x, 0)
    b.record()
    b.synchronize()
    fp_s = a.elapsed_time(b) * 1e-3
    a.record()
    p.backwardProject(yy, 0)
    b.record()
    b.synchronize()
    bp_s = a.elapsed_time(b) * 1e-3
    # algorithmic bytes of one FP: 4 B per summed-area-table tap. Each detector pixel sums over the slabs of
    # the volume's dominant axis; grouping 8 slabs per lane as the reference does (projectorType5.cl:64-337)
    # costs (4 + 2 * 7) / 8 = 2.25 taps per pixel and slab; plus the 4 B output per pixel.
    slabs = max(p.grid.Nx, p.grid.Ny)
    taps = yy.numel() * slabs * (4.0 + 2.0 * 7) / 8.0
    fp_bytes = 4.0 * taps + 4.0 * yy.numel()
    cpu = None
    if not args.no_cpu:
        _progress("cgls: CPU baseline")
        cpu = _cpu_cgls(p, mu, args.cpu_seconds)
    _common(args, {
        "metric": "CBCT CGLS iterations/s (branchless distance-driven FP/BP)", "value": round(n_iter * args.steps / dt, 3),
        "unit": "CGLS-iters/s", "ms_per_step": round(dt / args.steps * 1e3, 2),
        "data": "synthetic (12 seeded attenuation ellipsoids, line integrals + 1% Gaussian noise)",
        "config": {"workload": "ct_cgls_bdd", "volume": [256, 256, 256], "voxel_mm": 0.5, "views": 360,
                   "detector": [512, 384], "pixel_mm": 0.8, "sod_mm": 700.0, "sdd_mm": 1000.0, "iterations": n_iter,
                   "image_rel_l2_vs_phantom": round(rel, 4), "fp_ms": round(fp_s * 1e3, 2),
                   "bp_ms": round(bp_s * 1e3, 2)},
        "roofline": {"bound": "hbm", "achieved": round(fp_bytes / fp_s / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(fp_bytes / fp_s / 1e9 / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "forward projection (k_sat_* + k_bdd_fp)", "avg_launch_us": round(fp_s * 1e6, 1),
                     "algorithmic_bytes_per_launch": int(fp_bytes), "taps_per_fp": int(taps)},
        "cpu_baseline": cpu,
    })


def _threads():
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def _cpu_model():
    try:
        return [l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name")][0]
    except Exception:
        return "unknown"


def _cpu_cgls(p, mu, budget_s):
    """The oracle's distance-driven FP and BP (OpenMP) on the first views of the same CBCT geometry, scaled
    to one CGLS iteration (one FP and one BP of all views)."""
    from oracle import oracle as O
    threads = _threads()
    nv = 2
    while True:
        case = O.BDDCase(p.grid, p.nRowsD, p.nColsD, p.x[: nv * 6], p.z[: nv * 2], p.dPitchX, p.dPitchY)
        t = time.perf_counter()
        y = case.fp(mu, nthreads=threads)
        t_fp = time.perf_counter() - t
        t = time.perf_counter()
        case.bp(y, nthreads=threads)
        t_bp = time.perf_counter() - t
        if (t_fp + t_bp) * 4 > budget_s or nv >= 32:
            break
        nv *= 2
    per_iter = (t_fp + t_bp) * p.nProjections / nv
    return {"value": round(1.0 / per_iter, 5), "unit": "CGLS-iters/s", "cores": threads, "kind": "port",
            "sample": f"oracle distance-driven FP ({t_fp:.2f} s) and BP ({t_bp:.2f} s) of {nv} of the "
                      f"{p.nProjections} views at 256^3, scaled to all views; OpenMP threads={threads}; {_cpu_model()}"}


def _cpu_subset_iteration(p, meas, case, n, budget_s, tof=None):
    """One OSEM subset-iteration (FP, ratio, BP with sensitivity, update) of an oracle case on the host cores;
    repeated while the budget allows. ``tof``: (bins, sigma, centres, epps) for a TOF orthogonal / volume case.
    Returns (subset-iterations/s, seconds per subset-iteration, repeats)."""
    import numpy as np
    from oracle import oracle as O
    threads = _threads()
    x = np.ones(p.grid.N, np.float32)
    reps, t0 = 0, time.perf_counter()
    while True:
        yhat = case.fp(x, n, nthreads=threads) if tof is None else case.tof_fp(x, n, tof, nthreads=threads)
        r = O.ratio(meas, yhat, 1e-5)
        if tof is None:
            rhs, sens = case.bp(r, n, sens=True, nthreads=threads)
        else:
            rhs, sens = case.tof_bp(r, n, tof, sens=True, nthreads=threads)
        x = O.osem_update(x, O.sens_clamp(sens, 1e-5), rhs, 1e-5)
        reps += 1
        el = time.perf_counter() - t0
        if el * (reps + 1) / reps > budget_s:
            break
    return reps / el, el / reps, reps


def bench_orth(args, volume=False, tof=False):
    """2-D plane PET OSEM with the orthogonal distance-based projector (tube width 2.5 mm, transaxial
    and axial sweeps) on the pet2d_osem geometry. value = subset-iterations/s (8 subsets x 1 iteration per step).
    volume: the volume-of-intersection projector instead (projector type 3, the operator API's defaults
    tube_radius = 1 mm and voxel_radius = 1: voxels count out to 2.77 mm from the LOR). tof: with time of
    flight, 17 bins of 140 ps at 400 ps FWHM (as the pet2d_tof workload)."""
    import numpy as np
    import torch
    from omega_amd import _lib, geometry as G, workloads
    from omega_amd.projector import projectorClass
    from omega_amd.recon import OSEMRunner
    dev = torch.device("cuda:0")
    _lib.load()
    grid = G.Grid.from_fov(128, 128, 63, 320.0, 320.0, 157.5)
    sc = G.pet_2d_planes(grid, radius=400.0, n_radial=128, n_angles=160, radial_pitch=2.5)
    p = projectorClass()
    p.Nx, p.Ny, p.Nz = 128, 128, 63
    p.FOVa_x, p.FOVa_y, p.axial_fov = 320.0, 320.0, 157.5
    if volume:
        p.projector_type = 3
        p.tube_radius, p.voxel_radius = 1.0, 1.0
    else:
        p.projector_type = 2
        p.tube_width_xy = p.tube_width_z = 2.5
    p.x, p.z = sc.xy_table(), sc.z
    p.Ndist, p.Nang = sc.n_radial, sc.n_angles
    p.subsets, p.subsetType = 8, 8
    if tof:
        p.TOF_bins, p.TOF_bins_used, p.TOF_width, p.TOF_FWHM = 17, 17, 140e-12, 400e-12
    p.addProjector()
    p.initProj()
    nb = p.nBins
    ph, y = workloads.pet2d_data(p, seed=0)
    runner = OSEMRunner(p, y)
    x = torch.ones(p.grid.N, device=dev)
    for _ in range(args.warmup):
        for k in range(p.subsets):
            runner.subset_step(x, k)
        runner.sens_done = True
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for k in range(p.subsets):
            runner.subset_step(x, k)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ntube = _trace_counts(p, 0, dev)
    M0 = int(p.nMeasSubset[0, 0])
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    runner.subset_step(x, 0)
    b.record()
    b.synchronize()
    s_s = a.elapsed_time(b) * 1e-3
    byt = 12.0 * ntube + 8.0 * M0 * nb
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        _progress("orth: CPU baseline")
        vol = (p.V, p.bmin, p.bmax, p.Vmax) if volume else None
        width = float(np.float32(p.tube_radius)) if volume else 2.5
        case = O.OrthCase(p.grid, p.nRowsD, p.nColsD, p.x, p.z, width, flags=3, plane_offset=int(p.nMeas[0]),
                          volume=vol)
        tofp = (nb, float(p.sigma_x), p.TOFCenter, float(p.epps)) if tof else None
        rate, per, reps = _cpu_subset_iteration(p, y[:M0 * nb].cpu().numpy(), case, M0, args.cpu_seconds, tofp)
        cpu = {"value": round(rate, 3), "unit": "subset-iters/s", "cores": _threads(), "kind": "port",
               "sample": f"{reps} OSEM subset-iterations of subset 0 ({per:.2f} s each) with the oracle's "
                         f"{'volume-of-intersection' if volume else 'orthogonal'} projector"
                         f"{' with TOF' if tof else ''}; OpenMP threads={_threads()}; {_cpu_model()}"}
    if tof:
        kern = "k_orth_tof_osem (one lane per LOR, tubes traced) + update"
    elif volume:
        kern = "k_tube_vol_planar (tube lists kept on the device) + update"
    else:
        kern = "k_orth_osem + update"
    _common(args, {
        "metric": "OSEM subset-iters/s, " + ("volume-of-intersection" if volume else "orthogonal distance-based")
                  + " projector" + (" with TOF" if tof else ""), "value": round(p.subsets * args.steps / dt, 2),
        "unit": "subset-iters/s", "ms_per_step": round(dt / args.steps * 1e3, 2),
        "data": "synthetic (seeded ellipsoid phantom, Poisson data at 5e6 counts)",
        "config": {"workload": ("pet2d_osem_volume" + ("_tof" if tof else "")) if volume else "pet2d_osem_orth",
                   "image": [128, 128, 63], "sinogram": [128, 160, 63],
                   **({"tube_radius_mm": 1.0, "voxel_radius": 1.0, "table_entries": int(p.V.size)} if volume
                      else {"tube_width_mm": 2.5}),
                   **({"tof_bins": nb, "tof_fwhm_ps": 400, "tof_bin_ps": 140} if tof else {}),
                   "sweeps": "transaxial+axial", "subsets": 8, "tube_visits_subset0": ntube},
        "roofline": {"bound": "hbm", "achieved": round(byt / s_s / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(byt / s_s / 1e9 / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": kern, "avg_launch_us": round(s_s * 1e6, 1),
                     "algorithmic_bytes_per_launch": int(byt)},
        "cpu_baseline": cpu,
    })


def bench_interp(args, tof=False):
    """2-D plane PET OSEM with the interpolation projector (projector type 4, the operator API's default
    interpolation length dL = dx / 2) on the pet2d_osem geometry. value = subset-iterations/s (8 subsets x 1
    iteration per step). tof: with time of flight, 17 bins of 140 ps at 400 ps FWHM. Algorithmic bytes per
    subset step: 12 per sample (the forward projection's interpolated read, the backprojection's atomic
    read-modify-write) plus 8 per measurement and bin; the samples are counted from the LORs' chords through
    the image (chord / dL + 1 per LOR and plane)."""
    import numpy as np
    import torch
    from omega_amd import _lib, geometry as G, workloads
    from omega_amd.projector import projectorClass
    from omega_amd.recon import OSEMRunner
    dev = torch.device("cuda:0")
    _lib.load()
    grid = G.Grid.from_fov(128, 128, 63, 320.0, 320.0, 157.5)
    sc = G.pet_2d_planes(grid, radius=400.0, n_radial=128, n_angles=160, radial_pitch=2.5)
    p = projectorClass()
    p.Nx, p.Ny, p.Nz = 128, 128, 63
    p.FOVa_x, p.FOVa_y, p.axial_fov = 320.0, 320.0, 157.5
    p.projector_type = 4
    p.x, p.z = sc.xy_table(), sc.z
    p.Ndist, p.Nang = sc.n_radial, sc.n_angles
    p.subsets, p.subsetType = 8, 8
    if tof:
        p.TOF_bins, p.TOF_bins_used, p.TOF_width, p.TOF_FWHM = 17, 17, 140e-12, 400e-12
    p.addProjector()
    p.initProj()
    nb = p.nBins
    dl = float(p._interp_scalars()[0])
    ph, y = workloads.pet2d_data(p, seed=0)
    runner = OSEMRunner(p, y)
    x = torch.ones(p.grid.N, device=dev)
    for _ in range(args.warmup):
        for k in range(p.subsets):
            runner.subset_step(x, k)
        runner.sens_done = True
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for k in range(p.subsets):
            runner.subset_step(x, k)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    M0 = int(p.nMeasSubset[0, 0])
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    runner.subset_step(x, 0)
    b.record()
    b.synchronize()
    s_s = a.elapsed_time(b) * 1e-3
    # samples: every LOR's chord through the transaxial box, per plane of subset 0
    xy = np.asarray(sc.xy_table(), np.float64).reshape(-1, 4)
    s0, v = xy[:, 0:2], xy[:, 2:4] - xy[:, 0:2]
    lo, hi = np.array([grid.bx, grid.by]), np.array([grid.bx + 320.0, grid.by + 320.0])
    with np.errstate(divide="ignore", invalid="ignore"):
        ta, tb = (lo - s0) / v, (hi - s0) / v
    tmin = np.nanmax(np.where(np.isfinite(ta), np.minimum(ta, tb), -np.inf), axis=1)
    tmax = np.nanmin(np.where(np.isfinite(ta), np.maximum(ta, tb), np.inf), axis=1)
    chord = np.clip(tmax - tmin, 0, None) * np.linalg.norm(v, axis=1)
    per_plane = float(np.sum(np.where(chord > 0, np.floor(chord / dl) + 1, 0)))
    samples = per_plane * (M0 // (sc.n_radial * sc.n_angles))
    byt = 12.0 * samples + 8.0 * M0 * nb
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        _progress("interp: CPU baseline")
        case = O.InterpCase(p.grid, p.nRowsD, p.nColsD, p.x, p.z, dl, plane_offset=int(p.nMeas[0]))
        tofp = (nb, float(p.sigma_x), p.TOFCenter) if tof else None
        rate, per, reps = _cpu_subset_iteration(p, y[:M0 * nb].cpu().numpy(), case, M0, args.cpu_seconds, tofp)
        cpu = {"value": round(rate, 3), "unit": "subset-iters/s", "cores": _threads(), "kind": "port",
               "sample": f"{reps} OSEM subset-iterations of subset 0 ({per:.2f} s each) with the oracle's "
                         f"interpolation projector{' with TOF' if tof else ''}; OpenMP threads={_threads()}; "
                         f"{_cpu_model()}"}
    _common(args, {
        "metric": "OSEM subset-iters/s, interpolation projector" + (" with TOF" if tof else ""),
        "value": round(p.subsets * args.steps / dt, 2), "unit": "subset-iters/s",
        "ms_per_step": round(dt / args.steps * 1e3, 2),
        "data": "synthetic (seeded ellipsoid phantom, Poisson data at 5e6 counts)",
        "config": {"workload": "pet2d_osem_interp" + ("_tof" if tof else ""), "image": [128, 128, 63],
                   "sinogram": [128, 160, 63], "interp_dl_mm": round(dl, 4),
                   **({"tof_bins": nb, "tof_fwhm_ps": 400, "tof_bin_ps": 140} if tof else {}),
                   "subsets": 8, "samples_subset0": int(samples)},
        "roofline": {"bound": "hbm", "achieved": round(byt / s_s / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(byt / s_s / 1e9 / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": ("k_interp_tof_fp + ratio + k_interp_tof_bp + update" if tof
                                else "k_interp_fp_planar + ratio + k_interp_bp_planar (slices in LDS) + update"),
                     "avg_launch_us": round(s_s * 1e6, 1), "algorithmic_bytes_per_launch": int(byt)},
        "cpu_baseline": cpu,
    })


def bench_tof(args):
    """2-D plane PET OSEM with time of flight and attenuation on the pet2d_osem geometry: 17 TOF bins of
    140 ps (400 ps FWHM), a water cylinder (0.0096 / mm, radius 120 mm) as attenuation image.
    value = subset-iterations/s (8 subsets x 1 iteration per step). The TOF kernels are bound by the
    Gaussian's double-precision exponentials (3 per bin and voxel visit: the trapezoid's last point is
    the next voxel's first), reported as exps/s."""
    import ctypes as C
    import numpy as np
    import torch
    from omega_amd import _lib, workloads
    from omega_amd.recon import OSEMRunner
    dev = torch.device("cuda:0")
    _lib.load()
    p = workloads.pet2d_projector(subsets=8)
    p.TOF_bins, p.TOF_bins_used, p.TOF_width, p.TOF_FWHM = 17, 17, 140e-12, 400e-12
    g = p.grid
    xc = (np.arange(g.Nx) + 0.5) * g.dx + g.bx
    yc = (np.arange(g.Ny) + 0.5) * g.dy + g.by
    mu2 = np.where(xc[:, None] ** 2 + yc[None, :] ** 2 < 120.0 ** 2, 0.0096, 0.0).astype(np.float32)
    p.attenuation_correction = True
    p.vaimennus = np.repeat(mu2[:, :, None], g.Nz, axis=2)
    torch.cuda.synchronize()
    t_set = time.perf_counter()
    p.initProj()
    torch.cuda.synchronize()
    setup_ms = (time.perf_counter() - t_set) * 1e3
    ph, y = workloads.pet2d_data(p, seed=0)
    runner = OSEMRunner(p, y)
    x = torch.ones(p.grid.N, device=dev)
    for _ in range(args.warmup):
        for k in range(p.subsets):
            runner.subset_step(x, k)
        runner.sens_done = True
    torch.cuda.synchronize()
    _progress("pet2d_tof warm-up done")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for k in range(p.subsets):
            runner.subset_step(x, k)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    # one subset step timed alone; its visits give the exponential count (FP and BP each: 3 per bin and
    # visit, the first point shared with the previous voxel; the BP's weights reused from its sum)
    ntr = _trace_counts(p, 0, dev)
    M0 = int(p.nMeasSubset[0, 0])
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    runner.subset_step(x, 0)
    b.record()
    b.synchronize()
    s_s = a.elapsed_time(b) * 1e-3
    exps = 2.0 * 3.0 * p.nBins * ntr
    byt = 12.0 * ntr + 8.0 * M0 * p.nBins
    cpu = None
    if not args.no_cpu:
        _progress("pet2d_tof: CPU baseline")
        tof = dict(tof_bins=p.nBins, tof_sigma=np.float32(p.sigma_x), tof_centers=p.TOFCenter, tof_epps=float(p.epps))
        cpu = cpu_baseline(p, y, torch.ones(p.grid.N, device=dev), args.cpu_seconds, runner, 1,
                           osem_kw=dict(atten=p._attenuation()[0], atten_mode=1, tof=tof))
    _common(args, {
        "metric": "OSEM subset-iters/s, improved Siddon with TOF and attenuation", "value": round(p.subsets * args.steps / dt, 2),
        "unit": "subset-iters/s", "ms_per_step": round(dt / args.steps * 1e3, 2),
        "data": "synthetic (seeded ellipsoid phantom, Poisson TOF data at 5e6 counts)",
        "config": {"workload": "pet2d_osem_tof", "image": [128, 128, 63], "sinogram": [128, 160, 63],
                   "tof_bins": p.nBins, "tof_fwhm_ps": 400, "tof_bin_ps": 140, "attenuation": "water cylinder",
                   "subsets": 8, "voxel_visits_subset0": ntr, "geometry_setup_ms": round(setup_ms, 1)},
        # bound: the double-precision exponentials. exp(double) compiles for gfx950 to 19 FP64 VALU
        # instructions (13 FMA, mul, rndne, ldexp, cvt, 2 compares); at the FP64 vector rate of 78.6 TFLOP/s
        # (39.3 T lane-instructions/s) that is 2.07e12 exp/s with no other work
        "roofline": {"bound": "fp64", "achieved": round(exps / s_s / 1e9, 1), "peak": round(EXP64_PEAK / 1e9, 1),
                     "unit": "G exp/s", "frac": round(exps / s_s / EXP64_PEAK, 4), "traffic": None,
                     "peak_basis": f"78.6 TFLOP/s FP64 / 2 = 39.3e12 lane-instructions/s over {EXP64_INSTS} "
                                   "FP64 instructions per exp(double)",
                     "kernel": "k_tof_osem + update", "avg_launch_us": round(s_s * 1e6, 1),
                     "exps_per_launch": int(exps),
                     "hbm": {"algorithmic_bytes_per_launch": int(byt), "achieved": round(byt / s_s / 1e9, 1),
                             "frac": round(byt / s_s / 1e9 / HBM_PEAK_GBS, 4)}},
        "cpu_baseline": cpu,
    })


def bench_autograd(args):
    """Differentiable 3-D PET projector on the full ring geometry (224 x 224 x 1024 sinogram, 128 x 128 x 63
    image): adjoint identity on 20 seeded (x, y) pairs, central differences of -L(x) + R(x) (y = Poisson(A x),
    2e7 counts, gamma 2, beta 0.1) at 64 random voxels with h = 1e-3 against the autograd gradient, and the
    time of one forward + backward. value = forward+backward passes per second."""
    import numpy as np
    import torch
    from omega_amd import _lib, autograd as AG, phantoms, workloads
    dev = torch.device("cuda:0")
    _lib.load()
    p = workloads.pet3d_projector(subsets=1)
    p.initProj()
    ph = phantoms.ellipsoids(p.grid, seed=0) + 0.2
    meas, _ = workloads.poisson_data(p, ph, 2e7, seed=1)
    _progress("autograd: data ready")
    worst = 0.0
    for seed in range(20):
        g = torch.Generator(device=dev)
        g.manual_seed(seed)
        x = torch.rand(p.grid.N, device=dev, generator=g)
        yv = torch.rand(int(p.nMeasSubset[0, 0]), device=dev, generator=g)
        lhs = float((p.forwardProject(x, 0).double() * yv.double()).sum())
        rhs = float((x.double() * p.backwardProject(yv, 0).double()).sum())
        worst = max(worst, abs(lhs - rhs) / abs(lhs))
    _progress(f"autograd: adjoint worst {worst:.2e}")
    x = torch.as_tensor(ph, device=dev).clone().requires_grad_(True)
    for _ in range(args.warmup):   # untimed passes (allocations, first launches)
        x.grad = None
        AG.penalised_nll(x, p, meas, gamma=2.0, beta=0.1).backward()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reps = max(1, args.steps)
    for _ in range(reps):
        x.grad = None
        loss = AG.penalised_nll(x, p, meas, gamma=2.0, beta=0.1)
        loss.backward()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    grad = x.grad.double().cpu().numpy()
    rng = np.random.default_rng(7)
    h = 1e-3
    errs = []
    for j in rng.choice(p.grid.N, 64, replace=False):
        xp = x.detach().clone()
        xm = x.detach().clone()
        xp[j] += h
        xm[j] -= h
        fd = (float(AG.penalised_nll(xp, p, meas)) - float(AG.penalised_nll(xm, p, meas))) / (2 * h)
        errs.append(abs(fd - grad[j]) / max(1.0, abs(grad[j])))
    _progress(f"autograd: finite differences max {max(errs):.2e}")
    # roofline of the pass: one FP and one BP of the whole sinogram; algorithmic bytes 4 B per (LOR, voxel)
    # visit for the FP gather, 8 B for the BP update, 8 B per LOR (measurement and table)
    import ctypes as C
    M = int(p.nMeasSubset[0, 0])
    visits = _trace_counts(p, 0, dev)
    byt = 12.0 * visits + 8.0 * M
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        _progress("autograd: CPU baseline")
        per = p.nRowsD * p.nColsD
        planes = 16
        case = O.SiddonCase(p.grid, p.nRowsD, p.nColsD, p.x, p.z)
        xs = np.asarray(ph, np.float32)
        t = time.perf_counter()
        yc = case.fp(xs, planes * per, nthreads=_threads())
        case.bp(yc, planes * per, nthreads=_threads())
        t_s = (time.perf_counter() - t) * (p.nProjections / planes)
        cpu = {"value": round(1.0 / t_s, 5), "unit": "fwd+bwd/s", "cores": _threads(), "kind": "port",
               "sample": f"oracle Siddon FP + BP of {planes} of the {p.nProjections} sinogram planes, scaled to all "
                         f"planes (the RDP terms are negligible); OpenMP threads={_threads()}; {_cpu_model()}"}
    _common(args, {
        "metric": "differentiable 3-D PET projector: forward+backward of -L(x)+R(x) per second",
        "value": round(1.0 / dt, 3), "unit": "fwd+bwd/s", "ms_per_step": round(dt * 1e3, 2),
        "data": "synthetic (ellipsoid phantom + 0.2, Poisson data at 2e7 counts)",
        "config": {"workload": "pet3d_autograd", "image": [128, 128, 63], "sinogram": [224, 224, 1024],
                   "rings": 32, "crystals": 448, "max_ring_difference": 31, "gamma": 2.0, "beta": 0.1},
        "adjoint_pairs": 20, "adjoint_worst_rel_mismatch": worst,
        "fd_voxels": 64, "fd_h": h, "fd_max_rel_error": max(errs), "fd_median_rel_error": float(np.median(errs)),
        "roofline": {"bound": "hbm", "achieved": round(byt / dt / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(byt / dt / 1e9 / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "forward + backward projection of the pass (k_siddon_* )",
                     "avg_launch_us": round(dt * 1e6, 1), "algorithmic_bytes_per_launch": int(byt),
                     "visits": visits},
        "cpu_baseline": cpu,
    })


def bench_unet(args):
    """The warm start's 3-D U-Net on the 3-D PET image (128 x 128 x 63, z padded to 64): one training step
    (forward, masked MSE, backward, Adam) per step. value = training steps/s. Each forward convolution is
    timed alone with HIP events on the launch stream; the roofline is that of the convolution kernel over
    all ten forward layers (2 * 27 * Cin * Cout FLOP per voxel, BF16 dense MFMA peak), its launches queued
    behind a device-side sleep so that host launch overhead does not count."""
    import torch
    from omega_amd import _lib
    from omega_amd.unet import LAYERS, UNet3D
    dev = torch.device("cuda:0")
    _lib.load()
    net = UNet3D(128, 128, 63, seed=0)
    g = torch.Generator(device=dev)
    g.manual_seed(0)
    img = torch.rand(128 * 128 * 63, generator=g, device=dev)
    tgt = torch.rand(128 * 128 * 63, generator=g, device=dev)
    mask = net.mask()

    def step():
        net.forward(img, 1.0)
        net.loss(tgt, mask, 1.0)
        net.zero_grad()
        net.backward(net.act["gout"])
        net.adam_step(1e-4)

    for _ in range(max(1, args.warmup)):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    # forward pass alone, and each forward convolution alone
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 10

    def queued(fn):
        """Seconds per call of fn, its launches queued behind a device-side sleep so that the host's
        launch overhead does not open gaps between them."""
        torch.cuda._sleep(int(5e6))
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        b.synchronize()
        return a.elapsed_time(b) * 1e-3 / reps

    fwd_s = queued(lambda: net.forward(img, 1.0))
    inputs = {"e1a": ("x0", None), "e1b": ("a1", None), "e2a": ("p1", None), "e2b": ("a2", None),
              "bta": ("p2", None), "btb": ("a3", None), "d2a": ("u2", "s2"), "d2b": ("a4", None),
              "d1a": ("u1", "s1"), "d1b": ("a5", None)}
    outputs = {"e1a": "a1", "e1b": "s1", "e2a": "a2", "e2b": "s2", "bta": "a3", "btb": "b3", "d2a": "a4",
               "d2b": "d2", "d1a": "a5", "d1b": "d1"}
    layers = {}
    tot_flop, tot_s = 0.0, 0.0
    for name, cins, cout, lvl in LAYERS:
        x1, x2 = inputs[name]
        s_l = queued(lambda: net._conv(name, net.act[x1], None if x2 is None else net.act[x2], net.act[outputs[name]]))
        flop = 2.0 * 27 * sum(cins) * cout * net.V[lvl]
        layers[name] = {"cin": sum(cins), "cout": cout, "voxels": net.V[lvl], "us": round(s_l * 1e6, 1),
                        "tflops": round(flop / s_l / 1e12, 1)}
        tot_flop += flop
        tot_s += s_l
    peak = 2500.0
    achieved = tot_flop / tot_s / 1e12
    _common(args, {
        "metric": "3-D U-Net training steps/s (BF16 MFMA convolutions)", "value": round(args.steps / dt, 2),
        "unit": "train-steps/s", "ms_per_step": round(dt / args.steps * 1e3, 3), "dtype": "bf16",
        "data": "synthetic (uniform random image and target)",
        "config": {"workload": "unet3d_warmstart", "image": [128, 128, 63], "padded_depth": 64,
                   "channels": [16, 32, 64], "forward_ms": round(fwd_s * 1e3, 3), "layers": layers},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 1), "peak": peak, "unit": "TFLOP/s",
                     "frac": round(achieved / peak, 4), "traffic": None,
                     "kernel": "k_conv3d_fwd_tile (ten forward convolutions)",
                     "avg_launch_us": round(tot_s / len(LAYERS) * 1e6, 1),
                     "algorithmic_flop_per_forward": int(tot_flop)},
        "cpu_baseline": None,
    })


def read_traffic(path):
    """HBM bytes per launch of the dominant kernel from a committed rocprofv3 --pmc summary, if present."""
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("hbm_bytes_per_launch:"):
            return int(float(line.split(":", 1)[1]))
    return None


def read_issue(path):
    """VALU issue figures of the dominant kernel from a committed rocprofv3 --pmc summary, if present:
    lines 'valu_busy: f' (rocprof's VALUBusy), 'valu_busy_vs_sq_busy: f', 'valu_insts_per_launch: n',
    'waves_per_simd: w'."""
    if not os.path.exists(path):
        return None
    out = {}
    for line in open(path):
        for key in ("valu_busy", "valu_busy_vs_sq_busy", "valu_insts_per_launch", "waves_per_simd"):
            if line.startswith(key + ":"):
                out[key] = float(line.split(":", 1)[1])
    if "valu_busy" not in out:
        return None
    out["bound"] = "valu_issue"
    out["source"] = os.path.relpath(path, ROOT)
    return out


def cpu_baseline(p, y, x0, budget_s, runner, n_iter, osem_kw=None):
    """Reference CPU OSEM on a bounded number of subset-iterations; image compared with the GPU's.
    ``osem_kw``: corrections of the run (attenuation, TOF) passed to the oracle's OSEM driver."""
    import numpy as np
    import torch

    from oracle import oracle as O

    use_ref = O.ref() is not None
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    per = p.nRowsD * p.nColsD
    meas = y.cpu().numpy()
    nmeas = np.diff(p.nMeas).astype(np.int64)
    # probe one subset-iteration to size the sample, then run as many as fit the budget
    kw = dict(osem_kw or {})
    t = time.perf_counter()
    O.osem(p.grid, p.nRowsD, p.nColsD, p.x, p.z, nmeas, meas, 1, x0.cpu().numpy(), use_ref=use_ref,
           nthreads=threads, subsets_to_run=1, **kw)
    t1 = time.perf_counter() - t
    n_run = int(max(2, min(p.subsets * n_iter, budget_s / max(t1, 1e-6))))
    t = time.perf_counter()
    x_cpu = O.osem(p.grid, p.nRowsD, p.nColsD, p.x, p.z, nmeas, meas, n_iter, x0.cpu().numpy(), use_ref=use_ref,
                   nthreads=threads, subsets_to_run=n_run, **kw)
    dt = time.perf_counter() - t
    # GPU image after the same number of subset-iterations
    runner.sens_done = False
    xg = x0.clone()
    done = 0
    for it in range(n_iter):
        for k in range(p.subsets):
            if done < n_run:
                runner.subset_step(xg, k)
                done += 1
        runner.sens_done = True
    torch.cuda.synchronize()
    xg = xg.cpu().numpy().astype(np.float64)
    rel = float(np.linalg.norm(xg - x_cpu) / np.linalg.norm(x_cpu))
    try:
        cpu_model = [l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name")][0]
    except Exception:
        cpu_model = "unknown"
    return {"value": round(n_run / dt, 3), "unit": "subset-iters/s", "cores": threads,
            "kind": "reference" if use_ref else "port",
            "sample": f"first {n_run} subset-iterations (of {p.subsets * n_iter}) of the same OSEM run, incl. "
                      f"sensitivity images; OpenMP threads={threads}; {cpu_model}",
            "image_rel_l2_vs_gpu": rel}


if __name__ == "__main__":
    main()
