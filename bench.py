"""Benchmark: OSEM subset-iterations/s on the 2-D plane PET Siddon workload, one MI355X.

Workload (one "step"): a full OSEM reconstruction of the 2-D plane PET data set -- 8 subsets
(type 8) x 4 iterations = 32 subset-iterations, each = fused Siddon FP -> ratio -> Siddon BP
(+ sensitivity image in iteration 0) and the EM update -- from x0 = 1, with the data already
resident in device memory. value = 32 * steps / wall-seconds of the timed region.

Also reported:
  roofline      for the dominant kernel (the fused FP/ratio/BP projector launch of iterations >= 1):
                algorithmic bytes per launch / its average duration, timed with HIP events on the
                launch stream inside the timed region (the library's omg_timing_* hooks).
                Algorithmic bytes of one such launch = 12 * sum(Np) + 8 * M_k  (4 B gather of x per intersected voxel, 8 B atomic
                read-modify-write of rhs per voxel, 4 B measurement + 4 B LOR table per LOR);
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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "siddon_planar_pet2d_traffic.txt"),
                    help="summary of the rocprofv3 --pmc HBM traffic of the dominant kernel, if present")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1).",
              file=sys.stderr)
        sys.exit(2)

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
    p.initProj()
    ph, y = workloads.pet2d_data(p, seed=0)
    n_iter = 4
    sub_iters = p.subsets * n_iter

    # exact intersection counts per subset (GPU traversal; equal to the oracle's, see tests)
    np_sum = []
    for k in range(p.subsets):
        n = int(p.nMeasSubset[k, 0])
        counts = torch.zeros(n, dtype=torch.int32, device=dev)
        _lib.call("omg_siddon_trace", p.geometry.handle, int(p.nMeas[k]), n, C.c_void_p(counts.data_ptr()),
                  None, None, None, None)
        np_sum.append(int(counts.to(torch.int64).sum().item()))
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

    # dominant kernel: the fused FP -> ratio -> BP projector launch of iterations >= 1
    # (k_siddon_planar<640, true, true, false>), timed with HIP events on its stream
    avg_s = tot.value * 1e-3 / max(cnt.value, 1)
    bytes_per_launch = float(np.mean([12.0 * np_sum[k] + 8.0 * M[k] for k in range(p.subsets)]))
    achieved = bytes_per_launch / avg_s / 1e9
    traffic = read_traffic(args.traffic)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "k_siddon_planar<640,true,true,false>", "launches_timed": int(cnt.value),
                "avg_launch_us": round(avg_s * 1e6, 2), "algorithmic_bytes_per_launch": int(bytes_per_launch)}

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


def read_traffic(path):
    """HBM bytes per launch of the dominant kernel from a committed rocprofv3 --pmc summary, if present."""
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("hbm_bytes_per_launch:"):
            return int(float(line.split(":", 1)[1]))
    return None


def cpu_baseline(p, y, x0, budget_s, runner, n_iter):
    """Reference CPU OSEM on a bounded number of subset-iterations; image compared with the GPU's."""
    import numpy as np
    import torch

    from oracle import oracle as O

    use_ref = O.ref() is not None
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    per = p.nRowsD * p.nColsD
    meas = y.cpu().numpy()
    nmeas = np.diff(p.nMeas).astype(np.int64)
    # probe one subset-iteration to size the sample, then run as many as fit the budget
    t = time.perf_counter()
    O.osem(p.grid, p.nRowsD, p.nColsD, p.x, p.z, nmeas, meas, 1, x0.cpu().numpy(), use_ref=use_ref,
           nthreads=threads, subsets_to_run=1)
    t1 = time.perf_counter() - t
    n_run = int(max(2, min(p.subsets * n_iter, budget_s / max(t1, 1e-6))))
    t = time.perf_counter()
    x_cpu = O.osem(p.grid, p.nRowsD, p.nColsD, p.x, p.z, nmeas, meas, n_iter, x0.cpu().numpy(), use_ref=use_ref,
                   nthreads=threads, subsets_to_run=n_run)
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
            "sample": f"first {n_run} subset-iterations (of 32) of the same OSEM run, incl. sensitivity "
                      f"images; OpenMP threads={threads}; {cpu_model}",
            "image_rel_l2_vs_gpu": rel}


if __name__ == "__main__":
    main()
