#!/usr/bin/env python3
"""Benchmark: Metropolis+Wolff site-updates/s on one MI355X.

Workload (default): the batched disorder average -- L=32, 256 disorder
realizations x 8 temperatures T in {1.10, 1.15, ..., 1.45} = 2048 replicas,
J_xy[z] ~ U[0.6, 1.8] iid per layer from the realization's seed, J_perp = 1.0,
proposal width pi, one Wolff cluster per sweep, FP32 sweeps with FP64
accumulators, one measurement (13 series) per sweep binned on the device.
A step is one sweep of all 2048 replicas (Metropolis + Wolff + measurement).

Also reported: the plain single-replica GPU path (one L=32 replica on the HBM
kernels) and its ratio to the batched form; a roofline object for the dominant
kernel; and a CPU baseline (the oracle's reference-mode restatement of the
reference CPU path, timed on this host's cores on a bounded sample).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "slsf-rs_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, public spec


def splitmix_uniform_layers(lz, seed, lo=0.6, hi=1.8):
    """J_xy[z] = lo + (hi-lo) * u, u = SplitMix64(seed) f64 draws (the job
    config's uniform-disorder rule)."""
    g = 0x9E3779B97F4A7C15
    m = (1 << 64) - 1
    s = seed & m
    out = np.zeros(lz)
    for z in range(lz):
        s = (s + g) & m
        x = s
        x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & m
        x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & m
        x ^= x >> 31
        out[z] = lo + (hi - lo) * ((x >> 11) * 2.0 ** -53)
    return out


def c3_replicas(slsf, n_real=256, temps=None, L=32):
    temps = temps if temps is not None else [1.10 + 0.05 * k for k in range(8)]
    reps = []
    for r in range(n_real):
        jxy = splitmix_uniform_layers(L, 1000 + r)
        jz = np.ones(L)
        for t in temps:
            seed = 1000 + 100000 * L + 1000 * (r + 1) + int(round(100 * t))
            reps.append(slsf.Replica(L, L, L, t, jxy, jz, seed, float(np.pi), 1))
    return reps


def cpu_baseline(budget_s=12.0):
    """Reference-mode oracle (SplitMix64 stream, red/black with 8-lane batches,
    Wolff DFS with trig cache, libm sincos/exp) on one L=32 replica per core."""
    import multiprocessing as mp
    cores = min(os.cpu_count() or 1, 16)
    # calibrate one core
    rate1 = _cpu_worker((0, 3))
    sweeps = max(2, int(budget_s * rate1 / (32 ** 3)))
    t0 = time.perf_counter()
    with mp.get_context("fork").Pool(cores) as pool:
        rates = pool.map(_cpu_worker, [(i, sweeps) for i in range(cores)])
    wall = time.perf_counter() - t0
    total = cores * sweeps * 32 ** 3
    return {"value": total / wall, "unit": "site-updates/s", "cores": cores,
            "kind": "port",
            "sample": f"{cores} independent L=32 replicas (T=1.3, one per core), "
                      f"{sweeps} Metropolis+Wolff sweeps each, FP64, reference draw order",
            "per_core": float(np.mean(rates))}


def _cpu_worker(arg):
    idx, sweeps = arg
    import oracle_lib as orc
    L = orc.load(fast=True)
    lat = orc.Lattice(32, 32, 32, lib_=L)
    lat.set_couplings(splitmix_uniform_layers(32, 1000 + idx), np.ones(32))
    r = orc.FastRng()
    import ctypes
    L.or_rng_seed(ctypes.byref(r), 7 + idx)
    L.or_initialize_angles(lat.ref, 0, ctypes.byref(r))
    lat.refresh(False)
    t0 = time.perf_counter()
    for _ in range(sweeps):
        lat.metro_ref(1.3, np.pi, r, True)
        lat.wolff_ref(1.3, r, True)
    return sweeps * 32 ** 3 / (time.perf_counter() - t0)


def copy_calibration(nbytes=1 << 30, reps=5):
    """Device-to-device copy of a known size (torch as plumbing): the
    FETCH_SIZE / WRITE_SIZE counters of this dispatch calibrate the per-launch
    HBM traffic the profiler reports for the lattice kernels."""
    import torch
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    print(json.dumps({"copy_bytes_each_way": nbytes, "seconds": dt,
                      "GB_per_s_read_plus_write": 2 * nbytes / dt / 1e9}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--binsize", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--realizations", type=int, default=256)
    ap.add_argument("--sweeps-per-launch", type=int, default=0)
    ap.add_argument("--mode", choices=["bench", "copy"], default="bench",
                    help="copy: a plain 1 GiB device copy, to calibrate HBM byte counters")
    args = ap.parse_args()
    if args.mode == "copy":
        copy_calibration()
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the replica batch is sharded across GPUs "
              "by whole replicas, which this build does not launch", file=sys.stderr)
        sys.exit(2)

    import slsf_amd as slsf

    reps = c3_replicas(slsf, args.realizations)
    b = slsf.Batch(reps, slsf.FP32)
    b.init_angles()
    b.run(args.warmup, 0, 1)
    st = b.run(0, args.steps, args.binsize, sweeps_per_launch=args.sweeps_per_launch)
    V = 32 ** 3
    value = len(reps) * V * args.steps / st["seconds"]
    ms_per_step = 1e3 * st["seconds"] / args.steps

    plain = None
    if not args.no_plain:
        p = slsf.Batch(reps[:1], slsf.FP32)
        p.init_angles()
        p.run(5, 0, 1)
        nplain = max(20, min(200, args.steps))
        sp = p.run(0, nplain, 10)
        plain = V * nplain / sp["seconds"]

    kernel_ms = st.get("kernel_ms", 0.0)
    bytes_per_update = 8.0  # FP32 angle read + write per site update
    if kernel_ms > 0:
        achieved = len(reps) * V * args.steps * bytes_per_update / (kernel_ms * 1e-3) / 1e9
    else:
        achieved = value * bytes_per_update / 1e9
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None}
    cpu = None if args.no_cpu_baseline else cpu_baseline()
    out = {
        "metric": "Metropolis+Wolff site-updates/s (batched vs plain, 1 GPU)",
        "value": value, "unit": "site-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32 (fp64 accumulators)",
        "data": "synthetic: J_xy[z] ~ U[0.6,1.8] per realization seed, J_perp=1.0, "
                "random initial angles",
        "config": {"workload": f"L=32 batched disorder average: {args.realizations} "
                               "realizations x 8 T in [1.10,1.45], Metropolis + 1 Wolff "
                               "+ 13-series measurement per sweep",
                   "replicas": len(reps), "L": 32, "binsize": args.binsize},
        "plain_single_replica": plain,
        "batched_over_plain": (value / plain) if plain else None,
        "acceptance": st["acceptance_mean"], "mean_cluster_size": st["mean_cluster_size"],
        "roofline": roofline, "cpu_baseline": cpu,
        "kernel": {"name": "k_resident<32, paired>", "launches": st["launches"],
                   "avg_launch_ms": kernel_ms / max(st["launches"], 1),
                   "sweeps_per_launch": args.sweeps_per_launch or 50,
                   "algorithmic_bytes_per_launch": len(reps) * V * bytes_per_update *
                   (args.sweeps_per_launch or 50)},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
