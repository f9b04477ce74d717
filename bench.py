#!/usr/bin/env python3
"""Benchmark: Metropolis+Wolff site-updates/s on one MI355X.

Default workload (the headline line) -- the batched disorder average: L=32,
256 disorder realizations x 8 temperatures T in {1.10, 1.15, ..., 1.45} = 2048
replicas, J_xy[z] ~ U[0.6, 1.8] iid per layer from the realization's seed,
J_perp = 1.0, proposal width pi, one Wolff cluster per sweep, FP32 sweeps with
FP64 accumulators, one measurement (13 series) per sweep binned on the device;
2000 thermalization sweeps (the untimed warmup) + 10000 timed sweeps.  A step
is one sweep of all 2048 replicas (Metropolis + Wolff + measurement).

The same line reports the plain single-replica GPU path (one L=32 replica on
the HBM kernels) and the batched/plain ratio, Metropolis-only and Wolff-only
rates, the mean cluster size, a roofline object for the dominant kernel and a
CPU baseline (the oracle's reference-mode restatement of the reference CPU
path, timed on this host's cores on a bounded sample).

Other workloads, one JSON line each (--workload):
  c1  plain method: L=16, T=1.3, 1 replica, FP64 replay, 1000 + 5000 sweeps,
      measured every sweep, C(r) along x up to r = 8 (run_theta_simulation)
  c2  single replica at scale: L=64, T=1.3, 1000 sweeps, FP64 replay then FP32
  c4  uneven batch: L in {8,12,16,24,32} x T in {1.1,...,1.5} x 64 realizations,
      autotuned, 1000 + 4000 sweeps; disorder-averaged Binder ratio and helicity
      moduli with FP64 jackknife errors and Binder crossing temperatures
  c5  serving: Poisson arrivals (50 jobs/s, seed 42) of jobs with L in
      {16,32,48,64,96,128}, T ~ U[0.3,2.2], sweeps ~ U{1000..10000}; jobs are
      admitted by device-memory footprint; p50/p99 latency and site-updates/s
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "slsf-rs_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, public spec
LDS_PEAK_NOMINAL_GBS = 78643.2  # 256 CUs x 128 B/clk x 2.4 GHz
METRIC = "Metropolis+Wolff site-updates/s (batched vs plain, 1 GPU); job p50/p99 latency"
DATA = ("synthetic: J_xy[z] ~ U[0.6,1.8] per layer from the realization seed, J_perp=1.0, "
        "random initial angles")
# per-launch HBM bytes of the headline kernel, measured with rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE (calibrated against a device copy of known size); written by the profiling run
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_resident_L32_2048replicas.json")


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


def task_seed(base, L, sample, t):
    """Replica seed of the job config's task expansion (sample counts from 1)."""
    return base + 100000 * L + 1000 * sample + int(round(100 * t))


def disorder_batch(slsf, Ls, temps, n_real, base_seed, wolff_steps=1):
    """Replicas (L, realization, T) with the realization's couplings shared across T."""
    reps, keys = [], []
    for L in Ls:
        for r in range(n_real):
            jxy = splitmix_uniform_layers(L, base_seed + r)
            jz = np.ones(L)
            for t in temps:
                reps.append(slsf.Replica(L, L, L, t, jxy, jz, task_seed(base_seed, L, r + 1, t),
                                         float(np.pi), wolff_steps))
                keys.append((L, r, t))
    return reps, keys


def c3_replicas(slsf, n_real=256, temps=None, L=32, wolff_steps=1):
    temps = temps if temps is not None else [1.10 + 0.05 * k for k in range(8)]
    return disorder_batch(slsf, [L], temps, n_real, 1000, wolff_steps)[0]


def cpu_cores():
    """Host cores the CPU baseline may use: the CPUs this process may run on,
    capped by OMP_NUM_THREADS when set (a GPU box allots its GPU a share of the
    host's cores and states it there)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cap = os.environ.get("OMP_NUM_THREADS", "")
    if cap.isdigit() and int(cap) > 0:
        n = min(n, int(cap))
    return max(1, n)


def cpu_sample_plan(budget_s=12.0):
    """Sweeps per replica of the CPU sample: about budget_s of work on each core,
    calibrated on this host.  The first fifth thermalizes; the rest is measured
    every sweep in bins of up to 100 measurements (at least 20 bins)."""
    rate1 = _cpu_worker((0, 3, 0, 1))[0]  # calibrate one core
    sweeps = max(500, int(budget_s * rate1 / (32 ** 3)))
    therm = sweeps // 5
    measured = sweeps - therm
    binsize = max(1, min(100, measured // 20))
    return therm, measured, binsize


def cpu_baseline(plan):
    """Reference-mode oracle (SplitMix64 stream, red/black with 8- or 4-lane
    batches and vector sincos/exp as the host's width allows, Wolff DFS with
    trig cache) on one L=32 replica per core."""
    import multiprocessing as mp
    import oracle_lib as orc
    therm, measured, binsize = plan
    cores = cpu_cores()
    sweeps = therm + measured
    t0 = time.perf_counter()
    with mp.get_context("fork").Pool(cores) as pool:
        out = pool.map(_cpu_worker, [(i, sweeps, therm, binsize) for i in range(cores)])
    wall = time.perf_counter() - t0
    total = cores * sweeps * 32 ** 3
    lanes = orc.host_batch_lanes()
    return {"value": total / wall, "unit": "site-updates/s", "cores": cores,
            "kind": "port",
            "sample": f"{cores} independent L=32 replicas (realizations 0..{cores - 1} of the "
                      f"batched workload at T=1.3, one per core), {sweeps} Metropolis+Wolff "
                      f"sweeps each ({therm} thermalization, then measured every sweep), FP64, "
                      f"reference draw order, {lanes}-lane batches with vector sincos/exp",
            "per_core": float(np.mean([o[0] for o in out])),
            "cpu_model": _cpu_model(), "smt_active": _smt_active(),
            "host_cpus": os.cpu_count(), "batch_lanes": lanes,
            "bins": {i: o[1] for i, o in enumerate(out)}}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def _smt_active():
    try:
        with open("/sys/devices/system/cpu/smt/active") as f:
            return f.read().strip() == "1"
    except OSError:
        return None


def _cpu_worker(arg):
    """One replica: `sweeps` reference-mode sweeps, those after the first
    `therm` measured every sweep (the 13 series, binned by `binsize`)."""
    idx, sweeps, therm, binsize = arg
    import ctypes
    import oracle_lib as orc
    L = orc.load(fast=True)
    batch8 = orc.host_batch_lanes() == 8
    lat = orc.Lattice(32, 32, 32, lib_=L)
    lat.set_couplings(splitmix_uniform_layers(32, 1000 + idx), np.ones(32))
    r = orc.FastRng()
    L.or_rng_seed(ctypes.byref(r), 7 + idx)
    L.or_initialize_angles(lat.ref, 0, ctypes.byref(r))
    lat.refresh(False)
    rows = []
    t0 = time.perf_counter()
    for k in range(sweeps):
        lat.metro_ref(1.3, np.pi, r, batch8)
        lat.wolff_ref(1.3, r, True)
        if k >= therm:
            o = lat.observables()
            e, m2 = o["energy"], o["magnetization_squared"]
            c = [o["cos_x"], o["cos_y"], o["cos_z"]]
            sn = [o["sin_x"], o["sin_y"], o["sin_z"]]
            rows.append([e, e * e, m2, m2 * m2] + c + sn + [v * v for v in sn])
    rate = sweeps * 32 ** 3 / (time.perf_counter() - t0)
    if not rows:
        return rate, None
    series = np.array(rows).T  # 13 x measurements, in the order of the GPU series
    nb = series.shape[1] // binsize
    bins = series[:, :nb * binsize].reshape(13, nb, binsize).mean(axis=2)
    return rate, bins


def fp64_agreement(slsf, gpu_bins, cpu_bins, volume, beta):
    """FP32 production bins against FP64 CPU runs of the same couplings and
    temperature (independent chains): per observable the z-scores
    (gpu - cpu) / sqrt(err_gpu^2 + err_cpu^2) over the replicas, and the
    disorder average of the differences with its combined error.  Replicas
    with fewer than 2 bins on either side have no error bar and are left out;
    with none left the comparison is None."""
    names = ["Energy", "RhoXY", "RhoZ", "Binder", "Chi"]
    usable = [i for i, cb in cpu_bins.items()
              if cb is not None and cb.shape[1] >= 2 and i in gpu_bins
              and gpu_bins[i].shape[1] >= 2]
    if not usable:
        return None
    out = {"replicas": len(usable), "reference": "oracle reference mode, FP64, CPU",
           "observables": {}}
    for name in names:
        z, d, var = [], [], 0.0
        for i in usable:
            g = slsf.evaluate(gpu_bins[i], volume, beta)[name]
            c = slsf.evaluate(cpu_bins[i], volume, beta)[name]
            err = math.hypot(g[1], c[1])
            z.append((g[0] - c[0]) / err if err > 0 else 0.0)
            d.append(g[0] - c[0])
            var += err * err
        n = len(z)
        z = np.array(z)
        out["observables"][name] = {
            "max_abs_z": float(np.max(np.abs(z))),
            "within_2_sigma": float(np.mean(np.abs(z) <= 2.0)),
            "mean_z_squared": float(np.mean(z * z)),  # 1 +- sqrt(2 / n) when they agree
            "mean_difference": float(np.mean(d)),
            "mean_difference_error": math.sqrt(var) / n}
    return out


def headline_binsize(steps, requested):
    """Bin length of the timed run: the requested one, shortened so that the
    timed sweeps fill at least 20 bins (at least 1)."""
    return max(1, min(requested, steps // 20))


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


def line(value, unit, steps, warmup, ms_per_step, workload, dtype, higher=True, **extra):
    out = {"metric": METRIC, "value": value, "unit": unit, "n_gpus": 1, "steps": steps,
           "warmup": warmup, "ms_per_step": ms_per_step, "higher_is_better": higher,
           "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": DATA,
           "config": workload}
    out.update(extra)
    print(json.dumps(out), flush=True)


def profile_record(spl, nrep):
    """The committed counter run of k_resident (HBM bytes and VALU instructions
    per launch) for launches of `spl` sweeps over `nrep` L=32 replicas; None when
    no committed run has that launch shape."""
    try:
        with open(TRAFFIC_FILE) as f:
            recs = json.load(f)
    except (OSError, ValueError):
        return None
    for t in recs if isinstance(recs, list) else [recs]:
        if t.get("sweeps_per_launch") == spl and t.get("replicas") == nrep:
            return t
    return None


def instruction_record(nrep):
    """VALU instructions per update and SIMD busy share of k_resident from any
    committed counter run over `nrep` L=32 replicas (they do not depend on the
    launch length)."""
    try:
        with open(TRAFFIC_FILE) as f:
            recs = json.load(f)
    except (OSError, ValueError):
        return None
    for t in recs if isinstance(recs, list) else [recs]:
        if t.get("replicas") == nrep and t.get("valu_per_update"):
            return t
    return None


def _wrapped_diff(a, b):
    d = np.abs(a - b) % (2 * np.pi)
    return np.minimum(d, 2 * np.pi - d)


def fp32_accuracy(slsf, n_metro=32, n_full=8):
    """FP32 production arithmetic against FP64 on the same uniforms: one sweep
    from the same keyed initial angles by the FP32 resident kernel (its
    replay-compatible stream) and by the FP64 replay kernels, over the headline
    batch's first realizations (all 8 temperatures).  Metropolis only: a site
    whose angles differ by more than 1e-3 rad took the other decision; max
    |dtheta| is over the agreeing sites.  Metropolis + Wolff: replicas whose
    sweep differs anywhere by more than the FP32 rounding of an angle."""
    out = {}
    for key, n_real, wolff in (("metropolis", n_metro, 0), ("metropolis_wolff", n_full, 1)):
        reps = c3_replicas(slsf, n_real, wolff_steps=wolff)
        b64 = slsf.Batch(reps, slsf.FP64_REPLAY)
        b32 = slsf.Batch(reps, slsf.FP32)
        b64.init_angles()
        b32.init_angles()
        b64.sweep(1)
        b32.run(1, 0, 1, replay_stream=True)
        bad, maxd, differ = 0, 0.0, 0
        for r in range(len(reps)):
            d = _wrapped_diff(b64.get_theta(r), b32.get_theta(r))
            mism = d > 1e-3
            bad += int(mism.sum())
            if (~mism).any():
                maxd = max(maxd, float(d[~mism].max()))
            differ += int(mism.any())
        b64.close()
        b32.close()
        updates = len(reps) * 32 ** 3
        if wolff:
            out[key] = {"replicas": len(reps), "replicas_differing": differ,
                        "max_abs_dtheta_rad": maxd}
        else:
            out[key] = {"replicas": len(reps), "site_updates": updates,
                        "decision_mismatches": bad,
                        "mismatches_per_million": bad / updates * 1e6,
                        "max_abs_dtheta_rad": maxd}
    out["run"] = (f"one sweep from the keyed initial angles, L=32, realizations 0..{n_metro - 1} "
                  f"(Metropolis only) and 0..{n_full - 1} (with the Wolff step) x 8 T of the "
                  "batched workload; FP32 resident kernel on the replay stream vs FP64 replay")
    return out


# ---------------------------------------------------------------------------
def bench_c3(args, slsf):
    reps = c3_replicas(slsf, args.realizations)
    V = 32 ** 3
    binsize = headline_binsize(args.steps, args.binsize)
    b = slsf.Batch(reps, slsf.FP32)
    b.init_angles()
    b.run(args.warmup, 0, 1, sweeps_per_launch=args.sweeps_per_launch)
    b.reserve_bins(args.steps // binsize + 1)  # bin storage allocated before the clock starts
    # the timed region: exactly args.steps sweeps of every replica, with the
    # measurement of every sweep binned on the device; run() returns after the
    # batch's stream has drained, and the warmup run before it has too
    t0 = time.perf_counter()
    st = b.run(0, args.steps, binsize, sweeps_per_launch=args.sweeps_per_launch)
    wall = time.perf_counter() - t0
    b.close()
    updates = len(reps) * V * args.steps
    value = updates / wall

    # Metropolis-only rate: the same batch with no cluster step; the Wolff
    # share is the difference of the two times per sweep
    msteps = max(100, args.steps // 10)
    m = slsf.Batch(c3_replicas(slsf, args.realizations, wolff_steps=0), slsf.FP32)
    m.init_angles()
    m.run(min(args.warmup, 200), 0, 1, sweeps_per_launch=args.sweeps_per_launch)
    sm = m.run(0, msteps, headline_binsize(msteps, args.binsize),
               sweeps_per_launch=args.sweeps_per_launch)
    m.close()
    t_full = st["seconds"] / args.steps
    t_metro = sm["seconds"] / msteps
    metro_rate = len(reps) * V / t_metro
    wolff_rate = len(reps) * V / (t_full - t_metro) if t_full > t_metro else None

    plain = None
    if not args.no_plain:
        p = slsf.Batch(reps[:1], slsf.FP32)
        p.init_angles()
        p.run(5, 0, 1)
        nplain = 200
        sp = p.run(0, nplain, 10)
        p.close()
        plain = V * nplain / sp["seconds"]

    # the dominant kernel: k_resident, timed by HIP events on its launch stream
    launches = max(st["launches"], 1)
    kernel_ms = st.get("kernel_ms", 0.0)
    avg_launch_ms = kernel_ms / launches
    spl = -(-args.steps // launches)  # sweeps per launch of the timed run
    per_launch_updates = len(reps) * V * spl
    kernel_rate = updates / (kernel_ms * 1e-3) if kernel_ms > 0 else value
    prof = profile_record(spl, len(reps))
    ceil = slsf.measure_ceilings()
    # LDS form of the roofline: 32 B of LDS traffic per site update (6 neighbour
    # reads, the site's own read and its write, FP32), against the nominal LDS
    # bandwidth of the part; the measured figure beside it
    lds_bytes = 32.0
    achieved = kernel_rate * lds_bytes / 1e9
    roofline = {"bound": "lds", "achieved": achieved, "peak": LDS_PEAK_NOMINAL_GBS,
                "unit": "GB/s", "frac": achieved / LDS_PEAK_NOMINAL_GBS,
                "traffic": (prof["fetch_bytes"] + prof["write_bytes"]) if prof else None,
                "bytes_per_update": lds_bytes,
                "algorithmic_bytes_per_launch": per_launch_updates * lds_bytes,
                "peak_measured": ceil["lds_bytes_per_s"] / 1e9,
                "frac_of_measured": kernel_rate * lds_bytes / ceil["lds_bytes_per_s"]}
    # the binding ceiling: vector instruction issue (instructions per update
    # from the committed counter run of this kernel over the same replicas)
    ins = instruction_record(len(reps))
    valu = None
    if ins:
        issue = kernel_rate * ins["valu_per_update"] / 64.0
        valu = {"instr_per_update": ins["valu_per_update"], "wave_instr_per_s": issue,
                "fma_issue_measured": ceil["fma_wave_instr_per_s"],
                "frac": issue / ceil["fma_wave_instr_per_s"],
                # share of the SIMDs' cycles with a vector instruction issued
                # (SQ_ACTIVE_INST_VALU over SQ_WAVE_CYCLES, 4 waves per SIMD)
                "simd_valu_busy": ins.get("simd_valu_busy"),
                "source": ins.get("command")}
    # the HBM-equivalent figure: 8 B per update (the angle's read and write),
    # what the same rate would move if every sweep streamed the replicas
    hbm_equiv = {"bytes_per_update": 8.0, "achieved_GBs": kernel_rate * 8.0 / 1e9,
                 "peak": HBM_PEAK_GBS, "frac": kernel_rate * 8.0 / 1e9 / HBM_PEAK_GBS}

    accuracy = fp32_accuracy(slsf)
    cpu, agreement, agreement_note = None, None, None
    if not args.no_cpu_baseline:
        plan = cpu_sample_plan()
        cpu = cpu_baseline(plan)
        cpu_bins = cpu.pop("bins")
        # the FP64 comparison: the same realizations at T = 1.30 on the GPU,
        # with the CPU sample's thermalization, measured sweeps and bins (a
        # replica's chain does not depend on its batch, so these are the
        # batch's replicas 8 r + 4)
        therm, measured, cbin = plan
        sub = [reps[8 * r + 4] for r in range(min(len(cpu_bins), args.realizations))]
        g = slsf.Batch(sub, slsf.FP32)
        g.init_angles()
        g.run(therm, measured, cbin)
        gpu_bins = {r: g.all_bins(r) for r in range(len(sub))}
        g.close()
        agreement = fp64_agreement(slsf, gpu_bins, cpu_bins, V, 1.0 / 1.3)
        agreement_note = (f"separate sub-run: realizations 0..{len(sub) - 1} at T=1.30 on the "
                          f"GPU (FP32) and on the CPU (FP64 reference mode), {therm} "
                          f"thermalization + {measured} measured sweeps each, bins of {cbin}")
    line(value, "site-updates/s", args.steps, args.warmup, 1e3 * wall / args.steps,
         {"workload": f"L=32 batched disorder average: {args.realizations} realizations x "
                      f"8 T in [1.10,1.45], Metropolis + 1 Wolff + 13-series measurement "
                      f"per sweep, {args.warmup} warmup (thermalization) + {args.steps} "
                      f"timed measured sweeps, bins of {binsize}",
          "replicas": len(reps), "L": 32, "binsize": binsize},
         "fp32 (fp64 accumulators)",
         plain_single_replica=plain,
         batched_over_plain=(value / plain) if plain else None,
         metropolis_only_rate=metro_rate, wolff_only_rate=wolff_rate,
         acceptance=st["acceptance_mean"], mean_cluster_size=st["mean_cluster_size"],
         roofline=roofline, valu_issue=valu, hbm_equivalent=hbm_equiv, ceilings=ceil,
         cpu_baseline=cpu, fp32_same_uniforms=accuracy, fp64_cpu_agreement=agreement,
         fp64_cpu_agreement_run=agreement_note,
         kernel={"name": "k_resident<32, 1, paired>", "launches": st["launches"],
                 "avg_launch_ms": avg_launch_ms, "sweeps_per_launch": spl,
                 "site_updates_per_launch": per_launch_updates,
                 "rate_in_kernel": kernel_rate})


def bench_c1(args, slsf):
    L, T = 16, 1.3
    jxy = splitmix_uniform_layers(L, 1)
    # untimed warmup: a short run of the same driver loads the kernels
    slsf.run_theta_simulation(L, L, L, 1.2, 1.0, 0.0, T, thermal_sweeps=10, measure_sweeps=10,
                              measure_interval=1, correlation_rmax=8, seed=2, j_xy_array=jxy,
                              j_z_array=np.ones(L), precision=slsf.FP64_REPLAY)
    t0 = time.perf_counter()
    r = slsf.run_theta_simulation(L, L, L, 1.2, 1.0, 0.0, T, thermal_sweeps=1000,
                                  measure_sweeps=5000, measure_interval=1, correlation_rmax=8,
                                  seed=1, j_xy_array=jxy, j_z_array=np.ones(L),
                                  precision=slsf.FP64_REPLAY)
    dt = time.perf_counter() - t0
    line(6000 * L ** 3 / dt, "site-updates/s", 5000, 1000, 1e3 * dt / 6000,
         {"workload": "plain method: L=16, T=1.3, one replica, FP64 replay, 1000 + 5000 "
                      "sweeps, measured every sweep, C(r) along x up to r=8"},
         "fp64", energy=r["energy"], rho_sx=r["rho_sx"], rho_sz=r["rho_sz"],
         magnetization_squared=r["magnetization_squared"], acceptance=r["acceptance"],
         corr_x=list(map(float, r["corr_x"])))


def bench_c2(args, slsf):
    """Replicas at scale on the HBM kernels: one L=64 replica by default;
    --c2-l / --c2-replicas size it otherwise (realization r has the couplings
    of seed 1 + r), --c2-fp32-only leaves out the FP64 replay run."""
    L, T, n, sweeps = args.c2_l, 1.3, args.c2_replicas, args.c2_sweeps
    reps = [slsf.Replica(L, L, L, T, splitmix_uniform_layers(L, 1 + r), np.ones(L),
                         task_seed(1, L, 1 + r, T)) for r in range(n)]
    res = {}
    precs = (("fp32", slsf.FP32),) if args.c2_fp32_only else \
        (("fp64_replay", slsf.FP64_REPLAY), ("fp32", slsf.FP32))
    for name, prec in precs:
        b = slsf.Batch(reps, prec)
        b.init_angles()
        b.run(20, 0, 1)
        st = b.run(0, sweeps, 10)
        b.close()
        res[name] = {"rate": n * L ** 3 * sweeps / st["seconds"], "seconds": st["seconds"],
                     "acceptance": st["acceptance_mean"],
                     "mean_cluster_size": st["mean_cluster_size"]}
    f = res["fp32"]
    what = "single replica" if n == 1 else f"{n} replicas"
    line(f["rate"], "site-updates/s", sweeps, 20, 1e3 * f["seconds"] / sweeps,
         {"workload": f"{what} at scale: L={L}, T=1.3, {sweeps} Metropolis + Wolff sweeps "
                      "with measurement, HBM kernels", "L": L, "replicas": n},
         "fp32 (fp64 accumulators)", fp64_replay=res.get("fp64_replay"), fp32=f)


def bench_c4(args, slsf):
    Ls = [8, 12, 16, 24, 32]
    temps = [1.1, 1.2, 1.3, 1.4, 1.5]
    n_real = 64
    reps, keys = disorder_batch(slsf, Ls, temps, n_real, 1)
    b = slsf.Batch(reps, slsf.FP32)
    b.init_angles()
    tuned = b.autotune(16)
    therm, sweeps, binsize = 1000, 4000, 100
    b.run(therm, 0, 1)
    st = b.run(0, sweeps, binsize)
    updates = sum(r.volume for r in reps) * (therm + sweeps)
    value = sum(r.volume for r in reps) * sweeps / st["seconds"]
    # the same batch on the library's default plan (no autotuning), same sweeps
    u = slsf.Batch(reps, slsf.FP32)
    u.init_angles()
    u.run(therm, 0, 1)
    su = u.run(0, sweeps, binsize)
    u.close()
    untuned = sum(r.volume for r in reps) * sweeps / su["seconds"]
    # disorder averages per (L, T) with jackknife errors over realizations
    # (slsf_disorder_average: Binder = [<m^4>]/[<m^2>]^2, helicity moduli
    # averaged over realizations), and the crossings of the Binder ratio and of
    # L rho_xy between successive sizes (slsf_crossing, jackknifed)
    means = {L: np.zeros((len(temps), n_real, len(slsf.SERIES))) for L in Ls}
    for i, (L, r, t) in enumerate(keys):
        means[L][temps.index(t), r] = b.all_bins(i).mean(axis=1)
    b.close()
    table = {}
    for L in Ls:
        for k, t in enumerate(temps):
            av = slsf.disorder_average(means[L][k], float(L ** 3), 1.0 / t)
            table.setdefault(str(L), {})[f"{t:.1f}"] = {
                "binder": av["Binder"][0], "binder_err": av["Binder"][1],
                "rho_xy": av["RhoXY"][0], "rho_xy_err": av["RhoXY"][1],
                "rho_z": av["RhoZ"][0], "rho_z_err": av["RhoZ"][1]}
    crossings = {}
    for L1, L2 in zip(Ls[:-1], Ls[1:]):
        for name, ev, s1, s2 in (("binder", "Binder", 1.0, 1.0),
                                 ("L_rho_xy", "RhoXY", float(L1), float(L2))):
            c = slsf.crossing(temps, means[L1], float(L1 ** 3), means[L2], float(L2 ** 3), ev,
                              s1, s2)
            if c is not None:
                crossings[f"{name} {L1}/{L2}"] = {"T": c[0], "err": c[1]}
    line(value, "site-updates/s", sweeps, therm, 1e3 * st["seconds"] / sweeps,
         {"workload": "uneven batch, autotuned: L in {8,12,16,24,32} x 5 T in [1.1,1.5] x 64 "
                      "realizations (1600 replicas), 1000 + 4000 sweeps, binsize 100",
          "replicas": len(reps)},
         "fp32 (fp64 accumulators)", autotune=tuned, untuned_rate=untuned,
         tuned_over_untuned=value / untuned, binder_rho=table,
         crossings=crossings, acceptance=st["acceptance_mean"],
         mean_cluster_size=st["mean_cluster_size"], total_site_updates=updates)


def bench_c5(args, slsf):
    free, _total = slsf.device_memory()
    budget = int(free * 0.9)
    rng = np.random.default_rng(42)
    lam, window = 50.0, args.serve_window
    arrivals, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / lam)
        if t > window:
            break
        L = int(rng.choice([16, 32, 48, 64, 96, 128]))
        T = float(rng.uniform(0.3, 2.2))
        sw = int(rng.integers(1000, 10001))
        seed = int(rng.integers(0, 2 ** 63))
        arrivals.append((t, L, T, sw, seed))
    srv = slsf.Server(budget, 256, chunk_sweeps=100)
    i, res, steps = 0, [], 0
    t0 = time.monotonic()
    while True:
        now = time.monotonic() - t0
        while i < len(arrivals) and arrivals[i][0] <= now:
            _, L, T, sw, seed = arrivals[i]
            therm = sw // 10
            srv.submit(l=L, temperature=T, seed=seed, disorder_seed=seed ^ 0x5DEECE66D,
                       sweeps=sw - therm, thermalization=therm, binsize=max(1, (sw - therm) // 50))
            i += 1
        st = srv.stats()
        if i == len(arrivals) and st["queued"] == 0 and st["active"] == 0:
            break
        if st["queued"] == 0 and st["active"] == 0:
            time.sleep(max(0.0, arrivals[i][0] - now))
            continue
        srv.step()
        steps += 1
        res += srv.poll()
    wall = time.monotonic() - t0
    srv.close()
    lat = np.array([r["finish_s"] - r["submit_s"] for r in res])
    su = sum(r["site_updates"] for r in res)
    by_l = {}
    for r in res:
        by_l.setdefault(r["l"], []).append(r["finish_s"] - r["submit_s"])
    line(su / wall, "site-updates/s", steps, 0, 1e3 * wall / max(steps, 1),
         {"workload": f"serving: Poisson arrivals 50 jobs/s for {window:g} s (seed 42), L in "
                      "{16,32,48,64,96,128}, T ~ U[0.3,2.2], sweeps ~ U{1000..10000} "
                      "(10% thermalization), chunks of 100 sweeps",
          "jobs": len(arrivals), "device_budget_bytes": budget},
         "fp32 (fp64 accumulators)",
         latency_p50_s=float(np.percentile(lat, 50)), latency_p99_s=float(np.percentile(lat, 99)),
         latency_p50_by_L={str(k): float(np.median(v)) for k, v in sorted(by_l.items())},
         jobs_done=len(res), wall_s=wall,
         offered_site_updates_per_s=sum(a[1] ** 3 * a[3] for a in arrivals) / window)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=2000)
    ap.add_argument("--binsize", type=int, default=100)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--realizations", type=int, default=256)
    ap.add_argument("--sweeps-per-launch", type=int, default=0)
    ap.add_argument("--serve-window", type=float, default=3.0)
    ap.add_argument("--c2-l", type=int, default=64)
    ap.add_argument("--c2-replicas", type=int, default=1)
    ap.add_argument("--c2-sweeps", type=int, default=1000)
    ap.add_argument("--c2-fp32-only", action="store_true")
    ap.add_argument("--workload", choices=["c1", "c2", "c3", "c4", "c5"], default="c3")
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
    {"c1": bench_c1, "c2": bench_c2, "c3": bench_c3, "c4": bench_c4,
     "c5": bench_c5}[args.workload](args, slsf)


if __name__ == "__main__":
    main()
