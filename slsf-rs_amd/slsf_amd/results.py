"""Job runner and result files of the GPU path, in the reference's formats.

* ``run_job`` plays the carlo-mc ``Runner`` for one rank (src/results.rs:187-239):
  the rank's tasks of a TOML job (``idx % world == rank``, src/model.rs:62-69)
  run as one FP32 batch on the GPU -- thermalization, then ``sweeps`` measured
  sweeps binned ``binsize`` at a time on the device -- and each task gets the
  estimates ``finalize_theta_estimates`` produces (src/model.rs:906-934): every
  measured series as a binned mean with its error, plus the evaluables
  Magnetization, Chi, SpecificHeat, RhoXY, RhoZ (src/model.rs:857-888) and the
  Binder ratio, with jackknife errors.
* ``write_job_result`` / ``read_job_result`` / ``merge_job_results`` follow
  ``ThetaJobResult`` (src/model.rs:71-106) and src/results.rs:14-36,241-272:
  ``{job_name}.rank{rank}.results.json``, pretty-printed, merge sorted by task
  name with the reference's messages.
* ``dataframe_from_result`` / ``write_dataframe`` restate src/result_tools.rs:47-
  112,160-251 (columns, ``{:.16}`` floats, quoting of delimited fields).
* ``run_job_resumable`` is run_job under the job's run_time / checkpoint_time
  budget, resumable from a checkpoint image of the batch.
* ``write_job_measurements`` / ``write_task_measurements_to_path`` follow
  src/results.rs:38-129: one ``task{index+1:04}/run0001.meas.h5`` per task with
  ``observables/<series>/{bin_length, samples}`` (scalar int64; float64 bin
  means, chunks of 1000, unlimited) and ``version/{carlo_version, mc_package,
  mc_version}`` (fixed-length strings), written through ``.h5.tmp`` and a
  rename, with the reference's messages.  The device bins of a task travel in
  its ``measurement_bins``, which, like the reference's ``#[serde(skip)]``
  field, the JSON result leaves out.

An estimate is written as ``{"mean", "error", "internal_bin_len"}``, the
fields of carlo-mc's ``Estimate`` the reference itself reads
(src/results.rs:74-83; src/result_tools.rs:200-207).  Tasks with correlations
enabled also get the reference's ``CorrX_r*``, ``CorrY_r*``, ``CorrXY_r*``,
``CorrXY_z{z}_r*`` and ``CorrZ_r*`` series (src/model.rs:410-435): measured on
the device on every ``correlation_interval``-th measurement (the first
included), the batch stepping to each of those sweeps, and binned like every
series.  The reference's timers ``_ll_sweep_time`` and ``_ll_measure_time``
(src/model.rs:380,398,403,438-439) are binned like the series too: a batch runs
all its tasks' sweeps in the same launches, so a task's sweep time is its share
of the site updates of the run's wall time per sweep (the measurement on the
device is fused into the sweep kernels and counted there), and its measure
time is the host time spent finishing its measurements (reading its bins back,
taking its correlation samples) per measurement.
"""
from __future__ import annotations

import json
import math
import os
import time
from typing import Sequence

import numpy as np

from . import (EVALS, FP32, SERIES, Batch, Job, Replica, SlsfError, correlation_series_names,
               evaluate)
from . import hdf5

# fields of ThetaTask (src/model.rs:20-47), in declaration order
TASK_FIELDS = ["name", "l", "l_x", "l_y", "l_z", "temperature", "j_xy", "delta_j_xy",
               "j_z_mean", "delta_j_z", "disorder_seed", "seed", "sample", "sweeps",
               "thermalization", "binsize", "proposal_width", "wolff_steps",
               "correlation_rmax", "correlation_rmax_xy", "correlation_rmax_z",
               "correlation_interval", "j_xy_array", "j_z_array"]


def _task_json(t: dict) -> dict:
    out = {}
    for k in TASK_FIELDS:
        v = t[k]
        if isinstance(v, np.ndarray):
            v = [float(x) for x in v]
        elif isinstance(v, (np.floating, float)):
            v = float(v)
        elif isinstance(v, (np.integer, int)):
            v = int(v)
        out[k] = v
    return out


def _binned(bins: np.ndarray, bin_len: int) -> dict:
    """BinnedEstimate semantics (src/tests.rs:436-446): mean of the bins and
    std(bins) / sqrt(nbins)."""
    n = len(bins)
    mean = float(np.mean(bins)) if n else float("nan")
    err = float(np.std(bins, ddof=1) / math.sqrt(n)) if n > 1 else float("nan")
    return {"mean": mean, "error": err, "internal_bin_len": int(bin_len)}


def _rank_batch(job: Job, rank: int, world_size: int, precision: int):
    if world_size < 1 or not 0 <= rank < world_size:
        raise SlsfError(1, "rank must be below world_size")
    idx = job.selected_tasks(rank, world_size)
    tasks = [job.task(i) for i in idx]
    if not tasks:
        return idx, tasks, None
    reps = [Replica(t["l_x"], t["l_y"], t["l_z"], t["temperature"], t["j_xy_array"],
                    t["j_z_array"], int(t["seed"]), t["proposal_width"], int(t["wolff_steps"]))
            for t in tasks]
    return idx, tasks, Batch(reps, precision)


def _corr_spec(t: dict):
    """Shifts and interval of a task's correlation measurement (src/model.rs:360-365)."""
    rxy = min(int(t["correlation_rmax_xy"]), int(t["l_x"]) // 2, int(t["l_y"]) // 2)
    rz = min(int(t["correlation_rmax_z"]), int(t["l_z"]) // 2)
    return rxy, rz, max(1, int(t["correlation_interval"]))


_corr_names = correlation_series_names


class _Timers:
    """Wall time of a batch's measured runs, per measured sweep, and host time
    spent per task on finishing its measurements."""

    def __init__(self, tasks):
        vol = np.array([float(t["l_x"] * t["l_y"] * t["l_z"]) for t in tasks])
        self.share = vol / vol.sum() if len(vol) else vol
        self.segments = []  # (measured sweeps, seconds) per measured run
        self.measure_seconds = np.zeros(len(tasks))

    def run(self, b: Batch, n: int, binsize: int, **kw):
        t0 = time.perf_counter()
        b.run(0, n, binsize, **kw)
        self.segments.append((n, time.perf_counter() - t0))

    def sweep_times(self, k: int, measured: int) -> np.ndarray:
        """Task k's time per measured sweep, in measurement order."""
        per = [np.full(n, sec / n * self.share[k]) for n, sec in self.segments if n > 0]
        v = np.concatenate(per) if per else np.zeros(0)
        return v[:measured] if len(v) >= measured else np.concatenate(
            [np.full(measured - len(v), np.nan), v])

    def state(self) -> dict:
        return {"timer_segments": np.array(self.segments, dtype=np.float64).reshape(-1, 2),
                "timer_measure": self.measure_seconds}

    def restore(self, arrays) -> None:
        if arrays is not None and "timer_segments" in arrays:
            self.segments = [(int(n), float(t)) for n, t in arrays["timer_segments"]]
            self.measure_seconds = np.array(arrays["timer_measure"], dtype=np.float64)


def _bin_means(values: np.ndarray, binsize: int) -> np.ndarray:
    nb = len(values) // binsize
    return values[:nb * binsize].reshape(nb, binsize).mean(axis=1) if nb else np.zeros(0)


class _Correlations:
    """Correlation samples of the tasks that ask for them, taken on the
    measured sweeps with index m % interval == 0 (m counts from 0)."""

    def __init__(self, tasks):
        self.spec = [_corr_spec(t) for t in tasks]
        self.on = [rxy > 0 or rz > 0 for rxy, rz, _ in self.spec]
        ivals = {c for (_, _, c), on in zip(self.spec, self.on) if on}
        if len(ivals) > 1:
            raise SlsfError(1, "tasks of one batch must share correlation_interval")
        self.interval = ivals.pop() if ivals else 0
        self.samples = [[] for _ in tasks]

    def sample(self, b: Batch, timers: "_Timers | None" = None):
        for k, ((rxy, rz, _), on) in enumerate(zip(self.spec, self.on)):
            if not on:
                continue
            t0 = time.perf_counter()
            c = b.correlations(k, -1, rxy, rz)
            self.samples[k].append(np.concatenate(
                [c["corr_x"], c["corr_y"], c["corr_xy"], c["corr_xy_by_z"].ravel(), c["corr_z"]]))
            if timers is not None:
                timers.measure_seconds[k] += time.perf_counter() - t0

    def run(self, b: Batch, start: int, n: int, binsize: int, timers: "_Timers | None" = None,
            **kw):
        """Measured sweeps start .. start+n-1 (indices of measurements)."""
        timers = timers if timers is not None else _Timers([])
        if not self.interval:
            timers.run(b, n, binsize, **kw)
            return
        done = 0
        while done < n:
            m = start + done
            nxt = m + (-m) % self.interval  # next measurement taking correlations
            if nxt >= start + n:
                timers.run(b, start + n - m, binsize, **kw)
                break
            timers.run(b, nxt - m + 1, binsize, **kw)
            self.sample(b, timers)
            done = nxt - start + 1

    def state(self) -> dict:
        return {f"corr{k}": np.array(v) for k, v in enumerate(self.samples) if v}

    def restore(self, arrays, measured: int) -> None:
        """Samples saved beside a checkpoint of a batch that has measured
        `measured` sweeps: one per measurement m < measured with m % interval
        == 0, for every task that takes them."""
        want = -(-measured // self.interval) if self.interval else 0
        for k in range(len(self.samples)):
            key = f"corr{k}"
            # the samples are saved before the image, so they may run ahead of it
            self.samples[k] = (list(arrays[key])[:want]
                               if arrays is not None and key in arrays else [])
            if self.on[k] and len(self.samples[k]) != want:
                raise SlsfError(1, f"the checkpoint's correlation samples of task {k} "
                                   f"({len(self.samples[k])}) do not match its {measured} "
                                   "measured sweeps")

    def series(self, k: int, lz: int, binsize: int) -> dict:
        """Bin means of task k's correlation series (complete bins only)."""
        if not self.on[k]:
            return {}
        rxy, rz, _ = self.spec[k]
        names = _corr_names(rxy, rz, lz)
        v = np.array(self.samples[k]).reshape(-1, len(names))
        nb = v.shape[0] // binsize
        bins = v[:nb * binsize].reshape(nb, binsize, len(names)).mean(axis=1)
        return {name: bins[:, j] for j, name in enumerate(names)}


def _task_results(b: Batch, idx, tasks, measured: int, binsize: int,
                  corr: "_Correlations | None" = None, timers: "_Timers | None" = None) -> list:
    """ThetaTaskResult entries of a batch that has measured `measured` sweeps."""
    out = []
    acceptance = b.acceptance()
    for k, (i, t) in enumerate(zip(idx, tasks)):
        t0 = time.perf_counter()
        bins = b.all_bins(k) if measured else np.zeros((len(SERIES), 0))
        if timers is not None:
            timers.measure_seconds[k] += time.perf_counter() - t0
        obs = {name: _binned(bins[s], binsize) for s, name in enumerate(SERIES)}
        if bins.shape[1] > 0:
            ev = evaluate(bins, float(t["l_x"] * t["l_y"] * t["l_z"]), 1.0 / t["temperature"])
            for name in EVALS:
                if name in obs:
                    continue
                mean, err = ev[name]
                obs[name] = {"mean": float(mean), "error": float(err),
                             "internal_bin_len": binsize}
        mbins = {name: np.array(bins[s], dtype=np.float64)
                 for s, name in enumerate(SERIES) if bins.shape[1] > 0}
        if corr is not None:
            for name, v in corr.series(k, int(t["l_z"]), binsize).items():
                obs[name] = _binned(v, binsize)
                if len(v):
                    mbins[name] = v
        if timers is not None and measured:
            for name, v in (("_ll_sweep_time", timers.sweep_times(k, measured)),
                            ("_ll_measure_time",
                             np.full(measured, timers.measure_seconds[k] / measured))):
                tb = _bin_means(v, binsize)
                obs[name] = _binned(tb, binsize)
                if len(tb):
                    mbins[name] = tb
        entry = {"task": _task_json(t)}
        if i != 0:
            entry["task_index"] = int(i)  # serialised only when nonzero
        entry.update({"observables": dict(sorted(obs.items())),
                      "acceptance": float(acceptance[k]), "measurements": measured,
                      "measurement_bins": mbins})
        out.append(entry)
    return out


def run_job(job: Job, rank: int = 0, world_size: int = 1, precision: int = FP32,
            sweeps_per_launch: int = 0, use_graph: bool = False) -> dict:
    """Runs this rank's tasks of `job` on the GPU and returns a ThetaJobResult."""
    idx, tasks, b = _rank_batch(job, rank, world_size, precision)
    result = {"job_name": job.name, "rank": rank, "world_size": world_size, "tasks": []}
    if b is None:
        return result
    therm, sweeps, binsize = (int(tasks[0][k]) for k in ("thermalization", "sweeps", "binsize"))
    corr = _Correlations(tasks)
    timers = _Timers(tasks)
    try:
        b.init_angles()
        if therm:
            b.run(therm, 0, 1, sweeps_per_launch=sweeps_per_launch, use_graph=use_graph)
        if sweeps and corr.interval:
            corr.run(b, 0, sweeps, binsize, timers)
        elif sweeps:
            timers.run(b, sweeps, binsize, sweeps_per_launch=sweeps_per_launch,
                       use_graph=use_graph)
        result["tasks"] = _task_results(b, idx, tasks, sweeps, binsize, corr, timers)
    finally:
        b.close()
    return result


def checkpoint_path(job_name: str, rank: int, output_dir: str | None = None) -> str:
    """Where a rank's checkpoint lives: in the job's measurement directory."""
    return os.path.join(measurement_dir(job_name, output_dir), f"rank{rank}.checkpoint")


def _save_image(b: Batch, path: str, corr: "_Correlations | None" = None,
                timers: "_Timers | None" = None) -> None:
    parent = os.path.dirname(path)
    if parent:
        os.makedirs(parent, exist_ok=True)
    side = {}  # host-side correlation samples and timers beside the image
    if corr is not None and corr.interval:
        side.update(corr.state())
    if timers is not None:
        side.update(timers.state())
    if side:
        tmp = path + ".corr.tmp.npz"
        np.savez(tmp, **side)
        os.replace(tmp, path + ".corr.npz")
    tmp = path + ".tmp"
    b.checkpoint().tofile(tmp)
    os.replace(tmp, path)


def run_job_resumable(job: Job, path: str, rank: int = 0, world_size: int = 1,
                      run_time: float | None = None, checkpoint_time: float | None = None,
                      chunk_sweeps: int = 1000, precision: int = FP32) -> tuple:
    """run_job under a wall-time budget, resumable from a checkpoint file.

    The reference parses run_time / checkpoint_time (src/model.rs:567-574,
    1033-1060) but its runner does not use them (src/results.rs:195-198); here
    they are wired: the rank's batch resumes from `path` when that file exists
    (a checkpoint image: angles, keyed-stream counters, tallies, open bins),
    runs in chunks of `chunk_sweeps`, writes `path` every `checkpoint_time`
    seconds, and stops once `run_time` seconds have passed, writing `path`
    last.  A resumed run continues the very chain and bins of the interrupted
    one.  Defaults come from the job (12:00:00 and 00:30:00 unless it says
    otherwise).  Returns (ThetaJobResult of the sweeps done so far, summary
    with the fields of JobRunSummary: task_count, elapsed_seconds,
    checkpoint_paths, stopped_early)."""
    started = time.monotonic()
    if run_time is None or checkpoint_time is None:
        rt, ct = job.durations()
        run_time = rt if run_time is None else run_time
        checkpoint_time = ct if checkpoint_time is None else checkpoint_time
    if chunk_sweeps < 1:
        raise SlsfError(1, "chunk_sweeps must be positive")
    idx, tasks, b = _rank_batch(job, rank, world_size, precision)
    result = {"job_name": job.name, "rank": rank, "world_size": world_size, "tasks": []}
    summary = {"task_count": len(tasks), "elapsed_seconds": 0.0, "checkpoint_paths": [],
               "stopped_early": False}
    if b is None:
        return result, summary
    therm, sweeps, binsize = (int(tasks[0][k]) for k in ("thermalization", "sweeps", "binsize"))
    corr = _Correlations(tasks)
    timers = _Timers(tasks)
    try:
        if os.path.exists(path):
            b.restore(np.fromfile(path, dtype=np.uint8))
        else:
            b.init_angles()
        done = int(b.counters()[0][0])
        side = path + ".corr.npz"
        arrays = dict(np.load(side)) if os.path.exists(side) else None
        if corr.interval:
            corr.restore(arrays, max(0, done - therm))
        timers.restore(arrays)
        last = time.monotonic()
        while done < therm + sweeps:
            if time.monotonic() - started >= run_time:
                summary["stopped_early"] = True
                break
            if done < therm:
                n = min(chunk_sweeps, therm - done)
                b.run(n, 0, 1)
            else:
                n = min(chunk_sweeps, therm + sweeps - done)
                corr.run(b, done - therm, n, binsize, timers)
            done += n
            if time.monotonic() - last >= checkpoint_time:
                _save_image(b, path, corr, timers)
                last = time.monotonic()
        _save_image(b, path, corr, timers)
        summary["checkpoint_paths"] = [path]
        result["tasks"] = _task_results(b, idx, tasks, max(0, done - therm), binsize, corr,
                                        timers)
    finally:
        b.close()
    summary["elapsed_seconds"] = time.monotonic() - started
    return result, summary


def _file_stem(job_name: str) -> str:
    return os.path.basename(job_name) or job_name


def _output_dir(job_name: str, output_dir: str | None) -> str:
    """src/model.rs:948-995: the given directory, else $XY_OUTPUT_DIR, else the
    job name's parent directory, else ``results``."""
    if output_dir is not None:
        return output_dir
    env = os.environ.get("XY_OUTPUT_DIR")
    if env is not None:
        return env
    return os.path.dirname(job_name) or "results"


def result_path(job_name: str, rank: int, output_dir: str | None = None) -> str:
    return os.path.join(_output_dir(job_name, output_dir),
                        f"{_file_stem(job_name)}.rank{rank}.results.json")


def merged_result_path(job_name: str, output_dir: str | None = None) -> str:
    return os.path.join(_output_dir(job_name, output_dir), f"{_file_stem(job_name)}.results.json")


def measurement_dir(job_name: str, output_dir: str | None = None) -> str:
    return os.path.join(_output_dir(job_name, output_dir), f"{_file_stem(job_name)}.data")


def _finite(x):
    """JSON has no NaN/Inf: serde_json writes them as null."""
    if isinstance(x, float) and not math.isfinite(x):
        return None
    if isinstance(x, dict):
        return {k: _finite(v) for k, v in x.items()}
    if isinstance(x, list):
        return [_finite(v) for v in x]
    return x


def write_job_result(result: dict, output_dir: str) -> str:
    os.makedirs(output_dir, exist_ok=True)
    return write_job_result_to_path(result, os.path.join(
        output_dir, f"{result['job_name']}.rank{result['rank']}.results.json"))


def write_job_result_to_path(result: dict, path: str) -> str:
    parent = os.path.dirname(path)
    if parent:
        os.makedirs(parent, exist_ok=True)
    tasks = [{k: v for k, v in t.items() if k != "measurement_bins"} for t in result["tasks"]]
    with open(path, "w") as f:
        json.dump(_finite(dict(result, tasks=tasks)), f, indent=2)
    return path


def write_job_measurements(result: dict, output_dir: str) -> list:
    """src/results.rs:38-54: one measurement file per task."""
    paths = []
    for tr in result["tasks"]:
        task_dir = os.path.join(output_dir, f"task{tr.get('task_index', 0) + 1:04d}")
        os.makedirs(task_dir, exist_ok=True)
        path = os.path.join(task_dir, "run0001.meas.h5")
        write_task_measurements_to_path(tr, path)
        paths.append(path)
    return paths


def write_task_measurements_to_path(task_result: dict, path: str) -> str:
    """src/results.rs:56-129, the same groups, datasets and error messages."""
    bins = task_result.get("measurement_bins") or {}
    if not bins:
        raise ValueError("theta task result does not contain measurement bins")
    observables = {}
    for name in sorted(bins):
        samples = np.asarray(bins[name], dtype=np.float64)
        est = task_result["observables"].get(name)
        if est is None:
            raise ValueError(f"missing observable estimate for {name}")
        if int(est["internal_bin_len"]) == 0:
            raise ValueError(f"{name}: internal bin length must be positive")
        if samples.size == 0:
            raise ValueError(f"{name}: no complete measurement bins to write")
        observables[name] = {
            "bin_length": hdf5.Dataset("i8", int(est["internal_bin_len"]), shape=()),
            "samples": hdf5.Dataset("f8", samples, maxshape=(hdf5.UNDEF,), chunks=(1000,)),
        }
    version = {
        "carlo_version": hdf5.Dataset("S", os.environ.get("XY_CARLO_VERSION", "0.3.4")),
        "mc_package": hdf5.Dataset("S", "SLSF.XYCarlo"),
        "mc_version": hdf5.Dataset("S", os.environ.get("XY_MC_VERSION", "0.1.0")),
    }
    parent = os.path.dirname(path)
    if parent:
        os.makedirs(parent, exist_ok=True)
    stem, ext = os.path.splitext(path)
    tmp = (stem if ext else path) + ".h5.tmp"
    hdf5.write_file(tmp, {"observables": observables, "version": version})
    os.replace(tmp, path)
    return path


def read_job_result(path: str) -> dict:
    with open(path) as f:
        return json.load(f)


def merge_job_results(results: Sequence[dict]) -> dict:
    """src/results.rs:246-272."""
    if not results:
        raise ValueError("at least one rank result is required for merge")
    first = results[0]
    tasks = list(first["tasks"])
    for r in results[1:]:
        if r["job_name"] != first["job_name"]:
            raise ValueError("cannot merge results from different jobs")
        if r["world_size"] != first["world_size"]:
            raise ValueError("cannot merge results with different world sizes")
        tasks.extend(r["tasks"])
    tasks = sorted(tasks, key=lambda t: t["task"]["name"])
    for k, t in enumerate(tasks):
        if t.get("task_index", 0) == 0:
            t = dict(t)
            if k != 0:
                t["task_index"] = k
            tasks[k] = t
    return {"job_name": first["job_name"], "rank": 0, "world_size": first["world_size"],
            "tasks": tasks}


def format_float(v) -> str:
    """src/result_tools.rs:241-251 ({:.16} for finite values)."""
    if v is None:
        return "NaN"
    v = float(v)
    if math.isfinite(v):
        return f"{v:.16f}"
    if math.isnan(v):
        return "NaN"
    return "-Inf" if v < 0 else "Inf"


def dataframe_from_result(result: dict) -> tuple[list[str], list[dict]]:
    """Columns and rows of src/result_tools.rs:54-96,160-205."""
    names = sorted({n for t in result["tasks"] for n in t["observables"]})
    columns = ["job_name", "rank", "world_size", "task_index", "task_name", "L", "Lx", "Ly",
               "Lz", "T", "Jxy", "DeltaJxy", "Jz", "DeltaJz", "sample", "seed",
               "disorder_seed", "sweeps", "thermalization", "binsize", "measurements",
               "acceptance"]
    for n in names:
        columns += [n, f"{n}_error", f"{n}_measurement"]
    rows = []
    for tr in result["tasks"]:
        t = tr["task"]
        row = {"job_name": result["job_name"], "rank": str(result["rank"]),
               "world_size": str(result["world_size"]),
               "task_index": str(tr.get("task_index", 0)), "task_name": t["name"],
               "L": str(t["l"]), "Lx": str(t["l_x"]), "Ly": str(t["l_y"]), "Lz": str(t["l_z"]),
               "T": format_float(t["temperature"]), "Jxy": format_float(t["j_xy"]),
               "DeltaJxy": format_float(t["delta_j_xy"]), "Jz": format_float(t["j_z_mean"]),
               "DeltaJz": format_float(t["delta_j_z"]), "sample": str(t["sample"]),
               "seed": str(t["seed"]), "disorder_seed": str(t["disorder_seed"]),
               "sweeps": str(t["sweeps"]), "thermalization": str(t["thermalization"]),
               "binsize": str(t["binsize"]), "measurements": str(tr["measurements"]),
               "acceptance": format_float(tr["acceptance"])}
        for n, e in tr["observables"].items():
            m, s = format_float(e["mean"]), format_float(e["error"])
            row[n] = m
            row[f"{n}_error"] = s
            row[f"{n}_measurement"] = f"{m} +/- {s}"
        rows.append(row)
    return columns, rows


def _escape(field: str, delim: str) -> str:
    if delim in field or "\n" in field or "\r" in field or '"' in field:
        return '"' + field.replace('"', '""') + '"'
    return field


def write_dataframe(result_json: str, output_path: str) -> str:
    """Comma-separated for a .csv path, tab-separated otherwise."""
    columns, rows = dataframe_from_result(read_job_result(result_json))
    delim = "," if output_path.endswith(".csv") else "\t"
    parent = os.path.dirname(output_path)
    if parent:
        os.makedirs(parent, exist_ok=True)
    with open(output_path, "w") as f:
        f.write(delim.join(_escape(c, delim) for c in columns) + "\n")
        for r in rows:
            f.write(delim.join(_escape(r.get(c, ""), delim) for c in columns) + "\n")
    return output_path
