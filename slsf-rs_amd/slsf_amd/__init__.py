"""ctypes binding of libslsf.so, the C ABI declared in include/slsf.h.

This is a thin host-side mirror of the reference's interface for tests, the
benchmark and Python callers: ``Batch`` stands for a set of ``ThetaModel``
instances (src/model.rs:254-477) whose lattices live on the GPU, and its methods
are the free functions the reference calls (src/updates.rs:440,551;
src/observables.rs:80,166).  Every call goes through the compiled library; there
is no Python or CPU fallback -- a missing library raises at import time.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from typing import Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libslsf.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is missing: build it with `make -C slsf-rs_amd` "
        "(or __graft_entry__.build()) -- there is no fallback implementation")

_lib = ctypes.CDLL(LIB_PATH)

FP64_REPLAY = 0
FP32 = 1

SERIES = ["Energy", "EnergySquared", "MagnetizationSquared", "MagnetizationFourth",
          "_helicity_cos_x", "_helicity_cos_y", "_helicity_cos_z",
          "_helicity_sin_x", "_helicity_sin_y", "_helicity_sin_z",
          "_helicity_sin2_x", "_helicity_sin2_y", "_helicity_sin2_z"]
EVALS = ["Magnetization", "Chi", "SpecificHeat", "RhoXY", "RhoZ", "Binder", "Energy"]


def correlation_series_names(rxy: int, rz: int, lz: int) -> list:
    """Correlation series names in the order src/model.rs:413-435 measures them."""
    r = range(1, rxy + 1)
    return ([f"CorrX_r{k}" for k in r] + [f"CorrY_r{k}" for k in r] +
            [f"CorrXY_r{k}" for k in r] +
            [f"CorrXY_z{z}_r{k}" for z in range(lz) for k in r] +
            [f"CorrZ_r{k}" for k in range(1, rz + 1)])


class SlsfError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code
        self.message = message


class ReplicaDesc(ctypes.Structure):
    _fields_ = [("l_x", ctypes.c_int32), ("l_y", ctypes.c_int32), ("l_z", ctypes.c_int32),
                ("wolff_steps", ctypes.c_int32), ("temperature", ctypes.c_double),
                ("proposal_width", ctypes.c_double),
                ("j_xy", ctypes.POINTER(ctypes.c_double)),
                ("j_z", ctypes.POINTER(ctypes.c_double)), ("seed", ctypes.c_uint64)]


class Observables(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "energy", "magnetization_squared", "cos_x", "cos_y", "cos_z",
        "sin_x", "sin_y", "sin_z")]


class RunOptions(ctypes.Structure):
    _fields_ = [("thermalization", ctypes.c_int64), ("sweeps", ctypes.c_int64),
                ("binsize", ctypes.c_int32), ("sweeps_per_launch", ctypes.c_int32),
                ("use_graph", ctypes.c_int32), ("stream", ctypes.c_int32),
                ("hbm_form", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class RunStats(ctypes.Structure):
    _fields_ = [("seconds", ctypes.c_double), ("kernel_ms", ctypes.c_double),
                ("site_updates", ctypes.c_int64), ("launches", ctypes.c_int64),
                ("acceptance_mean", ctypes.c_double),
                ("mean_cluster_size", ctypes.c_double)]


class TuneRecord(ctypes.Structure):
    _fields_ = [("l", ctypes.c_int32), ("columns_per_thread", ctypes.c_int32),
                ("waves_per_replica", ctypes.c_int32), ("sweeps_per_launch", ctypes.c_int32),
                ("site_updates_per_s", ctypes.c_double), ("chosen", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


class JobRequest(ctypes.Structure):
    _fields_ = [("l", ctypes.c_int32), ("wolff_steps", ctypes.c_int32),
                ("temperature", ctypes.c_double), ("proposal_width", ctypes.c_double),
                ("j_xy_min", ctypes.c_double), ("j_xy_max", ctypes.c_double),
                ("j_perp", ctypes.c_double), ("disorder_seed", ctypes.c_uint64),
                ("seed", ctypes.c_uint64), ("thermalization", ctypes.c_int64),
                ("sweeps", ctypes.c_int64), ("binsize", ctypes.c_int32),
                ("correlation_interval", ctypes.c_int32),
                ("correlation_rmax_xy", ctypes.c_int32), ("correlation_rmax_z", ctypes.c_int32)]


class JobResult(ctypes.Structure):
    _fields_ = [("job_id", ctypes.c_int64), ("l", ctypes.c_int32), ("nbins", ctypes.c_int32),
                ("submit_s", ctypes.c_double), ("start_s", ctypes.c_double),
                ("finish_s", ctypes.c_double), ("acceptance", ctypes.c_double),
                ("value", ctypes.c_double * 7), ("error", ctypes.c_double * 7),
                ("series_mean", ctypes.c_double * 13), ("site_updates", ctypes.c_int64),
                ("corr_rxy", ctypes.c_int32), ("corr_rz", ctypes.c_int32),
                ("corr_nbins", ctypes.c_int32), ("corr_nseries", ctypes.c_int32)]


class ServerStats(ctypes.Structure):
    _fields_ = [("queued", ctypes.c_int64), ("active", ctypes.c_int64),
                ("finished", ctypes.c_int64), ("device_bytes", ctypes.c_int64),
                ("site_updates", ctypes.c_int64), ("clock_s", ctypes.c_double)]


class Task(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 128),
                ("l", ctypes.c_int32), ("l_x", ctypes.c_int32), ("l_y", ctypes.c_int32),
                ("l_z", ctypes.c_int32), ("temperature", ctypes.c_double),
                ("j_xy", ctypes.c_double), ("delta_j_xy", ctypes.c_double),
                ("j_z_mean", ctypes.c_double), ("delta_j_z", ctypes.c_double),
                ("disorder_seed", ctypes.c_uint64), ("seed", ctypes.c_uint64),
                ("sample", ctypes.c_int32), ("sweeps", ctypes.c_int64),
                ("thermalization", ctypes.c_int64), ("binsize", ctypes.c_int32),
                ("proposal_width", ctypes.c_double), ("wolff_steps", ctypes.c_int32),
                ("correlation_rmax", ctypes.c_int32), ("correlation_rmax_xy", ctypes.c_int32),
                ("correlation_rmax_z", ctypes.c_int32),
                ("correlation_interval", ctypes.c_int32), ("disorder_kind", ctypes.c_int32)]


class SimOptions(ctypes.Structure):
    _fields_ = [("thermal_sweeps", ctypes.c_int32), ("measure_sweeps", ctypes.c_int32),
                ("measure_interval", ctypes.c_int32), ("proposal_width", ctypes.c_double),
                ("wolff_steps", ctypes.c_int32), ("correlation_interval", ctypes.c_int32),
                ("correlation_rmax", ctypes.c_int32), ("correlation_rmax_xy", ctypes.c_int32),
                ("correlation_rmax_z", ctypes.c_int32), ("cold_init", ctypes.c_int32)]


class SimResult(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "energy", "rho_sx", "rho_sy", "rho_sz", "std_rho_sx", "std_rho_sy", "std_rho_sz",
        "magnetization_squared", "chi", "acceptance")] + [
        ("num_measurements", ctypes.c_int32), ("num_correlation_measurements", ctypes.c_int32),
        ("rxy", ctypes.c_int32), ("rz", ctypes.c_int32)] + [
        (n, ctypes.POINTER(ctypes.c_double)) for n in (
            "corr_x", "corr_y", "corr_xy", "corr_xy_by_z", "corr_z")]


_P = ctypes.POINTER
_dp = _P(ctypes.c_double)
_u8p = _P(ctypes.c_uint8)
_i64p = _P(ctypes.c_int64)
_i32p = _P(ctypes.c_int32)
_u32p = _P(ctypes.c_uint32)

_SIGS = {
    "slsf_last_error": (ctypes.c_char_p, []),
    "slsf_abi_version": (ctypes.c_int32, []),
    "slsf_batch_create": (ctypes.c_int, [_P(ReplicaDesc), ctypes.c_int32, ctypes.c_int32,
                                         _P(ctypes.c_void_p)]),
    "slsf_batch_destroy": (None, [ctypes.c_void_p]),
    "slsf_batch_size": (ctypes.c_int32, [ctypes.c_void_p]),
    "slsf_batch_precision": (ctypes.c_int32, [ctypes.c_void_p]),
    "slsf_batch_device_bytes": (ctypes.c_int64, [ctypes.c_void_p]),
    "slsf_init_angles": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "slsf_get_theta": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, _dp]),
    "slsf_set_theta": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, _dp]),
    "slsf_get_counters": (ctypes.c_int, [ctypes.c_void_p, _u32p, _u32p]),
    "slsf_set_counters": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32]),
    "slsf_checkpoint_size": (ctypes.c_int, [ctypes.c_void_p, _i64p]),
    "slsf_checkpoint_save": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]),
    "slsf_checkpoint_load": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]),
    "slsf_metropolis_sweep": (ctypes.c_int, [ctypes.c_void_p, _dp, _u8p]),
    "slsf_wolff_cluster_step": (ctypes.c_int, [ctypes.c_void_p, _i64p, _u8p]),
    "slsf_sweep": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, _dp]),
    "slsf_measure_observables": (ctypes.c_int, [ctypes.c_void_p, _P(Observables)]),
    "slsf_correlation_rmax": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                             ctypes.c_int32, ctypes.c_int32, _i32p, _i32p]),
    "slsf_measure_correlations": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32,
                                                 ctypes.c_int32, ctypes.c_int32,
                                                 ctypes.c_int32, _dp, _dp, _dp, _dp, _dp]),
    "slsf_measure_correlations_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32,
                                                         ctypes.c_int32, ctypes.c_int32,
                                                         ctypes.c_int32, _dp, _dp, _dp, _dp,
                                                         _dp]),
    "slsf_series_name": (ctypes.c_char_p, [ctypes.c_int32]),
    "slsf_run": (ctypes.c_int, [ctypes.c_void_p, _P(RunOptions), _P(RunStats)]),
    "slsf_get_bins": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, _dp,
                                     ctypes.c_int32, _i32p]),
    "slsf_get_acceptance": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "slsf_reserve_bins": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "slsf_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, _P(TuneRecord),
                                     ctypes.c_int32, _i32p]),
    "slsf_binned_estimate": (ctypes.c_int, [_dp, ctypes.c_int64, ctypes.c_int32, _dp, _dp,
                                            _dp, _i64p]),
    "slsf_blocking_stderr": (ctypes.c_double, [_dp, ctypes.c_int64]),
    "slsf_eval_name": (ctypes.c_char_p, [ctypes.c_int32]),
    "slsf_evaluate": (ctypes.c_int, [_dp, ctypes.c_int64, ctypes.c_double, ctypes.c_double,
                                     _dp, _dp]),
    "slsf_jackknife": (ctypes.c_int, [_dp, ctypes.c_int64, ctypes.c_double, _dp, _dp]),
    "slsf_disorder_average": (ctypes.c_int, [_dp, ctypes.c_int32, ctypes.c_double,
                                             ctypes.c_double, _dp, _dp]),
    "slsf_crossing": (ctypes.c_int, [_dp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _dp,
                                     ctypes.c_double, ctypes.c_double, _dp, ctypes.c_double,
                                     ctypes.c_double, _i32p, _dp, _dp]),
    "slsf_job_from_toml": (ctypes.c_int, [ctypes.c_char_p, _P(ctypes.c_void_p)]),
    "slsf_job_destroy": (None, [ctypes.c_void_p]),
    "slsf_job_name": (ctypes.c_char_p, [ctypes.c_void_p]),
    "slsf_job_num_tasks": (ctypes.c_int64, [ctypes.c_void_p]),
    "slsf_job_task": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, _P(Task)]),
    "slsf_job_task_couplings": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, _dp, _dp]),
    "slsf_job_durations": (ctypes.c_int, [ctypes.c_void_p, _i64p, _i64p]),
    "slsf_job_selected_tasks": (ctypes.c_int64, [ctypes.c_void_p, ctypes.c_int32,
                                                 ctypes.c_int32, _i64p, ctypes.c_int64]),
    "slsf_parse_duration": (ctypes.c_int, [ctypes.c_char_p, _i64p]),
    "slsf_server_create": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                          _P(ctypes.c_void_p)]),
    "slsf_server_destroy": (None, [ctypes.c_void_p]),
    "slsf_server_submit": (ctypes.c_int, [ctypes.c_void_p, _P(JobRequest), _i64p]),
    "slsf_server_step": (ctypes.c_int, [ctypes.c_void_p, _i32p]),
    "slsf_server_poll": (ctypes.c_int, [ctypes.c_void_p, _P(JobResult), ctypes.c_int32, _i32p]),
    "slsf_server_get_stats": (ctypes.c_int, [ctypes.c_void_p, _P(ServerStats)]),
    "slsf_server_correlations": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, _dp,
                                                ctypes.c_int64, _i64p]),
    "slsf_server_slot_bytes": (ctypes.c_int64, [ctypes.c_int32]),
    "slsf_device_memory": (ctypes.c_int, [_i64p, _i64p]),
    "slsf_measure_ceilings": (ctypes.c_int, [ctypes.c_void_p]),
    "slsf_check_exp2_floor": (ctypes.c_int, [ctypes.POINTER(ctypes.c_uint64)]),
    "slsf_run_theta_simulation": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                                 ctypes.c_double, ctypes.c_double,
                                                 ctypes.c_double, ctypes.c_double, _dp, _dp,
                                                 _P(SimOptions), ctypes.c_uint64,
                                                 ctypes.c_int32, _P(SimResult)]),
}

for _name, (_res, _args) in _SIGS.items():
    _fn = getattr(_lib, _name)
    _fn.restype = _res
    _fn.argtypes = _args


def lib():
    return _lib


def last_error() -> str:
    return _lib.slsf_last_error().decode("utf-8")


def _check(rc: int):
    if rc != 0:
        raise SlsfError(rc, last_error())


def _ptr(a: np.ndarray | None, ctype):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctype))


@dataclass
class Replica:
    """One lattice: sizes, temperature, proposal width, Wolff steps, per-layer
    couplings j_xy[z] (in-plane) and j_z[z] (bond z -> z+1), and the key seed."""
    l_x: int
    l_y: int
    l_z: int
    temperature: float
    j_xy: Sequence[float]
    j_z: Sequence[float]
    seed: int
    proposal_width: float = float(np.pi)
    wolff_steps: int = 1

    @property
    def volume(self) -> int:
        return self.l_x * self.l_y * self.l_z


class Batch:
    """Device-resident replicas behind one slsf_batch handle."""

    def __init__(self, replicas: Sequence[Replica], precision: int = FP64_REPLAY):
        self.replicas = list(replicas)
        self._keep = []
        descs = (ReplicaDesc * len(self.replicas))()
        for d, r in zip(descs, self.replicas):
            jxy = np.ascontiguousarray(r.j_xy, dtype=np.float64)
            jz = np.ascontiguousarray(r.j_z, dtype=np.float64)
            if len(jxy) < max(r.l_z, 0) or len(jz) < max(r.l_z, 0):
                raise ValueError("coupling arrays must have l_z entries")
            self._keep += [jxy, jz]
            d.l_x, d.l_y, d.l_z = r.l_x, r.l_y, r.l_z
            d.wolff_steps = r.wolff_steps
            d.temperature = r.temperature
            d.proposal_width = r.proposal_width
            d.j_xy = _ptr(jxy, ctypes.c_double)
            d.j_z = _ptr(jz, ctypes.c_double)
            d.seed = r.seed & 0xFFFFFFFFFFFFFFFF
        h = ctypes.c_void_p()
        self._h = None
        _check(_lib.slsf_batch_create(descs, len(self.replicas), precision, ctypes.byref(h)))
        self._h = h
        self.precision = precision
        self.volumes = np.array([r.volume for r in self.replicas], dtype=np.int64)

    def close(self):
        if self._h:
            _lib.slsf_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self.replicas)

    @property
    def device_bytes(self) -> int:
        return int(_lib.slsf_batch_device_bytes(self._h))

    def init_angles(self, cold: bool = False):
        _check(_lib.slsf_init_angles(self._h, int(cold)))

    def get_theta(self, replica: int = 0) -> np.ndarray:
        out = np.empty(self.replicas[replica].volume, dtype=np.float64)
        _check(_lib.slsf_get_theta(self._h, replica, _ptr(out, ctypes.c_double)))
        return out

    def set_theta(self, replica: int, theta: np.ndarray):
        t = np.ascontiguousarray(theta, dtype=np.float64)
        assert t.size == self.replicas[replica].volume
        _check(_lib.slsf_set_theta(self._h, replica, _ptr(t, ctypes.c_double)))

    def counters(self):
        m = np.zeros(len(self), dtype=np.uint32)
        w = np.zeros(len(self), dtype=np.uint32)
        _check(_lib.slsf_get_counters(self._h, _ptr(m, ctypes.c_uint32),
                                      _ptr(w, ctypes.c_uint32)))
        return m, w

    def set_counters(self, metro: int, wolff: int):
        _check(_lib.slsf_set_counters(self._h, metro, wolff))

    def checkpoint(self) -> np.ndarray:
        """The batch's whole run state as a checkpoint image (uint8 array)."""
        n = ctypes.c_int64()
        _check(_lib.slsf_checkpoint_size(self._h, ctypes.byref(n)))
        img = np.empty(n.value, dtype=np.uint8)
        _check(_lib.slsf_checkpoint_save(self._h, img.ctypes.data_as(ctypes.c_void_p), n.value))
        return img

    def restore(self, image) -> None:
        """Loads a checkpoint image saved from a batch of the same replicas."""
        img = np.frombuffer(image, dtype=np.uint8) if not isinstance(image, np.ndarray) \
            else np.ascontiguousarray(image, dtype=np.uint8)
        _check(_lib.slsf_checkpoint_load(self._h, img.ctypes.data_as(ctypes.c_void_p),
                                         img.size))

    def metropolis_sweep(self, want_decisions: bool = False):
        acc = np.zeros(len(self), dtype=np.float64)
        dec = np.zeros(int(self.volumes.sum()), dtype=np.uint8) if want_decisions else None
        _check(_lib.slsf_metropolis_sweep(self._h, _ptr(acc, ctypes.c_double),
                                          _ptr(dec, ctypes.c_uint8)))
        return (acc, dec) if want_decisions else acc

    def wolff_step(self, want_members: bool = False):
        cs = np.zeros(len(self), dtype=np.int64)
        mem = np.zeros(int(self.volumes.sum()), dtype=np.uint8) if want_members else None
        _check(_lib.slsf_wolff_cluster_step(self._h, _ptr(cs, ctypes.c_int64),
                                            _ptr(mem, ctypes.c_uint8)))
        return (cs, mem) if want_members else cs

    def sweep(self, nsweeps: int = 1) -> np.ndarray:
        acc = np.zeros(len(self), dtype=np.float64)
        _check(_lib.slsf_sweep(self._h, nsweeps, _ptr(acc, ctypes.c_double)))
        return acc

    def observables(self) -> list[dict]:
        out = (Observables * len(self))()
        _check(_lib.slsf_measure_observables(self._h, out))
        return [{n: getattr(o, n) for n, _ in Observables._fields_} for o in out]

    def correlation_rmax(self, replica=0, rmax=-1, rmax_xy=-1, rmax_z=-1):
        a, b = ctypes.c_int32(), ctypes.c_int32()
        _check(_lib.slsf_correlation_rmax(self._h, replica, rmax, rmax_xy, rmax_z,
                                          ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def correlations(self, replica=0, rmax=-1, rmax_xy=-1, rmax_z=-1,
                     noalloc: bool = False) -> dict:
        """C(r) of one replica (measure_theta_correlations_with_scratch); with
        noalloc=True through the entry point that allocates no device memory."""
        rxy, rz = self.correlation_rmax(replica, rmax, rmax_xy, rmax_z)
        lz = self.replicas[replica].l_z
        cx, cy, cxy = (np.zeros(max(rxy, 1)) for _ in range(3))
        cbz = np.zeros(max(lz * rxy, 1))
        cz = np.zeros(max(rz, 1))
        fn = (_lib.slsf_measure_correlations_noalloc if noalloc
              else _lib.slsf_measure_correlations)
        _check(fn(
            self._h, replica, rmax, rmax_xy, rmax_z, *(_ptr(a, ctypes.c_double)
                                                       for a in (cx, cy, cxy, cbz, cz))))
        return {"corr_x": cx[:rxy], "corr_y": cy[:rxy], "corr_xy": cxy[:rxy],
                "corr_xy_by_z": cbz[:lz * rxy].reshape(lz, rxy), "corr_z": cz[:rz]}

    def run(self, thermalization: int, sweeps: int, binsize: int,
            sweeps_per_launch: int = 0, use_graph: bool = False,
            replay_stream: bool = False, hbm_form: int = 0) -> dict:
        """slsf_run: thermalization + measured sweeps with on-device bins.
        hbm_form (replicas on the HBM kernels): 0 by launch size, 1 slab
        sweeps, 2 per-phase sweeps -- the same chains and bins either way."""
        opt = RunOptions(thermalization, sweeps, binsize, sweeps_per_launch, int(use_graph),
                         int(replay_stream), int(hbm_form), 0)
        st = RunStats()
        _check(_lib.slsf_run(self._h, ctypes.byref(opt), ctypes.byref(st)))
        return {n: getattr(st, n) for n, _ in RunStats._fields_}

    def bins(self, replica: int, series: int | str) -> np.ndarray:
        s = SERIES.index(series) if isinstance(series, str) else series
        n = ctypes.c_int32()
        _check(_lib.slsf_get_bins(self._h, replica, s, None, 0, ctypes.byref(n)))
        out = np.zeros(max(n.value, 1))
        _check(_lib.slsf_get_bins(self._h, replica, s, _ptr(out, ctypes.c_double), n.value,
                                  ctypes.byref(n)))
        return out[:n.value]

    def reserve_bins(self, nbins: int):
        """Device storage for nbins completed bins per replica and series, so
        that measured runs up to that many bins allocate nothing."""
        _check(_lib.slsf_reserve_bins(self._h, nbins))

    def all_bins(self, replica: int) -> np.ndarray:
        rows = [self.bins(replica, s) for s in range(len(SERIES))]
        n = min(len(r) for r in rows)
        return np.stack([r[:n] for r in rows])

    def autotune(self, trial_sweeps: int = 16) -> list[dict]:
        """Pick waves per replica and sweeps per launch for every replica size;
        the chains are left exactly as they were."""
        recs = (TuneRecord * 64)()
        n = ctypes.c_int32()
        _check(_lib.slsf_autotune(self._h, trial_sweeps, recs, 64, ctypes.byref(n)))
        return [{k: getattr(r, k) for k, _ in TuneRecord._fields_ if k != "reserved"}
                for r in recs[:n.value]]

    def acceptance(self) -> np.ndarray:
        out = np.zeros(len(self))
        _check(_lib.slsf_get_acceptance(self._h, _ptr(out, ctypes.c_double)))
        return out


# ---- plain driver (src/simulation.rs:13) --------------------------------------
def run_theta_simulation(l_x, l_y, l_z, j_xy, j_z_mean, delta_j_z, temperature, *,
                         thermal_sweeps=1000, measure_sweeps=5000, measure_interval=10,
                         proposal_width=float(np.pi), wolff_steps=1, correlation_interval=1,
                         correlation_rmax=None, correlation_rmax_xy=None,
                         correlation_rmax_z=None, cold_init=False, seed=0,
                         j_xy_array=None, j_z_array=None, precision=FP64_REPLAY) -> dict:
    """GPU run_theta_simulation; defaults are ThetaSimulationOptions::default
    (src/types.rs:280-296)."""
    none = lambda v: -1 if v is None else int(v)
    o = SimOptions(thermal_sweeps, measure_sweeps, measure_interval, proposal_width,
                   wolff_steps, correlation_interval, none(correlation_rmax),
                   none(correlation_rmax_xy), none(correlation_rmax_z), int(cold_init))
    cap = max(1, max(l_x, l_y, 1) // 2)
    bufs = [np.zeros(cap), np.zeros(cap), np.zeros(cap), np.zeros(max(1, l_z) * cap),
            np.zeros(max(1, l_z // 2))]
    r = SimResult()
    r.corr_x, r.corr_y, r.corr_xy, r.corr_xy_by_z, r.corr_z = (_ptr(b, ctypes.c_double)
                                                               for b in bufs)
    ax = None if j_xy_array is None else np.ascontiguousarray(j_xy_array, dtype=np.float64)
    az = None if j_z_array is None else np.ascontiguousarray(j_z_array, dtype=np.float64)
    _check(_lib.slsf_run_theta_simulation(l_x, l_y, l_z, j_xy, j_z_mean, delta_j_z, temperature,
                                          _ptr(ax, ctypes.c_double), _ptr(az, ctypes.c_double),
                                          ctypes.byref(o), seed & 0xFFFFFFFFFFFFFFFF,
                                          precision, ctypes.byref(r)))
    out = {n: getattr(r, n) for n, t in SimResult._fields_ if t in (ctypes.c_double,
                                                                    ctypes.c_int32)}
    rxy, rz = r.rxy, r.rz
    out.update(corr_x=bufs[0][:rxy], corr_y=bufs[1][:rxy], corr_xy=bufs[2][:rxy],
               corr_xy_by_z=bufs[3][:l_z * rxy].reshape(l_z, rxy), corr_z=bufs[4][:rz])
    return out


# ---- estimators --------------------------------------------------------------
def binned_estimate(samples, binsize: int):
    x = np.ascontiguousarray(samples, dtype=np.float64)
    bins = np.zeros(max(len(x) // max(binsize, 1), 1))
    m, s, nb = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
    _check(_lib.slsf_binned_estimate(_ptr(x, ctypes.c_double), len(x), binsize,
                                     ctypes.byref(m), ctypes.byref(s),
                                     _ptr(bins, ctypes.c_double), ctypes.byref(nb)))
    return m.value, s.value, bins[:nb.value]


def blocking_stderr(x) -> float:
    a = np.ascontiguousarray(x, dtype=np.float64)
    return float(_lib.slsf_blocking_stderr(_ptr(a, ctypes.c_double), len(a)))


def evaluate(bins: np.ndarray, volume: float, beta: float) -> dict:
    b = np.ascontiguousarray(bins, dtype=np.float64)
    assert b.shape[0] == len(SERIES)
    v = np.zeros(len(EVALS))
    e = np.zeros(len(EVALS))
    _check(_lib.slsf_evaluate(_ptr(b, ctypes.c_double), b.shape[1], volume, beta,
                              _ptr(v, ctypes.c_double), _ptr(e, ctypes.c_double)))
    return {n: (v[i], e[i]) for i, n in enumerate(EVALS)}


def jackknife(loo, f_all: float):
    a = np.ascontiguousarray(loo, dtype=np.float64)
    est, err = ctypes.c_double(), ctypes.c_double()
    _check(_lib.slsf_jackknife(_ptr(a, ctypes.c_double), len(a), f_all, ctypes.byref(est),
                               ctypes.byref(err)))
    return est.value, err.value


def disorder_average(means, volume: float, beta: float) -> dict:
    """Disorder averages of the evaluables over realizations with jackknife
    errors; means[r][s] = realization r's mean of series s (slsf_disorder_average)."""
    m = np.ascontiguousarray(means, dtype=np.float64)
    assert m.ndim == 2 and m.shape[1] == len(SERIES)
    v = np.zeros(len(EVALS))
    e = np.zeros(len(EVALS))
    _check(_lib.slsf_disorder_average(_ptr(m, ctypes.c_double), m.shape[0], volume, beta,
                                      _ptr(v, ctypes.c_double), _ptr(e, ctypes.c_double)))
    return {n: (v[i], e[i]) for i, n in enumerate(EVALS)}


def crossing(temps, means_a, volume_a: float, means_b, volume_b: float,
             evaluable: str = "Binder", scale_a: float = 1.0, scale_b: float = 1.0):
    """Crossing temperature of scale * (a disorder-averaged evaluable) between
    two sizes with its jackknife error over realizations (slsf_crossing);
    means_x[t][r][s].  Returns (temperature, error) or None when the curves do
    not cross."""
    t = np.ascontiguousarray(temps, dtype=np.float64)
    a = np.ascontiguousarray(means_a, dtype=np.float64)
    b = np.ascontiguousarray(means_b, dtype=np.float64)
    assert a.shape == b.shape and a.shape[0] == len(t) and a.shape[2] == len(SERIES)
    found = ctypes.c_int32()
    tc, te = ctypes.c_double(), ctypes.c_double()
    _check(_lib.slsf_crossing(_ptr(t, ctypes.c_double), len(t), a.shape[1],
                              EVALS.index(evaluable), _ptr(a, ctypes.c_double), volume_a,
                              scale_a, _ptr(b, ctypes.c_double), volume_b, scale_b,
                              ctypes.byref(found),
                              ctypes.byref(tc), ctypes.byref(te)))
    return (tc.value, te.value) if found.value else None


# ---- serving -----------------------------------------------------------------------
class Ceilings(ctypes.Structure):
    _fields_ = [("lds_bytes_per_s", ctypes.c_double), ("fma_wave_instr_per_s", ctypes.c_double),
                ("sin_wave_instr_per_s", ctypes.c_double),
                ("exp_wave_instr_per_s", ctypes.c_double), ("compute_units", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


def measure_ceilings() -> dict:
    """Measured LDS bandwidth and FMA / v_sin / v_exp issue rates of the device."""
    c = Ceilings()
    _check(_lib.slsf_measure_ceilings(ctypes.byref(c)))
    return {n: getattr(c, n) for n, _ in Ceilings._fields_ if n != "reserved"}


def check_exp2_floor() -> int:
    """Floats x >= 0 (and -0) for which the hardware 2^x returns less than 1."""
    n = ctypes.c_uint64()
    _check(_lib.slsf_check_exp2_floor(ctypes.byref(n)))
    return n.value


def device_memory() -> tuple[int, int]:
    """(free, total) bytes of the current device."""
    f, t = ctypes.c_int64(), ctypes.c_int64()
    _check(_lib.slsf_device_memory(ctypes.byref(f), ctypes.byref(t)))
    return f.value, t.value


class Server:
    """Continuous-batching job server (slsf_server_*).  Without a budget it
    admits against 90% of the device's free memory."""

    def __init__(self, device_budget_bytes: int | None = None, max_slots_per_size: int = 256,
                 chunk_sweeps: int = 100):
        h = ctypes.c_void_p()
        self._h = None
        if device_budget_bytes is None:
            device_budget_bytes = int(0.9 * device_memory()[0])
        _check(_lib.slsf_server_create(int(device_budget_bytes), max_slots_per_size,
                                       chunk_sweeps, ctypes.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            _lib.slsf_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def slot_bytes(l: int) -> int:
        return int(_lib.slsf_server_slot_bytes(l))

    def submit(self, l, temperature, seed, disorder_seed, sweeps, thermalization=0,
               binsize=10, j_xy_min=0.6, j_xy_max=1.8, j_perp=1.0,
               proposal_width=float(np.pi), wolff_steps=1, correlation_interval=1,
               correlation_rmax_xy=0, correlation_rmax_z=0) -> int:
        """Queues a job; with a correlation radius above 0 its result carries
        the binned correlation series under "correlations"."""
        q = JobRequest(l, wolff_steps, temperature, proposal_width, j_xy_min, j_xy_max, j_perp,
                       disorder_seed & 0xFFFFFFFFFFFFFFFF, seed & 0xFFFFFFFFFFFFFFFF,
                       thermalization, sweeps, binsize, correlation_interval,
                       correlation_rmax_xy, correlation_rmax_z)
        jid = ctypes.c_int64()
        _check(_lib.slsf_server_submit(self._h, ctypes.byref(q), ctypes.byref(jid)))
        return jid.value

    def step(self) -> int:
        n = ctypes.c_int32()
        _check(_lib.slsf_server_step(self._h, ctypes.byref(n)))
        return n.value

    def poll(self) -> list[dict]:
        out = []
        buf = (JobResult * 256)()
        while True:
            n = ctypes.c_int32()
            _check(_lib.slsf_server_poll(self._h, buf, 256, ctypes.byref(n)))
            for r in buf[:n.value]:
                d = {k: getattr(r, k) for k, _ in JobResult._fields_}
                d["value"] = dict(zip(EVALS, list(r.value)))
                d["error"] = dict(zip(EVALS, list(r.error)))
                d["series_mean"] = dict(zip(SERIES, list(r.series_mean)))
                if r.corr_nseries > 0:
                    d["correlations"] = self._take_correlations(r)
                out.append(d)
            if n.value < 256:
                return out

    def _take_correlations(self, r) -> dict:
        """{series name: bin means} of a finished job (slsf_server_correlations)."""
        count = ctypes.c_int64()
        _check(_lib.slsf_server_correlations(self._h, r.job_id, None, 0, ctypes.byref(count)))
        v = np.zeros(max(count.value, 1))
        _check(_lib.slsf_server_correlations(self._h, r.job_id, _ptr(v, ctypes.c_double),
                                             count.value, ctypes.byref(count)))
        v = v[:count.value].reshape(r.corr_nseries, r.corr_nbins)
        names = correlation_series_names(r.corr_rxy, r.corr_rz, r.l)
        return {name: v[k].copy() for k, name in enumerate(names)}

    def stats(self) -> dict:
        st = ServerStats()
        _check(_lib.slsf_server_get_stats(self._h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def run_until_idle(self, max_steps: int = 1_000_000) -> list[dict]:
        res = []
        for _ in range(max_steps):
            st = self.stats()
            if st["queued"] == 0 and st["active"] == 0:
                break
            self.step()
            res += self.poll()
        return res + self.poll()


# ---- job configuration -----------------------------------------------------------
def parse_duration(text: str) -> int:
    out = ctypes.c_int64()
    _check(_lib.slsf_parse_duration(text.encode(), ctypes.byref(out)))
    return out.value


class Job:
    """A job TOML expanded into tasks (src/model.rs:549-810)."""

    def __init__(self, toml_text: str):
        h = ctypes.c_void_p()
        self._h = None
        _check(_lib.slsf_job_from_toml(toml_text.encode("utf-8"), ctypes.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.slsf_job_destroy(self._h)
            self._h = None

    @property
    def name(self) -> str:
        return _lib.slsf_job_name(self._h).decode()

    def __len__(self):
        return int(_lib.slsf_job_num_tasks(self._h))

    def task(self, i: int) -> dict:
        t = Task()
        _check(_lib.slsf_job_task(self._h, i, ctypes.byref(t)))
        d = {n: getattr(t, n) for n, _ in Task._fields_}
        d["name"] = t.name.decode()
        jxy = np.zeros(t.l_z)
        jz = np.zeros(t.l_z)
        _check(_lib.slsf_job_task_couplings(self._h, i, _ptr(jxy, ctypes.c_double),
                                            _ptr(jz, ctypes.c_double)))
        d["j_xy_array"] = jxy
        d["j_z_array"] = jz
        return d

    def tasks(self) -> list[dict]:
        return [self.task(i) for i in range(len(self))]

    def durations(self):
        a, b = ctypes.c_int64(), ctypes.c_int64()
        _check(_lib.slsf_job_durations(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def selected_tasks(self, rank: int, world: int) -> list[int]:
        n = _lib.slsf_job_selected_tasks(self._h, rank, world, None, 0)
        if n < 0:
            raise SlsfError(1, "invalid rank/world")
        out = np.zeros(max(n, 1), dtype=np.int64)
        _lib.slsf_job_selected_tasks(self._h, rank, world, _ptr(out, ctypes.c_int64), n)
        return out[:n].tolist()

    def replica(self, i: int, disorder_override=None) -> Replica:
        t = self.task(i)
        return Replica(t["l_x"], t["l_y"], t["l_z"], t["temperature"], t["j_xy_array"],
                       t["j_z_array"], t["seed"], t["proposal_width"], t["wolff_steps"])
