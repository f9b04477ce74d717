"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) -- test infrastructure.

Only tests, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, and only as the checker.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_PATH = os.path.join(ROOT, "oracle", "liboracle.so")
ORACLE_FAST_PATHS = {8: os.path.join(ROOT, "oracle", "liboracle_fast_v4.so"),
                     4: os.path.join(ROOT, "oracle", "liboracle_fast_v3.so")}


def host_batch_lanes() -> int:
    """8 on a host with AVX-512 (F and DQ), else 4: the reference's run-time
    choice of its batch width (src/fast_math.rs:28-37, src/updates.rs:454)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    flags = set(line.split(":", 1)[1].split())
                    return 8 if {"avx512f", "avx512dq"} <= flags else 4
    except OSError:
        pass
    return 4

_P = ctypes.POINTER
_dp = _P(ctypes.c_double)


class FastRng(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("state", ctypes.c_uint64),
                ("draws", ctypes.c_uint64)]


class OLattice(ctypes.Structure):
    _fields_ = [("lx", ctypes.c_int), ("ly", ctypes.c_int), ("lz", ctypes.c_int),
                ("theta", _dp), ("j_xy", _dp), ("j_z", _dp), ("cos_t", _dp), ("sin_t", _dp),
                ("in_cluster", _P(ctypes.c_uint8)), ("stack", _P(ctypes.c_int)),
                ("cluster", _P(ctypes.c_int))]


class OObs(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "energy", "magnetization_squared", "cos_x", "cos_y", "cos_z", "sin_x", "sin_y", "sin_z")]


class OSimOptions(ctypes.Structure):
    _fields_ = [("thermal_sweeps", ctypes.c_int), ("measure_sweeps", ctypes.c_int),
                ("measure_interval", ctypes.c_int), ("proposal_width", ctypes.c_double),
                ("wolff_steps", ctypes.c_int), ("correlation_interval", ctypes.c_int),
                ("correlation_rmax", ctypes.c_int), ("correlation_rmax_xy", ctypes.c_int),
                ("correlation_rmax_z", ctypes.c_int), ("cold_init", ctypes.c_int),
                ("use_batch8", ctypes.c_int)]


class OSimResult(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "energy", "rho_sx", "rho_sy", "rho_sz", "std_rho_sx", "std_rho_sy", "std_rho_sz",
        "magnetization_squared", "chi", "acceptance")] + [
        ("num_measurements", ctypes.c_int), ("num_correlation_measurements", ctypes.c_int),
        ("rxy", ctypes.c_int), ("rz", ctypes.c_int),
        ("corr_x", _dp), ("corr_y", _dp), ("corr_xy", _dp), ("corr_xy_by_z", _dp),
        ("corr_z", _dp)]


def load(fast: bool = False):
    """The strict checker library, or with fast=True the timed CPU-baseline build
    for this host's batch width (host_batch_lanes)."""
    path = ORACLE_FAST_PATHS[host_batch_lanes()] if fast else ORACLE_PATH
    if not os.path.exists(path):
        raise ImportError(f"{path} missing: run `make -C oracle`")
    L = ctypes.CDLL(path)
    sig = {
        "or_rng_seed": (None, [_P(FastRng), ctypes.c_uint64]),
        "or_rng_next_u64": (ctypes.c_uint64, [_P(FastRng)]),
        "or_rng_next_u32": (ctypes.c_uint32, [_P(FastRng)]),
        "or_rng_set_position": (None, [_P(FastRng), ctypes.c_uint64]),
        "or_rng_f64": (ctypes.c_double, [_P(FastRng)]),
        "or_rng_bool": (ctypes.c_int, [_P(FastRng)]),
        "or_rng_range": (ctypes.c_uint64, [_P(FastRng), ctypes.c_uint64]),
        "or_philox4x32_10": (None, [_P(ctypes.c_uint32), _P(ctypes.c_uint32),
                                    _P(ctypes.c_uint32)]),
        "or_sincos": (None, [ctypes.c_double, _dp, _dp]),
        "or_exp": (ctypes.c_double, [ctypes.c_double]),
        "or_lattice_new": (ctypes.c_int, [_P(OLattice), ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "or_lattice_free": (None, [_P(OLattice)]),
        "or_scratch_refresh": (None, [_P(OLattice), ctypes.c_int]),
        "or_last_error": (ctypes.c_char_p, []),
        "or_two_point_disorder": (ctypes.c_int, [_dp, ctypes.c_int, ctypes.c_double,
                                                 ctypes.c_double, _P(FastRng), ctypes.c_char_p]),
        "or_uniform_disorder": (ctypes.c_int, [_dp, ctypes.c_int, ctypes.c_double,
                                               ctypes.c_double, _P(FastRng)]),
        "or_initialize_disorder": (ctypes.c_int, [_P(OLattice), ctypes.c_double,
                                                  ctypes.c_double, ctypes.c_double, _P(FastRng)]),
        "or_initialize_angles": (None, [_P(OLattice), ctypes.c_int, _P(FastRng)]),
        "or_initialize_angles_keyed": (None, [_P(OLattice), ctypes.c_uint64]),
        "or_vector_exp_min_uphill": (ctypes.c_int, [ctypes.c_int]),
        "or_metropolis_accept_probabilities": (None, [_dp, ctypes.c_int, ctypes.c_double, _dp]),
        "or_metropolis_sweep_ref": (ctypes.c_int, [_P(OLattice), ctypes.c_double,
                                                   ctypes.c_double, ctypes.c_int, _P(FastRng),
                                                   _dp]),
        "or_local_metropolis_step_ref": (ctypes.c_int, [_P(OLattice), ctypes.c_double,
                                                        ctypes.c_double, _P(FastRng),
                                                        _P(ctypes.c_int)]),
        "or_metropolis_sweep_replay": (ctypes.c_int, [_P(OLattice), ctypes.c_double,
                                                      ctypes.c_double, ctypes.c_uint64,
                                                      ctypes.c_uint32, _P(ctypes.c_uint8), _dp]),
        "or_wolff_add_probability": (ctypes.c_double, [ctypes.c_double] * 4),
        "or_wolff_reflect_angle": (ctypes.c_double, [ctypes.c_double] * 2),
        "or_wolff_step_ref": (ctypes.c_int, [_P(OLattice), ctypes.c_double, ctypes.c_int,
                                             _P(FastRng), _P(ctypes.c_int64)]),
        "or_wolff_step_replay": (ctypes.c_int, [_P(OLattice), ctypes.c_double, ctypes.c_uint64,
                                                ctypes.c_uint32, _P(ctypes.c_uint8),
                                                _P(ctypes.c_int64)]),
        "or_measure_observables": (None, [_P(OLattice), _P(OObs)]),
        "or_measure_energy_direct": (ctypes.c_double, [_P(OLattice)]),
        "or_measure_magnetization_squared_direct": (ctypes.c_double, [_P(OLattice)]),
        "or_helicity_sums_direct": (None, [_P(OLattice), _dp]),
        "or_measure_helicity_modulus": (ctypes.c_int, [_P(OLattice), ctypes.c_double, _dp]),
        "or_correlation_rmax": (None, [_P(OLattice), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       _P(ctypes.c_int), _P(ctypes.c_int)]),
        "or_measure_correlations": (None, [_P(OLattice), ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int, _dp, _dp, _dp, _dp, _dp]),
        "or_run_theta_simulation": (ctypes.c_int, [_P(OLattice), ctypes.c_double,
                                                   ctypes.c_double, ctypes.c_double,
                                                   ctypes.c_double, _P(OSimOptions),
                                                   _P(FastRng), _P(OSimResult)]),
        "or_run_theta_simulation_replay": (ctypes.c_int, [_P(OLattice), ctypes.c_double,
                                                          ctypes.c_double, ctypes.c_double,
                                                          ctypes.c_double, _dp, _dp,
                                                          _P(OSimOptions), ctypes.c_uint64,
                                                          _P(OSimResult)]),
        "or_mean": (ctypes.c_double, [_dp, ctypes.c_int]),
        "or_sample_std": (ctypes.c_double, [_dp, ctypes.c_int]),
        "or_blocking_stderr": (ctypes.c_double, [_dp, ctypes.c_int]),
        "or_binned_estimate": (ctypes.c_int, [_dp, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp,
                                              _P(ctypes.c_int)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = load()
    return _LIB


class OracleError(RuntimeError):
    pass


def err() -> str:
    return lib().or_last_error().decode("utf-8")


def _chk(rc):
    if rc:
        raise OracleError(err())


def rng(seed: int) -> FastRng:
    r = FastRng()
    lib().or_rng_seed(ctypes.byref(r), seed & 0xFFFFFFFFFFFFFFFF)
    return r


def philox(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().or_philox4x32_10(c, k, o)
    return list(o)


def sincos(x: float):
    s, c = ctypes.c_double(), ctypes.c_double()
    lib().or_sincos(x, ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


class Lattice:
    """Oracle lattice: theta, couplings and its trig / Wolff scratch."""

    def __init__(self, lx, ly, lz, lib_=None):
        self.L = lib_ or lib()
        self.lat = OLattice()
        rc = self.L.or_lattice_new(ctypes.byref(self.lat), lx, ly, lz)
        if rc:
            raise OracleError(self.L.or_last_error().decode())
        self.lx, self.ly, self.lz = lx, ly, lz
        self.V = lx * ly * lz

    def __del__(self):
        try:
            self.L.or_lattice_free(ctypes.byref(self.lat))
        except Exception:
            pass

    @property
    def ref(self):
        return ctypes.byref(self.lat)

    def _arr(self, ptr, n):
        return np.ctypeslib.as_array(ptr, shape=(n,))

    @property
    def theta(self):
        return self._arr(self.lat.theta, self.V)

    @property
    def j_xy(self):
        return self._arr(self.lat.j_xy, self.lz)

    @property
    def j_z(self):
        return self._arr(self.lat.j_z, self.lz)

    @property
    def cos_t(self):
        return self._arr(self.lat.cos_t, self.V)

    @property
    def sin_t(self):
        return self._arr(self.lat.sin_t, self.V)

    def set_couplings(self, j_xy, j_z):
        self.j_xy[:] = j_xy
        self.j_z[:] = j_z

    def refresh(self, replay: bool):
        self.L.or_scratch_refresh(self.ref, int(replay))

    def init_keyed(self, seed):
        self.L.or_initialize_angles_keyed(self.ref, seed & 0xFFFFFFFFFFFFFFFF)
        self.refresh(True)

    def metro_replay(self, T, width, seed, sweep, want_decisions=True):
        dec = np.zeros(self.V, dtype=np.uint8)
        acc = ctypes.c_double()
        _chk(self.L.or_metropolis_sweep_replay(self.ref, T, width, seed & 0xFFFFFFFFFFFFFFFF,
                                               sweep, dec.ctypes.data_as(_P(ctypes.c_uint8)),
                                               ctypes.byref(acc)))
        return acc.value, dec

    def wolff_replay(self, T, seed, step):
        mem = np.zeros(self.V, dtype=np.uint8)
        cs = ctypes.c_int64()
        _chk(self.L.or_wolff_step_replay(self.ref, T, seed & 0xFFFFFFFFFFFFFFFF, step,
                                         mem.ctypes.data_as(_P(ctypes.c_uint8)),
                                         ctypes.byref(cs)))
        return cs.value, mem

    def metro_ref(self, T, width, r: FastRng, batch8=True):
        acc = ctypes.c_double()
        _chk(self.L.or_metropolis_sweep_ref(self.ref, T, width, int(batch8), ctypes.byref(r),
                                            ctypes.byref(acc)))
        return acc.value

    def wolff_ref(self, T, r: FastRng, use_scratch=True):
        cs = ctypes.c_int64()
        _chk(self.L.or_wolff_step_ref(self.ref, T, int(use_scratch), ctypes.byref(r),
                                      ctypes.byref(cs)))
        return cs.value

    def observables(self) -> dict:
        o = OObs()
        self.L.or_measure_observables(self.ref, ctypes.byref(o))
        return {n: getattr(o, n) for n, _ in OObs._fields_}

    def correlations(self, rmax=-1, rmax_xy=-1, rmax_z=-1) -> dict:
        a, b = ctypes.c_int(), ctypes.c_int()
        self.L.or_correlation_rmax(self.ref, rmax, rmax_xy, rmax_z, ctypes.byref(a),
                                   ctypes.byref(b))
        rxy, rz = a.value, b.value
        cx, cy, cxy = (np.zeros(max(rxy, 1)) for _ in range(3))
        cbz = np.zeros(max(self.lz * rxy, 1))
        cz = np.zeros(max(rz, 1))
        self.L.or_measure_correlations(self.ref, rmax, rmax_xy, rmax_z,
                                       *(x.ctypes.data_as(_dp) for x in (cx, cy, cxy, cbz, cz)))
        return {"corr_x": cx[:rxy], "corr_y": cy[:rxy], "corr_xy": cxy[:rxy],
                "corr_xy_by_z": cbz[:self.lz * rxy].reshape(self.lz, rxy), "corr_z": cz[:rz]}


def uniform_layers(lz: int, lo: float, hi: float, seed: int) -> np.ndarray:
    out = np.zeros(lz)
    r = rng(seed)
    _chk(lib().or_uniform_disorder(out.ctypes.data_as(_dp), lz, lo, hi, ctypes.byref(r)))
    return out
