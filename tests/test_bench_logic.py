"""Host-side logic of bench.py's headline leg (no GPU): bin lengths of short
timed runs, the FP32-vs-FP64 comparison with too few bins, and the CPU
baseline's worker on a tiny sample."""
import numpy as np

import bench
import slsf_amd as slsf


def test_headline_binsize_leaves_bins_for_short_runs():
    assert bench.headline_binsize(20, 100) == 1      # the driver's --steps 20
    assert bench.headline_binsize(10000, 100) == 100
    assert bench.headline_binsize(1000, 100) == 50
    assert bench.headline_binsize(0, 100) == 1
    for steps in (1, 5, 19, 20, 21, 399, 2000, 10000):
        b = bench.headline_binsize(steps, 100)
        assert b >= 1 and (steps < 20 or steps // b >= 20)


def _bins(nbins, seed):
    rng = np.random.default_rng(seed)
    e = -2.0 + 0.01 * rng.standard_normal(nbins)
    m2 = 0.5 + 0.01 * rng.standard_normal(nbins)
    c = 300.0 + rng.standard_normal((3, nbins))
    s = rng.standard_normal((3, nbins))
    return np.vstack([e, e * e, m2, m2 * m2, c, s, s * s])


def test_fp64_agreement_without_enough_bins_is_none():
    gpu = {0: _bins(1, 1), 1: _bins(0, 2)}
    cpu = {0: _bins(5, 3), 1: _bins(5, 4)}
    assert bench.fp64_agreement(slsf, gpu, cpu, 32768.0, 1 / 1.3) is None
    assert bench.fp64_agreement(slsf, {0: _bins(5, 1)}, {0: None}, 32768.0, 1 / 1.3) is None
    assert bench.fp64_agreement(slsf, {}, {}, 32768.0, 1 / 1.3) is None


def test_fp64_agreement_uses_replicas_with_bins():
    gpu = {0: _bins(20, 1), 1: _bins(1, 2), 2: _bins(20, 5)}
    cpu = {0: _bins(20, 3), 1: _bins(20, 4), 2: _bins(20, 6)}
    out = bench.fp64_agreement(slsf, gpu, cpu, 32768.0, 1 / 1.3)
    assert out["replicas"] == 2
    assert set(out["observables"]) == {"Energy", "RhoXY", "RhoZ", "Binder", "Chi"}
    for v in out["observables"].values():
        assert np.isfinite(v["mean_z_squared"]) and 0.0 <= v["within_2_sigma"] <= 1.0


def test_cpu_worker_bins_and_zero_measurements():
    rate, bins = bench._cpu_worker((0, 4, 2, 1))
    assert rate > 0 and bins.shape == (13, 2)
    assert np.all(bins[0] < 0) and np.allclose(bins[1], bins[0] ** 2)
    rate, bins = bench._cpu_worker((1, 2, 2, 1))
    assert bins is None


def test_cpu_cores_respects_thread_cap(monkeypatch):
    monkeypatch.setenv("OMP_NUM_THREADS", "1")
    assert bench.cpu_cores() == 1
    monkeypatch.delenv("OMP_NUM_THREADS")
    assert bench.cpu_cores() >= 1
