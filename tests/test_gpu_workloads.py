"""The batched workloads at their full shapes on the GPU.

* The L=32 disorder average: 256 realizations x 8 temperatures = 2048
  replicas in one FP32 batch (the resident kernel), a few thousand sweeps.
* The uneven batch: L in {8, 12, 16, 24, 32} x 5 temperatures x 64
  realizations = 1600 replicas, autotuned.
* FP32 production against FP64 CPU runs of the oracle in the reference's own
  mode (SplitMix64 stream, reference draw order, Wolff DFS) on the L=32
  workload's couplings at T = 1.30: the disorder averages of the energy,
  helicity moduli, Binder ratio and susceptibility agree within 2 sigma (one
  of the five may lie between 2 and 3 sigma).

Every replica is checked for finite bins, an energy per site between its
ground-state bound and 0, and acceptance and cluster size in range; a few
replicas per size must be bit-identical alone and in the batch, and one FP64
replay replica per size must be bit-identical with the oracle.
"""
import math
import multiprocessing as mp

import numpy as np
import pytest

import bench
import oracle_lib as orc

pytestmark = pytest.mark.gpu
slsf = pytest.importorskip("slsf_amd")


def _check_replicas(b, reps, keys=None):
    """finite bins; -(2 max J_xy + max J_z) <= E/V <= 0 for every bin"""
    for i, r in enumerate(reps):
        bins = b.all_bins(i)
        assert bins.shape[1] >= 1 and np.all(np.isfinite(bins)), (i, keys and keys[i])
        e = bins[0]
        lo = -(2 * np.max(r.j_xy) + np.max(r.j_z))
        assert np.all(e <= 0.0) and np.all(e >= lo), (i, e.min(), e.max(), lo)
        m2 = bins[2]
        assert np.all(m2 >= 0.0) and np.all(m2 <= 1.0 + 1e-6), i


def _alone_equals_batched(batch, reps, idx, therm, sweeps, binsize):
    for i in idx:
        one = slsf.Batch([reps[i]], slsf.FP32)
        one.init_angles()
        one.run(therm, sweeps, binsize)
        assert np.array_equal(batch.get_theta(i), one.get_theta(0)), i
        assert np.array_equal(batch.all_bins(i), one.all_bins(0)), i
        one.close()


def _replay_equals_oracle(rep, sweeps=2):
    b = slsf.Batch([rep], slsf.FP64_REPLAY)
    b.init_angles()
    b.sweep(sweeps)
    lat = orc.Lattice(rep.l_x, rep.l_y, rep.l_z)
    lat.set_couplings(rep.j_xy, rep.j_z)
    lat.init_keyed(rep.seed)
    for s in range(sweeps):
        lat.metro_replay(rep.temperature, rep.proposal_width, rep.seed, s)
        lat.wolff_replay(rep.temperature, rep.seed, s)
    assert np.array_equal(b.get_theta(0), lat.theta), (rep.l_x, rep.seed)
    b.close()


def test_l32_disorder_average_full_batch():
    reps = bench.c3_replicas(slsf)
    assert len(reps) == 2048
    b = slsf.Batch(reps, slsf.FP32)
    b.init_angles()
    st = b.run(1000, 1000, 50)
    assert st["site_updates"] == 2048 * 32 ** 3 * 2000
    assert 0.1 < st["acceptance_mean"] < 0.6
    assert 1.0 <= st["mean_cluster_size"] <= 32 ** 3
    acc = b.acceptance()
    assert np.all((acc > 0.05) & (acc < 0.7))
    _check_replicas(b, reps)
    # colder replicas are more ordered: the disorder-averaged m^2 falls with T
    m2 = np.array([b.bins(i, "MagnetizationSquared").mean() for i in range(2048)])
    by_t = m2.reshape(256, 8).mean(axis=0)
    assert np.all(np.diff(by_t) < 0), by_t
    _alone_equals_batched(b, reps, (0, 777, 2047), 1000, 1000, 50)
    b.close()
    _replay_equals_oracle(reps[4])


def test_uneven_batch_full_shape_autotuned():
    Ls = [8, 12, 16, 24, 32]
    temps = [1.1, 1.2, 1.3, 1.4, 1.5]
    reps, keys = bench.disorder_batch(slsf, Ls, temps, 64, 1)
    assert len(reps) == 1600
    b = slsf.Batch(reps, slsf.FP32)
    b.init_angles()
    recs = b.autotune(16)
    assert {r["l"] for r in recs} == set(Ls)
    st = b.run(300, 400, 20)
    assert st["site_updates"] == sum(r.volume for r in reps) * 700
    assert 0.1 < st["acceptance_mean"] < 0.6
    _check_replicas(b, reps, keys)
    # one replica of every size alone, and one FP64 replay replica per size
    # against the oracle
    first = {L: keys.index((L, 0, 1.3)) for L in Ls}
    _alone_equals_batched(b, reps, [first[L] for L in Ls], 300, 400, 20)
    b.close()
    for L in Ls:
        _replay_equals_oracle(reps[first[L]])


def test_fp32_matches_fp64_cpu_reference_mode_at_l32():
    """Realizations 0..15 of the L=32 workload at T = 1.30: the GPU's FP32
    production chains against the oracle's FP64 reference-mode chains on the
    host (independent streams), 500 + 2500 sweeps each, bins of 100.  Tolerance
    (the FP32 production accuracy bar): the disorder-averaged observables within
    2 combined sigma.
    Five observables held to 2 sigma each would fail about one seed set in four
    by chance alone (1 - 0.95^5), so one of the five may lie between 2 and 3
    sigma (a chance of about 3% to fail when they agree)."""
    n, therm, measured, binsize = 16, 500, 2500, 100
    reps = bench.c3_replicas(slsf)
    sub = [reps[8 * r + 4] for r in range(n)]
    assert all(abs(r.temperature - 1.3) < 1e-12 for r in sub)
    g = slsf.Batch(sub, slsf.FP32)
    g.init_angles()
    g.run(therm, measured, binsize)
    gpu_bins = {r: g.all_bins(r) for r in range(n)}
    g.close()
    with mp.get_context("fork").Pool(min(n, bench.cpu_cores())) as pool:
        out = pool.map(bench._cpu_worker, [(i, therm + measured, therm, binsize)
                                           for i in range(n)])
    cpu_bins = {i: o[1] for i, o in enumerate(out)}
    res = bench.fp64_agreement(slsf, gpu_bins, cpu_bins, 32.0 ** 3, 1 / 1.3)
    assert res["replicas"] == n
    beyond_2 = []
    for name, v in res["observables"].items():
        z = v["mean_difference"] / v["mean_difference_error"]
        print(f"{name}: disorder average FP32 - FP64 = {v['mean_difference']:.3e} "
              f"+- {v['mean_difference_error']:.3e} ({z:+.2f} sigma); replicas within 2 "
              f"sigma {v['within_2_sigma']:.3f}, mean z^2 {v['mean_z_squared']:.2f}")
        assert abs(z) <= 3.0, name
        if abs(z) > 2.0:
            beyond_2.append(name)
        assert v["within_2_sigma"] >= 0.75, name
        assert 0.2 < v["mean_z_squared"] < 2.5, name
    assert len(beyond_2) <= 1, beyond_2
