"""GPU tests of the FP32 batched (register/LDS-resident) kernel.

* Replay stream: the FP32 kernel consumes the FP64 replay path's per-site
  uniforms; from the same state, one sweep differs only by FP32 rounding and
  rare decisions taken the other way (reported per 10^6 updates).
* Alone vs in a batch: a replica's chain is bit-identical whatever it is
  batched with (counter-based RNG keyed by replica, per-workgroup arithmetic).
* Statistics: energy of decoupled disordered z-rings against its exact Bessel
  value; infinite temperature moments; FP32 production vs FP64 replay runs on
  the same couplings within 4 combined sigma.
"""
import math

import numpy as np
import pytest

import oracle_lib as orc
from test_oracle import _ring_energy

pytestmark = pytest.mark.gpu
slsf = pytest.importorskip("slsf_amd")


def reps_for(L, n, T=1.3, seed0=100, jxy=None, jz=None, wolff=1, width=math.pi):
    jxy = orc.uniform_layers(L, 0.6, 1.8, 7) if jxy is None else jxy
    jz = np.ones(L) if jz is None else jz
    return [slsf.Replica(L, L, L, T if np.isscalar(T) else T[i], jxy, jz, seed0 + i, width,
                         wolff) for i in range(n)]


def wrapped_diff(a, b):
    d = np.abs(a - b) % (2 * math.pi)
    return np.minimum(d, 2 * math.pi - d)


# FP32 against FP64 on the same uniforms.  A decision can go the other way only
# when a uniform lies within the FP32 error of its threshold (about 1e-7 of its
# range), so at most one is allowed over a test's 0.004-0.9 M updates; an
# agreeing site differs by the FP32 rounding of its angle in turns (measured
# below 1e-6 rad).
MAX_DECISION_MISMATCHES = 1
MAX_ABS_DTHETA = 2e-6


@pytest.mark.parametrize("L", [8, 12, 16, 24, 32, 48])
def test_replay_stream_metropolis_decisions(L):
    """Metropolis only (wolff_steps=0): decision mismatches between the FP32
    kernels and the FP64 replay kernel on the same uniforms (L <= 32: the
    resident kernel's replay stream; L = 48: the FP32 HBM kernels, which always
    draw the per-site stream)."""
    n = 8
    reps = reps_for(L, n, wolff=0)
    b64 = slsf.Batch(reps, slsf.FP64_REPLAY)
    b32 = slsf.Batch(reps, slsf.FP32)
    b64.init_angles()
    b32.init_angles()
    b64.sweep(1)
    b32.run(1, 0, 1, replay_stream=True)
    bad = 0
    maxd = 0.0
    for r in range(n):
        d = wrapped_diff(b64.get_theta(r), b32.get_theta(r))
        mism = d > 1e-3
        bad += int(mism.sum())
        if (~mism).any():
            maxd = max(maxd, float(d[~mism].max()))
    rate = bad / (n * L ** 3) * 1e6
    print(f"L={L}: {bad} decision mismatches in {n * L ** 3} updates ({rate:.2f} per 1e6), "
          f"max |dtheta| = {maxd:.2e}")
    assert bad <= MAX_DECISION_MISMATCHES
    assert maxd < MAX_ABS_DTHETA


@pytest.mark.parametrize("L", [8, 12, 16, 24, 32, 48])
def test_replay_stream_full_sweep(L):
    """Metropolis + Wolff on the same uniforms: every replica reproduces the
    FP64 sweep to FP32 rounding (a cluster flips differently only if a bond's
    uniform falls within FP32 error of its probability); at most one replica
    may differ by a decision (one site, or one cluster)."""
    n = 16
    reps = reps_for(L, n, T=1.1)
    b64 = slsf.Batch(reps, slsf.FP64_REPLAY)
    b32 = slsf.Batch(reps, slsf.FP32)
    b64.init_angles()
    b32.init_angles()
    b64.sweep(1)
    b32.run(1, 0, 1, replay_stream=True)
    differ, maxd = 0, 0.0
    for r in range(n):
        d = wrapped_diff(b64.get_theta(r), b32.get_theta(r))
        if d.max() < MAX_ABS_DTHETA:
            maxd = max(maxd, float(d.max()))
        else:
            differ += 1
    print(f"L={L}: {differ} of {n} replicas differ by a decision, max |dtheta| of the "
          f"others {maxd:.2e}")
    assert differ <= 1, differ
    m64, w64 = b64.counters()
    m32, w32 = b32.counters()
    assert list(m64) == list(m32) and list(w64) == list(w32)


def test_alone_vs_batched_bit_identical():
    L = 16
    reps = reps_for(L, 5, T=[0.9, 1.2, 1.5, 2.0, 2.5])
    big = slsf.Batch(reps, slsf.FP32)
    one = slsf.Batch([reps[2]], slsf.FP32)
    for b in (big, one):
        b.init_angles()
        b.run(15, 20, 5)
    assert np.array_equal(big.get_theta(2), one.get_theta(0))
    assert np.array_equal(big.all_bins(2), one.all_bins(0))


def test_uneven_sizes_in_one_batch():
    sizes = [8, 12, 16, 24, 32]
    reps = []
    for i, L in enumerate(sizes):
        reps += reps_for(L, 2, T=1.2 + 0.05 * i, seed0=1000 * L)
    big = slsf.Batch(reps, slsf.FP32)
    big.init_angles()
    st = big.run(10, 10, 5)
    assert st["site_updates"] == sum(r.volume for r in reps) * 20
    for i in (0, 3, 5, 9):
        one = slsf.Batch([reps[i]], slsf.FP32)
        one.init_angles()
        one.run(10, 10, 5)
        assert np.array_equal(big.get_theta(i), one.get_theta(0)), i
        e = big.bins(i, "Energy")
        assert len(e) == 2 and np.all(e < 0) and np.all(np.isfinite(e))


def test_bins_match_observables_of_final_state():
    L = 16
    b = slsf.Batch(reps_for(L, 3), slsf.FP32)
    b.init_angles()
    b.run(5, 0, 1)
    b.run(0, 1, 1)
    obs = b.observables()
    for r in range(3):
        bins = b.all_bins(r)[:, -1]
        o = obs[r]
        assert abs(bins[0] - o["energy"]) < 2e-5 * abs(o["energy"]) + 1e-6
        assert abs(bins[2] - o["magnetization_squared"]) < 2e-5
        for k, name in zip(range(4, 10), ["cos_x", "cos_y", "cos_z", "sin_x", "sin_y", "sin_z"]):
            assert abs(bins[k] - o[name]) < 2e-5 * L ** 3 * 0.05 + 1e-2, name


def test_decoupled_rings_exact_energy():
    """J_xy = 0: independent periodic z-rings with disordered J_z; the exact
    <E> from Bessel functions (Z = sum_n prod_z I_n(beta J_z))."""
    L = 16
    jz = orc.uniform_layers(L, 0.4, 2.0, 11)
    T = 0.9
    n = 32
    b = slsf.Batch(reps_for(L, n, T=T, jxy=np.zeros(L), jz=jz), slsf.FP32)
    b.init_angles()
    b.run(300, 2000, 100)
    e = np.array([b.bins(r, "Energy").mean() for r in range(n)])
    exact = _ring_energy(jz, 1.0 / T)
    err = e.std(ddof=1) / math.sqrt(n)
    print(f"rings: {e.mean():.6f} +- {err:.6f}, exact {exact:.6f}")
    assert abs(e.mean() - exact) < 4 * err + 2e-5


def test_infinite_temperature():
    L = 8
    n = 16
    b = slsf.Batch(reps_for(L, n, T=1e7), slsf.FP32)
    b.init_angles()
    b.run(10, 2000, 100)
    e = np.array([b.bins(r, "Energy").mean() for r in range(n)])
    m2 = np.array([b.bins(r, "MagnetizationSquared").mean() for r in range(n)])
    assert abs(e.mean()) < 4 * e.std(ddof=1) / math.sqrt(n) + 1e-4
    assert abs(m2.mean() * L ** 3 - 1.0) < 0.05


def test_fp32_statistics_match_fp64_replay():
    """Layered disorder at L=8, T=1.3: FP32 production runs and FP64 replay
    runs (independent streams) agree within 4 combined standard errors for the
    energy, |m|^2 and the helicity sums."""
    L, n, therm, sweeps = 8, 48, 100, 400
    reps = reps_for(L, n, T=1.3, seed0=5000)
    b32 = slsf.Batch(reps, slsf.FP32)
    b64 = slsf.Batch(reps, slsf.FP64_REPLAY)
    for b in (b32, b64):
        b.init_angles()
        b.run(therm, sweeps, 100)
    for name in ("Energy", "MagnetizationSquared", "_helicity_cos_x", "_helicity_cos_z"):
        a = np.array([b32.bins(r, name).mean() for r in range(n)])
        c = np.array([b64.bins(r, name).mean() for r in range(n)])
        se = math.sqrt(a.var(ddof=1) / n + c.var(ddof=1) / n)
        print(f"{name}: fp32 {a.mean():.6f} fp64 {c.mean():.6f} se {se:.2e}")
        assert abs(a.mean() - c.mean()) < 4 * se, name


def test_fp32_evaluables_match_fp64_replay_at_l32():
    """The headline lattice, L=32 at T=1.3 with the batched workload's
    couplings (realizations 0..63): FP32 production runs (resident kernel)
    and FP64 replay runs (HBM kernels, the oracle's arithmetic), independent
    streams, 1000 + 5000 sweeps each.  Per replica the jackknife estimates of
    the energy, helicity moduli, Binder ratio and susceptibility agree within
    2 combined sigma for at least 85% of the replicas, the mean squared z-score
    is near 1, and the disorder averages agree within 4 combined sigma."""
    L, T, n = 32, 1.3, 64
    reps = [slsf.Replica(L, L, L, T, _splitmix_layers(L, 1000 + r), np.ones(L), 77 + r)
            for r in range(n)]
    out = {}
    for prec in (slsf.FP32, slsf.FP64_REPLAY):
        b = slsf.Batch(reps, prec)
        b.init_angles()
        b.run(1000, 0, 1)
        b.run(0, 5000, 100)
        out[prec] = [b.all_bins(r) for r in range(n)]
    V, beta = L ** 3, 1.0 / T
    for name in ("Energy", "RhoXY", "RhoZ", "Binder", "Chi"):
        z, d, var = [], [], 0.0
        for r in range(n):
            g = slsf.evaluate(out[slsf.FP32][r], V, beta)[name]
            c = slsf.evaluate(out[slsf.FP64_REPLAY][r], V, beta)[name]
            e = math.hypot(g[1], c[1])
            z.append((g[0] - c[0]) / e)
            d.append(g[0] - c[0])
            var += e * e
        z = np.array(z)
        avg, err = float(np.mean(d)), math.sqrt(var) / n
        print(f"{name}: within 2 sigma {np.mean(np.abs(z) <= 2):.3f}, mean z^2 "
              f"{np.mean(z * z):.2f}, average difference {avg:.3e} +- {err:.3e}")
        assert np.mean(np.abs(z) <= 2.0) >= 0.85, name
        assert 0.4 < np.mean(z * z) < 2.0, name
        assert abs(avg) < 4 * err, name


def _splitmix_layers(lz, seed, lo=0.6, hi=1.8):
    """U[lo, hi] per layer from SplitMix64(seed), the batched workload's rule."""
    out, state = [], seed & (2 ** 64 - 1)
    for _ in range(lz):
        state = (state + 0x9E3779B97F4A7C15) & (2 ** 64 - 1)
        x = state
        x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & (2 ** 64 - 1)
        x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & (2 ** 64 - 1)
        x ^= x >> 31
        out.append(lo + (hi - lo) * ((x >> 11) * 2.0 ** -53))
    return np.array(out)


def test_launch_split_invariance():
    """The same run split into launches of 7 or 64 sweeps is bit-identical:
    the FP32 state (angles in turns) is stored exactly between launches."""
    L = 16
    reps = reps_for(L, 3)
    out = []
    for spl in (7, 64):
        b = slsf.Batch(reps, slsf.FP32)
        b.init_angles()
        b.run(20, 30, 5, sweeps_per_launch=spl)
        out.append((np.stack([b.get_theta(r) for r in range(3)]),
                    np.stack([b.all_bins(r) for r in range(3)])))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_checkpoint_resume_equivalence(precision):
    """N sweeps == k sweeps, checkpoint (theta + counters), restore into a new
    handle, N - k sweeps (the analogue of ThetaModel's serde checkpoint)."""
    prec = slsf.FP32 if precision == "fp32" else slsf.FP64_REPLAY
    L = 8
    rep = reps_for(L, 1, T=1.2)
    a = slsf.Batch(rep, prec)
    a.init_angles()
    a.run(30, 0, 1)
    b = slsf.Batch(rep, prec)
    b.init_angles()
    b.run(12, 0, 1)
    theta = b.get_theta(0)
    m, w = b.counters()
    c = slsf.Batch(rep, prec)
    c.set_theta(0, theta)
    c.set_counters(int(m[0]), int(w[0]))
    c.run(18, 0, 1)
    assert np.array_equal(a.get_theta(0), c.get_theta(0))
    assert list(a.counters()[0]) == list(c.counters()[0]) == [30]


def _uneven_reps():
    reps = []
    for i, L in enumerate([8, 12, 16, 24, 32]):
        reps += reps_for(L, 3, T=1.1 + 0.1 * i, seed0=77 * L)
    return reps


def test_autotune_picks_and_leaves_chains_unchanged():
    """Tuning times candidates on the batch's own replicas and restores them:
    angles after tune + run equal those after run alone (waves per replica and
    sweeps per launch do not change any site's arithmetic)."""
    reps = _uneven_reps()
    a = slsf.Batch(reps, slsf.FP32)
    b = slsf.Batch(reps, slsf.FP32)
    a.init_angles()
    b.init_angles()
    recs = a.autotune(trial_sweeps=8)
    assert {r["l"] for r in recs} == {8, 12, 16, 24, 32}
    for L in (8, 12, 16, 24, 32):
        chosen = [r for r in recs if r["l"] == L and r["chosen"]]
        assert len(chosen) == 1 and all(r["site_updates_per_s"] > 0 for r in recs)
    assert all(r["sweeps_per_launch"] == 1000 for r in recs)  # the default launch length
    for L in (8, 12, 16, 24, 32):
        best = max(r["site_updates_per_s"] for r in recs if r["l"] == L)
        assert [r["site_updates_per_s"] for r in recs if r["l"] == L and r["chosen"]] == [best]
    assert list(a.counters()[0]) == [0] * len(reps)
    a.run(10, 10, 5)
    b.run(10, 10, 5)
    for i in range(len(reps)):
        assert np.array_equal(a.get_theta(i), b.get_theta(i)), i
        assert np.allclose(a.all_bins(i), b.all_bins(i), rtol=2e-5, atol=1e-6)


def test_autotuned_batch_is_not_slower_than_the_default():
    """The uneven batch at 16 replicas per size: a timed run after autotuning
    against the same run with the default plan (same replicas, same sweeps);
    the tuned run may not be slower beyond timing noise (3%)."""
    reps = []
    for i, L in enumerate([8, 12, 16, 24, 32]):
        reps += reps_for(L, 16, T=1.1 + 0.1 * i, seed0=91 * L)
    rates = {}
    for tuned in (False, True, False, True):
        b = slsf.Batch(reps, slsf.FP32)
        b.init_angles()
        if tuned:
            b.autotune(32)
        b.run(20, 0, 1)
        st = b.run(0, 400, 20)
        rate = sum(r.volume for r in reps) * 400 / st["seconds"]
        rates[tuned] = max(rates.get(tuned, 0.0), rate)
        b.close()
    print(f"default {rates[False]:.3e}, autotuned {rates[True]:.3e} site-updates/s")
    assert rates[True] >= 0.97 * rates[False]


def test_graph_capture_matches_stream_launches():
    reps = _uneven_reps()
    out = []
    for graph in (False, True):
        b = slsf.Batch(reps, slsf.FP32)
        b.init_angles()
        st = b.run(12, 24, 6, sweeps_per_launch=5, use_graph=graph)
        assert st["launches"] == 5 * (3 + 5)
        out.append((np.concatenate([b.get_theta(i) for i in range(len(reps))]),
                    np.stack([b.all_bins(i) for i in range(len(reps))])))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("precision", [slsf.FP64_REPLAY, slsf.FP32])
def test_hbm_graph_replay_matches_stream_launches(precision):
    """HBM-path runs (sizes the resident kernel does not take, mixed in one
    batch) give the same angles and bins with and without graph replay, also
    when the graph's sweep chunk does not divide the run (the replicas' windows
    cut the last replay)."""
    reps = reps_for(20, 2, T=1.1, seed0=40) + reps_for(10, 3, T=1.6, seed0=50)
    reps += [slsf.Replica(6, 10, 14, 0.9, np.linspace(0.6, 1.8, 14), np.ones(14), 77,
                          wolff_steps=2)]
    out = []
    for graph in (False, True):
        b = slsf.Batch(reps, precision)
        b.init_angles()
        b.run(7, 0, 1, sweeps_per_launch=3, use_graph=graph)
        b.run(5, 16, 4, sweeps_per_launch=6, use_graph=graph)
        out.append((np.concatenate([b.get_theta(i) for i in range(len(reps))]),
                    np.stack([b.all_bins(i) for i in range(len(reps))]), b.counters()))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(out[0][2][0], out[1][2][0]) and np.array_equal(out[0][2][1], out[1][2][1])
    assert list(out[0][2][0]) == [28] * len(reps)


def test_mixed_resident_and_hbm_sizes_follow_their_own_chains():
    """FP32 batches that mix sizes the resident kernel takes with sizes it does
    not: each replica runs the path of its own size, so its angles and bins are
    bit-identical to the same replica run alone."""
    reps = (reps_for(8, 2, T=1.2, seed0=300) + reps_for(10, 2, T=1.4, seed0=310) +
            reps_for(16, 1, T=0.9, seed0=320) + reps_for(20, 1, T=1.6, seed0=330))
    reps.append(slsf.Replica(8, 8, 12, 1.1, np.linspace(0.6, 1.8, 12), np.ones(12), 340))
    mix = slsf.Batch(reps, slsf.FP32)
    mix.init_angles()
    st = mix.run(6, 12, 4)
    assert st["site_updates"] == sum(r.volume for r in reps) * 18
    m, w = mix.counters()
    assert list(m) == [18] * len(reps) and list(w) == [18] * len(reps)
    for i, r in enumerate(reps):
        one = slsf.Batch([r], slsf.FP32)
        one.init_angles()
        one.run(6, 12, 4)
        assert np.array_equal(mix.get_theta(i), one.get_theta(0)), i
        assert np.array_equal(mix.all_bins(i), one.all_bins(0)), i


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_large_lattice_invariants(precision):
    """One L=256 replica (16.8 M sites, 4096 union-find tiles) on the HBM
    kernels: after thermalization and measured sweeps the angles stay in
    [0, 2 pi) (FP32 turns) or in [0, 3 pi) (FP64 replay, which keeps the
    reference's single-shift wrap of the Wolff reflection), the counters
    advance, acceptance and cluster size are inside their ranges, and the bins
    agree with a direct measurement of the final state (binsize 1)."""
    prec = slsf.FP32 if precision == "fp32" else slsf.FP64_REPLAY
    L = 256
    b = slsf.Batch(reps_for(L, 1, T=1.4, seed0=31), prec)
    b.init_angles()
    st = b.run(2, 3, 1)
    assert st["site_updates"] == L ** 3 * 5
    th = b.get_theta(0)
    top = 2 * math.pi if precision == "fp32" else 3 * math.pi
    assert np.all(th >= 0.0) and np.all(th < top)
    assert list(b.counters()[0]) == [5] and list(b.counters()[1]) == [5]
    assert 0.0 < st["acceptance_mean"] < 1.0
    assert 1.0 <= st["mean_cluster_size"] <= L ** 3
    o = b.observables()[0]
    e = b.all_bins(0)[0]
    assert len(e) == 3 and np.all(np.isfinite(e))
    tol = 2e-5 if precision == "fp32" else 1e-10
    assert abs(e[-1] - o["energy"]) < tol * abs(o["energy"]) + 1e-6


def test_hardware_exp2_is_at_least_one_for_nonnegative_arguments():
    """The resident kernel accepts a Metropolis move iff u < exp(-beta dE),
    which also covers dE <= 0 (src/updates.rs:255-264) because u < 1 and the
    hardware 2^x returns at least 1 for every x >= 0: checked here over all
    such floats (and -0)."""
    assert slsf.check_exp2_floor() == 0
