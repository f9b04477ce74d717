"""Pins of the CPU oracle (CPU only).

What pins what:
  * SplitMix64 FastRng (src/types.rs:9-82): an independent Python-integer
    restatement (tests/golden/splitmix64.json) and the published first outputs
    for seed 1234567; set_position skip-ahead.
  * Philox4x32-10: the published known-answer vectors of Salmon et al. (Random123).
  * sincos / exp: libm within 1e-14 on the reference's points (src/tests.rs:11-23)
    and on a dense grid.
  * The reference's own tests restated (src/tests.rs): acceptance probabilities,
    disorder statistics and error strings, cold-lattice observables, width-0
    updates, helicity of the cold state, Wolff probability/reflection/cluster
    invariants, the plain driver's counts and messages, BinnedEstimate.
  * Closed forms: beta -> 0 moments, decoupled z-rings (Bessel functions) for
    Metropolis + Wolff with disorder, ordered state Binder = 1, a brute-force
    local-energy identity, and the Wolff replay cluster against an independent
    numpy/scipy connected-component computation.
"""
import ctypes
import json
import math
import os

import numpy as np
import pytest
from scipy import special
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components

import oracle_lib as orc

HERE = os.path.dirname(os.path.abspath(__file__))
L = orc.lib()


# ---------------------------------------------------------------------------
# RNGs
# ---------------------------------------------------------------------------
def test_splitmix64_golden():
    gold = json.load(open(os.path.join(HERE, "golden", "splitmix64.json")))
    for seed, vals in gold.items():
        r = orc.rng(int(seed))
        got = [L.or_rng_next_u64(ctypes.byref(r)) for _ in vals]
        assert got == [int(v) for v in vals], seed
    # published test values of SplitMix64 for seed 1234567
    assert gold["1234567"][:3] == ["6457827717110365317", "3203168211198807973",
                                   "9817491932198370423"]


def test_splitmix64_set_position_and_adapters():
    r = orc.rng(42)
    words = [L.or_rng_next_u64(ctypes.byref(r)) for _ in range(10)]
    r2 = orc.rng(42)
    L.or_rng_set_position(ctypes.byref(r2), 6)
    assert L.or_rng_next_u64(ctypes.byref(r2)) == words[6]
    assert r.draws == 10
    # next_u32 is the low half of next_u64 (src/types.rs:60-62)
    r3 = orc.rng(42)
    assert L.or_rng_next_u32(ctypes.byref(r3)) == words[0] & 0xFFFFFFFF
    # f64 in [0,1), range in [0,n): distribution sanity (adapters unpinned)
    r4 = orc.rng(9)
    u = np.array([L.or_rng_f64(ctypes.byref(r4)) for _ in range(20000)])
    assert u.min() >= 0.0 and u.max() < 1.0 and abs(u.mean() - 0.5) < 0.01
    k = np.array([L.or_rng_range(ctypes.byref(r4), 7) for _ in range(7000)])
    assert k.min() == 0 and k.max() == 6
    counts = np.bincount(k, minlength=7)
    chi2 = ((counts - 1000.0) ** 2 / 1000.0).sum()
    assert chi2 < 30.0


PHILOX_KAT = [  # Salmon et al. SC'11 / Random123 kat_vectors, philox4x32 10 rounds
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]


@pytest.mark.parametrize("ctr,key,expect", PHILOX_KAT)
def test_philox_kat(ctr, key, expect):
    assert orc.philox(ctr, key) == expect


# ---------------------------------------------------------------------------
# math (src/tests.rs:11-23)
# ---------------------------------------------------------------------------
def test_sincos_exp_against_libm():
    pts = [-2.0, -1.25, -0.5, -0.125, 0.0, 0.125, 0.5, 1.75]
    for x in pts:
        s, c = orc.sincos(x)
        assert abs(s - math.sin(x)) <= 1e-14 and abs(c - math.cos(x)) <= 1e-14
        assert abs(L.or_exp(x) - math.exp(x)) <= 1e-14
    xs = np.linspace(-4 * math.pi, 4 * math.pi, 20001)
    err = max(max(abs(orc.sincos(x)[0] - math.sin(x)), abs(orc.sincos(x)[1] - math.cos(x)))
              for x in xs)
    assert err <= 2.3e-16
    for x in np.linspace(-700.0, 0.0, 7001):
        assert abs(L.or_exp(x) - math.exp(x)) <= 3e-16 * math.exp(x)
    assert L.or_exp(-709.5) == 0.0


def test_accept_probabilities():  # src/tests.rs:25-43
    delta = np.array([-2.0, 0.0, 0.125, 0.5, 1.0, 1.5, 2.0, 3.0])
    prob = np.zeros(8)
    L.or_metropolis_accept_probabilities(delta.ctypes.data_as(orc._dp), 8, 0.75,
                                         prob.ctypes.data_as(orc._dp))
    for d, p in zip(delta, prob):
        assert abs(p - (1.0 if d <= 0 else math.exp(-d * 0.75))) <= 1e-14
    assert L.or_vector_exp_min_uphill(4) == 2 and L.or_vector_exp_min_uphill(8) == 3


# ---------------------------------------------------------------------------
# lattice and disorder (src/tests.rs:45-91)
# ---------------------------------------------------------------------------
def test_lattice_and_disorder():
    for dims, msg in [((0, 2, 2), "L_x must be positive"), ((2, 0, 2), "L_y must be positive"),
                      ((2, 2, 0), "L_z must be positive")]:
        with pytest.raises(orc.OracleError, match=msg):
            orc.Lattice(*dims)
    lat = orc.Lattice(3, 4, 5)
    assert lat.theta.size == 60 and lat.j_xy.size == 5
    r = orc.rng(12345)
    assert L.or_initialize_disorder(lat.ref, 1.0, 0.7, 0.0, ctypes.byref(r)) == 0
    assert np.all(lat.j_z == 0.7) and np.all(lat.j_xy == 1.0)
    big = orc.Lattice(1, 1, 200000)
    assert L.or_initialize_disorder(big.ref, 1.0, 1.0, 0.5, ctypes.byref(r)) == 0
    jz = big.j_z
    assert set(np.unique(jz)) <= {0.5, 1.5}
    assert abs(jz.mean() - 1.0) < 5e-3 and abs(jz.std(ddof=1) - 0.5) < 5e-3
    assert L.or_initialize_disorder(lat.ref, 1.0, 1.0, -0.1, ctypes.byref(r)) != 0
    assert orc.err() == "δJ_z must be nonnegative"
    assert L.or_initialize_disorder(lat.ref, 1.0, 0.05, 0.1, ctypes.byref(r)) != 0
    assert orc.err() == "two-point layer disorder requires J_z_mean - δJ_z >= 0"
    L.or_initialize_angles(lat.ref, 1, ctypes.byref(r))
    assert np.all(lat.theta == 0.0)
    L.or_initialize_angles(lat.ref, 0, ctypes.byref(r))
    assert np.all((lat.theta >= 0) & (lat.theta < 2 * math.pi))


def test_uniform_disorder_extension():
    v = orc.uniform_layers(100000, 0.6, 1.8, 5)
    assert v.min() >= 0.6 and v.max() < 1.8
    assert abs(v.mean() - 1.2) < 5e-3 and abs(v.std() - 1.2 / math.sqrt(12)) < 5e-3


# ---------------------------------------------------------------------------
# observables (src/tests.rs:93-153, 195-206)
# ---------------------------------------------------------------------------
def _cold(lx, ly, lz, jxy, jz):
    lat = orc.Lattice(lx, ly, lz)
    lat.set_couplings(np.full(lz, jxy), np.full(lz, jz))
    lat.theta[:] = 0.0
    lat.refresh(False)
    return lat


def test_cold_observables_and_correlations():
    lat = _cold(2, 2, 2, 1.5, 0.5)
    assert abs(L.or_measure_energy_direct(lat.ref) + (2 * 1.5 + 0.5)) <= 1e-12
    assert abs(L.or_measure_magnetization_squared_direct(lat.ref) - 1.0) <= 1e-12
    o = lat.observables()
    h = np.zeros(6)
    L.or_helicity_sums_direct(lat.ref, h.ctypes.data_as(orc._dp))
    assert abs(o["energy"] - L.or_measure_energy_direct(lat.ref)) <= 1e-12
    for i, k in enumerate(["cos_x", "cos_y", "cos_z", "sin_x", "sin_y", "sin_z"]):
        assert abs(o[k] - h[i]) <= 1e-12
    c = lat.correlations(rmax=1)
    for k in ("corr_x", "corr_y", "corr_xy", "corr_z"):
        assert c[k].shape == (1,) and abs(c[k][0] - 1.0) <= 1e-12
    assert np.allclose(c["corr_xy_by_z"], 1.0, atol=1e-12)
    # checkerboard layer at z=1 (src/tests.rs:135-148)
    for y in range(2):
        for x in range(2):
            lat.theta[x + 2 * (y + 2 * 1)] = -1.0 if (x + y) % 2 == 0 else 1.0
    lat.refresh(False)
    c = lat.correlations(rmax=1)
    assert abs(c["corr_xy_by_z"][0, 0] - 1.0) <= 1e-12 and c["corr_xy_by_z"][1, 0] < 0
    assert abs(c["corr_xy"][0] - c["corr_xy_by_z"].mean()) <= 1e-12


def test_cached_equals_direct_random_lattice():
    lat = orc.Lattice(6, 4, 8)
    lat.set_couplings(orc.uniform_layers(8, 0.6, 1.8, 3), orc.uniform_layers(8, 0.5, 1.5, 4))
    lat.init_keyed(77)
    o = lat.observables()
    h = np.zeros(6)
    L.or_helicity_sums_direct(lat.ref, h.ctypes.data_as(orc._dp))
    assert abs(o["energy"] - L.or_measure_energy_direct(lat.ref)) <= 1e-12
    assert abs(o["magnetization_squared"] -
               L.or_measure_magnetization_squared_direct(lat.ref)) <= 1e-12
    for i, k in enumerate(["cos_x", "cos_y", "cos_z", "sin_x", "sin_y", "sin_z"]):
        assert abs(o[k] - h[i]) <= 1e-11


def test_cold_helicity_equals_coupling():
    lat = _cold(3, 3, 3, 1.2, 0.6)
    out = np.zeros(3)
    assert L.or_measure_helicity_modulus(lat.ref, 0.5, out.ctypes.data_as(orc._dp)) == 0
    assert np.allclose(out, [1.2, 1.2, 0.6], atol=1e-12)
    for bad in (0.0, -1.0, math.inf, math.nan):
        assert L.or_measure_helicity_modulus(lat.ref, bad, out.ctypes.data_as(orc._dp)) != 0
        assert orc.err() == "temperature T must be positive and finite"


# ---------------------------------------------------------------------------
# Metropolis (src/tests.rs:155-193)
# ---------------------------------------------------------------------------
def test_metropolis_validation_and_wrapping():
    r = orc.rng(12)
    lat = orc.Lattice(3, 3, 3)
    lat.set_couplings(np.ones(3), np.full(3, 0.8))
    L.or_initialize_angles(lat.ref, 0, ctypes.byref(r))
    old = lat.theta.copy()
    acc = ctypes.c_int()
    assert L.or_local_metropolis_step_ref(lat.ref, 2.0, 0.0, ctypes.byref(r), ctypes.byref(acc)) == 0
    assert acc.value == 1 and np.array_equal(lat.theta, old)
    lat.refresh(False)
    with pytest.raises(orc.OracleError,
                       match="red-black Metropolis sweep requires even lattice dimensions"):
        lat.metro_ref(2.0, 0.5, r)
    assert L.or_local_metropolis_step_ref(lat.ref, 2.0, -0.1, ctypes.byref(r), ctypes.byref(acc))
    assert orc.err() == "proposal_width must be nonnegative and finite"
    ev = orc.Lattice(4, 4, 4)
    ev.set_couplings(np.ones(4), np.full(4, 0.8))
    L.or_initialize_angles(ev.ref, 0, ctypes.byref(r))
    ev.refresh(False)
    with pytest.raises(orc.OracleError, match="proposal_width must be nonnegative and finite"):
        ev.metro_ref(2.0, math.inf, r)
    for batch8 in (True, False):
        a = ev.metro_ref(2.0, 0.5, r, batch8)
        assert 0.0 <= a <= 1.0
        assert np.all((ev.theta >= 0) & (ev.theta < 2 * math.pi))
    for bad in (0.0, -1.0, math.inf, math.nan):
        with pytest.raises(orc.OracleError, match="temperature T must be positive and finite"):
            ev.metro_ref(bad, math.pi, r)


def test_replay_sweep_matches_direct_energy_rule():
    """Brute force: every replay decision equals the Metropolis rule applied to
    V * (E_after - E_before) from the direct-trig energy (src/observables.rs:45-61)
    of that single move, with the keyed proposal/acceptance uniforms recomputed
    from Philox -- independent of the oracle's local-field code."""
    lat = orc.Lattice(4, 4, 4)
    lat.set_couplings(orc.uniform_layers(4, 0.6, 1.8, 2), np.ones(4))
    seed = 5
    lat.init_keyed(seed)
    V, T, width = 64, 0.9, math.pi
    th = lat.theta.copy()
    _, dec = lat.metro_replay(T, width, seed, 0)
    probe = orc.Lattice(4, 4, 4)
    probe.set_couplings(lat.j_xy, lat.j_z)
    u53 = lambda hi, lo: ((hi << 21) | (lo >> 11)) * 2.0 ** -53
    checked = 0
    for parity in range(2):
        for z in range(4):
            for y in range(4):
                for x in range((parity + y + z) & 1, 4, 2):
                    i = x + 4 * (y + 4 * z)
                    w = orc.philox([i, 0, 2, 0], [seed, 0])
                    t_new = th[i] + width * (2.0 * u53(w[0], w[1]) - 1.0)
                    t_new = t_new - 2 * math.pi if t_new >= 2 * math.pi else (
                        t_new + 2 * math.pi if t_new < 0 else t_new)
                    probe.theta[:] = th
                    e0 = L.or_measure_energy_direct(probe.ref) * V
                    probe.theta[i] = t_new
                    de = L.or_measure_energy_direct(probe.ref) * V - e0
                    u_acc = u53(w[2], w[3])
                    p = 1.0 if de <= 0 else math.exp(-de / T)
                    if abs(de) < 1e-9 or abs(u_acc - p) < 1e-9:
                        continue  # too close to the threshold for a direct sum
                    pred = de <= 0 or u_acc < p
                    assert dec[i] == pred, (i, de, u_acc, p)
                    if pred:
                        th[i] = t_new
                    checked += 1
    assert checked >= 60
    assert np.array_equal(th, lat.theta)


# ---------------------------------------------------------------------------
# Wolff (src/tests.rs:372-432)
# ---------------------------------------------------------------------------
def test_wolff_probability_and_reflection():
    assert abs(L.or_wolff_add_probability(1, 1, 1, 1) - (1 - math.exp(-2))) <= 1e-12
    assert abs(L.or_wolff_add_probability(1, 1, 1, -1)) <= 1e-12
    assert abs(L.or_wolff_add_probability(1, 1, -1, 1)) <= 1e-12
    assert abs(L.or_wolff_add_probability(1, 1, 0, 1)) <= 1e-12
    t = L.or_wolff_reflect_angle(0.37, 1.20)
    assert abs(L.or_wolff_reflect_angle(t, 1.20) - 0.37) <= 1e-12
    assert 0 <= t < 2 * math.pi
    # The single-shift wrap (src/types.rs:335-343) brings 2 phi + pi - theta
    # into [0, 2 pi) only from (-2 pi, 4 pi); with phi in [0, 2 pi) the argument
    # reaches up to 5 pi, so a reflection may leave an angle in [2 pi, 3 pi),
    # which the next Metropolis move wraps back.  The restatement keeps this.
    t = L.or_wolff_reflect_angle(0.1, 6.0)
    assert t == (2.0 * 6.0 + math.pi - 0.1) - 2 * math.pi
    assert 2 * math.pi <= t < 3 * math.pi


@pytest.mark.parametrize("use_scratch", [True, False])
def test_wolff_reference_mode_invariants(use_scratch):
    r = orc.rng(15)
    lat = orc.Lattice(4, 4, 4)
    lat.set_couplings(np.ones(4), np.ones(4))
    L.or_initialize_angles(lat.ref, 0, ctypes.byref(r))
    lat.refresh(False)
    for _ in range(2):
        cs = lat.wolff_ref(1.0, r, use_scratch)
        assert 1 <= cs <= 64
        assert np.all((lat.theta >= 0) & (lat.theta < 2 * math.pi))
        assert not np.ctypeslib.as_array(lat.lat.in_cluster, shape=(64,)).any()


def _replay_cluster_independent(lat, T, seed, step):
    """Seed component of the activated-bond graph, computed with numpy/scipy
    from the keyed uniforms: all bonds at once, no traversal order."""
    lx, ly, lz = lat.lx, lat.ly, lat.lz
    V = lat.V
    w = orc.philox([0, step, 3, 0], [seed & 0xFFFFFFFF, seed >> 32])
    phi = 2 * math.pi * ((((w[0] << 21) | (w[1] >> 11))) * 2.0 ** -53)
    seed_site = (w[2] * V) >> 32
    sphi, cphi = orc.sincos(phi)
    r = lat.cos_t * cphi + lat.sin_t * sphi
    idx = np.arange(V)
    x, y, z = idx % lx, (idx // lx) % ly, idx // (lx * ly)
    nbrs = [((x + 1) % lx) + lx * (y + ly * z), x + lx * (((y + 1) % ly) + ly * z),
            x + lx * (y + ly * ((z + 1) % lz))]
    J = [lat.j_xy[z], lat.j_xy[z], lat.j_z[z]]
    words = np.array([orc.philox([i, step, 4, 0], [seed & 0xFFFFFFFF, seed >> 32])
                      for i in range(V)], dtype=np.float64)
    rows, cols = [], []
    beta = 1.0 / T
    for a in range(3):
        prod = J[a] * r * r[nbrs[a]]
        p = 1.0 - np.exp(np.minimum(-2.0 * beta * J[a] * r * r[nbrs[a]], 0.0))
        act = (prod > 0) & (words[:, a] * 2.0 ** -32 < p)
        rows.append(idx[act])
        cols.append(nbrs[a][act])
    rows = np.concatenate(rows)
    cols = np.concatenate(cols)
    g = coo_matrix((np.ones(len(rows)), (rows, cols)), shape=(V, V))
    _, lab = connected_components(g, directed=False)
    return (lab == lab[seed_site]).astype(np.uint8)


@pytest.mark.parametrize("dims,T", [((4, 4, 4), 1.0), ((2, 2, 2), 2.0), ((6, 4, 2), 0.6),
                                    ((8, 8, 8), 2.4)])
def test_wolff_replay_equals_connected_component(dims, T):
    """The replay DFS cluster (traversal order of the reference) equals the
    seed's connected component of the active-bond graph -- the fact that lets a
    GPU BFS reproduce it.  exp here is libm's; p differs from the oracle's own
    exp by <= 1 ulp, far below the uniforms' 2^-32 resolution."""
    lat = orc.Lattice(*dims)
    lz = dims[2]
    lat.set_couplings(orc.uniform_layers(lz, 0.6, 1.8, 8), np.full(lz, 1.0))
    seed = 0x1234_5678_9ABC
    lat.init_keyed(seed)
    for step in range(6):
        expect = _replay_cluster_independent(lat, T, seed, step)
        cs, mem = lat.wolff_replay(T, seed, step)
        assert cs == expect.sum()
        assert np.array_equal(mem, expect)
        lat.metro_replay(T, math.pi, seed, step)


# ---------------------------------------------------------------------------
# closed forms
# ---------------------------------------------------------------------------
def _replay_run(dims, jxy, jz, T, seed, therm, sweeps, wolff=1):
    lat = orc.Lattice(*dims)
    lat.set_couplings(jxy, jz)
    lat.init_keyed(seed)
    k = 0
    e, m2 = [], []
    for s in range(therm + sweeps):
        lat.metro_replay(T, math.pi, seed, s)
        for _ in range(wolff):
            lat.wolff_replay(T, seed, k)
            k += 1
        if s >= therm:
            o = lat.observables()
            e.append(o["energy"])
            m2.append(o["magnetization_squared"])
    return np.array(e), np.array(m2)


def _err(x, nb=20):
    b = x[: len(x) // nb * nb].reshape(nb, -1).mean(axis=1)
    return b.std(ddof=1) / math.sqrt(nb)


def test_infinite_temperature_moments():
    """beta -> 0: <E> = 0, <m2> = 1/V, <m4>/<m2>^2 = 2 - 1/V."""
    dims = (4, 4, 2)
    V = 32
    e, m2 = _replay_run(dims, np.ones(2), np.ones(2), 1e9, 3, 10, 4000)
    assert abs(e.mean()) < 4 * _err(e) + 1e-3
    assert abs(m2.mean() - 1.0 / V) < 4 * _err(m2)
    binder = (m2 ** 2).mean() / m2.mean() ** 2
    assert abs(binder - (2 - 1.0 / V)) < 0.1


def _ring_energy(jz, beta):
    """Exact <E>/V of independent periodic z-rings (J_xy = 0):
    Z = sum_n prod_z I_n(beta J_z), E_ring = -d ln Z / d beta."""
    ns = np.arange(-60, 61)

    def lnZ(b):
        terms = np.ones_like(ns, dtype=float)
        for j in jz:
            terms = terms * special.iv(ns, b * j)
        return math.log(terms.sum())

    h = 1e-5
    e_ring = -(lnZ(beta + h) - lnZ(beta - h)) / (2 * h)
    return e_ring / len(jz)


@pytest.mark.parametrize("mode", ["replay", "reference"])
def test_decoupled_rings_exact_energy(mode):
    """J_xy = 0 with disordered J_z[z]: exact <E> from Bessel functions;
    Metropolis + Wolff agree within 4 sigma (both uniform sources)."""
    lz = 6
    jz = np.array([0.4, 1.3, 0.8, 2.0, 0.6, 1.1])
    T = 0.8
    dims = (4, 4, lz)
    if mode == "replay":
        e, _ = _replay_run(dims, np.zeros(lz), jz, T, 99, 200, 6000)
    else:
        lat = orc.Lattice(*dims)
        lat.set_couplings(np.zeros(lz), jz)
        r = orc.rng(99)
        L.or_initialize_angles(lat.ref, 0, ctypes.byref(r))
        lat.refresh(False)
        e = []
        for s in range(6200):
            lat.metro_ref(T, math.pi, r, True)
            lat.wolff_ref(T, r, True)
            if s >= 200:
                e.append(lat.observables()["energy"])
        e = np.array(e)
    exact = _ring_energy(jz, 1.0 / T)
    assert abs(e.mean() - exact) < 4 * _err(e), (e.mean(), exact, _err(e))


def test_ordered_state_binder_is_one():
    e, m2 = _replay_run((4, 4, 4), np.full(4, 1.0), np.full(4, 1.0), 0.05, 2, 50, 300)
    assert abs((m2 ** 2).mean() / m2.mean() ** 2 - 1.0) < 1e-3
    assert m2.mean() > 0.95


# ---------------------------------------------------------------------------
# plain driver and estimators (src/tests.rs:241-370,434-446)
# ---------------------------------------------------------------------------
def _opts(**kw):
    base = dict(thermal_sweeps=1000, measure_sweeps=5000, measure_interval=10,
                proposal_width=math.pi, wolff_steps=1, correlation_interval=1,
                correlation_rmax=-1, correlation_rmax_xy=-1, correlation_rmax_z=-1,
                cold_init=0, use_batch8=1)
    base.update(kw)
    return orc.OSimOptions(**base)


def _run_sim(dims, params, opts, r):
    lat = orc.Lattice(*dims)
    res = orc.OSimResult()
    bufs = [np.zeros(dims[0] * dims[2] + 64) for _ in range(5)]
    res.corr_x, res.corr_y, res.corr_xy, res.corr_xy_by_z, res.corr_z = (
        b.ctypes.data_as(orc._dp) for b in bufs)
    rc = L.or_run_theta_simulation(lat.ref, *params, ctypes.byref(opts), ctypes.byref(r),
                                   ctypes.byref(res))
    if rc:
        raise orc.OracleError(orc.err())
    return res, bufs


def test_plain_driver():
    r = orc.rng(14)
    res, bufs = _run_sim((4, 4, 4), (1.0, 1.0, 0.0, 0.5),
                         _opts(thermal_sweeps=5, measure_sweeps=10, measure_interval=2,
                               proposal_width=0.4, correlation_rmax=1, cold_init=1), r)
    for v in (res.energy, res.rho_sx, res.rho_sy, res.rho_sz, res.chi, res.corr_z[0]):
        assert math.isfinite(v)
    assert 0 <= res.magnetization_squared <= 1 and 0 <= res.acceptance <= 1
    assert res.rxy == 1 and res.rz == 1
    assert res.num_correlation_measurements == res.num_measurements == 5
    msgs = [
        (_opts(measure_interval=0), "measure_interval must be positive"),
        (_opts(thermal_sweeps=0, measure_sweeps=1, measure_interval=2),
         "measure_sweeps must include at least one measurement; require measure_sweeps >= "
         "measure_interval"),
        (_opts(correlation_interval=0), "correlation_interval must be positive"),
        (_opts(thermal_sweeps=0, measure_sweeps=1, measure_interval=1,
               proposal_width=math.nan), "proposal_width must be nonnegative and finite"),
    ]
    for o, msg in msgs:
        with pytest.raises(orc.OracleError) as e:
            _run_sim((2, 2, 2), (1.0, 1.0, 0.0, 0.5), o, r)
        assert str(e.value) == msg
    res, bufs = _run_sim((2, 2, 2), (1.0, 1.0, 0.0, 0.5),
                         _opts(thermal_sweeps=0, measure_sweeps=5, measure_interval=1,
                               proposal_width=0.0, wolff_steps=0, correlation_rmax=1,
                               correlation_interval=2, cold_init=1), r)
    assert res.num_measurements == 5 and res.num_correlation_measurements == 3
    assert abs(bufs[0][0] - 1.0) <= 1e-12
    with pytest.raises(orc.OracleError) as e:
        _run_sim((2, 2, 2), (1.0, -0.1, 0.0, 1.0),
                 _opts(thermal_sweeps=0, measure_sweeps=1, measure_interval=1), r)
    assert str(e.value) == "two-point layer disorder requires J_z_mean - δJ_z >= 0"


def test_binned_estimate_and_blocking():
    x = np.array([1.0, 3.0, 5.0, 7.0])
    m, s, bins, nb = ctypes.c_double(), ctypes.c_double(), np.zeros(2), ctypes.c_int()
    assert L.or_binned_estimate(x.ctypes.data_as(orc._dp), 4, 2, ctypes.byref(m),
                                ctypes.byref(s), bins.ctypes.data_as(orc._dp),
                                ctypes.byref(nb)) == 0
    assert m.value == 4.0 and abs(s.value - 2.0) <= 1e-12 and list(bins) == [2.0, 6.0]
    assert L.or_binned_estimate(x.ctypes.data_as(orc._dp), 1, 2, ctypes.byref(m),
                                ctypes.byref(s), bins.ctypes.data_as(orc._dp), ctypes.byref(nb))
    assert orc.err() == "binsize is larger than the sample series"
    # blocking_stderr by hand: 16 values -> one halving level (16 -> 8)
    y = np.arange(16, dtype=float) ** 2
    se0 = y.std(ddof=1) / 4.0
    h = y.reshape(8, 2).mean(axis=1)
    se1 = h.std(ddof=1) / math.sqrt(8)
    assert abs(L.or_blocking_stderr(y.ctypes.data_as(orc._dp), 16) - max(se0, se1)) <= 1e-12
    assert L.or_blocking_stderr(y.ctypes.data_as(orc._dp), 1) == 0.0


def _corr_numpy(theta, lx, ly, lz, rxy, rz):
    """C(r) of src/observables.rs:166-265 from the angles alone, by np.roll of
    the (z, y, x) array: cos(theta_i - theta_{i + r e_a}) averaged over sites
    (per plane for corr_xy_by_z), periodic in every direction."""
    t = theta.reshape(lz, ly, lx)
    cx = [np.cos(t - np.roll(t, -r, axis=2)) for r in range(1, rxy + 1)]
    cy = [np.cos(t - np.roll(t, -r, axis=1)) for r in range(1, rxy + 1)]
    cz = [np.cos(t - np.roll(t, -r, axis=0)) for r in range(1, rz + 1)]
    corr_x = np.array([c.mean() for c in cx])
    corr_y = np.array([c.mean() for c in cy])
    by_z = np.array([[(cx[r][z].sum() + cy[r][z].sum()) / (2.0 * lx * ly)
                      for r in range(rxy)] for z in range(lz)]).reshape(lz, rxy)
    return {"corr_x": corr_x, "corr_y": corr_y, "corr_xy": (corr_x + corr_y) / 2,
            "corr_xy_by_z": by_z, "corr_z": np.array([c.mean() for c in cz])}


@pytest.mark.parametrize("shape,radii", [
    ((6, 4, 8), (-1, -1, -1)),
    ((10, 6, 4), (-1, -1, -1)),
    ((8, 8, 12), (-1, -1, -1)),
    ((4, 12, 6), (-1, -1, -1)),
    ((12, 10, 8), (-1, 3, 2)),
    ((16, 16, 16), (5, -1, -1)),
    ((8, 8, 8), (-1, 0, 4)),
])
def test_correlations_equal_numpy_roll(shape, radii):
    """Every output of or_measure_correlations on random, non-cubic lattices for
    all r up to L/2, against an independent computation from the angles (an
    x shift taken along y, or a wrong wrap of the second x run for r >= 2,
    fails here)."""
    lx, ly, lz = shape
    lat = orc.Lattice(lx, ly, lz)
    lat.set_couplings(np.ones(lz), np.ones(lz))
    lat.init_keyed(1000 + lx * 31 + ly * 7 + lz)
    lat.refresh(False)
    c = lat.correlations(*radii)
    rxy, rz = c["corr_x"].shape[0], c["corr_z"].shape[0]
    rmax, rmax_xy, rmax_z = radii
    exp_xy = min(rmax_xy if rmax_xy >= 0 else (rmax if rmax >= 0 else min(lx, ly) // 2),
                 lx // 2, ly // 2)
    exp_z = min(rmax_z if rmax_z >= 0 else (rmax if rmax >= 0 else lz // 2), lz // 2)
    assert (rxy, rz) == (exp_xy, exp_z)
    ref = _corr_numpy(lat.theta, lx, ly, lz, rxy, rz)
    for k in ref:
        assert c[k].shape == ref[k].shape, k
        assert np.allclose(c[k], ref[k], rtol=0, atol=1e-12), (k, np.max(np.abs(c[k] - ref[k])))
    # the random lattice leaves no two shifts alike, so a swapped axis shows
    if rxy >= 1 and rz >= 1:
        assert abs(ref["corr_x"][0] - ref["corr_y"][0]) > 1e-6
        assert abs(ref["corr_x"][0] - ref["corr_z"][0]) > 1e-6


@pytest.mark.parametrize("shape,k", [((8, 6, 10), (1, 2, 3)), ((12, 12, 4), (5, 0, 1)),
                                     ((16, 8, 8), (3, 1, 2))])
def test_correlations_of_a_spin_wave_closed_form(shape, k):
    """theta = 2 pi (kx x / Lx + ky y / Ly + kz z / Lz): every site sees the
    same neighbour angle difference, so C_a(r) = cos(2 pi k_a r / L_a) exactly
    and every plane has corr_xy_by_z = (C_x + C_y) / 2."""
    lx, ly, lz = shape
    lat = orc.Lattice(lx, ly, lz)
    lat.set_couplings(np.ones(lz), np.ones(lz))
    z, y, x = np.meshgrid(np.arange(lz), np.arange(ly), np.arange(lx), indexing="ij")
    th = np.mod(2 * np.pi * (k[0] * x / lx + k[1] * y / ly + k[2] * z / lz), 2 * np.pi)
    lat.theta[:] = th.ravel()
    lat.refresh(False)
    c = lat.correlations()
    rxy, rz = c["corr_x"].shape[0], c["corr_z"].shape[0]
    r = np.arange(1, rxy + 1)
    cx, cy = np.cos(2 * np.pi * k[0] * r / lx), np.cos(2 * np.pi * k[1] * r / ly)
    assert np.allclose(c["corr_x"], cx, atol=1e-12)
    assert np.allclose(c["corr_y"], cy, atol=1e-12)
    assert np.allclose(c["corr_xy"], (cx + cy) / 2, atol=1e-12)
    assert np.allclose(c["corr_xy_by_z"], np.tile((cx + cy) / 2, (lz, 1)), atol=1e-12)
    rzv = np.arange(1, rz + 1)
    assert np.allclose(c["corr_z"], np.cos(2 * np.pi * k[2] * rzv / lz), atol=1e-12)
