"""GPU parity: the FP64 replay kernels against the CPU oracle.

Both sides consume the same keyed Philox uniforms, so every Metropolis
accept/reject decision, every Wolff cluster and every angle must agree
bit-for-bit; observables (FP64 sums in a different order) agree to 1e-12.
Cases follow the reference tests' shapes (src/tests.rs) plus the
L=16, T=1.3 layered-disorder configuration.
"""
import numpy as np
import pytest

import oracle_lib as orc

pytestmark = pytest.mark.gpu

slsf = pytest.importorskip("slsf_amd")


def uniform_disorder(lz, seed, lo=0.6, hi=1.8):
    return orc.uniform_layers(lz, lo, hi, seed)


def make_pair(lx, ly, lz, T, seed, j_xy, j_z, width=np.pi, wolff_steps=1):
    rep = slsf.Replica(lx, ly, lz, T, j_xy, j_z, seed, width, wolff_steps)
    lat = orc.Lattice(lx, ly, lz)
    lat.set_couplings(j_xy, j_z)
    lat.init_keyed(seed)
    return rep, lat


CASES = [
    # (lx, ly, lz, T, seed, width)
    (2, 2, 2, 2.0, 11, np.pi),
    (4, 4, 4, 0.5, 14, 0.4),
    (2, 4, 6, 1.0, 99, np.pi),
    (8, 8, 8, 1.3, 12345, np.pi),
    (16, 16, 16, 1.3, 1, np.pi),
]


@pytest.mark.parametrize("lx,ly,lz,T,seed,width", CASES)
def test_init_angles_bitexact(lx, ly, lz, T, seed, width):
    jxy = uniform_disorder(lz, seed)
    jz = np.ones(lz)
    rep, lat = make_pair(lx, ly, lz, T, seed, jxy, jz, width)
    b = slsf.Batch([rep])
    b.init_angles()
    assert np.array_equal(b.get_theta(0), lat.theta)


@pytest.mark.parametrize("lx,ly,lz,T,seed,width", CASES)
def test_metropolis_decisions_bitexact(lx, ly, lz, T, seed, width):
    jxy = uniform_disorder(lz, seed)
    jz = np.full(lz, 1.0)
    rep, lat = make_pair(lx, ly, lz, T, seed, jxy, jz, width)
    b = slsf.Batch([rep])
    b.init_angles()
    nsweeps = 30 if lx < 16 else 10
    for s in range(nsweeps):
        acc_o, dec_o = lat.metro_replay(T, width, seed, s)
        acc_g, dec_g = b.metropolis_sweep(want_decisions=True)
        assert np.array_equal(dec_g, dec_o), f"decision mismatch at sweep {s}"
        assert acc_g[0] == acc_o
        assert np.array_equal(b.get_theta(0), lat.theta), f"theta differs after sweep {s}"


@pytest.mark.parametrize("lx,ly,lz,T,seed,width", CASES)
def test_wolff_cluster_bitexact(lx, ly, lz, T, seed, width):
    jxy = uniform_disorder(lz, seed)
    jz = np.full(lz, 1.0)
    rep, lat = make_pair(lx, ly, lz, T, seed, jxy, jz, width)
    b = slsf.Batch([rep])
    b.init_angles()
    for k in range(20):
        # a Metropolis sweep between clusters decorrelates the start states
        lat.metro_replay(T, width, seed, k)
        b.metropolis_sweep()
        cs_o, mem_o = lat.wolff_replay(T, seed, k)
        cs_g, mem_g = b.wolff_step(want_members=True)
        assert cs_g[0] == cs_o, f"cluster size differs at step {k}"
        assert np.array_equal(mem_g, mem_o), f"cluster membership differs at step {k}"
        assert np.array_equal(b.get_theta(0), lat.theta)


def test_chain_c1_bitexact():
    """L=16, T=1.3, J_xy[z] ~ U[0.6,1.8], J_perp=1: 200 full sweeps
    (Metropolis + 1 Wolff) stay bit-identical; observables agree to 1e-12."""
    L, T, seed = 16, 1.3, 1
    jxy = uniform_disorder(L, 1)
    jz = np.ones(L)
    rep, lat = make_pair(L, L, L, T, seed, jxy, jz)
    b = slsf.Batch([rep])
    b.init_angles()
    for s in range(200):
        lat.metro_replay(T, np.pi, seed, s)
        lat.wolff_replay(T, seed, s)
    b.sweep(200)
    assert np.array_equal(b.get_theta(0), lat.theta)
    m, w = b.counters()
    assert m[0] == 200 and w[0] == 200
    og = b.observables()[0]
    oo = lat.observables()
    for k in oo:
        assert abs(og[k] - oo[k]) <= 1e-12 * max(1.0, abs(oo[k])), k


def test_batch_matches_each_replica_alone():
    """Replicas of different shapes, temperatures, seeds and wolff_steps in one
    batch follow exactly the chains they follow alone."""
    specs = [(4, 4, 4, 0.7, 3, 2), (8, 6, 4, 1.9, 77, 1), (6, 6, 6, 1.2, 5, 0),
             (2, 2, 2, 3.0, 8, 3)]
    reps, lats = [], []
    for lx, ly, lz, T, seed, ws in specs:
        jxy = uniform_disorder(lz, seed)
        jz = np.full(lz, 0.7)
        rep, lat = make_pair(lx, ly, lz, T, seed, jxy, jz, np.pi, ws)
        reps.append(rep)
        lats.append(lat)
    b = slsf.Batch(reps)
    b.init_angles()
    nsw = 25
    b.sweep(nsw)
    for i, ((lx, ly, lz, T, seed, ws), lat) in enumerate(zip(specs, lats)):
        k = 0
        for s in range(nsw):
            lat.metro_replay(T, np.pi, seed, s)
            for _ in range(ws):
                lat.wolff_replay(T, seed, k)
                k += 1
        assert np.array_equal(b.get_theta(i), lat.theta), f"replica {i}"
    m, w = b.counters()
    assert list(m) == [nsw] * 4
    assert list(w) == [nsw * s[5] for s in specs]


@pytest.mark.parametrize("L,rmax", [(8, -1), (16, 8), (6, 2)])
def test_correlations_match_oracle(L, rmax):
    T, seed = 1.3, 21
    jxy = uniform_disorder(L, seed)
    jz = np.ones(L)
    rep, lat = make_pair(L, L, L, T, seed, jxy, jz)
    b = slsf.Batch([rep])
    b.init_angles()
    for s in range(5):
        lat.metro_replay(T, np.pi, seed, s)
        lat.wolff_replay(T, seed, s)
    b.sweep(5)
    cg = b.correlations(0, rmax=rmax)
    co = lat.correlations(rmax=rmax)
    for k in co:
        assert cg[k].shape == co[k].shape, k
        assert np.allclose(cg[k], co[k], rtol=0, atol=1e-12), k


def test_run_bins_match_oracle_series():
    """slsf_run with on-device binning: bins equal the oracle's per-sweep
    observables averaged over the same bins (FP64 replay, 1e-11 relative)."""
    L, T, seed = 8, 1.3, 4
    jxy = uniform_disorder(L, seed)
    jz = np.ones(L)
    rep, lat = make_pair(L, L, L, T, seed, jxy, jz)
    b = slsf.Batch([rep])
    b.init_angles()
    therm, sweeps, bs = 10, 40, 5
    st = b.run(therm, sweeps, bs)
    assert st["site_updates"] == L ** 3 * (therm + sweeps)
    series = []
    for s in range(therm + sweeps):
        lat.metro_replay(T, np.pi, seed, s)
        lat.wolff_replay(T, seed, s)
        if s >= therm:
            o = lat.observables()
            e, m2 = o["energy"], o["magnetization_squared"]
            series.append([e, e * e, m2, m2 * m2, o["cos_x"], o["cos_y"], o["cos_z"],
                           o["sin_x"], o["sin_y"], o["sin_z"], o["sin_x"] ** 2,
                           o["sin_y"] ** 2, o["sin_z"] ** 2])
    series = np.array(series)
    expect = series.reshape(sweeps // bs, bs, 13).mean(axis=1).T
    got = b.all_bins(0)
    assert got.shape == expect.shape
    assert np.allclose(got, expect, rtol=1e-11, atol=1e-12)


