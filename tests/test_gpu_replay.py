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
    # several 16^3 union-find tiles, clipped tiles and wrapped faces, large clusters
    (20, 18, 34, 1.1, 7, np.pi),
    (36, 4, 2, 0.8, 21, np.pi),
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


def test_cluster_tiling_does_not_change_the_chain():
    """The cluster kernels take 16x16x8 tiles when a launch has fewer 16-deep
    tiles than the device has compute units (a replica alone) and 16x16x16
    tiles otherwise (here five L=64 replicas: 320 tiles).  The cluster is the
    seed's component either way, so replica 0 follows the same FP64 chain in
    both, and that chain is the oracle's."""
    L, T, seed = 64, 1.3, 21
    jxy = uniform_disorder(L, 4)
    jz = np.ones(L)
    one = slsf.Batch([slsf.Replica(L, L, L, T, jxy, jz, seed, np.pi, 1)])
    many = slsf.Batch([slsf.Replica(L, L, L, T, jxy, jz, seed + i, np.pi, 1) for i in range(5)])
    for b in (one, many):
        b.init_angles()
        b.sweep(4)
    assert np.array_equal(one.get_theta(0), many.get_theta(0))
    _, lat = make_pair(L, L, L, T, seed, jxy, jz)
    for s in range(4):
        lat.metro_replay(T, np.pi, seed, s)
        lat.wolff_replay(T, seed, s)
    assert np.array_equal(one.get_theta(0), lat.theta)


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




def _oracle_driver(dims, T, jxy, jz, seed, **opt):
    lat = orc.Lattice(*dims)
    o = orc.OSimOptions(opt.get("thermal_sweeps", 1000), opt.get("measure_sweeps", 5000),
                        opt.get("measure_interval", 10), opt.get("proposal_width", np.pi),
                        opt.get("wolff_steps", 1), opt.get("correlation_interval", 1),
                        opt.get("correlation_rmax", -1), -1, -1, int(opt.get("cold_init", 0)), 0)
    res = orc.OSimResult()
    # capacities the driver requires: l_x/2, l_z * l_x/2, l_z/2 (generously sized)
    cap = dims[0] * dims[2] + 64
    bufs = [np.zeros(cap) for _ in range(5)]
    res.corr_x, res.corr_y, res.corr_xy, res.corr_xy_by_z, res.corr_z = (
        b.ctypes.data_as(orc._dp) for b in bufs)
    import ctypes
    rc = orc.lib().or_run_theta_simulation_replay(
        lat.ref, 1.0, 1.0, 0.0, T, jxy.ctypes.data_as(orc._dp), jz.ctypes.data_as(orc._dp),
        ctypes.byref(o), seed, ctypes.byref(res))
    assert rc == 0, orc.err()
    return res, bufs


def test_driver_c1_matches_oracle_replay():
    """The plain driver on the GPU (FP64 replay) against the oracle's replay
    driver: L=16, T=1.3, J_xy[z] ~ U[0.6,1.8], J_perp = 1, C(r) up to r=8,
    measured every sweep.  Every estimator agrees to 1e-10 (relative)."""
    L, T, seed = 16, 1.3, 1
    jxy = orc.uniform_layers(L, 0.6, 1.8, 1)
    jz = np.ones(L)
    opt = dict(thermal_sweeps=50, measure_sweeps=150, measure_interval=1, correlation_rmax=8)
    g = slsf.run_theta_simulation(L, L, L, 1.0, 1.0, 0.0, T, seed=seed, j_xy_array=jxy,
                                  j_z_array=jz, **opt)
    o, bufs = _oracle_driver((L, L, L), T, jxy, jz, seed, **opt)
    for k in ("energy", "rho_sx", "rho_sy", "rho_sz", "std_rho_sx", "std_rho_sy", "std_rho_sz",
              "magnetization_squared", "chi", "acceptance"):
        ov = getattr(o, k)
        assert abs(g[k] - ov) <= 1e-10 * max(1.0, abs(ov)), k
    assert g["num_measurements"] == o.num_measurements == 150
    assert g["num_correlation_measurements"] == o.num_correlation_measurements
    assert np.allclose(g["corr_x"], bufs[0][:8], atol=1e-12)
    assert np.allclose(g["corr_z"], bufs[4][:8], atol=1e-12)
    assert np.allclose(g["corr_xy_by_z"].ravel(), bufs[3][:L * 8], atol=1e-12)


# Shapes of the one-workgroup replay (Lx a divisor of 64 or not, Lx = 32, plane
# sums over x from lane groups of 2, 4, 8 and 16 or from the scratch fold) and
# one too large for it (HBM kernels).
@pytest.mark.parametrize("dims,mi,ci", [((12, 12, 12), 1, 1), ((32, 4, 8), 2, 3),
                                        ((6, 8, 10), 3, 2), ((4, 4, 4), 1, 1),
                                        ((8, 8, 8), 1, 1), ((16, 8, 12), 1, 2),
                                        ((2, 4, 32), 2, 1), ((4, 8, 16), 1, 1),
                                        ((16, 16, 8), 1, 2),
                                        ((18, 18, 18), 2, 2)])
def test_driver_shapes_match_oracle_replay(dims, mi, ci):
    lx, ly, lz = dims
    T, seed = 1.1, 7
    jxy = orc.uniform_layers(lz, 0.6, 1.8, 3)
    jz = np.linspace(0.5, 1.5, lz)
    opt = dict(thermal_sweeps=20, measure_sweeps=60, measure_interval=mi, correlation_rmax=4,
               correlation_interval=ci, wolff_steps=2)
    g = slsf.run_theta_simulation(lx, ly, lz, 1.0, 1.0, 0.0, T, seed=seed, j_xy_array=jxy,
                                  j_z_array=jz, **opt)
    o, bufs = _oracle_driver(dims, T, jxy, jz, seed, **opt)
    for k in ("energy", "rho_sx", "rho_sy", "rho_sz", "std_rho_sx", "std_rho_sy", "std_rho_sz",
              "magnetization_squared", "chi", "acceptance"):
        ov = getattr(o, k)
        assert abs(g[k] - ov) <= 1e-10 * max(1.0, abs(ov)), (dims, k)
    assert g["num_measurements"] == o.num_measurements == 60 // mi
    assert g["num_correlation_measurements"] == o.num_correlation_measurements
    rxy, rz = len(g["corr_x"]), len(g["corr_z"])
    assert np.allclose(g["corr_x"], bufs[0][:rxy], atol=1e-12)
    assert np.allclose(g["corr_y"], bufs[1][:rxy], atol=1e-12)
    assert np.allclose(g["corr_z"], bufs[4][:rz], atol=1e-12)
    assert np.allclose(g["corr_xy_by_z"].ravel(), bufs[3][:lz * rxy], atol=1e-12)


def test_driver_chunked_run_matches_oracle_replay():
    """A run longer than one launch of the one-workgroup replay (1024 sweeps):
    measurement and correlation indices carry across the launch boundaries."""
    dims, T, seed = (4, 4, 6), 0.9, 5
    jxy = orc.uniform_layers(dims[2], 0.6, 1.8, 9)
    jz = np.full(dims[2], 0.8)
    opt = dict(thermal_sweeps=700, measure_sweeps=1500, measure_interval=3,
               correlation_rmax=2, correlation_interval=2)
    g = slsf.run_theta_simulation(*dims, 1.0, 1.0, 0.0, T, seed=seed, j_xy_array=jxy,
                                  j_z_array=jz, **opt)
    o, bufs = _oracle_driver(dims, T, jxy, jz, seed, **opt)
    for k in ("energy", "rho_sx", "rho_sz", "std_rho_sx", "std_rho_sz",
              "magnetization_squared", "acceptance"):
        ov = getattr(o, k)
        assert abs(g[k] - ov) <= 1e-10 * max(1.0, abs(ov)), k
    assert g["num_measurements"] == o.num_measurements == 500
    assert g["num_correlation_measurements"] == o.num_correlation_measurements == 250
    assert np.allclose(g["corr_x"], bufs[0][:2], atol=1e-12)
    assert np.allclose(g["corr_z"], bufs[4][:2], atol=1e-12)


@pytest.mark.parametrize("precision", [0, 1])
def test_driver_reference_assertions(precision):  # src/tests.rs:241-298,333-353
    r = slsf.run_theta_simulation(4, 4, 4, 1.0, 1.0, 0.0, 0.5, thermal_sweeps=5,
                                  measure_sweeps=10, measure_interval=2, proposal_width=0.4,
                                  correlation_rmax=1, cold_init=True, seed=14,
                                  precision=precision)
    for k in ("energy", "rho_sx", "rho_sy", "rho_sz", "chi"):
        assert np.isfinite(r[k])
    assert 0 <= r["magnetization_squared"] <= 1 and 0 <= r["acceptance"] <= 1
    assert len(r["corr_x"]) == 1 and len(r["corr_z"]) == 1
    assert r["num_correlation_measurements"] == r["num_measurements"] == 5
    t = slsf.run_theta_simulation(2, 2, 2, 1.0, 1.0, 0.0, 0.5, thermal_sweeps=0,
                                  measure_sweeps=5, measure_interval=1, proposal_width=0.0,
                                  wolff_steps=0, correlation_rmax=1, correlation_interval=2,
                                  cold_init=True, precision=precision)
    assert t["num_measurements"] == 5 and t["num_correlation_measurements"] == 3
    assert abs(t["corr_x"][0] - 1.0) < 1e-6
