"""GPU tests of the continuous-batching server (slsf_server_*).

* A job's result is bit-identical alone and in a mix of jobs of other sizes,
  temperatures and lengths, even with a different chunk size (keyed uniforms,
  per-replica windows and reductions), as long as its pool keeps one form of
  the HBM path; where the number of active jobs moves an HBM pool across the
  slab / per-phase crossover, the chain is the same and the bins differ in
  the order of their FP64 sums only.
* Served correlation series equal Batch's samples bit for bit, on resident,
  slab and per-phase pools.
* Jobs that arrive while others run are admitted as slots free up and overlap
  on the device; a device-memory budget that fits one slot serialises jobs;
  a job whose slot can never fit is refused with a clear message.
"""
import ctypes
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
slsf = pytest.importorskip("slsf_amd")

GB = 1 << 30


def job_args(l, seed, sweeps=300, therm=50, T=1.3, binsize=10):
    return dict(l=l, temperature=T, seed=seed, disorder_seed=seed * 7 + 1, sweeps=sweeps,
                thermalization=therm, binsize=binsize)


def same_result(a, b):
    assert a["nbins"] == b["nbins"] and a["acceptance"] == b["acceptance"]
    for k in a["series_mean"]:
        assert a["series_mean"][k] == b["series_mean"][k], k
    for k in a["value"]:
        va, vb = a["value"][k], b["value"][k]
        assert (va == vb) or (np.isnan(va) and np.isnan(vb)), k


def test_job_result_same_alone_and_in_mix():
    probe_small = job_args(16, 11, sweeps=400, T=1.2)
    probe_big = job_args(48, 12, sweeps=120, therm=20, T=1.6)
    alone = {}
    for name, args in (("small", probe_small), ("big", probe_big)):
        srv = slsf.Server(8 * GB, 16, chunk_sweeps=50)
        srv.submit(**args)
        res = srv.run_until_idle()
        assert len(res) == 1
        alone[name] = res[0]
    mix = slsf.Server(8 * GB, 16, chunk_sweeps=37)
    ids = {}
    others = [job_args(32, 100 + i, sweeps=200 + 50 * i, T=0.5 + 0.3 * i) for i in range(4)]
    others += [job_args(16, 200 + i, sweeps=150, T=2.0) for i in range(3)]
    others += [job_args(48, 300, sweeps=60, therm=10)]
    for i, a in enumerate(others[:4]):
        mix.submit(**a)
    ids["small"] = mix.submit(**probe_small)
    mix.step()
    for a in others[4:]:
        mix.submit(**a)
    ids["big"] = mix.submit(**probe_big)
    res = {r["job_id"]: r for r in mix.run_until_idle()}
    assert len(res) == len(others) + 2
    same_result(alone["small"], res[ids["small"]])
    same_result(alone["big"], res[ids["big"]])
    for r in res.values():
        assert r["nbins"] > 0 and np.isfinite(r["value"]["Energy"]) and r["value"]["Energy"] < 0


def test_continuous_admission_and_overlap():
    srv = slsf.Server(8 * GB, 4, chunk_sweeps=25)
    first = [srv.submit(**job_args(16, 500 + i, sweeps=400)) for i in range(4)]
    srv.step()
    late = [srv.submit(**job_args(16, 600 + i, sweeps=100)) for i in range(3)]
    other = srv.submit(**job_args(32, 700, sweeps=100))
    assert srv.stats()["queued"] == 4 and srv.stats()["active"] == 4
    srv.step()  # admits the L=32 job into its own pool; L=16 has 4 slots, all busy
    st = srv.stats()
    assert st["queued"] == 3 and st["active"] == 5
    res = {r["job_id"]: r for r in srv.run_until_idle()}
    assert len(res) == 8
    # the L=32 job ran while the first L=16 jobs were running
    assert res[other]["start_s"] < min(res[j]["finish_s"] for j in first)
    # late L=16 jobs waited for a free slot, then ran
    for j in late:
        assert res[j]["start_s"] >= min(res[f]["finish_s"] for f in first) - 1e-9


def test_memory_budget_admission():
    one = slsf.Server.slot_bytes(16)
    srv = slsf.Server(one + one // 2, 8, chunk_sweeps=50)  # room for one L=16 slot
    a = srv.submit(**job_args(16, 1, sweeps=100))
    b = srv.submit(**job_args(16, 2, sweeps=100))
    srv.step()
    assert srv.stats()["active"] == 1 and srv.stats()["queued"] == 1
    res = {r["job_id"]: r for r in srv.run_until_idle()}
    assert res[b]["start_s"] >= res[a]["finish_s"] - 1e-9
    with pytest.raises(slsf.SlsfError) as e:
        srv.submit(**job_args(64, 3))
    assert e.value.code == 3 and "more than the device budget" in e.value.message
    # a budget that holds one slot but not its pool's correlation sums refuses
    # the job at once instead of queueing it for ever
    tight = slsf.Server(one + 100, 8, chunk_sweeps=50)
    with pytest.raises(slsf.SlsfError) as e:
        tight.submit(**job_args(16, 5, sweeps=100))
    assert e.value.code == 3 and "more than the device budget" in e.value.message
    tight.close()
    with pytest.raises(slsf.SlsfError) as e:
        srv.submit(**dict(job_args(16, 4), temperature=0.0))
    assert e.value.message == "temperature T must be positive and finite"
    with pytest.raises(slsf.SlsfError) as e:
        srv.submit(**job_args(15, 4))
    assert e.value.message == "red-black Metropolis sweep requires even lattice dimensions"


def test_idle_pools_of_other_sizes_give_their_memory_back():
    """A budget with room for one L=16 slot or one L=32 slot, not both: the
    L=16 job's pool takes the budget first; once that job is done, its idle
    pool is freed so that the queued L=32 job is admitted, and both finish."""
    b16, b32 = slsf.Server.slot_bytes(16), slsf.Server.slot_bytes(32)
    budget = max(b16, b32) + min(b16, b32) // 2
    srv = slsf.Server(budget, 8, chunk_sweeps=50)
    a = srv.submit(**job_args(16, 11, sweeps=100))
    b = srv.submit(**job_args(32, 12, sweeps=100))
    srv.step()
    assert srv.stats()["active"] == 1 and srv.stats()["queued"] == 1
    res = {r["job_id"]: r for r in srv.run_until_idle(max_steps=1000)}
    assert set(res) == {a, b}, res.keys()
    assert res[b]["start_s"] >= res[a]["finish_s"] - 1e-9
    st = srv.stats()
    assert st["queued"] == 0 and st["active"] == 0 and st["device_bytes"] <= budget
    srv.close()


def test_served_job_equals_batch_run():
    """A served job (resident pool for L=16, HBM pool for L=48, chunks of 30
    sweeps) gives exactly the bins of the same replica run through Batch.run:
    same keyed stream, same windows, couplings from SplitMix64 U[0.6, 1.8]."""
    import oracle_lib as orc
    srv = slsf.Server(8 * GB, 4, chunk_sweeps=30)
    specs = {L: job_args(L, 4242 + L, sweeps=120, therm=40, T=1.25) for L in (16, 48)}
    for a in specs.values():
        srv.submit(**a)
    res = {r["l"]: r for r in srv.run_until_idle()}
    for L, a in specs.items():
        jxy = orc.uniform_layers(L, 0.6, 1.8, a["disorder_seed"])
        b = slsf.Batch([slsf.Replica(L, L, L, a["temperature"], jxy, np.ones(L), a["seed"])],
                       slsf.FP32)
        b.init_angles()
        b.run(a["thermalization"], 0, 1)
        b.run(0, a["sweeps"], a["binsize"])
        bins = b.all_bins(0)
        r = res[L]
        assert r["nbins"] == bins.shape[1] == a["sweeps"] // a["binsize"]
        for k, name in enumerate(slsf.SERIES):
            m = 0.0
            for v in bins[k]:
                m += v
            assert r["series_mean"][name] == m / bins.shape[1], (L, name)


def _batch_correlation_bins(a, rxy, rz):
    """Bin means of the correlation samples of a served job's chain, run
    through Batch.run in segments that end on the sampled sweeps (the
    measurements m with m % interval == 0, src/model.rs:410-412)."""
    import oracle_lib as orc
    L, c = a["l"], max(1, a["correlation_interval"])
    jxy = orc.uniform_layers(L, 0.6, 1.8, a["disorder_seed"])
    b = slsf.Batch([slsf.Replica(L, L, L, a["temperature"], jxy, np.ones(L), a["seed"])],
                   slsf.FP32)
    b.init_angles()
    if a["thermalization"]:
        b.run(a["thermalization"], 0, 1)
    samples, done = [], 0
    for m in range(0, a["sweeps"], c):
        b.run(0, m + 1 - done, 1)
        done = m + 1
        o = b.correlations(0, -1, rxy, rz)
        samples.append(np.concatenate([o["corr_x"], o["corr_y"], o["corr_xy"],
                                       o["corr_xy_by_z"].ravel(), o["corr_z"]]))
    b.close()
    bs = a["binsize"]
    nb = len(samples) // bs
    out = np.zeros((nb, len(samples[0])))
    for i in range(nb):
        acc = np.zeros(len(samples[0]))
        for v in samples[i * bs:(i + 1) * bs]:
            acc = acc + v
        out[i] = acc / bs
    return out


def test_served_correlation_series_equal_batch_samples():
    """Correlation series of served jobs (src/model.rs:410-435): the server
    ends chunks on each job's sampled sweeps, so a resident-pool job (L=16,
    every third measurement, radii clamped to L/2) and an HBM-pool job (L=48,
    every second measurement, in-plane only) give the bin means of the
    samples Batch.correlations takes on the same chains, bit for bit (the
    plane sums add the same row dots in the same order).  A job without
    correlations in the same pool keeps its result (the shorter chunks do
    not change its chain), and its result carries no correlation series."""
    with_corr = [dict(job_args(16, 900, sweeps=40, therm=10, T=1.1, binsize=4),
                      correlation_interval=3, correlation_rmax_xy=20, correlation_rmax_z=5),
                 dict(job_args(48, 901, sweeps=30, therm=7, T=1.4, binsize=3),
                      correlation_interval=2, correlation_rmax_xy=3, correlation_rmax_z=0)]
    plain = job_args(16, 902, sweeps=60, therm=5, T=1.3)
    alone = slsf.Server(8 * GB, 4, chunk_sweeps=25)
    alone.submit(**plain)
    plain_alone = alone.run_until_idle()[0]
    assert "correlations" not in plain_alone
    srv = slsf.Server(8 * GB, 4, chunk_sweeps=25)
    ids = [srv.submit(**a) for a in with_corr]
    pid = srv.submit(**plain)
    res = {r["job_id"]: r for r in srv.run_until_idle()}
    same_result(plain_alone, res[pid])
    assert "correlations" not in res[pid]
    for jid, a in zip(ids, with_corr):
        L = a["l"]
        rxy, rz = min(a["correlation_rmax_xy"], L // 2), min(a["correlation_rmax_z"], L // 2)
        r = res[jid]
        assert (r["corr_rxy"], r["corr_rz"]) == (rxy, rz)
        expect = _batch_correlation_bins(a, rxy, rz)
        names = slsf.correlation_series_names(rxy, rz, L)
        assert sorted(r["correlations"]) == sorted(names)
        assert r["corr_nbins"] == expect.shape[0] > 0
        got = np.stack([r["correlations"][n] for n in names], axis=1)
        exact = int(np.sum(got == expect))
        print(f"L={L}: {got.size} bin means, {exact} bit-identical, max |diff| "
              f"{np.max(np.abs(got - expect)):.2e}")
        # same chain, same dots, same fold order: bit for bit
        assert exact == got.size, L
    # the series were taken by poll: a second take is refused
    cnt = ctypes.c_int64()
    with pytest.raises(slsf.SlsfError, match="no correlation series"):
        slsf._check(slsf.lib().slsf_server_correlations(srv._h, ids[0], None, 0,
                                                        ctypes.byref(cnt)))


def test_served_correlations_on_a_per_phase_pool():
    """An L=48 pool with twenty active jobs (540 16-deep tiles, above the
    slab form's 128) runs the per-phase kernels; no FP32 kernel keeps the trig
    cache, so each correlation sample refreshes its slot's (cos, sin) pairs: a
    correlation job among them (z and in-plane radii, every measurement
    sampled) still gives the bin means of Batch's samples on the same chain,
    bit for bit, and its neighbours keep their results."""
    corr = dict(job_args(48, 950, sweeps=12, therm=3, T=1.2, binsize=3),
                correlation_interval=1, correlation_rmax_xy=2, correlation_rmax_z=3)
    others = [job_args(48, 960 + i, sweeps=12 + i % 3, therm=3, T=1.0 + 0.05 * i)
              for i in range(19)]
    srv = slsf.Server(8 * GB, 32, chunk_sweeps=5)
    cid = srv.submit(**corr)
    oids = [srv.submit(**a) for a in others]
    res = {r["job_id"]: r for r in srv.run_until_idle()}
    assert len(res) == 20
    r = res[cid]
    expect = _batch_correlation_bins(corr, 2, 3)
    names = slsf.correlation_series_names(2, 3, 48)
    got = np.stack([r["correlations"][n] for n in names], axis=1)
    assert r["corr_nbins"] == expect.shape[0] == 4
    assert np.array_equal(got, expect)
    # one neighbour against its chain run alone, where its pool takes the slab
    # form: the same chain (acceptance counts equal), bins summed in the other
    # form's order (equal to FP64 rounding)
    alone = slsf.Server(8 * GB, 4, chunk_sweeps=50)
    alone.submit(**others[7])
    a, b = alone.run_until_idle()[0], res[oids[7]]
    assert a["nbins"] == b["nbins"] and a["acceptance"] == b["acceptance"]
    for k in a["series_mean"]:
        assert math.isclose(a["series_mean"][k], b["series_mean"][k], rel_tol=1e-12,
                            abs_tol=1e-14), k


def test_correlation_requests_and_takes_are_checked():
    """Negative radii are refused at submit; a take into a buffer that is too
    small is refused and leaves the series for a proper take."""
    srv = slsf.Server(8 * GB, 4, chunk_sweeps=10)
    with pytest.raises(slsf.SlsfError, match="correlation radii must be nonnegative"):
        srv.submit(**dict(job_args(16, 970, sweeps=8), correlation_rmax_xy=-1))
    a = dict(job_args(16, 971, sweeps=8, therm=0, binsize=2), correlation_interval=2,
             correlation_rmax_xy=2, correlation_rmax_z=1)
    jid = srv.submit(**a)
    res = []
    for _ in range(100):
        srv.step()
        n = ctypes.c_int32()
        buf = (slsf.JobResult * 4)()
        slsf._check(slsf.lib().slsf_server_poll(srv._h, buf, 4, ctypes.byref(n)))
        res += list(buf[:n.value])
        if res:
            break
    assert len(res) == 1 and res[0].job_id == jid and res[0].corr_nbins == 2
    count = ctypes.c_int64()
    slsf._check(slsf.lib().slsf_server_correlations(srv._h, jid, None, 0, ctypes.byref(count)))
    assert count.value == res[0].corr_nseries * 2 == (3 * 2 + 16 * 2 + 1) * 2
    small = np.zeros(count.value - 1)
    with pytest.raises(slsf.SlsfError, match="output buffer too small"):
        slsf._check(slsf.lib().slsf_server_correlations(
            srv._h, jid, small.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(small),
            ctypes.byref(count)))
    full = np.zeros(count.value)
    slsf._check(slsf.lib().slsf_server_correlations(
        srv._h, jid, full.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(full),
        ctypes.byref(count)))
    assert np.all(np.isfinite(full)) and np.all(np.abs(full) <= 1.0 + 1e-12)
    srv.close()
