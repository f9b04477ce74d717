"""GPU parity of the slab sweeps of the HBM path (slsf_run on replicas the
resident kernel does not take, and on every FP64 replay replica).

A slab sweep updates each tile's halo again from the keyed uniforms, so its
FP64 chains must equal the oracle's replay chains bit for bit on every shape:
clipped tiles, lattices smaller than the halo (sites appearing several times in
a tile's region), wide and flat lattices, 0 to 3 cluster rounds per sweep,
replicas of different sizes in one batch.  Bins agree with the oracle's series
to 1e-11; the trig cache is restored after a run; FP32 slab chains are held to
the FP64 ones on the same uniforms.
"""
import numpy as np
import pytest

import oracle_lib as orc

pytestmark = pytest.mark.gpu
slsf = pytest.importorskip("slsf_amd")


def _pair(lx, ly, lz, T, seed, ws=1, jzv=1.0):
    jxy = orc.uniform_layers(lz, 0.6, 1.8, seed)
    jz = np.full(lz, jzv)
    rep = slsf.Replica(lx, ly, lz, T, jxy, jz, seed, np.pi, ws)
    lat = orc.Lattice(lx, ly, lz)
    lat.set_couplings(jxy, jz)
    lat.init_keyed(seed)
    return rep, lat


def _oracle_sweeps(lat, rep, nsweeps, m0=0, w0=0):
    k = w0
    for s in range(m0, m0 + nsweeps):
        lat.metro_replay(rep.temperature, rep.proposal_width, rep.seed, s)
        for _ in range(rep.wolff_steps):
            lat.wolff_replay(rep.temperature, rep.seed, k)
            k += 1


SHAPES = [
    # (lx, ly, lz, T, seed, wolff_steps, sweeps)
    (2, 2, 2, 2.0, 11, 1, 12),
    (4, 4, 4, 0.5, 14, 2, 10),
    (2, 4, 6, 1.0, 99, 1, 10),
    (8, 8, 8, 1.3, 12345, 1, 10),
    (6, 6, 6, 1.2, 5, 0, 10),
    (16, 16, 16, 1.3, 1, 1, 6),
    (20, 18, 34, 1.1, 7, 1, 4),
    (36, 4, 2, 0.8, 21, 3, 4),
    (10, 30, 12, 1.6, 8, 1, 4),
    (48, 48, 48, 1.3, 3, 1, 2),
    (64, 64, 64, 1.3, 1, 1, 2),
]


@pytest.mark.parametrize("lx,ly,lz,T,seed,ws,nsw", SHAPES)
def test_slab_chain_equals_oracle(lx, ly, lz, T, seed, ws, nsw):
    rep, lat = _pair(lx, ly, lz, T, seed, ws)
    b = slsf.Batch([rep], slsf.FP64_REPLAY)
    b.init_angles()
    st = b.run(nsw, 0, 1)
    _oracle_sweeps(lat, rep, nsw)
    assert np.array_equal(b.get_theta(0), lat.theta)
    m, w = b.counters()
    assert m[0] == nsw and w[0] == nsw * ws
    assert st["launches"] > 0
    b.close()


def test_slab_batch_of_shapes_equals_oracle_and_per_phase_kernels():
    """Replicas of different shapes and cluster rounds in one batch, run in two
    pieces (the second continues the first): every chain is the oracle's, and
    the per-phase kernels (slsf_sweep) give the same chains."""
    specs = [(4, 4, 4, 0.7, 3, 2), (8, 6, 4, 1.9, 77, 1), (6, 6, 6, 1.2, 5, 0),
             (2, 2, 2, 3.0, 8, 3), (20, 18, 34, 1.1, 9, 1), (32, 8, 16, 1.4, 10, 1)]
    pairs = [_pair(lx, ly, lz, T, seed, ws, 0.7) for lx, ly, lz, T, seed, ws in specs]
    reps = [p[0] for p in pairs]
    a = slsf.Batch(reps, slsf.FP64_REPLAY)
    c = slsf.Batch(reps, slsf.FP64_REPLAY)
    for b in (a, c):
        b.init_angles()
    a.run(5, 0, 1)
    a.run(4, 3, 1)
    c.sweep(12)
    for i, (rep, lat) in enumerate(pairs):
        _oracle_sweeps(lat, rep, 12)
        assert np.array_equal(a.get_theta(i), lat.theta), i
        assert np.array_equal(c.get_theta(i), lat.theta), i
    assert list(a.counters()[0]) == [12] * len(specs)
    assert list(a.counters()[1]) == [12 * s[5] for s in specs]
    assert np.allclose(a.acceptance(), c.acceptance(), rtol=0, atol=1e-15)


@pytest.mark.parametrize("shape,ws", [((8, 8, 8), 1), ((20, 18, 34), 2), ((48, 48, 48), 1)])
def test_slab_bins_match_oracle_series(shape, ws):
    """Measurements ride on the next sweep's slab and the run's closing
    measurement: the bins equal the oracle's per-sweep series binned alike."""
    lx, ly, lz = shape
    T, seed = 1.3, 4
    rep, lat = _pair(lx, ly, lz, T, seed, ws)
    b = slsf.Batch([rep], slsf.FP64_REPLAY)
    b.init_angles()
    therm, sweeps, bs = (6, 12, 3) if lx < 40 else (2, 4, 2)
    b.run(therm, sweeps, bs)
    series = []
    k = 0
    for s in range(therm + sweeps):
        lat.metro_replay(T, np.pi, seed, s)
        for _ in range(ws):
            lat.wolff_replay(T, seed, k)
            k += 1
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
    # the trig cache is restored for readers of it after the run
    og = b.observables()[0]
    oo = lat.observables()
    for key in oo:
        assert abs(og[key] - oo[key]) <= 1e-12 * max(1.0, abs(oo[key])), key
    b.close()


def test_slab_fp32_follows_fp64_on_the_same_uniforms():
    """FP32 slab sweeps (replay stream: the FP64 path's per-site uniforms)
    against FP64 slab sweeps from the same angles: one sweep of 8 L=48
    replicas agrees to FP32 rounding, as the per-phase kernels do."""
    reps = [_pair(48, 48, 48, 1.1, 100 + i)[0] for i in range(8)]
    b64 = slsf.Batch(reps, slsf.FP64_REPLAY)
    b32 = slsf.Batch(reps, slsf.FP32)
    for b in (b64, b32):
        b.init_angles()
    b64.run(1, 0, 1)
    b32.run(1, 0, 1, replay_stream=True)
    differ = 0
    for r in range(len(reps)):
        d = np.abs(b64.get_theta(r) - b32.get_theta(r)) % (2 * np.pi)
        d = np.minimum(d, 2 * np.pi - d)
        differ += int(d.max() >= 2e-6)
    assert differ <= 1, differ


def test_slab_replica_alone_and_in_a_large_batch_bit_identical():
    """Slab sweeps of a replica alone take 8-deep tiles, of the same replica in
    a batch of 256 or more 16-deep tiles 16-deep ones: its FP32 chain and its
    bins are bit-identical either way (the measurement is summed per 8-deep half
    tile in both).  With the form chosen by launch size (per-phase sweeps for
    the batch) the chain is still bit-identical and the bins agree to 1e-12."""
    L = 64
    reps = [_pair(L, L, L, 1.2 + 0.05 * i, 500 + i)[0] for i in range(4)]  # 4 x 64 tiles
    for form in (1, 0):
        big = slsf.Batch(reps, slsf.FP32)
        one = slsf.Batch([reps[2]], slsf.FP32)
        for b in (big, one):
            b.init_angles()
            b.run(3, 6, 2, hbm_form=form)
        assert np.array_equal(big.get_theta(2), one.get_theta(0)), form
        if form == 1:
            assert np.array_equal(big.all_bins(2), one.all_bins(0))
        else:
            assert np.allclose(big.all_bins(2), one.all_bins(0), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("precision", ["FP64_REPLAY", "FP32"])
def test_slab_and_per_phase_forms_give_the_same_chains_and_bins(precision):
    """The two forms of the HBM path (slab sweeps, per-phase sweeps) on a batch
    of shapes and cluster rounds, also switching form between runs of one
    batch: angles and counters are bit-identical; the bins, FP64 sums over
    differently shaped blocks, agree to 1e-12 (relative)."""
    prec = getattr(slsf, precision)
    specs = [(48, 48, 48, 1.3, 1, 1), (20, 18, 34, 1.1, 2, 2), (10, 30, 12, 1.6, 3, 0),
             (64, 64, 64, 1.25, 4, 1)]
    reps = [_pair(lx, ly, lz, T, seed, ws)[0] for lx, ly, lz, T, seed, ws in specs]
    out = []
    for forms in ((1, 1), (2, 2), (1, 2), (2, 1)):
        b = slsf.Batch(reps, prec)
        b.init_angles()
        b.run(3, 6, 2, hbm_form=forms[0])
        b.run(0, 4, 2, hbm_form=forms[1])
        out.append((forms, [b.get_theta(i) for i in range(len(reps))],
                    [b.all_bins(i) for i in range(len(reps))], b.counters()))
        b.close()
    ref = out[0]
    for forms, th, bins, ctr in out[1:]:
        for i in range(len(reps)):
            assert np.array_equal(th[i], ref[1][i]), (forms, i)
            assert np.allclose(bins[i], ref[2][i], rtol=1e-12, atol=1e-12), (forms, i)
        assert np.array_equal(ctr[0], ref[3][0]) and np.array_equal(ctr[1], ref[3][1])


def test_forms_agree_on_metropolis_only_batches():
    """A batch whose replicas take no cluster rounds ends its per-phase sweeps
    with the fold kernel of its own instead of the reflection kernel's
    bookkeeping: chains and counters equal the slab form's, bins to 1e-12."""
    specs = [(48, 48, 48, 1.3, 5, 0), (10, 30, 12, 1.6, 6, 0)]
    reps = [_pair(lx, ly, lz, T, seed, ws)[0] for lx, ly, lz, T, seed, ws in specs]
    out = []
    for form in (1, 2):
        b = slsf.Batch(reps, slsf.FP32)
        b.init_angles()
        b.run(2, 6, 2, hbm_form=form)
        out.append(([b.get_theta(i) for i in range(len(reps))],
                    [b.all_bins(i) for i in range(len(reps))], b.counters()))
        b.close()
    for i in range(len(reps)):
        assert np.array_equal(out[0][0][i], out[1][0][i]), i
        assert out[0][1][i].shape[1] == 3
        assert np.allclose(out[0][1][i], out[1][1][i], rtol=1e-12, atol=1e-12), i
    for k in range(2):
        assert np.array_equal(out[0][2][k], out[1][2][k])
    assert list(out[1][2][0]) == [8, 8] and list(out[1][2][1]) == [0, 0]


def test_hbm_form_is_checked():
    b = slsf.Batch([_pair(48, 48, 48, 1.3, 1, 1)[0]], slsf.FP32)
    b.init_angles()
    with pytest.raises(slsf.SlsfError, match="hbm_form must be 0"):
        b.run(1, 0, 1, hbm_form=3)
    b.close()
