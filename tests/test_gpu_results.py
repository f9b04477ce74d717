"""GPU job runner (slsf_amd.results.run_job): a TOML job's tasks run as one
batch; each task's estimates equal those of the same replica run alone (keyed
uniforms), the JSON and dataframe files are written, and ranks split the tasks
as the reference does (idx % world == rank) and merge back to the full job."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
slsf = pytest.importorskip("slsf_amd")
from slsf_amd import results as R  # noqa: E402

JOB = ('name = "gpu_results_unit"\n[model]\nL=[8,12]\nT=[1.1,1.6]\nsamples=2\nbase_seed=3\n'
       'disorder="uniform"\nj_xy_min=0.6\nj_xy_max=1.8\nj_z=1.0\n'
       '[run]\nsweeps=60\nthermalization=20\nbinsize=10\n')


def test_run_job_files_and_alone_equivalence(tmp_path):
    job = slsf.Job(JOB)
    res = R.run_job(job)
    assert res["job_name"] == "gpu_results_unit" and len(res["tasks"]) == 8
    for k, tr in enumerate(res["tasks"]):
        obs = tr["observables"]
        for name in ("Energy", "MagnetizationSquared", "_helicity_cos_x", "RhoXY", "RhoZ",
                     "Chi", "SpecificHeat", "Magnetization", "Binder"):
            assert name in obs and np.isfinite(obs[name]["mean"]), name
        assert obs["Energy"]["internal_bin_len"] == 10 and tr["measurements"] == 60
        assert 0 < tr["acceptance"] < 1
        assert tr.get("task_index", 0) == k
    # task 5 alone: identical bins, so identical estimates
    t = job.task(5)
    b = slsf.Batch([job.replica(5)], slsf.FP32)
    b.init_angles()
    b.run(20, 0, 1)
    b.run(0, 60, 10)
    bins = b.all_bins(0)
    e = res["tasks"][5]["observables"]["Energy"]
    assert e["mean"] == float(np.mean(bins[0]))
    assert res["tasks"][5]["task"]["name"] == t["name"]
    p = R.write_job_result(res, str(tmp_path))
    frame = R.write_dataframe(p, str(tmp_path / "frame.tsv"))
    lines = open(frame).read().splitlines()
    assert len(lines) == 9 and "RhoXY_measurement" in lines[0].split("\t")


def test_ranks_split_and_merge():
    job = slsf.Job(JOB)
    parts = [R.run_job(job, rank=r, world_size=3) for r in range(3)]
    assert sorted(len(p["tasks"]) for p in parts) == [2, 3, 3]
    merged = R.merge_job_results(parts)
    full = R.run_job(job)
    assert [t["task"]["name"] for t in merged["tasks"]] == sorted(
        t["task"]["name"] for t in full["tasks"])
    by_name = {t["task"]["name"]: t for t in full["tasks"]}
    for t in merged["tasks"]:
        assert t["observables"]["Energy"] == by_name[t["task"]["name"]]["observables"]["Energy"]
    with pytest.raises(slsf.SlsfError):
        R.run_job(job, rank=3, world_size=3)


def test_run_job_measurement_files(tmp_path):
    """A GPU run's device bins go to the reference's measurement files
    (task{n}/run0001.meas.h5); read back bit-exactly when h5dump is present."""
    import os
    import shutil
    import subprocess
    job = slsf.Job(JOB)
    res = R.run_job(job, rank=1, world_size=4)
    paths = R.write_job_measurements(res, R.measurement_dir(job.name, str(tmp_path)))
    assert len(paths) == 2
    for tr, p in zip(res["tasks"], paths):
        assert p.endswith(os.path.join(f"task{tr['task_index'] + 1:04d}", "run0001.meas.h5"))
        # the 13 series and, correlations being on by default (interval 1), the Corr series
        t = tr["task"]
        corr = R._corr_names(t["l_x"] // 2, t["l_z"] // 2, t["l_z"])
        assert sorted(tr["measurement_bins"]) == sorted(
            list(slsf.SERIES) + corr + ["_ll_sweep_time", "_ll_measure_time"])
        assert all(len(v) == 6 for v in tr["measurement_bins"].values())
    h5dump = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
    if os.path.exists(h5dump):
        out = tmp_path / "raw.bin"
        subprocess.run([h5dump, "-d", "/observables/Energy/samples", "-b", "LE", "-o", str(out),
                        paths[1]], check=True, capture_output=True)
        assert out.read_bytes() == res["tasks"][1]["measurement_bins"]["Energy"].tobytes()


JOB_CORR = ('name = "corr_unit"\n[model]\nL=[8]\nT=[1.3]\nsamples=1\nbase_seed=9\n'
            '[run]\nsweeps=20\nthermalization=10\nbinsize=2\n'
            '[measure]\ncorr_rmax_xy=[3]\ncorr_rmax_z=[2]\ncorrelation_interval=3\n')


def test_correlation_series_follow_the_reference_cadence():
    """CorrX/CorrY/CorrXY/CorrXY_z*/CorrZ series (src/model.rs:410-435): taken
    on measurements 0, 3, 6, ... of the task's own chain and binned like every
    series.  An independent stepping of the same replica, one measured sweep
    at a time, gives exactly the same bins."""
    job = slsf.Job(JOB_CORR)
    t0 = job.task(0)
    assert (t0["correlation_rmax_xy"], t0["correlation_rmax_z"], t0["correlation_interval"]) \
        == (3, 2, 3)
    res = R.run_job(job)
    tr = res["tasks"][0]
    corr = sorted(n for n in tr["observables"] if n.startswith("Corr"))
    assert corr == sorted(R._corr_names(3, 2, 8))
    assert len(corr) == 3 * 3 + 8 * 3 + 2
    b = slsf.Batch([job.replica(0)], slsf.FP32)
    b.init_angles()
    b.run(10, 0, 1)
    samples = []
    for m in range(20):
        b.run(0, 1, 2)
        if m % 3 == 0:
            c = b.correlations(0, -1, 3, 2)
            samples.append((c["corr_x"][1], c["corr_xy_by_z"][5, 2], c["corr_z"][0]))
    s = np.array(samples[:6]).reshape(3, 2, 3).mean(axis=1)  # 7 samples: 3 full bins of 2
    for j, name in enumerate(("CorrX_r2", "CorrXY_z5_r3", "CorrZ_r1")):
        assert np.array_equal(tr["measurement_bins"][name], s[:, j]), name
        assert tr["observables"][name]["mean"] == float(np.mean(s[:, j])), name
        assert tr["observables"][name]["internal_bin_len"] == 2
    assert np.array_equal(b.all_bins(0)[0], tr["measurement_bins"]["Energy"])


JOB_ORACLE = ('name = "oracle_unit"\n[model]\nL=[6,8]\nT=[1.2,1.7]\nsamples=2\nbase_seed=5\n'
              'disorder="uniform"\nj_xy_min=0.6\nj_xy_max=1.8\nj_z=1.0\n'
              '[run]\nsweeps=24\nthermalization=6\nbinsize=4\n'
              '[measure]\ncorr_rmax_xy=[2,3]\ncorr_rmax_z=[2,2]\ncorrelation_interval=2\n')


def test_run_job_fp64_replay_equals_the_oracle_series():
    """A TOML job through run_job in FP64 replay: every task's binned series,
    its correlation series (on measurements 0, 2, 4, ...) and its estimates
    equal the oracle's replay chain of the same task binned alike (1e-11); the
    timer series (_ll_sweep_time, _ll_measure_time) are positive and binned
    alike."""
    import oracle_lib as orc
    job = slsf.Job(JOB_ORACLE)
    res = R.run_job(job, precision=slsf.FP64_REPLAY)
    assert len(res["tasks"]) == 8
    for k, tr in enumerate(res["tasks"]):
        t = tr["task"]
        rep = job.replica(k)
        therm, sweeps, bs = 6, 24, 4
        lat = orc.Lattice(rep.l_x, rep.l_y, rep.l_z)
        lat.set_couplings(rep.j_xy, rep.j_z)
        lat.init_keyed(rep.seed)
        rxy, rz, interval = R._corr_spec(t)
        series, corr = [], []
        kw = 0
        for sw in range(therm + sweeps):
            lat.metro_replay(rep.temperature, rep.proposal_width, rep.seed, sw)
            for _ in range(rep.wolff_steps):
                lat.wolff_replay(rep.temperature, rep.seed, kw)
                kw += 1
            if sw < therm:
                continue
            o = lat.observables()
            e, m2 = o["energy"], o["magnetization_squared"]
            series.append([e, e * e, m2, m2 * m2, o["cos_x"], o["cos_y"], o["cos_z"],
                           o["sin_x"], o["sin_y"], o["sin_z"], o["sin_x"] ** 2,
                           o["sin_y"] ** 2, o["sin_z"] ** 2])
            if (sw - therm) % interval == 0:
                c = lat.correlations(-1, rxy, rz)
                corr.append(np.concatenate([c["corr_x"], c["corr_y"], c["corr_xy"],
                                            c["corr_xy_by_z"].ravel(), c["corr_z"]]))
        expect = np.array(series).reshape(sweeps // bs, bs, 13).mean(axis=1).T
        for sidx, name in enumerate(slsf.SERIES):
            got = tr["measurement_bins"][name]
            assert np.allclose(got, expect[sidx], rtol=1e-11, atol=1e-12), (k, name)
            assert abs(tr["observables"][name]["mean"] - expect[sidx].mean()) <= \
                1e-11 * max(1.0, abs(expect[sidx].mean())), (k, name)
        corr = np.array(corr)
        nb = len(corr) // bs
        cexp = corr[:nb * bs].reshape(nb, bs, -1).mean(axis=1)
        for j, name in enumerate(R._corr_names(rxy, rz, rep.l_z)):
            assert np.allclose(tr["measurement_bins"][name], cexp[:, j], rtol=1e-11,
                               atol=1e-12), (k, name)
        ev = slsf.evaluate(expect, float(rep.volume), 1.0 / rep.temperature)
        for name in ("RhoXY", "RhoZ", "Chi", "SpecificHeat", "Binder"):
            assert abs(tr["observables"][name]["mean"] - ev[name][0]) <= \
                1e-10 * max(1.0, abs(ev[name][0])), (k, name)
        for name in ("_ll_sweep_time", "_ll_measure_time"):
            v = tr["measurement_bins"][name]
            assert len(v) == sweeps // bs and np.all(v > 0) and np.all(np.isfinite(v)), name
