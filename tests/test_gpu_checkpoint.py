"""Checkpoint images and the resumable job runner (the analogue of
ThetaModel's serde checkpoint, src/model.rs:255-271, and of the run_time /
checkpoint_time budget the reference parses, src/model.rs:567-574).

* A batch saved in the middle of a run (mid-bin) and loaded into a new handle
  continues bit-identically: angles, counters, bins and acceptance equal those
  of the batch that ran on, for the FP32 kernels (resident sizes and an HBM
  size in one batch) and for the FP64 replay path.
* Loading checks the image: wrong replicas, precision or size are refused
  with a message.
* run_job_resumable stopped by its time budget and resumed from its
  checkpoint file gives exactly the result of an uninterrupted run_job."""
import math
import os

import numpy as np
import pytest

import oracle_lib as orc

pytestmark = pytest.mark.gpu
slsf = pytest.importorskip("slsf_amd")
from slsf_amd import results as R  # noqa: E402


def _reps(sizes, seed0=500):
    out = []
    for i, L in enumerate(sizes):
        out.append(slsf.Replica(L, L, L, 1.1 + 0.1 * i, orc.uniform_layers(L, 0.6, 1.8, 3 + i),
                                np.ones(L), seed0 + i, math.pi, 1))
    return out


def _state(b):
    return ([b.get_theta(r) for r in range(len(b))], [c.tolist() for c in b.counters()],
            [b.all_bins(r) for r in range(len(b))], b.acceptance())


@pytest.mark.parametrize("precision,sizes", [("fp32", [8, 10, 16, 32]), ("fp64", [8, 10])])
def test_checkpoint_image_continues_the_chain(precision, sizes):
    prec = slsf.FP32 if precision == "fp32" else slsf.FP64_REPLAY
    reps = _reps(sizes)
    a = slsf.Batch(reps, prec)
    a.init_angles()
    a.run(7, 0, 1)
    a.run(0, 13, 5)            # two bins done, three measurements in the open one
    img = a.checkpoint()
    c = slsf.Batch(reps, prec)
    c.restore(img)
    assert np.array_equal(c.checkpoint(), img)
    for b in (a, c):
        b.run(0, 12, 5)
        b.run(3, 0, 1)
        b.run(0, 5, 5)
    sa, sc = _state(a), _state(c)
    for r in range(len(reps)):
        assert np.array_equal(sa[0][r], sc[0][r]), r
        assert np.array_equal(sa[2][r], sc[2][r]), r
        assert sa[2][r].shape[1] == 6
    assert sa[1] == sc[1] and np.array_equal(sa[3], sc[3])


def test_checkpoint_image_is_checked():
    reps = _reps([8, 10])
    a = slsf.Batch(reps, slsf.FP32)
    a.init_angles()
    a.run(2, 3, 1)
    img = a.checkpoint()
    other = _reps([8, 10], seed0=900)
    with pytest.raises(slsf.SlsfError, match="differs from the batch's"):
        slsf.Batch(other, slsf.FP32).restore(img)
    with pytest.raises(slsf.SlsfError, match="precision does not match"):
        slsf.Batch(reps, slsf.FP64_REPLAY).restore(img)
    with pytest.raises(slsf.SlsfError, match="replica count does not match"):
        slsf.Batch(reps[:1], slsf.FP32).restore(img)
    with pytest.raises(slsf.SlsfError, match="wrong size"):
        slsf.Batch(reps, slsf.FP32).restore(img[:-8])
    bad = img.copy()
    bad[0] ^= 0xFF
    with pytest.raises(slsf.SlsfError, match="not a checkpoint image"):
        slsf.Batch(reps, slsf.FP32).restore(bad)
    # a refused image leaves the batch usable
    b = slsf.Batch(reps, slsf.FP32)
    b.init_angles()
    b.run(1, 0, 1)


JOB = ('name = "ckpt_unit"\n[model]\nL=[8,12]\nT=[1.1,1.6]\nsamples=1\nbase_seed=5\n'
       'disorder="uniform"\nj_xy_min=0.6\nj_xy_max=1.8\nj_z=1.0\n'
       '[run]\nsweeps=40\nthermalization=15\nbinsize=4\n')


def test_resumable_run_stops_and_resumes_to_the_same_result(tmp_path, monkeypatch):
    job = slsf.Job(JOB)
    full = R.run_job(job)
    path = str(tmp_path / "rank0.checkpoint")
    # A clock that advances one second per reading: with a 9 s budget the run
    # does the chunks 7 + 7 + 1 thermalization sweeps and 7 measured sweeps
    # (one full bin of 4 and three measurements in the open bin), then stops.
    ticks = iter(range(10 ** 6))
    monkeypatch.setattr(R.time, "monotonic", lambda: float(next(ticks)))
    part, s1 = R.run_job_resumable(job, path, run_time=9, checkpoint_time=1000, chunk_sweeps=7)
    monkeypatch.undo()
    assert s1["stopped_early"] and s1["checkpoint_paths"] == [path] and os.path.exists(path)
    assert [t["measurements"] for t in part["tasks"]] == [7] * 4
    assert part["tasks"][1]["observables"]["Energy"]["internal_bin_len"] == 4
    assert len(part["tasks"][1]["measurement_bins"]["Energy"]) == 1
    done, s2 = R.run_job_resumable(job, path, run_time=3600, checkpoint_time=3600,
                                   chunk_sweeps=7)
    assert not s2["stopped_early"] and s2["task_count"] == len(full["tasks"]) == 4
    for tf, td in zip(full["tasks"], done["tasks"]):
        assert td["measurements"] == 40
        for name, est in tf["observables"].items():
            if name.startswith("_ll_"):
                # wall-clock series: same bins, values from this process's clock
                continue
            assert td["observables"][name]["mean"] == est["mean"], name
        for name, v in tf["measurement_bins"].items():
            if name.startswith("_ll_"):
                w = td["measurement_bins"][name]
                assert w.shape == v.shape and np.all(np.isfinite(w)) and np.all(w >= 0), name
                continue
            assert np.array_equal(td["measurement_bins"][name], v), name
        assert math.isclose(td["acceptance"], tf["acceptance"], rel_tol=1e-12)
    # a finished checkpoint resumes to the same result without running again
    again, s3 = R.run_job_resumable(job, path, run_time=0, checkpoint_time=3600)
    assert not s3["stopped_early"]
    assert again["tasks"][2]["observables"]["Energy"] == done["tasks"][2]["observables"]["Energy"]
