"""Result files of the GPU path in the reference's formats (CPU part):
ThetaJobResult JSON (src/model.rs:71-106; src/results.rs:14-36), rank merge
rules and messages (src/results.rs:246-272), and the dataframe restatement
(src/result_tools.rs:54-251): columns, {:.16} floats, quoting."""
import json
import math

import pytest

slsf = pytest.importorskip("slsf_amd")
from slsf_amd import results as R  # noqa: E402


def task(name, idx=0, **kw):
    t = {k: 0 for k in R.TASK_FIELDS}
    t.update(name=name, l=4, l_x=4, l_y=4, l_z=4, temperature=1.3, j_xy=1.2, delta_j_xy=0.0,
             j_z_mean=1.0, delta_j_z=0.0, disorder_seed=7, seed=11, sample=1, sweeps=20,
             thermalization=5, binsize=4, proposal_width=math.pi, wolff_steps=1,
             correlation_interval=1, j_xy_array=[1.2] * 4, j_z_array=[1.0] * 4)
    t.update(kw)
    e = {"task": t}
    if idx:
        e["task_index"] = idx
    e.update(observables={"Energy": {"mean": -1.5, "error": 0.25, "internal_bin_len": 4},
                          "Chi": {"mean": float("nan"), "error": float("inf"),
                                  "internal_bin_len": 4}},
             acceptance=0.5, measurements=20)
    return e


def test_format_float_matches_rust_fixed_16():
    assert R.format_float(1.0) == "1.0000000000000000"
    assert R.format_float(-0.125) == "-0.1250000000000000"
    assert R.format_float(1e20) == "100000000000000000000.0000000000000000"
    assert R.format_float(float("nan")) == "NaN"
    assert R.format_float(float("inf")) == "Inf"
    assert R.format_float(-float("inf")) == "-Inf"
    assert R.format_float(None) == "NaN"  # serde_json wrote a NaN as null


def test_merge_rules_and_messages():
    a = {"job_name": "j", "rank": 0, "world_size": 2, "tasks": [task("b"), task("d", 2)]}
    b = {"job_name": "j", "rank": 1, "world_size": 2, "tasks": [task("a", 1), task("c")]}
    m = R.merge_job_results([a, b])
    assert [t["task"]["name"] for t in m["tasks"]] == ["a", "b", "c", "d"]
    # a zero index takes the position in the merged order; nonzero ones are kept
    assert [t.get("task_index", 0) for t in m["tasks"]] == [1, 1, 2, 2]
    assert m["rank"] == 0 and m["world_size"] == 2
    with pytest.raises(ValueError, match="at least one rank result is required for merge"):
        R.merge_job_results([])
    with pytest.raises(ValueError, match="cannot merge results from different jobs"):
        R.merge_job_results([a, dict(b, job_name="k")])
    with pytest.raises(ValueError, match="cannot merge results with different world sizes"):
        R.merge_job_results([a, dict(b, world_size=3)])


def test_json_round_trip_and_dataframe(tmp_path):
    res = {"job_name": "unit job", "rank": 1, "world_size": 2,
           "tasks": [task("L4_T1.3", 0), task('odd,"name"', 3)]}
    p = R.write_job_result(res, str(tmp_path))
    assert p.endswith("unit job.rank1.results.json")
    back = R.read_job_result(p)
    assert back["tasks"][0]["observables"]["Chi"]["mean"] is None  # NaN -> null
    assert "task_index" not in back["tasks"][0] and back["tasks"][1]["task_index"] == 3
    cols, rows = R.dataframe_from_result(back)
    assert cols[:22] == ["job_name", "rank", "world_size", "task_index", "task_name", "L", "Lx",
                         "Ly", "Lz", "T", "Jxy", "DeltaJxy", "Jz", "DeltaJz", "sample", "seed",
                         "disorder_seed", "sweeps", "thermalization", "binsize", "measurements",
                         "acceptance"]
    assert cols[22:] == ["Chi", "Chi_error", "Chi_measurement", "Energy", "Energy_error",
                         "Energy_measurement"]
    r = rows[0]
    assert r["Energy"] == "-1.5000000000000000" and r["Energy_error"] == "0.2500000000000000"
    assert r["Energy_measurement"] == "-1.5000000000000000 +/- 0.2500000000000000"
    assert r["Chi"] == "NaN" and r["T"] == "1.3000000000000000" and r["task_index"] == "0"
    csv = R.write_dataframe(p, str(tmp_path / "frame.csv"))
    lines = open(csv).read().splitlines()
    assert lines[0].split(",")[:3] == ["job_name", "rank", "world_size"]
    assert '"odd,""name"""' in lines[2]  # quoted field with an escaped quote
    tsv = R.write_dataframe(p, str(tmp_path / "frame.tsv"))
    # a quote forces quoting whatever the delimiter (src/result_tools.rs:227-239)
    assert open(tsv).read().splitlines()[2].split("\t")[4] == '"odd,""name"""'
    json.loads(open(p).read())  # strict JSON (no NaN tokens)
