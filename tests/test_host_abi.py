"""Host-side logic of libslsf.so (CPU only, no kernel launches).

* The library loads and exports every entry point include/slsf.h declares.
* Replica validation returns the reference's exact messages before any device
  call (src/types.rs:122-130,373-385; src/updates.rs:435; src/model.rs:347).
* The TOML job loader against a restatement of the reference's schema and task
  expansion built on tomli (src/model.rs:159-213,549-644,697-828) for every
  example job file of the reference, plus the reference's own job-config test
  (src/tests.rs:448-566).
* Estimators: BinnedEstimate (src/tests.rs:436-446), blocking_stderr, jackknife
  closed forms, evaluables (src/model.rs:857-888).
"""
import glob
import math
import os
import re

import numpy as np
import pytest
import tomli

import slsf_amd as slsf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HERE = os.path.join(ROOT, "tests")


def test_library_exports_every_declared_symbol():
    text = open(os.path.join(ROOT, "include", "slsf.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(slsf_[a-z0-9_]+)\s*\(", text))
    assert len(names) > 30
    lib = slsf.lib()
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.slsf_abi_version() == 1


def _rep(**kw):
    base = dict(l_x=4, l_y=4, l_z=4, temperature=1.0, j_xy=np.ones(4), j_z=np.ones(4), seed=1)
    base.update(kw)
    return slsf.Replica(**base)


@pytest.mark.parametrize("kw,msg", [
    (dict(l_x=0), "L_x must be positive"),
    (dict(l_y=0), "L_y must be positive"),
    (dict(l_z=0, j_xy=np.ones(1), j_z=np.ones(1)), "L_z must be positive"),
    (dict(temperature=0.0), "temperature T must be positive and finite"),
    (dict(temperature=-1.0), "temperature T must be positive and finite"),
    (dict(temperature=math.inf), "temperature T must be positive and finite"),
    (dict(temperature=math.nan), "temperature T must be positive and finite"),
    (dict(proposal_width=-0.1), "proposal_width must be nonnegative and finite"),
    (dict(proposal_width=math.inf), "proposal_width must be nonnegative and finite"),
    (dict(l_x=3, l_y=3, l_z=3, j_xy=np.ones(3), j_z=np.ones(3)),
     "red-black Metropolis sweep requires even lattice dimensions"),
    (dict(j_z=-np.ones(4)), "theta simulation requires nonnegative layer couplings"),
    (dict(j_xy=np.array([1.0, -0.5, 1.0, 1.0])),
     "theta simulation requires nonnegative layer couplings"),
])
def test_validation_messages(kw, msg):
    with pytest.raises(slsf.SlsfError) as e:
        slsf.Batch([_rep(**kw)])
    assert e.value.message == msg
    assert e.value.code == 1


# ---------------------------------------------------------------------------
# job configuration
# ---------------------------------------------------------------------------
M64 = (1 << 64) - 1


class SplitMix:
    def __init__(self, seed):
        self.s = seed & M64

    def u64(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & M64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
        return z ^ (z >> 31)

    def bool(self):
        return (self.u64() & 0xFFFFFFFF) >= 0x80000000


def rust_round(x):
    r = math.floor(abs(x) + 0.5) * (1 if x >= 0 else -1)
    return max(0, int(r))


def expand_reference(doc):
    """Restatement of ThetaJobConfig::try_from_toml_spec + make_theta_tasks."""
    m = doc.get("model", {})
    r = doc.get("run", {})
    me = doc.get("measure")
    L = m.get("l", m.get("L", list(range(4, 21))))
    lx, ly, lz = m.get("l_x"), m.get("l_y"), m.get("l_z")
    temps = m.get("temperatures", m.get("T", m.get("t", [2.8, 2.9, 3.0, 3.1, 3.2, 3.3])))
    djz = m.get("delta_j_z", m.get("djz", [0.0]))
    djxy = m.get("delta_j_xy", m.get("djxy", [0.0]))
    samples = m.get("samples", 16)
    base = m.get("base_seed", 20260414)
    jxy = m.get("j_xy", 1.0)
    jzm = m.get("j_z_mean", m.get("j_z", 0.1))
    sweeps = r.get("sweeps", 20000)
    binsize = r.get("binsize", 50)
    rm = rmxy = rmz = None
    if me is not None:
        rm = me.get("corr_rmax")
        rmxy = me.get("corr_rmax_xy", rm)
        rmz = me.get("corr_rmax_z", rm)
    if lx is None and ly is None and lz is None:
        specs = [(v, v, v, v) for v in L]
    else:
        specs = [(x, y, z, z) for x in (lx or L) for y in (ly or L) for z in (lz or L)]
    tasks = []
    for li, (sx, sy, sz, sl) in enumerate(specs):
        for dxy in djxy:
            for dz in djz:
                for t in temps:
                    for s in range(1, samples + 1):
                        dseed = base + s - 1
                        seed = (base + 100000 * sz + 1000 * s + rust_round(100 * t)
                                + rust_round(1000 * dxy) * 10000 + rust_round(1000 * dz)) & M64
                        g = SplitMix(dseed)
                        ax = [jxy if dxy == 0 else (jxy + dxy if g.bool() else jxy - dxy)
                              for _ in range(sz)]
                        az = [jzm if dz == 0 else (jzm + dz if g.bool() else jzm - dz)
                              for _ in range(sz)]
                        tasks.append({
                            "name": f"L{sx}x{sy}x{sz}_T{t:.6f}_dJxy{dxy:.6f}_dJz{dz:.6f}_sample{s}",
                            "l": sl, "l_x": sx, "l_y": sy, "l_z": sz, "temperature": t,
                            "seed": seed, "disorder_seed": dseed, "sample": s,
                            "sweeps": sweeps, "thermalization": r.get("thermalization", 5000),
                            "binsize": binsize,
                            "proposal_width": r.get("proposal_width", math.pi),
                            "wolff_steps": r.get("wolff_steps", 1),
                            "correlation_rmax": rm[li] if rm else sz // 2,
                            "correlation_rmax_xy": rmxy[li] if rmxy else min(sx, sy) // 2,
                            "correlation_rmax_z": rmz[li] if rmz else sz // 2,
                            "j_xy_array": ax, "j_z_array": az})
    return tasks


EXAMPLES = sorted(glob.glob(os.path.join(HERE, "golden", "examples", "*.toml")))


@pytest.mark.parametrize("path", EXAMPLES, ids=[os.path.basename(p) for p in EXAMPLES])
def test_reference_example_configs_expand_identically(path):
    text = open(path).read()
    doc = tomli.loads(text)
    expect = expand_reference(doc)
    job = slsf.Job(text)
    assert job.name == doc["name"]
    assert len(job) == len(expect)
    step = max(1, len(expect) // 300)  # large grids: check a strided subset
    for i in list(range(0, len(expect), step)) + [len(expect) - 1]:
        got = job.task(i)
        exp = expect[i]
        for k, v in exp.items():
            if k.endswith("_array"):
                assert np.array_equal(got[k], np.array(v, dtype=float)), (i, k)
            elif isinstance(v, float):
                assert got[k] == pytest.approx(v, abs=0), (i, k)
            else:
                assert got[k] == v, (i, k, got[k], v)
    rt, ct = job.durations()
    assert rt == slsf.parse_duration(doc["run_time"])
    assert ct == slsf.parse_duration(doc["checkpoint_time"])


def test_reference_job_config_test():  # src/tests.rs:448-566 restated through TOML
    job = slsf.Job("""name = "unit_theta_job"
[model]
l = [2]
temperatures = [1.0, 1.5]
delta_j_z = [0.0]
samples = 2
base_seed = 17
[run]
sweeps = 4
thermalization = 1
binsize = 2
wolff_steps = 0
[measure]
corr_rmax = [0]
correlation_interval = 2
""")
    assert job.name == "unit_theta_job" and len(job) == 4
    t0 = job.task(0)
    assert (t0["l"], t0["l_x"], t0["l_y"], t0["l_z"], t0["binsize"]) == (2, 2, 2, 2, 2)
    assert t0["correlation_interval"] == 2 and len(t0["j_xy_array"]) == 2
    assert job.selected_tasks(1, 2) == [1, 3]
    assert job.selected_tasks(0, 1) == [0, 1, 2, 3]
    dims = slsf.Job("[model]\nl=[4]\nl_x=[2]\nl_y=[3]\nl_z=[5]\nT=[1.0]\nsamples=1\n"
                    "[run]\nsweeps=2\nbinsize=1\n")
    t = dims.task(0)
    assert (t["l"], t["l_x"], t["l_y"], t["l_z"]) == (5, 2, 3, 5)
    dur = slsf.Job('name = "carlo_mpi_params"\nrun_time = "1-02:03:04"\n'
                   'checkpoint_time = "05:06"\ncheckpoint = true\n')
    assert dur.durations() == (93784, 306)
    corr = slsf.Job("[model]\nL=[16,32]\nT=[1.0]\nsamples=1\n[run]\nsweeps=2\nbinsize=1\n"
                    "[measure]\ncorr_rmax=[8,16]\ncorr_rmax_xy=[7,15]\ncorr_rmax_z=[6,14]\n")
    a, b = corr.task(0), corr.task(1)
    assert (a["l"], a["correlation_rmax"], a["correlation_rmax_xy"], a["correlation_rmax_z"]) \
        == (16, 8, 7, 6)
    assert (b["l"], b["correlation_rmax"], b["correlation_rmax_xy"], b["correlation_rmax_z"]) \
        == (32, 16, 15, 14)
    with pytest.raises(slsf.SlsfError) as e:
        slsf.Job("[model]\nL=[16,32]\n[measure]\ncorr_rmax_xy=[16]\n")
    assert e.value.message == ("corr_rmax_xy must contain exactly one value per lattice size: "
                               "expected 2, got 1")
    with pytest.raises(slsf.SlsfError) as e:
        slsf.Job('run_time = "48:invalid:00"\n')
    assert "invalid run_time" in e.value.message


@pytest.mark.parametrize("text,msg", [
    ("[model]\nsamples=0\n", "samples must be positive"),
    ("[run]\nbinsize=0\n", "binsize must be positive"),
    ("[model]\nT=[]\n", "at least one temperature must be configured"),
    ("[model]\nL=[]\n", "at least one lattice size must be configured"),
    ("[model]\ndelta_j_xy=[]\n", "delta_j_xy and delta_j_z must not be empty"),
    ("[run]\nsweeps=10\nbinsize=20\n", "sweeps must be at least binsize"),
    ("[measure]\ncorrelation_interval=0\n", "correlation_interval must be positive"),
    ("[model]\nj_z=0.05\ndelta_j_z=[0.1]\n",
     "two-point layer disorder requires J_z_mean - δJ_z >= 0"),
    ("[model]\ndelta_j_xy=[-0.1]\n", "δJ_xy must be nonnegative"),
])
def test_task_expansion_errors(text, msg):
    with pytest.raises(slsf.SlsfError) as e:
        slsf.Job(text)
    assert e.value.message == msg


def test_uniform_disorder_extension_and_durations():
    job = slsf.Job('[model]\nL=[8]\nT=[1.3]\nsamples=3\nbase_seed=1\ndisorder="uniform"\n'
                   'j_xy_min=0.6\nj_xy_max=1.8\nj_z=1.0\n[run]\nsweeps=10\nbinsize=5\n')
    for i in range(3):
        t = job.task(i)
        g = SplitMix(t["disorder_seed"])
        expect = [0.6 + (1.8 - 0.6) * ((g.u64() >> 11) * 2.0 ** -53) for _ in range(8)]
        assert np.array_equal(t["j_xy_array"], expect)
        assert np.all(t["j_z_array"] == 1.0) and t["disorder_kind"] == 1
    assert slsf.parse_duration("45") == 45
    assert slsf.parse_duration("2-00:00:01") == 2 * 86400 + 1
    for bad, msg in [("1:2:3:4", "1:2:3:4 does not match [[HH:]MM:]SS"),
                     ("x-01", "failed to parse duration days in x-01: invalid digit found in string"),
                     ("01::02", "failed to parse duration component in 01::02: cannot parse "
                                "integer from empty string")]:
        with pytest.raises(slsf.SlsfError) as e:
            slsf.parse_duration(bad)
        assert e.value.message == msg


# ---------------------------------------------------------------------------
# estimators
# ---------------------------------------------------------------------------
def test_binned_estimate_kat():
    m, s, bins = slsf.binned_estimate([1.0, 3.0, 5.0, 7.0], 2)
    assert m == 4.0 and abs(s - 2.0) < 1e-12 and list(bins) == [2.0, 6.0]
    with pytest.raises(slsf.SlsfError) as e:
        slsf.binned_estimate([1.0], 2)
    assert e.value.message == "binsize is larger than the sample series"


def test_blocking_stderr_by_hand():
    y = np.arange(16, dtype=float) ** 2
    se0 = y.std(ddof=1) / 4.0
    se1 = y.reshape(8, 2).mean(axis=1).std(ddof=1) / math.sqrt(8)
    assert abs(slsf.blocking_stderr(y) - max(se0, se1)) < 1e-12
    assert slsf.blocking_stderr([1.0]) == 0.0


def test_jackknife_closed_forms():
    rng = np.random.default_rng(3)
    nb = 40
    bins = rng.normal(size=(len(slsf.SERIES), nb)) + 5.0
    V, beta = 1000.0, 0.8
    ev = slsf.evaluate(bins, V, beta)
    e = bins[0]
    # linear observable: the jackknife error is the naive error of the bin means
    assert abs(ev["Energy"][0] - e.mean()) < 1e-12
    assert abs(ev["Energy"][1] - e.std(ddof=1) / math.sqrt(nb)) < 1e-12
    assert abs(ev["Chi"][1] - beta * V * bins[2].std(ddof=1) / math.sqrt(nb)) < 1e-9
    # value formulas (src/model.rs:857-888)
    mm = bins.mean(axis=1)
    assert abs(ev["Magnetization"][0] - math.sqrt(mm[2])) < 1e-12
    assert abs(ev["SpecificHeat"][0] - beta ** 2 * V * (mm[1] - mm[0] ** 2)) < 1e-9
    rx = mm[4] / V - beta * (mm[10] - mm[7] ** 2) / V
    ry = mm[5] / V - beta * (mm[11] - mm[8] ** 2) / V
    rz = mm[6] / V - beta * (mm[12] - mm[9] ** 2) / V
    assert abs(ev["RhoXY"][0] - (rx + ry) / 2) < 1e-12 and abs(ev["RhoZ"][0] - rz) < 1e-12
    assert abs(ev["Binder"][0] - mm[3] / mm[2] ** 2) < 1e-12
    # generic jackknife: estimate of a linear f is unbiased, error as above
    loo = (e.sum() - e) / (nb - 1)
    est, err = slsf.jackknife(loo, e.mean())
    assert abs(est - e.mean()) < 1e-12 and abs(err - e.std(ddof=1) / math.sqrt(nb)) < 1e-12


@pytest.mark.parametrize("kw,msg", [
    (dict(measure_interval=0), "measure_interval must be positive"),
    (dict(thermal_sweeps=0, measure_sweeps=1, measure_interval=2),
     "measure_sweeps must include at least one measurement; require measure_sweeps >= "
     "measure_interval"),
    (dict(correlation_interval=0), "correlation_interval must be positive"),
    (dict(thermal_sweeps=0, measure_sweeps=1, measure_interval=1, proposal_width=math.nan),
     "proposal_width must be nonnegative and finite"),
])
def test_driver_validation_messages(kw, msg):  # src/tests.rs:300-352
    with pytest.raises(slsf.SlsfError) as e:
        slsf.run_theta_simulation(2, 2, 2, 1.0, 1.0, 0.0, 0.5, **kw)
    assert e.value.message == msg


def test_driver_disorder_messages():  # src/tests.rs:354-369
    with pytest.raises(slsf.SlsfError) as e:
        slsf.run_theta_simulation(2, 2, 2, 1.0, -0.1, 0.0, 1.0, thermal_sweeps=0,
                                  measure_sweeps=1, measure_interval=1)
    assert e.value.message == "two-point layer disorder requires J_z_mean - δJ_z >= 0"
    with pytest.raises(slsf.SlsfError) as e:
        slsf.run_theta_simulation(2, 2, 2, 1.0, 1.0, -0.5, 1.0, thermal_sweeps=0,
                                  measure_sweeps=1, measure_interval=1)
    assert e.value.message == "δJ_z must be nonnegative"
    with pytest.raises(slsf.SlsfError) as e:
        slsf.run_theta_simulation(2, 2, 2, 1.0, 1.0, 0.0, 0.0)
    assert e.value.message == "temperature T must be positive and finite"


def _synthetic_bins(rng, nb):
    """Bins of the 13 series with realistic means and correlated fluctuations:
    E and E^2, m^2 and m^4, and the helicity sums drawn jointly (the nonlinear
    evaluables mix them)."""
    e = -2.0 + 0.02 * rng.standard_normal(nb)
    m2 = 0.6 + 0.05 * rng.standard_normal(nb)
    m4 = m2 ** 2 + 0.01 + 0.01 * rng.standard_normal(nb)
    cos = 3000.0 + 20.0 * rng.standard_normal((3, nb))
    sin = 5.0 * rng.standard_normal((3, nb))
    sin2 = sin ** 2 + 400.0 + 30.0 * rng.standard_normal((3, nb))
    return np.vstack([e, e * e + 1e-4 * rng.standard_normal(nb), m2, m4, cos, sin, sin2])


def _evaluables_numpy(mean, V, beta):
    """The evaluables of src/model.rs:857-888 (plus the Binder ratio) from
    series means, for the error checks below."""
    rho = [mean[4 + a] / V - beta * (mean[10 + a] - mean[7 + a] ** 2) / V for a in range(3)]
    return {"Magnetization": math.sqrt(mean[2]), "Chi": beta * V * mean[2],
            "SpecificHeat": beta ** 2 * V * (mean[1] - mean[0] ** 2),
            "RhoXY": (rho[0] + rho[1]) / 2, "RhoZ": rho[2], "Binder": mean[3] / mean[2] ** 2,
            "Energy": mean[0]}


def test_jackknife_errors_of_nonlinear_evaluables_equal_leave_one_out():
    """slsf_evaluate's errors against a direct leave-one-out (each bin removed
    with np.delete, the evaluable recomputed, sqrt((n-1)/n sum (f_i - f.)^2))."""
    rng = np.random.default_rng(11)
    V, beta = 32768.0, 1 / 1.3
    for nb in (2, 3, 7, 40):
        bins = _synthetic_bins(rng, nb)
        ev = slsf.evaluate(bins, V, beta)
        loo = [_evaluables_numpy(np.delete(bins, i, axis=1).mean(axis=1), V, beta)
               for i in range(nb)]
        for name in slsf.EVALS:
            f = np.array([x[name] for x in loo])
            err = math.sqrt((nb - 1) / nb * np.sum((f - f.mean()) ** 2))
            assert abs(ev[name][1] - err) <= 1e-9 * max(abs(err), 1e-300), (nb, name)


def test_jackknife_errors_approach_the_delta_method():
    """At many bins the jackknife error of a smooth f equals the delta-method
    error sqrt(g^T C g / n), g the analytic gradient of f at the means and C the
    bins' covariance (an independent route to the same number, O(1/n) apart)."""
    rng = np.random.default_rng(5)
    V, beta, nb = 32768.0, 1 / 1.3, 4000
    bins = _synthetic_bins(rng, nb)
    ev = slsf.evaluate(bins, V, beta)
    mu = bins.mean(axis=1)
    C = np.cov(bins)
    g = {k: np.zeros(13) for k in ("SpecificHeat", "RhoXY", "RhoZ", "Binder", "Magnetization")}
    g["SpecificHeat"][0] = -2 * beta ** 2 * V * mu[0]
    g["SpecificHeat"][1] = beta ** 2 * V
    for a, w, name in ((0, 0.5, "RhoXY"), (1, 0.5, "RhoXY"), (2, 1.0, "RhoZ")):
        g[name][4 + a] += w / V
        g[name][10 + a] += -w * beta / V
        g[name][7 + a] += w * 2 * beta * mu[7 + a] / V
    g["Binder"][3] = 1 / mu[2] ** 2
    g["Binder"][2] = -2 * mu[3] / mu[2] ** 3
    g["Magnetization"][2] = 0.5 / math.sqrt(mu[2])
    for name, grad in g.items():
        delta = math.sqrt(grad @ C @ grad / nb)
        assert abs(ev[name][1] / delta - 1) < 0.01, (name, ev[name][1], delta)


def test_jackknife_error_matches_the_spread_over_independent_datasets():
    """The reported error of the Binder ratio and the specific heat against the
    actual standard deviation of the estimate over 400 independent synthetic
    runs of 60 bins each (within 10%)."""
    rng = np.random.default_rng(8)
    V, beta = 32768.0, 1 / 1.3
    vals = {"Binder": [], "SpecificHeat": [], "RhoZ": []}
    errs = {k: [] for k in vals}
    for _ in range(400):
        ev = slsf.evaluate(_synthetic_bins(rng, 60), V, beta)
        for k in vals:
            vals[k].append(ev[k][0])
            errs[k].append(ev[k][1])
    for k in vals:
        spread = np.std(vals[k], ddof=1)
        assert abs(np.mean(errs[k]) / spread - 1) < 0.1, (k, np.mean(errs[k]), spread)


def test_disorder_average_against_numpy():
    """Disorder averages over realizations: the mean of per-realization
    evaluables with the standard error of the mean (a linear average's
    jackknife error), and the Binder ratio as the ratio of disorder averages
    [<m^4>]/[<m^2>]^2 with a leave-one-realization-out jackknife."""
    rng = np.random.default_rng(21)
    V, beta, n = 4096.0, 1 / 1.2, 24
    means = _synthetic_bins(rng, n).T  # [n_real][13]
    out = slsf.disorder_average(means, V, beta)
    per = [_evaluables_numpy(m, V, beta) for m in means]
    for name in ("Energy", "Chi", "SpecificHeat", "RhoXY", "RhoZ", "Magnetization"):
        x = np.array([p[name] for p in per])
        assert abs(out[name][0] - x.mean()) <= 1e-12 * max(1.0, abs(x.mean())), name
        assert abs(out[name][1] - x.std(ddof=1) / math.sqrt(n)) <= 1e-10 * max(1.0, x.std()), name
    m2, m4 = means[:, 2], means[:, 3]
    g = m4.mean() / m2.mean() ** 2
    loo = np.array([np.delete(m4, r).mean() / np.delete(m2, r).mean() ** 2 for r in range(n)])
    err = math.sqrt((n - 1) / n * np.sum((loo - loo.mean()) ** 2))
    assert abs(out["Binder"][0] - g) <= 1e-14 and abs(out["Binder"][1] - err) <= 1e-12
    with pytest.raises(slsf.SlsfError, match="at least two realizations"):
        slsf.disorder_average(means[:1], V, beta)


def _binder_means(temps, n_real, slope, noise, rng, tc=1.3, b0=1.2):
    """Per-realization series means whose disorder-averaged Binder ratio is
    b0 - slope (T - tc) plus realization noise: m^2 = 1, m^4 = that ratio."""
    out = np.zeros((len(temps), n_real, len(slsf.SERIES)))
    out[:, :, 2] = 1.0
    for t, T in enumerate(temps):
        out[t, :, 3] = b0 - slope * (T - tc) + noise * rng.standard_normal(n_real)
    return out


def test_crossing_of_lines_with_a_known_intersection():
    temps = np.array([1.1, 1.2, 1.25, 1.35, 1.5])
    rng = np.random.default_rng(4)
    a = _binder_means(temps, 16, 0.5, 0.0, rng)
    b = _binder_means(temps, 16, 0.8, 0.0, rng)
    t, e = slsf.crossing(temps, a, 512.0, b, 4096.0)
    assert abs(t - 1.3) <= 1e-12 and e <= 1e-12
    # curves that do not cross
    c = _binder_means(temps, 16, 0.5, 0.0, rng, b0=1.5)
    assert slsf.crossing(temps, a, 512.0, c, 4096.0) is None
    # a crossing exactly on a grid temperature
    grid = np.array([1.2, 1.3, 1.4])
    t2, _ = slsf.crossing(grid, _binder_means(grid, 16, 0.5, 0.0, rng), 512.0,
                          _binder_means(grid, 16, 0.8, 0.0, rng), 4096.0)
    assert abs(t2 - 1.3) <= 1e-12
    with pytest.raises(slsf.SlsfError, match="at least two temperatures"):
        slsf.crossing(temps[:1], a[:1], 512.0, b[:1], 4096.0)
    # scaled curves (the L rho crossings): 1.02 (1.2 - 0.5 d) = 1.2 - 0.8 d at
    # d = T - 1.3 = -0.024 / 0.29
    t3, _ = slsf.crossing(temps, a, 512.0, b, 4096.0, scale_a=1.02, scale_b=1.0)
    assert abs(t3 - (1.3 - 0.024 / 0.29)) <= 1e-12


def test_crossing_jackknife_error_matches_the_spread_over_datasets():
    """With realization noise the crossing scatters around 1.3; its jackknife
    error over realizations matches the standard deviation of the crossing over
    300 independent synthetic datasets (within 15%)."""
    temps = np.array([1.1, 1.2, 1.3, 1.4, 1.5])
    rng = np.random.default_rng(9)
    ts, es = [], []
    for _ in range(300):
        a = _binder_means(temps, 32, 0.5, 0.02, rng)
        b = _binder_means(temps, 32, 0.9, 0.02, rng)
        res = slsf.crossing(temps, a, 512.0, b, 4096.0)
        assert res is not None
        ts.append(res[0])
        es.append(res[1])
    spread = np.std(ts, ddof=1)
    assert abs(np.mean(ts) - 1.3) < 4 * spread / math.sqrt(300) + 1e-3
    assert abs(np.mean(es) / spread - 1) < 0.15, (np.mean(es), spread)
