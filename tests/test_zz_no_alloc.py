"""The allocation-free correlation entry point (slsf_measure_correlations_noalloc)
against slsf_measure_correlations, through the C ABI: the same kernels and the
host fold's order, so every output is bit-identical.  Two calls in a row on one
device share the fixed scratch array, the second on a smaller input; sizes
beyond the array's limit are refused with a message."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

slsf = pytest.importorskip("slsf_amd")
import oracle_lib as orc  # noqa: E402

KEYS = ("corr_x", "corr_y", "corr_xy", "corr_xy_by_z", "corr_z")


def _batch(shapes, precision, sweeps=4):
    reps = [slsf.Replica(lx, ly, lz, 1.3, orc.uniform_layers(lz, 0.6, 1.8, 3 + k), np.ones(lz),
                         40 + k) for k, (lx, ly, lz) in enumerate(shapes)]
    b = slsf.Batch(reps, precision)
    b.init_angles()
    b.sweep(sweeps)
    return b


def _same(a, b, what):
    for k in KEYS:
        assert a[k].shape == b[k].shape, (what, k)
        assert np.array_equal(a[k], b[k]), (what, k, np.max(np.abs(a[k] - b[k])))


@pytest.mark.parametrize("precision", ["FP64_REPLAY", "FP32"])
def test_noalloc_correlations_equal_the_allocating_entry_point(precision):
    prec = getattr(slsf, precision)
    b = _batch([(24, 24, 24), (8, 6, 10), (16, 16, 16)], prec)
    # first call: the largest input; second: a smaller replica and radii
    for replica, radii in ((0, (-1, -1, -1)), (1, (-1, -1, -1)), (2, (3, -1, 2)),
                           (1, (-1, 2, 1)), (0, (-1, 0, 5))):
        rmax, rxy, rz = radii
        ref = b.correlations(replica, rmax, rxy, rz)
        got = b.correlations(replica, rmax, rxy, rz, noalloc=True)
        _same(got, ref, (precision, replica, radii))
    # after more sweeps the shared array holds the last call's values; a new
    # call still gives the allocating entry point's bits
    b.sweep(2)
    _same(b.correlations(2, noalloc=True), b.correlations(2), "after sweeps")
    b.close()


def test_noalloc_correlations_refuse_sizes_beyond_the_scratch():
    b = _batch([(128, 128, 128)], slsf.FP32, sweeps=0)
    with pytest.raises(slsf.SlsfError, match="no-allocation limit"):
        b.correlations(0, noalloc=True)
    # radii that fit are served
    got = b.correlations(0, rmax=4, noalloc=True)
    _same(got, b.correlations(0, rmax=4), "L=128, rmax 4")
    b.close()
