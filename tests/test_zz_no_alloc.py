"""The allocation-free BPNN evaluation (mdc_compute_noalloc) against the
allocating one (mdc_compute) through the C ABI: two calls in a row, the
second on a smaller system (the shared work area still holds the first
call's arrays), at the FP64 bar (1e-12 relative) and the mixed bar (force
RMS 1e-5; the two paths run the same kernels, so they agree far closer),
and a clear error for an evaluation whose arrays exceed the work area."""
import numpy as np
import pytest

from mdcraft_amd import workloads as W
from mdcraft_amd.atoms import mat, vec
from tests.helpers import force_rms_rel, max_rel

pytestmark = pytest.mark.gpu


def evaluate(ctx, w, precision, noalloc, what=7, lanes=8):
    from mdcraft_amd import engine as E
    s = E.System(ctx, [len(w["atoms"])], [w["box"]], lanes=lanes)
    s.upload(w["atoms"])
    p = E.MaiseBPNNPotential(ctx, w["pot"]["path"])
    s.build()
    if noalloc:
        s.compute_noalloc(p, what, precision)
    else:
        s.compute(p, what, precision)
    return s.download()


@pytest.fixture(scope="module")
def cells(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("noalloc"))
    return [W.bpnn_cu(reps=5, model_dir=d), W.bpnn_cu(reps=4, model_dir=d)]


@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_matches_compute_on_two_calls(ctx, cells, precision):
    big, small = cells
    assert len(small["atoms"]) < len(big["atoms"])
    for w in (big, small):
        print("noalloc precision=%d n=%d" % (precision, len(w["atoms"])), flush=True)
        want = evaluate(ctx, w, precision, False)
        got = evaluate(ctx, w, precision, True)
        f, fw = vec(got, "f"), vec(want, "f")
        if precision == 0:
            assert max_rel(f, fw) <= 1e-12
            assert max_rel(got["Ep"], want["Ep"]) <= 1e-12
            assert max_rel(mat(got), mat(want)) <= 1e-12
        else:
            assert force_rms_rel(f, fw) <= 1e-5
            assert max_rel(f, fw) <= 1e-6
            assert max_rel(got["Ep"], want["Ep"]) <= 1e-6
            assert max_rel(mat(got), mat(want)) <= 1e-6


def test_noalloc_needs_a_built_list_and_a_bpnn_potential(ctx, cells):
    from mdcraft_amd import engine as E
    w = cells[1]
    s = E.System(ctx, [len(w["atoms"])], [w["box"]])
    s.upload(w["atoms"])
    p = E.MaiseBPNNPotential(ctx, w["pot"]["path"])
    with pytest.raises(RuntimeError, match="build"):
        s.compute_noalloc(p, 7, 0)
    s.build()
    lj = E.LJsCut(ctx, 1.0, 0.34, 0.6)
    with pytest.raises(RuntimeError, match="BPNN"):
        s.compute_noalloc(lj, 7, 0)


def test_noalloc_at_13500_atoms_and_above_the_work_area(ctx, tmp_path):
    """The 13 500-atom BPNN workload: mixed precision with the virial fits
    the 64 MiB work area and matches mdc_compute; FP64 with the virial
    needs more and is refused with the size in the message."""
    w = W.bpnn_cu(reps=15, model_dir=str(tmp_path))
    print("noalloc n=%d mixed" % len(w["atoms"]), flush=True)
    want = evaluate(ctx, w, 1, False)
    got = evaluate(ctx, w, 1, True)
    assert force_rms_rel(vec(got, "f"), vec(want, "f")) <= 1e-5
    assert max_rel(vec(got, "f"), vec(want, "f")) <= 1e-6
    assert max_rel(mat(got), mat(want)) <= 1e-6
    print("noalloc n=%d fp64 (refused)" % len(w["atoms"]), flush=True)
    with pytest.raises(RuntimeError, match="MiB"):
        evaluate(ctx, w, 0, True)
