"""The engine at a million atoms against the oracle: the fcc Cu crystal of
64^3 unit cells (1 048 576 atoms) under the EAM Cu tables and the argon
crystal of the same cell count under LJ.

Bars: pair sets equal (sorted (i, j) keys); FP64 within 1e-12 relative
(max |difference| over the max magnitude); mixed precision force RMS within
1e-5 and energy sums within 1e-6; 20 FP64 NVE steps through list rebuilds
within 1e-9 nm on positions and 1e-8 relative on velocities."""
import numpy as np
import pytest

from mdcraft_amd import workloads as W
from mdcraft_amd.atoms import vec
from tests.helpers import engine_potential, max_rel, oracle_compute
from tests.test_gpu_parity import compare, oracle_nve, run_engine

pytestmark = pytest.mark.gpu

REPS = 64


def E():
    from mdcraft_amd import engine
    return engine


def workload(which):
    return W.lj_argon(reps=REPS, rattle=0.05) if which == "lj" else W.eam_cu(reps=REPS)


@pytest.fixture(scope="module")
def oracle_results():
    cache = {}

    def get(which):
        if which not in cache:
            w = workload(which)
            ref, lists = oracle_compute(w["atoms"], w["box"], w["pot"], threads=16)
            cache[which] = (w, ref, lists)
        return cache[which]

    return get


def pair_keys(off, ids):
    n = len(off) - 1
    i = np.repeat(np.arange(n, dtype=np.int64), np.diff(off))
    return np.sort(i * n + ids)


@pytest.mark.parametrize("which", ["eam", "lj"])
@pytest.mark.parametrize("precision", [0, 1])
def test_million_atoms(ctx, oracle_results, which, precision):
    w, ref, lists = oracle_results(which)
    n = len(w["atoms"])
    assert n == 4 * REPS ** 3
    print(f"million-atom case: {which}, precision {precision}, {n} atoms", flush=True)
    got, s, _ = run_engine(ctx, w["atoms"], w["box"], w["pot"], precision)
    assert np.array_equal(pair_keys(*s.export()), pair_keys(*lists))
    compare(got, ref, precision, f"{which}{n}")


def test_million_atoms_nve_fp64(ctx):
    w = W.eam_cu(reps=REPS)
    steps = 20
    print(f"million-atom NVE: eam, {len(w['atoms'])} atoms, {steps} steps", flush=True)
    ref, reb = oracle_nve(w["atoms"], w["box"], w["pot"], w["dt"], steps)
    assert reb >= 1
    eng = E()
    s = eng.System(ctx, [len(w["atoms"])], [w["box"]])
    s.upload(w["atoms"])
    p = engine_potential(eng, ctx, w["pot"])
    s.build()
    s.compute(p, 1, 0)
    assert s.step(p, w["dt"], steps, 0) >= 1
    got = s.download()
    L = w["box"][1] - w["box"][0]
    d = vec(got, "r") - vec(ref, "r")
    d -= np.round(d / L) * L
    assert np.max(np.abs(d)) < 1e-9
    assert max_rel(vec(got, "v"), vec(ref, "v")) < 1e-8
