"""LJs and the tabulated Pair on the device against the oracle.

Bars: FP64 within 1e-12 relative (max |difference| over the max
magnitude); mixed precision force RMS within 1e-5 and energy sums within
1e-6; a 30-step FP64 trajectory within 1e-9 nm."""
import numpy as np
import pytest

from mdcraft_amd import workloads as W
from mdcraft_amd.atoms import mat, vec
from tests.helpers import engine_potential, force_rms_rel, max_rel, oracle_compute, oracle_mod, oracle_potential

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
MIXED_FORCE_RMS = 1e-5
MIXED_ENERGY = 1e-6


def E():
    from mdcraft_amd import engine
    return engine


def workload(kind):
    if kind == "ljs":
        w = W.lj_argon(reps=5, rattle=0.1)
        sigma = w["pot"]["s"]
        w["pot"] = {"kind": "ljs", "aVr": w["pot"]["a"], "rVr": sigma, "rc": 2.5 * sigma}
        return w
    w = W.eam_cu(reps=5, rattle=0.05)
    t = w["pot"]
    w["pot"] = {"kind": "pair", "rc": t["rc"], "rphi": t["rphi"], "dr": t["dr"]}
    return w


def run(ctx, w, precision, lanes):
    eng = E()
    s = eng.System(ctx, [len(w["atoms"])], [w["box"]], lanes=lanes)
    s.upload(w["atoms"])
    p = engine_potential(eng, ctx, w["pot"])
    s.build()
    s.compute(p, 7, precision)
    return s.download(), s, p


@pytest.mark.parametrize("kind", ["ljs", "pair"])
@pytest.mark.parametrize("precision", [0, 1])
def test_single_point(ctx, kind, precision):
    w = workload(kind)
    ref, _ = oracle_compute(w["atoms"], w["box"], w["pot"])
    for lanes in (1, 8, 16):
        got, _, _ = run(ctx, w, precision, lanes)
        if precision == 0:
            assert max_rel(vec(got, "f"), vec(ref, "f")) <= FP64_TOL, (kind, lanes)
            assert max_rel(got["Ep"], ref["Ep"]) <= FP64_TOL, (kind, lanes)
            assert max_rel(mat(got), mat(ref)) <= FP64_TOL, (kind, lanes)
        else:
            assert force_rms_rel(vec(got, "f"), vec(ref, "f")) <= MIXED_FORCE_RMS, (kind, lanes)
            assert abs(got["Ep"].sum() - ref["Ep"].sum()) <= MIXED_ENERGY * abs(ref["Ep"].sum()), (kind, lanes)


@pytest.mark.parametrize("kind", ["ljs", "pair"])
def test_nve_fp64(ctx, kind):
    w = workload(kind)
    dt, steps = w["dt"], 30
    got_at, s, p = run(ctx, w, 0, None)
    s.step(p, dt, steps, 0)
    got = s.download()
    R = oracle_mod()
    a = w["atoms"].copy()
    dom = R.Domain(*w["box"])
    pool = R.Pool(4)
    op = oracle_potential(R, w["pot"])
    nl = R.VerletList(pool)
    nl.update(a, dom)
    sol = R.Solver(op, pool, True)
    sol.prepare(a, dom, nl)
    sol.stages(a, dom, nl)
    sol.run_nve(a, dom, nl, dt, steps, initial_forces=False)
    L = w["box"][1] - w["box"][0]
    d = vec(got, "r") - vec(a, "r")
    d -= np.round(d / L) * L
    assert np.max(np.abs(d)) < 1e-9
    assert max_rel(vec(got, "v"), vec(a, "v")) < 1e-8


def test_checkpoint_and_resume(ctx, tmp_path):
    """A resident run saved to a snapshot after 15 steps and resumed in a
    new System reaches the uninterrupted run's state after 30 (FP64; the
    resumed run rebuilds its list at once, which moves rebuild steps but
    not the physics of boxes above 2 (rns + skin))."""
    from mdcraft_amd import snapshot
    eng = E()
    w = W.eam_cu(reps=5, rattle=0.05)
    p = engine_potential(eng, ctx, w["pot"])
    ref = eng.System(ctx, [len(w["atoms"])], [w["box"]])
    ref.upload(w["atoms"])
    ref.build()
    ref.compute(p, 1, 0)
    ref.step(p, w["dt"], 15, 0)
    path = str(tmp_path / "run000015.h5")
    snapshot.save_system(path, ref, [w["box"]], cycles=np.array([15]))
    ref.step(p, w["dt"], 15, 0)
    want = ref.download()
    s, atoms, boxes, extra = snapshot.load_system(path, eng, ctx)
    assert int(extra["cycles"][0]) == 15 and len(boxes) == 1
    s.build()
    s.step(p, w["dt"], 15, 0)
    got = s.download()
    L = w["box"][1] - w["box"][0]
    d = vec(got, "r") - vec(want, "r")
    d -= np.round(d / L) * L
    assert np.max(np.abs(d)) < 1e-9
    assert max_rel(vec(got, "v"), vec(want, "v")) < 1e-8


def test_snapshot_resume_matches_oracle_run(ctx, tmp_path):
    """The checkpoint layout of the reference's scripts
    (test/python/solver/test-EAM.py:193-197: the atom records and the domain
    in one file) written after 15 engine steps holds the oracle's state
    after 15 steps, and a new System resumed from it reaches the oracle's
    uninterrupted 30-step state (FP64: positions within 1e-9 nm,
    velocities within 1e-8 relative)."""
    from mdcraft_amd import snapshot
    from oracle import refcpu as R
    from tests.helpers import oracle_potential
    eng = E()
    w = W.eam_cu(reps=5, rattle=0.05)
    L = w["box"][1] - w["box"][0]
    p = engine_potential(eng, ctx, w["pot"])
    s0 = eng.System(ctx, [len(w["atoms"])], [w["box"]])
    s0.upload(w["atoms"])
    s0.build()
    s0.compute(p, 1, 0)
    s0.step(p, w["dt"], 15, 0)
    path = str(tmp_path / "run000015.h5")
    snapshot.save_system(path, s0, [w["box"]], cycles=np.array([15]))
    s0.close()
    # the oracle's make_step loop from the same start
    a = w["atoms"].copy()
    dom = R.Domain(*w["box"])
    pool = R.Pool(8)
    rp = oracle_potential(R, w["pot"])
    a["rcut"] = rp.rcut
    nl = R.VerletList(pool)
    nl.update(a, dom)
    sol = R.Solver(rp, pool, True)
    sol.run_nve(a, dom, nl, w["dt"], 15, initial_forces=True)
    with snapshot.File(path, "r") as f:
        saved = f["atoms"]
    d = vec(saved, "r") - vec(a, "r")
    d -= np.round(d / L) * L
    assert np.max(np.abs(d)) < 1e-9
    assert max_rel(vec(saved, "v"), vec(a, "v")) < 1e-8
    sol.run_nve(a, dom, nl, w["dt"], 15, initial_forces=False)
    s, atoms, boxes, extra = snapshot.load_system(path, eng, ctx)
    assert int(extra["cycles"][0]) == 15
    s.build()
    s.compute(p, 1, 0)
    s.step(p, w["dt"], 15, 0)
    got = s.download()
    d = vec(got, "r") - vec(a, "r")
    d -= np.round(d / L) * L
    assert np.max(np.abs(d)) < 1e-9
    assert max_rel(vec(got, "v"), vec(a, "v")) < 1e-8
