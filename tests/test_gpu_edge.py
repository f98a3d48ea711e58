"""Edge cases of batch shape through the C ABI: empty systems inside a
batch, a single atom, two atoms, a partial crystal, and a batch with no
atoms at all.

Bars: FP64 within 1e-12 relative (max |difference| over the max magnitude)
against each system's oracle, and mixed precision force RMS within 1e-5;
20 FP64 NVE steps within 1e-9 nm on positions and 1e-8 relative on
velocities; the forces of an isolated atom exactly zero; a batch with no
atoms builds, computes, steps and reports without error."""
import numpy as np
import pytest

from mdcraft_amd import workloads as W
from mdcraft_amd.atoms import vec
from tests.helpers import engine_potential, force_rms_rel, max_rel, oracle_compute, oracle_mod, oracle_potential
from tests.test_gpu_parity import compare

pytestmark = pytest.mark.gpu

SIZES = [0, 1, 2, 256, 0, 37]


def E():
    from mdcraft_amd import engine
    return engine


def ragged(kind):
    """Systems of SIZES atoms cut from the same 4x4x4 crystal (256 atoms),
    each in the crystal's periodic box (above 2 (rns + skin) per axis)."""
    w = W.lj_argon(reps=4, rattle=0.05) if kind == "lj" else W.eam_cu(reps=4, rattle=0.05)
    at, box = w["atoms"], w["box"]
    skin = at["rns"].max() - w["pot"]["rc"]
    assert np.min(box[1] - box[0]) > 2 * (at["rns"].max() + skin)
    return [(at[:k].copy(), box) for k in SIZES], w


def engine_batch(ctx, cs, pot):
    eng = E()
    s = eng.System(ctx, [len(a) for a, _ in cs], [b for _, b in cs])
    s.upload(np.concatenate([a for a, _ in cs]))
    p = engine_potential(eng, ctx, pot)
    s.build()
    return s, p


@pytest.mark.parametrize("kind", ["eam", "lj"])
@pytest.mark.parametrize("precision", [0, 1])
def test_ragged_batch_with_empty_systems(ctx, kind, precision):
    cs, w = ragged(kind)
    print(f"ragged batch: {kind}, precision {precision}, sizes {SIZES}", flush=True)
    s, p = engine_batch(ctx, cs, w["pot"])
    s.compute(p, 7, precision)
    got = s.download()
    assert len(got) == sum(SIZES)
    off = 0
    for q, (at, box) in enumerate(cs):
        n = len(at)
        if n:
            ref, _ = oracle_compute(at, box, w["pot"])
            if n == 1:
                assert np.array_equal(vec(got[off:off + 1], "f"), np.zeros((1, 3)))
                assert np.array_equal(vec(ref, "f"), np.zeros((1, 3)))
            else:
                compare(got[off:off + n], ref, precision, f"{kind}/system{q}/{n} atoms")
        off += n
    th = s.thermo()
    assert th.shape[0] == len(SIZES)
    for q, k in enumerate(SIZES):
        if k == 0:
            assert np.all(th[q] == 0.0), th[q]


@pytest.mark.parametrize("kind", ["eam", "lj"])
def test_ragged_batch_nve_fp64(ctx, kind):
    """The batch stepped together follows each system's own oracle run."""
    R = oracle_mod()
    cs, w = ragged(kind)
    dt, steps = w["dt"], 20
    print(f"ragged NVE: {kind}, sizes {SIZES}, {steps} steps", flush=True)
    s, p = engine_batch(ctx, cs, w["pot"])
    s.compute(p, 1, 0)
    s.step(p, dt, steps, 0)
    got = s.download()
    off = 0
    for q, (at, box) in enumerate(cs):
        n = len(at)
        if n == 0:
            continue
        a = at.copy()
        dom = R.Domain(*box)
        pool = R.Pool(4)
        op = oracle_potential(R, w["pot"])
        nl = R.VerletList(pool)
        nl.update(a, dom)
        sol = R.Solver(op, pool, True)
        sol.prepare(a, dom, nl)
        sol.stages(a, dom, nl)
        sol.run_nve(a, dom, nl, dt, steps, initial_forces=False)
        L = box[1] - box[0]
        d = vec(got[off:off + n], "r") - vec(a, "r")
        d -= np.round(d / L) * L
        assert np.max(np.abs(d)) < 1e-9, (q, n)
        g, r = vec(got[off:off + n], "v"), vec(a, "v")
        assert max_rel(g, r) < 1e-8, (q, n)
        off += n


@pytest.mark.parametrize("nsys", [1, 3])
def test_batch_without_atoms(ctx, nsys):
    eng = E()
    w = W.eam_cu(reps=4)
    print(f"batch without atoms: {nsys} systems", flush=True)
    s = eng.System(ctx, [0] * nsys, [w["box"]] * nsys)
    s.upload(w["atoms"][:0])
    p = engine_potential(eng, ctx, w["pot"])
    s.build()
    s.compute(p, 7, 0)
    s.step(p, w["dt"], 5, 1)
    assert len(s.download()) == 0
    assert np.all(s.thermo() == 0.0)
    assert s.stats()["rebuilds"] >= 0


@pytest.mark.parametrize("precision", [0, 1])
def test_bpnn_ragged_batch(ctx, tmp_path, precision):
    """BPNN over the same ragged shapes: FP64 forces and energies within
    1e-12, mixed force RMS within 1e-5 and energy sums within 1e-5 of the
    sum of |E_i| (the cut crystals' per-atom energies take both signs, so
    their sum cancels: 21 of 481 for the 37-atom cut)."""
    from tests.test_gpu_bpnn import oracle as bpnn_oracle
    eng = E()
    w = W.bpnn_cu(reps=4, model_dir=str(tmp_path))
    cs = [(w["atoms"][:k].copy(), w["box"]) for k in SIZES]
    print(f"BPNN ragged batch: precision {precision}, sizes {SIZES}", flush=True)
    s = eng.System(ctx, SIZES, [w["box"]] * len(SIZES))
    s.upload(np.concatenate([a for a, _ in cs]))
    p = eng.MaiseBPNNPotential(ctx, w["pot"]["path"])
    s.build()
    s.compute(p, 7, precision)
    got = s.download()
    off = 0
    for q, (at, box) in enumerate(cs):
        n = len(at)
        if n == 0:
            continue
        ref = bpnn_oracle({"atoms": at, "box": box, "pot": w["pot"]})[0]
        f, fr = vec(got[off:off + n], "f"), vec(ref, "f")
        if n == 1:
            assert np.array_equal(f, np.zeros((1, 3))) and np.array_equal(fr, np.zeros((1, 3)))
        elif precision == 0:
            assert max_rel(f, fr) < 1e-12, (q, n, max_rel(f, fr))
        else:
            assert force_rms_rel(f, fr) < 1e-5, (q, n, force_rms_rel(f, fr))
        e, er = got["Ep"][off:off + n], ref["Ep"]
        if precision == 0:
            assert max_rel(e, er) < 1e-12, (q, n)
        else:
            assert abs(e.sum() - er.sum()) <= 1e-5 * np.abs(er).sum(), (q, n, e.sum(), er.sum())
        off += n
