"""Engine (HIP, through the C ABI) against the oracle for solver::Multi over
EAM objects (the reference's NiAl set-up) and for batches of independent
systems in one launch.

Bars: FP64 within 1e-12 relative (max |difference| over the max magnitude);
mixed precision force RMS within 1e-5 and energy sums within 1e-6.
"""
import numpy as np
import pytest

from mdcraft_amd import workloads as W
from mdcraft_amd.atoms import mat, vec
from tests.helpers import engine_potential, force_rms_rel, max_rel, oracle_compute, oracle_mod

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
MIXED_FORCE_RMS = 1e-5
MIXED_ENERGY = 1e-6


def E():
    from mdcraft_amd import engine
    return engine


def cells(indices, targets=None):
    out = []
    for q, i in enumerate(indices):
        out.append(W.nial_cell(i, None if targets is None else targets[q]))
    return out


def run_batch(ctx, cs, pot, precision, lanes=None):
    eng = E()
    s = eng.System(ctx, [len(a) for a, _ in cs], [b for _, b in cs], lanes=lanes)
    s.upload(np.concatenate([a for a, _ in cs]))
    p = engine_potential(eng, ctx, pot)
    s.build()
    s.compute(p, 7, precision)
    return s.download(), s, p


def compare(got, ref, precision, label):
    if precision == 0:
        for k in ("nc", "Em", "Ep"):
            assert max_rel(got[k], ref[k]) <= FP64_TOL, (label, k, max_rel(got[k], ref[k]))
        assert max_rel(vec(got, "f"), vec(ref, "f")) <= FP64_TOL, (label, "f")
        assert max_rel(mat(got), mat(ref)) <= FP64_TOL, (label, "V")
    else:
        assert force_rms_rel(vec(got, "f"), vec(ref, "f")) <= MIXED_FORCE_RMS, (label, "f")
        for k in ("Ep", "Em"):
            s, sr = got[k].sum(), ref[k].sum()
            assert abs(s - sr) <= MIXED_ENERGY * abs(sr), (label, k, s, sr)


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("target", [2000, 32])
def test_multi_single_cell(ctx, precision, target):
    """A bulk cell (box > 2 rns) and a 32-atom cell (one image per pair)."""
    at, box = W.nial_cell(7, target)
    pot = W.nial_pots()
    ref, _ = oracle_compute(at, box, pot)
    for lanes in (1, 8):
        got, _, _ = run_batch(ctx, [(at, box)], pot, precision, lanes)
        compare(got, ref, precision, f"nial{target}/lanes{lanes}")


@pytest.mark.parametrize("precision", [0, 1])
def test_multi_batch_matches_cells_one_by_one(ctx, precision):
    cs = cells(range(12))
    pot = W.nial_pots()
    got, s, _ = run_batch(ctx, cs, pot, precision)
    off = 0
    for q, (at, box) in enumerate(cs):
        ref, _ = oracle_compute(at, box, pot)
        compare(got[off:off + len(at)], ref, precision, f"cell{q}")
        off += len(at)
    # per-system reductions in the same launch
    th = s.thermo()
    off = 0
    for q, (at, _) in enumerate(cs):
        part = got[off:off + len(at)]
        assert th[q, 0] == pytest.approx(part["Ep"].sum(), rel=1e-12)
        assert th[q, 1] == pytest.approx(part["Em"].sum(), rel=1e-12)
        off += len(at)


def _oracle_solver(R, pot, at, box, threads=4):
    from tests.helpers import oracle_potential
    a = at.copy()
    dom = R.Domain(*box)
    pool = R.Pool(threads)
    op = oracle_potential(R, pot)
    nl = R.VerletList(pool)
    nl.update(a, dom)
    return a, dom, nl, R.Solver(op, pool, True), (pool, op)


def _engine_batch(ctx, cs, pot):
    eng = E()
    s = eng.System(ctx, [len(a) for a, _ in cs], [b for _, b in cs])
    s.upload(np.concatenate([a for a, _ in cs]))
    p = engine_potential(eng, ctx, pot)
    s.build()
    return s, p


def test_batch_nve_fp64(ctx):
    """A batch stepped together, with skin-triggered rebuilds, follows each
    system's own oracle run.  Every box exceeds 2 (rns + skin) on each axis,
    where one image per pair is fixed between rebuilds and the rebuild step
    does not change the result."""
    R = oracle_mod()
    cs = cells(range(6), targets=[128, 250, 500, 1000, 432, 2048])
    for _, box in cs:
        assert np.min(box[1] - box[0]) > 2 * (W.EAM_RC + 0.6) * W.ANG
    pot = W.nial_pots()
    dt, steps = 1 * W.FS, 30
    s, p = _engine_batch(ctx, cs, pot)
    s.compute(p, 1, 0)
    s.step(p, dt, steps, 0)
    got = s.download()
    off = 0
    for q, (at, box) in enumerate(cs):
        a, dom, nl, sol, keep = _oracle_solver(R, pot, at, box)
        sol.prepare(a, dom, nl)
        sol.stages(a, dom, nl)
        sol.run_nve(a, dom, nl, dt, steps, initial_forces=False)
        L = box[1] - box[0]
        d = vec(got[off:off + len(a)], "r") - vec(a, "r")
        d -= np.round(d / L) * L
        assert np.max(np.abs(d)) < 1e-9, q
        assert max_rel(vec(got[off:off + len(a)], "v"), vec(a, "v")) < 1e-8, q
        off += len(a)


def test_batch_small_boxes_rebuild_every_step_fp64(ctx):
    """Boxes shorter than 2 rns (one image per pair, chosen from the wrapped
    positions) stepped with a rebuild after every step, as the reference's
    run_md_steps harness does (md_test_utils.h:255-279)."""
    R = oracle_mod()
    cs = cells(range(4), targets=[32, 64, 32, 48])
    pot = W.nial_pots()
    dt, steps = 1 * W.FS, 8
    s, p = _engine_batch(ctx, cs, pot)
    s.compute(p, 7, 0)
    hist = [s.thermo()]
    for _ in range(steps):
        s.step(p, dt, 1, 0)
        s.build()
        s.compute(p, 7, 0)
        hist.append(s.thermo())
    hist = np.array(hist)
    got = s.download()
    off = 0
    for q, (at, box) in enumerate(cs):
        a, dom, nl, sol, keep = _oracle_solver(R, pot, at, box, threads=1)
        ref = sol.run_md_steps(a, dom, nl, dt, steps, 1)
        assert np.max(np.abs(vec(got[off:off + len(a)], "r") - vec(a, "r"))) < 1e-12, q
        for k in (0, 1, 2, 4):
            assert np.allclose(hist[:, q, k], ref[:, k], rtol=1e-11, atol=1e-9), (q, k)
        off += len(a)


def _pair_member_spec(cross):
    ni_al = W.nial_pots()
    like = {(i, j): q for i, j, q in ni_al["pairs"] if i == j}
    return {"kind": "eam_multi", "nkinds": 2, "pairs": [(0, 0, like[(0, 0)]), (1, 1, like[(1, 1)]), (0, 1, cross)]}


def _write_pair_file(path, rphi, dr, rc):
    """The reference's Pair file layout (pair.cxx:36-56): four lines, then
    "Nr dr rcut" in Angstrom and Nr values of r*phi in eV*Angstrom."""
    Na, Qe = 6.02214076e+23, 1.602176634e-19
    UnEnrg = (1e-3 / Na) * (1e-9 / 1e-12) * (1e-9 / 1e-12)
    conv = Qe / (10 * UnEnrg)
    vals = np.asarray(rphi) / conv
    with open(path, "w") as f:
        f.write(" synthetic cross pair\n line 2\n line 3\n 2 Ni Al\n")
        f.write("%d %r %r\n" % (len(vals), dr * 10, rc * 10))
        for v in vals:
            f.write("%r\n" % float(v))
    back = np.array([float("%r" % float(v)) for v in vals]) * conv
    return back


@pytest.mark.parametrize("precision", [0, 1])
def test_multi_with_tabulated_pair_member(ctx, precision, tmp_path):
    """problems/NiAl/test-NiAl.py builds Multi from EAM objects for like
    kinds and a tabulated Pair read from a file for the cross kinds."""
    t = W.eam_cu_tables()
    path = str(tmp_path / "NiAlpair.txt")
    rphi = _write_pair_file(path, 0.9 * t["rphi"], t["dr"], t["rc"])
    oracle_spec = _pair_member_spec({"kind": "pair", "rc": t["rc"], "rphi": rphi, "dr": t["dr"]})
    eng = E()
    cross = eng.Pair(ctx, path=path)
    assert cross.rcut == pytest.approx(t["rc"], rel=1e-15)
    for idx, target in ((3, 500), (8, 32)):
        at, box = W.nial_cell(idx, target)
        ref, _ = oracle_compute(at, box, oracle_spec)
        like = {(i, j): engine_potential(eng, ctx, q) for i, j, q in oracle_spec["pairs"] if i == j}
        p = eng.Multi(ctx, 2, [(0, 0, like[(0, 0)]), (1, 1, like[(1, 1)]), (0, 1, cross)])
        s = eng.System(ctx, [len(at)], [box])
        s.upload(at)
        s.build()
        s.compute(p, 7, precision)
        compare(s.download(), ref, precision, f"pair-member cell{idx}")


def test_all_512_cells_of_the_batched_workload(ctx):
    """The batched NiAl workload as benchmarked (512 cells, 32 to 4096
    atoms, one Multi potential, one launch per pass for all cells), every
    cell against its own oracle evaluation: FP64 within 1e-12 and mixed
    precision force RMS within 1e-5 per cell."""
    nb = W.nial_batch(512)
    cs, pot = nb["cells"], nb["pot"]
    assert len(cs) == 512
    print("512 cells: engine", flush=True)
    got = {prec: run_batch(ctx, cs, pot, prec)[0] for prec in (0, 1)}
    print("512 cells: oracle", flush=True)
    off = 0
    for q, (at, box) in enumerate(cs):
        ref, _ = oracle_compute(at, box, pot, threads=4)
        for prec in (0, 1):
            compare(got[prec][off:off + len(at)], ref, prec, f"cell{q}/{len(at)}")
        off += len(at)
    assert off == sum(len(a) for a, _ in cs)
