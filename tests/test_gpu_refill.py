"""Box slots of a system refilled while it runs (mdc_sys_refill_box), on
the GPU against the oracle.

A refill leaves the box's spare records void; the sort puts every void
record after all live ones and the kernels run over the live records only.
These tests check that every live box still gives what the oracle gives for
its structure alone, through the paths that move records between live and
void:
* boxes refilled with fewer atoms than their records, and emptied, then a
  single point (FP64 1e-12) and NVE steps with skin-triggered rebuilds of
  the live records only (positions 1e-9 nm, velocities 1e-8 relative);
* a box refilled while the others are part-way through their runs;
* an upload of every record over a system holding void records;
* a BPNN system with an emptied box and a refilled one (FP64 1e-12, virial
  included).
Boxes exceed 2 (rns + skin) per axis, where one image per pair is fixed
between rebuilds, so rebuild timing does not change the physics."""
import numpy as np
import pytest

from mdcraft_amd import workloads as W
from mdcraft_amd.atoms import mat, vec
from tests.helpers import engine_potential, max_rel, oracle_compute, oracle_mod, oracle_potential

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


def E():
    from mdcraft_amd import engine
    return engine


def structure(kind, n, seed, model=None):
    at, box, pot, dt = W.served_structure(kind, n, W.structure_seed(17, seed), False, model)
    assert len(at) == n
    return at, box, pot, dt


def check_single_point(got, at, box, pot, label):
    ref, _ = oracle_compute(got, box, pot, threads=8)
    assert max_rel(vec(got, "f"), vec(ref, "f")) <= FP64_TOL, (label, "f")
    assert max_rel(got["Ep"], ref["Ep"]) <= FP64_TOL, (label, "Ep")
    assert max_rel(mat(got), mat(ref)) <= FP64_TOL, (label, "V")
    return ref


class OracleRun:
    """One box's uninterrupted FP64 NVE run on the oracle."""

    def __init__(self, at, box, pot):
        R = oracle_mod()
        self.a = at.copy()
        self.dom = R.Domain(*box)
        self.pool = R.Pool(8)
        self.pot = oracle_potential(R, pot)
        self.a["rcut"] = self.pot.rcut
        self.nl = R.VerletList(self.pool)
        self.nl.update(self.a, self.dom)
        self.sol = R.Solver(self.pot, self.pool, True)
        self.sol.prepare(self.a, self.dom, self.nl)
        self.sol.stages(self.a, self.dom, self.nl)
        self.box = box

    def run(self, dt, steps):
        self.sol.run_nve(self.a, self.dom, self.nl, dt, steps, initial_forces=False)

    def check(self, got, label):
        L = self.box[1] - self.box[0]
        d = vec(got, "r") - vec(self.a, "r")
        d -= np.round(d / L) * L
        assert np.max(np.abs(d)) < 1e-9, label
        assert max_rel(vec(got, "v"), vec(self.a, "v")) < 1e-8, label


def test_refill_empty_and_step_fp64(ctx):
    eng = E()
    caps = [500, 864, 500, 256]
    first = [structure("eam", n, q) for q, n in enumerate(caps)]
    pot = first[0][2]
    dt = first[0][3]
    s = eng.System(ctx, caps, [b for _, b, _, _ in first])
    s.upload(np.concatenate([a for a, _, _, _ in first]))
    p = engine_potential(eng, ctx, pot)
    off = np.concatenate([[0], np.cumsum(caps)])

    # box 1 (864 records) takes a 256-atom crystal, box 2 is emptied
    new1 = structure("eam", 256, 11)
    s.refill_box(1, new1[0], new1[1])
    s.refill_box(2, np.zeros(0, dtype=first[0][0].dtype), first[2][1])
    s.build()
    s.compute(p, eng.FORCE | eng.ENERGY | eng.VIRIAL, eng.FP64)
    th = s.thermo()
    live = {0: first[0], 1: new1, 3: first[3]}
    for q, (at, box, _, _) in live.items():
        got = s.download_range(int(off[q]), len(at))
        ref = check_single_point(got, at, box, pot, q)
        assert abs(th[q, 0] - ref["Ep"].sum()) <= FP64_TOL * abs(ref["Ep"].sum()), q
    assert np.all(th[2, :5] == 0.0)

    # NVE with rebuilds of the live records only, against each box's run
    runs = {q: OracleRun(at, box, pot) for q, (at, box, _, _) in live.items()}
    s.step(p, dt, 25, eng.FP64)
    assert s.stats()["rebuilds"] >= 2
    for q, run in runs.items():
        run.run(dt, 25)
        run.check(s.download_range(int(off[q]), len(live[q][0])), ("25 steps", q))

    # box 2 takes a crystal while the others are part-way through
    new2 = structure("eam", 500, 12)
    s.refill_box(2, new2[0], new2[1])
    s.build()
    s.compute(p, eng.FORCE | eng.ENERGY | eng.VIRIAL, eng.FP64)
    check_single_point(s.download_range(int(off[2]), 500), new2[0], new2[1], pot, "refill 2")
    runs[2] = OracleRun(new2[0], new2[1], pot)
    live[2] = new2
    s.step(p, dt, 15, eng.FP64)
    for q, run in runs.items():
        run.run(dt, 15)
        run.check(s.download_range(int(off[q]), len(live[q][0])), ("15 more steps", q))


def test_upload_over_void_records_fp64(ctx):
    """Records of an emptied box become live again through a full upload
    (the sort then takes all records again)."""
    eng = E()
    caps = [500, 864]
    cs = [structure("eam", n, 20 + q) for q, n in enumerate(caps)]
    pot = cs[0][2]
    s = eng.System(ctx, caps, [b for _, b, _, _ in cs])
    allrec = np.concatenate([a for a, _, _, _ in cs])
    s.upload(allrec)
    p = engine_potential(eng, ctx, pot)
    s.build()
    s.compute(p, eng.FORCE, eng.FP64)
    s.refill_box(1, np.zeros(0, dtype=allrec.dtype), cs[1][1])
    s.build()
    s.compute(p, eng.FORCE | eng.ENERGY | eng.VIRIAL, eng.FP64)
    check_single_point(s.download_range(0, 500), cs[0][0], cs[0][1], pot, "box 0 beside an empty box")
    s.upload(allrec)
    s.build()
    s.compute(p, eng.FORCE | eng.ENERGY | eng.VIRIAL, eng.FP64)
    got = s.download()
    check_single_point(got[:500], cs[0][0], cs[0][1], pot, "box 0 after upload")
    check_single_point(got[500:], cs[1][0], cs[1][1], pot, "box 1 after upload")


def test_bpnn_refill_fp64(ctx, tmp_path):
    eng = E()
    model = W.write_maise_model(str(tmp_path / "cu.maise"))
    caps = [256, 500, 256]
    cs = [structure("bpnn", n, 30 + q, model) for q, n in enumerate(caps)]
    pot = cs[0][2]
    s = eng.System(ctx, caps, [b for _, b, _, _ in cs])
    s.upload(np.concatenate([a for a, _, _, _ in cs]))
    p = engine_potential(eng, ctx, pot)
    new1 = structure("bpnn", 256, 40, model)
    s.refill_box(0, np.zeros(0, dtype=cs[0][0].dtype), cs[0][1])
    s.refill_box(1, new1[0], new1[1])
    s.build()
    s.compute(p, eng.FORCE | eng.ENERGY | eng.VIRIAL, eng.FP64)
    check_single_point(s.download_range(256, 256), new1[0], new1[1], pot, "refilled box")
    check_single_point(s.download_range(756, 256), cs[2][0], cs[2][1], pot, "untouched box")
