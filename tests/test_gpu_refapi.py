"""The reference's Python API (mdcraft_amd.refapi, importable as ``mdcraft``
from mdcraft_amd/compat) driving the engine, against the oracle driven the
same way.  The loops follow the reference's own scripts:
test/python/solver/test-EAM.py (Single + Periodic + Verlet, prepare /
forces / virials without a charge call first, list update every NCbuff
steps, virials every NCout steps), problems/Cu-bpnn/run.py (Ternary) and
the non-periodic argon ball of test/python/solver/run_ball_ar.py.

Bars: FP64 within 1e-12 relative (max |difference| over the max magnitude)
for single calls; after tens of steps 1e-9 nm on positions and 1e-8
relative on velocities and forces (FP64 sums in a different order, carried
through the dynamics)."""
import os
import sys

import numpy as np
import pytest

from mdcraft_amd import workloads as W
from mdcraft_amd.atoms import mat, vec
from tests.helpers import max_rel, oracle_mod, oracle_potential, pair_set

pytestmark = pytest.mark.gpu

CALL_TOL = 1e-12
RUN_R_ABS = 1e-9
RUN_REL = 1e-8

COMPAT = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mdcraft_amd", "compat")


def api():
    if COMPAT not in sys.path:
        sys.path.insert(0, COMPAT)
    import mdcraft.data
    import mdcraft.lattice
    import mdcraft.neibs
    import mdcraft.solver
    import mdcraft.solver.boundary
    import mdcraft.solver.potential
    import mdcraft.solver.stepper
    import mdcraft.tools
    return mdcraft


def make_domain(m, box):
    lo, hi, per = box
    d = m.lattice.Domain(lo[0], hi[0], lo[1], hi[1], lo[2], hi[2])
    for a in range(3):
        d.set_periodic(a, bool(per[a]))
    return d


def api_potential(m, pot, domain):
    P = m.solver.potential
    if pot["kind"] == "lj":
        return P.LJsCut(a=pot["a"], s=pot["s"], rcut=pot["rc"])
    if pot["kind"] == "eam":
        return P.EAM(Rcutoff=pot["rc"], pairPotential=pot["rphi"], chargeDensity=pot["rho"],
                     embeddingEnergy=pot["F"], NrPot=len(pot["rho"]), drPot=pot["dr"], NrEmb=len(pot["F"]),
                     drEmb=pot["drho"])
    if pot["kind"] == "bpnn":
        return P.MaiseBPNNPotential(pot["path"], domain)
    raise ValueError(pot["kind"])


def compare(got, ref, tol, label, fields=("f", "Ep", "Em", "nc", "V")):
    for k in fields:
        if k in ("r", "v", "f"):
            assert max_rel(vec(got, k), vec(ref, k)) <= tol, (label, k, max_rel(vec(got, k), vec(ref, k)))
        elif k == "V":
            assert max_rel(mat(got), mat(ref)) <= tol, (label, k)
        else:
            assert max_rel(got[k], ref[k]) <= tol, (label, k, max_rel(got[k], ref[k]))


class OracleScript:
    """The same script on the oracle."""

    def __init__(self, at, box, pot, periodic_boundary=True):
        R = oracle_mod()
        self.R = R
        self.a = at.copy()
        self.dom = R.Domain(*box)
        self.pool = R.Pool(1)
        self.pot = oracle_potential(R, pot, box)
        self.nl = R.VerletList(self.pool)
        self.sol = R.Solver(self.pot, self.pool, periodic_boundary)

    def update(self, kbuf=-1.0):
        self.nl.update(self.a, self.dom, kbuf)

    def __getattr__(self, name):
        fn = getattr(self.sol, name)
        return lambda *extra: fn(self.a, self.dom, self.nl, *extra)


def test_eam_script_matches_oracle():
    """test-EAM.py: prepare/forces/virials, then make_step with an update
    every 10 steps and virials every 5."""
    m = api()
    w = W.eam_cu(reps=5, rattle=0.05)
    at, box, pot, dt = w["atoms"], w["box"], w["pot"], w["dt"]
    at["rns"] = 1.25 * at["rcut"]
    atoms = m.data.Atoms(len(at))
    atoms.data[:] = at
    domain = make_domain(m, box)
    potential = api_potential(m, pot, domain)
    assert potential.rcut == pytest.approx(pot["rc"])
    periodic = m.solver.boundary.Periodic(domain)
    solver = m.solver.Single(potential=potential, boundary=periodic)
    stepper = m.solver.stepper.Verlet(solver=solver)
    nlist = m.neibs.VerletList(atoms=atoms, neighbors=atoms, domain=domain)
    nlist.update()
    o = OracleScript(at, box, pot)
    o.update()
    compare(atoms.data, o.a, 0.0, "wrap", fields=("r",))
    assert pair_set(*nlist.system.export()) == pair_set(*o.nl.export(len(at)))
    solver.prepare(atoms, atoms, nlist.get())
    o.prepare()
    compare(atoms.data, o.a, 0.0, "prepare", fields=("f", "Ep", "Em", "nc", "V", "rcut", "bc"))
    # forces without a charge call: the density is the zero prepare left
    solver.forces(atoms, atoms, nlist.get())
    o.forces()
    compare(atoms.data, o.a, CALL_TOL, "forces", fields=("f", "nc"))
    solver.virials(atoms, atoms, nlist.get())
    o.virials()
    compare(atoms.data, o.a, CALL_TOL, "virials", fields=("f", "Ep", "Em", "V"))
    for step in range(1, 31):
        stepper.make_step(atoms, nlist.get(), dt)
        o.make_step(dt)
        if step % 5 == 0:
            solver.virials(atoms, atoms, nlist.get())
            o.virials()
        if step % 10 == 0:
            assert nlist.decide() == o.nl.decide(o.a, o.dom)
            nlist.update()
            o.update()
    d = vec(atoms.data, "r") - vec(o.a, "r")
    assert np.max(np.abs(d)) < RUN_R_ABS
    compare(atoms.data, o.a, RUN_REL, "run", fields=("v", "f", "Ep", "Em", "nc", "V"))


def test_stage_calls_one_at_a_time():
    """charge then forces equals the stages of a step; virials adds to Ep
    and V and leaves f; forces adds to f; atoms with tag == 0 are left."""
    m = api()
    w = W.eam_cu(reps=4, rattle=0.05)
    at, box, pot = w["atoms"], w["box"], w["pot"]
    at["tag"][::7] = 0
    at["f"]["x"] = 3.0  # tag == 0 atoms keep this through prepare
    atoms = m.data.Atoms(len(at))
    atoms.data[:] = at
    domain = make_domain(m, box)
    solver = m.solver.Single(api_potential(m, pot, domain), m.solver.boundary.Periodic(domain))
    nlist = m.neibs.VerletList(atoms, atoms, domain)
    nlist.update()
    o = OracleScript(at, box, pot)
    o.update()
    seq = [("prepare", 0.0), ("charge", CALL_TOL), ("forces", CALL_TOL), ("virials", CALL_TOL),
           ("virials", CALL_TOL), ("forces", CALL_TOL)]
    for name, tol in seq:
        getattr(solver, name)(atoms, atoms, nlist.get())
        getattr(o, name)()
        compare(atoms.data, o.a, tol, name, fields=("f", "Ep", "Em", "nc", "V"))


def test_update_kbuf_and_decide():
    """update(kbuf > 0) rescales rns over the current list
    (verlet-list.cxx:71-92), with per-atom rcut; decide() follows the
    oracle as atoms move."""
    m = api()
    w = W.lj_argon(reps=4, rattle=0.1)
    at, box, pot = w["atoms"], w["box"], w["pot"]
    rng = np.random.default_rng(5)
    at["rcut"] = pot["rc"] * (1.0 + 0.05 * rng.random(len(at)))
    atoms = m.data.Atoms(len(at))
    atoms.data[:] = at
    domain = make_domain(m, box)
    nlist = m.neibs.VerletList(atoms, atoms, domain)
    o = OracleScript(at, box, pot)
    # the first update with kbuf: the lists are still empty, rns becomes 0
    nlist.update(1.1)
    o.update(1.1)
    assert np.array_equal(atoms.data["rns"], o.a["rns"])
    atoms.data["rns"] = o.a["rns"] = at["rns"]
    nlist.update()
    o.update()
    nlist.update(1.1)
    o.update(1.1)
    assert np.array_equal(atoms.data["rns"], o.a["rns"])
    assert pair_set(*nlist.system.export()) == pair_set(*o.nl.export(len(at)))
    assert nlist.decide() is False and o.nl.decide(o.a, o.dom) is False
    skin = float(np.min(o.a["rns"] - o.a["rcut"]))
    for shift, want in ((0.49 * skin, False), (0.51 * skin, True)):
        atoms.data["r"]["x"][3] += shift
        o.a["r"]["x"][3] += shift
        # uniform skin along the moved atom: the box's smallest skin is its own
        assert o.nl.decide(o.a, o.dom) == nlist.decide()
        atoms.data["r"]["x"][3] -= shift
        o.a["r"]["x"][3] -= shift


def test_lists_full_and_half():
    m = api()
    w = W.lj_argon(reps=4, rattle=0.1)
    at, box = w["atoms"], w["box"]
    atoms = m.data.Atoms(len(at))
    atoms.data[:] = at
    domain = make_domain(m, box)
    R = oracle_mod()
    for half in (False, True):
        nlist = m.neibs.VerletList(atoms, atoms, domain, half=half)
        assert nlist.half is half
        nlist.update()
        a = at.copy()
        nl = R.VerletList(R.Pool(1), half=half)
        nl.update(a, R.Domain(*box))
        ref = nl.lists(len(a))
        lists = nlist.get()
        assert len(lists) == len(a)
        for i in range(len(a)):
            assert sorted(lists[i].tolist()) == sorted(ref[i].tolist()), (half, i)


def test_argon_ball_without_boundary():
    """A non-periodic LJ cluster with the null boundary, 20 steps with an
    update every 5 (run_ball_ar.py without MPI)."""
    m = api()
    w = W.lj_argon(reps=4, rattle=0.1)
    at, pot, dt = w["atoms"], w["pot"], w["dt"]
    L = w["box"][1][0]
    box = (np.full(3, -0.5 * L), np.full(3, 1.5 * L), (False, False, False))
    atoms = m.data.Atoms(len(at))
    atoms.data[:] = at
    domain = make_domain(m, box)
    solver = m.solver.Single(api_potential(m, pot, domain))
    stepper = m.solver.stepper.Verlet(solver)
    nlist = m.neibs.VerletList(atoms, atoms, domain)
    nlist.update()
    o = OracleScript(at, box, pot, periodic_boundary=False)
    o.update()
    solver.prepare(atoms, atoms, nlist.get())
    solver.forces(atoms, atoms, nlist.get())
    o.prepare()
    o.forces()
    compare(atoms.data, o.a, CALL_TOL, "forces", fields=("f",))
    for step in range(1, 21):
        stepper.make_step(atoms, nlist.get(), dt)
        o.make_step(dt)
        if step % 5 == 0:
            nlist.update()
            o.update()
    solver.virials(atoms, atoms, nlist.get())
    o.virials()
    assert np.max(np.abs(vec(atoms.data, "r") - vec(o.a, "r"))) < RUN_R_ABS
    compare(atoms.data, o.a, RUN_REL, "run", fields=("v", "f", "Ep", "V"))


def test_bpnn_script_matches_oracle(tmp_path):
    """Cu-bpnn/run.py: Ternary prepare/forces/virials, then make_step with
    an update every 5 steps."""
    m = api()
    w = W.bpnn_cu(reps=4, model_dir=str(tmp_path))
    at, box, pot, dt = w["atoms"], w["box"], w["pot"], w["dt"]
    atoms = m.data.Atoms(len(at))
    atoms.data[:] = at
    domain = make_domain(m, box)
    potential = api_potential(m, pot, domain)
    atoms.data["rcut"] = potential.rcut
    atoms.data["rns"] = potential.rcut + 0.03
    solver = m.solver.Ternary(potential)
    stepper = m.solver.stepper.Verlet(solver)
    nlist = m.neibs.VerletList(atoms, atoms, domain=domain)
    nlist.update()
    at2 = atoms.data.copy()
    o = OracleScript(at2, box, pot)
    o.update()
    for name in ("prepare", "forces", "virials"):
        getattr(solver, name)(atoms, atoms, nlist.get())
        getattr(o, name)()
    compare(atoms.data, o.a, CALL_TOL, "initial", fields=("f", "Ep"))
    for step in range(1, 11):
        stepper.make_step(atoms, nlist.get(), dt)
        o.make_step(dt)
        if step % 5 == 0:
            nlist.update()
            o.update()
    assert np.max(np.abs(vec(atoms.data, "r") - vec(o.a, "r"))) < RUN_R_ABS
    compare(atoms.data, o.a, RUN_REL, "run", fields=("v", "f", "Ep"))


def test_multi_nial_stages():
    """solver::Multi with add_potential over the NiAl objects, stage by
    stage against the oracle."""
    m = api()
    at, box = W.nial_cell(3, 500)
    spec = W.nial_pots()
    atoms = m.data.Atoms(len(at))
    atoms.data[:] = at
    domain = make_domain(m, box)
    solver = m.solver.Multi(boundary=m.solver.boundary.Periodic(domain), nkinds=spec["nkinds"])
    for i, j, q in spec["pairs"]:
        solver.add_potential(i, j, api_potential(m, q, domain))
    nlist = m.neibs.VerletList(atoms, atoms, domain)
    nlist.update()
    o = OracleScript(at, box, spec)
    o.update()
    for name in ("prepare", "charge", "forces", "virials"):
        getattr(solver, name)(atoms, atoms, nlist.get())
        getattr(o, name)()
        compare(atoms.data, o.a, CALL_TOL, name, fields=("f", "Ep", "Em", "nc", "V"))


@pytest.mark.parametrize("kind", ["ljs", "pair"])
def test_ljs_and_tabulated_pair(kind):
    """LJs(aVr, rVr, Rcutoff) and Pair(Rcutoff, pairPotential, NrPot,
    drPot) through mdcraft.solver.potential, one prepare/forces/virials."""
    m = api()
    if kind == "ljs":
        w = W.lj_argon(reps=4, rattle=0.1)
        s = w["pot"]["s"]
        spec = {"kind": "ljs", "aVr": w["pot"]["a"], "rVr": s, "rc": 2.5 * s}
        potential = m.solver.potential.LJs(aVr=spec["aVr"], rVr=s, Rcutoff=spec["rc"])
    else:
        w = W.eam_cu(reps=4, rattle=0.05)
        t = w["pot"]
        spec = {"kind": "pair", "rc": t["rc"], "rphi": t["rphi"], "dr": t["dr"]}
        potential = m.solver.potential.Pair(Rcutoff=t["rc"], pairPotential=t["rphi"], NrPot=len(t["rphi"]),
                                            drPot=t["dr"])
    at, box = w["atoms"], w["box"]
    atoms = m.data.Atoms(len(at))
    atoms.data[:] = at
    domain = make_domain(m, box)
    solver = m.solver.Single(potential, m.solver.boundary.Periodic(domain))
    nlist = m.neibs.VerletList(atoms, atoms, domain)
    nlist.update()
    o = OracleScript(at, box, spec)
    o.update()
    for name in ("prepare", "forces", "virials"):
        getattr(solver, name)(atoms, atoms, nlist.get())
        getattr(o, name)()
        compare(atoms.data, o.a, CALL_TOL, name, fields=("f", "Ep", "V"))


def test_alloy_script_matches_oracle(tmp_path):
    """test-EAM.py's loop with the potential read from a three-element setfl
    file (EAMAlloy) on a NiAl cell with copper on some aluminium sites."""
    m = api()
    from tests.test_gpu_alloy import with_copper
    path = str(tmp_path / "NiAlCu.eam.alloy")
    W.write_setfl(path, ["Ni", "Al", "Cu"])
    at, box = W.nial_cell(4, 500)
    at = with_copper(at, 4)
    dt = 1 * W.FS
    atoms = m.data.Atoms(len(at))
    atoms.data[:] = at
    domain = make_domain(m, box)
    potential = m.solver.potential.EAMAlloy(path)
    solver = m.solver.Single(potential=potential, boundary=m.solver.boundary.Periodic(domain))
    stepper = m.solver.stepper.Verlet(solver=solver)
    nlist = m.neibs.VerletList(atoms=atoms, neighbors=atoms, domain=domain)
    nlist.update()
    o = OracleScript(at, box, {"kind": "alloy", "path": path})
    o.update()
    solver.prepare(atoms, atoms, nlist.get())
    o.prepare()
    solver.charge(atoms, atoms, nlist.get())
    o.charge()
    solver.forces(atoms, atoms, nlist.get())
    o.forces()
    solver.virials(atoms, atoms, nlist.get())
    o.virials()
    compare(atoms.data, o.a, CALL_TOL, "virials")
    for step in range(1, 21):
        stepper.make_step(atoms, nlist.get(), dt)
        o.make_step(dt)
        if step % 10 == 0:
            nlist.update()
            o.update()
    d = vec(atoms.data, "r") - vec(o.a, "r")
    assert np.max(np.abs(d)) < RUN_R_ABS
    assert max_rel(vec(atoms.data, "v"), vec(o.a, "v")) < RUN_REL
