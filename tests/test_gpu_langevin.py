"""The Langevin thermostat on the device against the oracle's, which draws
the same counter-based normal deviates (the stream itself is pinned in
tests/test_oracle_langevin.py).

Bars: one force evaluation within 1e-12 relative (FP64); trajectories of
20-40 steps within 1e-9 nm on positions and 1e-8 relative on velocities;
the bath temperature within 3 % over a long mixed-precision run."""
import os
import sys

import numpy as np
import pytest

from mdcraft_amd import workloads as W
from mdcraft_amd.atoms import KB, vec
from tests.helpers import engine_potential, max_rel, oracle_mod, oracle_potential

pytestmark = pytest.mark.gpu

CALL_TOL = 1e-12
RUN_R_ABS = 1e-9
RUN_REL = 1e-8


def E():
    from mdcraft_amd import engine
    return engine


def rows(T=300.0, beta=2.0, dt=0.002, window=(-1e15, 1e15), heat=(1, 1, 1), U=(0.0, 0.0, 0.0)):
    return [(beta, T, dt, *heat, *U, *window)]


def oracle_setup(at, box, pot, th_rows, seed):
    R = oracle_mod()
    a = at.copy()
    dom = R.Domain(*box)
    pool = R.Pool(1)
    p = oracle_potential(R, pot, box)
    nl = R.VerletList(pool)
    nl.update(a, dom)
    th = R.Thermostats(th_rows, seed)
    return a, dom, nl, R.Solver(p, pool, True, th), (p, pool, th)


@pytest.mark.parametrize("kind", ["lj", "eam"])
def test_force_evaluation_with_thermostats(ctx, kind):
    w = W.lj_argon(reps=5, rattle=0.1) if kind == "lj" else W.eam_cu(reps=5, rattle=0.05)
    at, box, pot = w["atoms"], w["box"], w["pot"]
    th = rows(T=300.0, beta=3.0, window=(0.2, 1.4)) + rows(T=50.0, beta=1.0, heat=(1, 0, 1), U=(0.1, 0, 0))
    eng = E()
    s = eng.System(ctx, [len(at)], [box])
    s.upload(at)
    s.set_thermostat(th, 42)
    p = engine_potential(eng, ctx, pot)
    s.build()
    s.compute(p, eng.FORCE, eng.FP64)
    got = s.download()
    a, dom, nl, sol, keep = oracle_setup(at, box, pot, th, 42)
    sol.prepare(a, dom, nl)
    sol.stages(a, dom, nl)
    assert max_rel(vec(got, "f"), vec(a, "f")) <= CALL_TOL


@pytest.mark.parametrize("kind", ["lj", "eam"])
def test_thermostatted_steps_with_rebuilds(ctx, kind):
    """Resident steps (rebuilds on the half-skin trigger) against the
    oracle's make_step loop with decide()/update(); the stream position is
    the step, so rebuild timing does not change the deviates."""
    w = W.lj_argon(reps=5, rattle=0.1) if kind == "lj" else W.eam_cu(reps=5, rattle=0.05)
    at, box, pot = w["atoms"], w["box"], w["pot"]
    dt, steps = w["dt"], 40
    th = rows(T=400.0, beta=5.0, dt=dt)
    eng = E()
    s = eng.System(ctx, [len(at)], [box])
    s.upload(at)
    s.set_thermostat(th, 7)
    p = engine_potential(eng, ctx, pot)
    s.build()
    s.compute(p, eng.FORCE, eng.FP64)
    s.step(p, dt, steps, eng.FP64)
    got = s.download()
    a, dom, nl, sol, keep = oracle_setup(at, box, pot, th, 7)
    sol.prepare(a, dom, nl)
    sol.stages(a, dom, nl)
    for _ in range(steps):
        if nl.decide(a, dom):
            nl.update(a, dom)
        sol.make_step(a, dom, nl, dt)
    L = box[1] - box[0]
    d = vec(got, "r") - vec(a, "r")
    d -= np.round(d / L) * L
    assert np.max(np.abs(d)) < RUN_R_ABS
    assert max_rel(vec(got, "v"), vec(a, "v")) < RUN_REL


def test_reference_script_with_langevin():
    """test-EAM.py with the thermostat attached to the stepping solver:
    Langevin(beta, temperature, time_step, ...) through the reference's
    Python API, with a seed, against the oracle with the same seed."""
    compat = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mdcraft_amd", "compat")
    if compat not in sys.path:
        sys.path.insert(0, compat)
    from mdcraft.data import Atoms
    from mdcraft.lattice import Domain
    from mdcraft.neibs import VerletList
    from mdcraft.solver import Single
    from mdcraft.solver.boundary import Periodic
    from mdcraft.solver.potential import EAM
    from mdcraft.solver.stepper import Verlet
    from mdcraft.solver.thermostat import Langevin
    w = W.eam_cu(reps=5, rattle=0.05)
    at, box, pot = w["atoms"], w["box"], w["pot"]
    dt = 0.002
    at["rns"] = 1.25 * at["rcut"]
    atoms = Atoms(len(at))
    atoms.data[:] = at
    lo, hi, _ = box
    domain = Domain(lo[0], hi[0], lo[1], hi[1], lo[2], hi[2])
    for k in range(3):
        domain.set_periodic(k, True)
    potential = EAM(Rcutoff=pot["rc"], pairPotential=pot["rphi"], chargeDensity=pot["rho"],
                    embeddingEnergy=pot["F"], NrPot=len(pot["rho"]), drPot=pot["dr"], NrEmb=len(pot["F"]),
                    drEmb=pot["drho"])
    thermostat = Langevin(beta=0.5, temperature=300.0, time_step=dt, heat_x=1, heat_y=1, heat_z=1, Ux=0, Uy=0, Uz=0,
                          seed=2024)
    periodic = Periodic(domain)
    solver = Single(potential=potential, boundary=periodic, thermostat=thermostat)
    solver_cold = Single(potential=potential, boundary=periodic)
    stepper = Verlet(solver=solver)
    nlist = VerletList(atoms=atoms, neighbors=atoms, domain=domain)
    nlist.update()
    solver_cold.prepare(atoms, atoms, nlist.get())
    solver_cold.forces(atoms, atoms, nlist.get())
    solver_cold.virials(atoms, atoms, nlist.get())
    R = oracle_mod()
    a = at.copy()
    dom = R.Domain(*box)
    pool = R.Pool(1)
    p = oracle_potential(R, pot)
    nl = R.VerletList(pool)
    nl.update(a, dom)
    th = R.Thermostats([thermostat.row()], 2024)
    cold, hot = R.Solver(p, pool, True), R.Solver(p, pool, True, th)
    cold.prepare(a, dom, nl)
    cold.forces(a, dom, nl)
    cold.virials(a, dom, nl)
    for step in range(1, 21):
        stepper.make_step(atoms, nlist.get(), dt)
        hot.make_step(a, dom, nl, dt)
        if step % 10 == 0:
            nlist.update()
            nl.update(a, dom)
    assert np.max(np.abs(vec(atoms.data, "r") - vec(a, "r"))) < RUN_R_ABS
    assert max_rel(vec(atoms.data, "v"), vec(a, "v")) < RUN_REL
    assert max_rel(vec(atoms.data, "f"), vec(a, "f")) < RUN_REL


def test_bath_temperature_mixed(ctx):
    """LJ argon (4000 atoms, 100 K start) under Langevin(beta = 5/ps,
    T = 160 K), mixed precision: the kinetic temperature over steps
    2000-4000 is 160 K within 3 %."""
    w = W.lj_argon(reps=10, T=100.0)
    at, box, pot, dt = w["atoms"], w["box"], w["pot"], w["dt"]
    eng = E()
    s = eng.System(ctx, [len(at)], [box])
    s.upload(at)
    s.set_thermostat(rows(T=160.0, beta=5.0, dt=dt), 99)
    p = engine_potential(eng, ctx, pot)
    s.build()
    s.compute(p, eng.FORCE, eng.MIXED)
    s.step(p, dt, 2000, eng.MIXED)
    temps = []
    n = len(at)
    for _ in range(20):
        s.step(p, dt, 100, eng.MIXED)
        th = s.thermo()[0]
        temps.append(2 * th[2] / (3 * n * KB))
    assert np.mean(temps) == pytest.approx(160.0, rel=0.03)
