"""List rebuilds inside step(): the list is rebuilt on the device when the
half-skin trigger fires, and a list that overflows there is grown when the
host reads the next step's counters (the overflowing build stops every
later step until then).

Bars: FP64 trajectories within 1e-9 nm on positions and 1e-8 relative on
velocities against a run whose list was large enough from the start (the
two differ only in the order of FP64 sums)."""
import numpy as np
import pytest

from mdcraft_amd import workloads as W
from mdcraft_amd.atoms import set_vec, vec
from tests.helpers import engine_potential, max_rel

pytestmark = pytest.mark.gpu


def E():
    from mdcraft_amd import engine
    return engine


def compressed(at, box, factor):
    """The atoms of a periodic box pressed into the lower part of it along
    z (a denser slab with vacuum above)."""
    dense = at.copy()
    r = vec(at, "r")
    r[:, 2] = box[0][2] + (r[:, 2] - box[0][2]) * factor
    set_vec(dense, "r", r)
    return dense


@pytest.mark.parametrize("kind", ["lj", "eam"])
def test_list_overflow_inside_steps_regrows(ctx, kind):
    eng = E()
    # LJ: 10^3 cells (the build's fast path, 5 cells an axis); EAM: 7^3
    # cells (4 cells an axis, the general image rule)
    w = W.lj_argon(reps=10, rattle=0.05) if kind == "lj" else W.eam_cu(reps=7, rattle=0.05)
    at, box, pot, dt = w["atoms"], w["box"], w["pot"], w["dt"]
    steps = 12
    p = engine_potential(eng, ctx, pot)
    # the list sized on the uniform lattice
    s = eng.System(ctx, [len(at)], [box])
    s.upload(at)
    s.build()
    cap0 = s.stats()["capacity"]
    # the reference: the list built on a compressed configuration with more
    # neighbours per atom than that capacity
    for factor in (0.7, 0.6, 0.5):
        ref = eng.System(ctx, [len(at)], [box])
        ref.upload(compressed(at, box, factor))
        ref.build()
        need = ref.stats()["max"]
        if need > cap0:
            break
    assert need > cap0, (cap0, need)
    ref.compute(p, eng.FORCE, eng.FP64)
    start = ref.download()
    ref.step(p, dt, steps, eng.FP64)
    want = ref.download()
    # the compressed atoms and their forces into the small list: the first
    # rebuild inside step() overflows it
    s.upload(start, eng.F_R | eng.F_V | eng.F_F)
    s.step(p, dt, steps, eng.FP64)
    got = s.download()
    assert s.stats()["capacity"] >= need
    L = box[1] - box[0]
    d = vec(got, "r") - vec(want, "r")
    d -= np.round(d / L) * L
    assert np.max(np.abs(d)) < 1e-9
    assert max_rel(vec(got, "v"), vec(want, "v")) < 1e-8
    assert max_rel(vec(got, "f"), vec(want, "f")) < 1e-8


def test_upload_of_rewrapped_positions_between_steps(ctx):
    """Positions uploaded between step sequences may put an atom at another
    periodic image than the list was built with (here the atom nearest the
    upper x face moves to its image below the lower face).  The list stays,
    and the next steps must use each pair's image from the new positions,
    not the shift recorded at the build; compared with the oracle's
    make_step loop from the same state."""
    from oracle import refcpu as R
    from tests.helpers import oracle_potential
    eng = E()
    w = W.eam_cu(reps=7, rattle=0.05)
    at, box, pot, dt = w["atoms"], w["box"], w["pot"], w["dt"]
    p = engine_potential(eng, ctx, pot)
    s = eng.System(ctx, [len(at)], [box])
    s.upload(at)
    s.build()
    s.compute(p, eng.FORCE, eng.FP64)
    s.step(p, dt, 3, eng.FP64)
    mid = s.download()
    L = np.asarray(box[1]) - np.asarray(box[0])
    r = vec(mid, "r")
    k = int(np.argmax(r[:, 0]))
    r[k, 0] -= L[0]
    set_vec(mid, "r", r)
    s.upload(mid, eng.F_R)
    s.step(p, dt, 5, eng.FP64)
    got = s.download()
    a = mid.copy()
    dom = R.Domain(*box)
    pool = R.Pool(8)
    rp = oracle_potential(R, pot)
    a["rcut"] = rp.rcut
    nl = R.VerletList(pool)
    nl.update(a, dom)
    R.Solver(rp, pool, True).run_nve(a, dom, nl, dt, 5, initial_forces=True)
    d = vec(got, "r") - vec(a, "r")
    d -= np.round(d / L) * L
    assert np.max(np.abs(d)) < 1e-9
    assert max_rel(vec(got, "v"), vec(a, "v")) < 1e-8
    assert max_rel(vec(got, "f"), vec(a, "f")) < 1e-8
