"""The oracle's LJs (LJs.cxx) and tabulated Pair (pair.cxx) potentials.

The reference has no tests of either; they are pinned by what the
mathematics fixes: LJs's tail is built so that V and V' vanish at the
cutoff and leaves the LJ minimum (X = 0) untouched, so V(2^(1/6) rVr) = -a;
forces are minus the energy gradient (central differences); a tabulated
Pair sampled from an analytic r*phi converges to it at the spline's order,
and a Pair equals an EAM object whose density and embedding tables vanish
(the same spline of r*phi, with F' = 0)."""
import numpy as np
import pytest

from mdcraft_amd import workloads as W
from mdcraft_amd.atoms import EV, make_atoms, set_vec, vec
from tests.helpers import max_rel, oracle_mod

R = oracle_mod()


def cluster(pos, rns=1.2):
    at = make_atoms(len(pos))
    set_vec(at, "r", np.asarray(pos, float))
    at["m"] = 39.948
    at["rns"] = rns
    return at


def evaluate(at, pot, box=None, periodic_boundary=False):
    a = at.copy()
    box = box or (np.full(3, -5.0), np.full(3, 5.0), (False, False, False))
    dom = R.Domain(*box)
    pool = R.Pool(1)
    nl = R.VerletList(pool)
    nl.update(a, dom)
    s = R.Solver(pot, pool, periodic_boundary)
    s.compute(a, dom, nl)
    return a


def test_ljs_tail_vanishes_at_cutoff_and_keeps_the_minimum():
    a, rv, rc = 1.3, 0.34, 0.85
    p = R.LJs(a, rv, rc)
    at = evaluate(cluster([[0, 0, 0], [rc, 0, 0]]), p)
    assert abs(at["Ep"].sum()) < 1e-13 * a
    assert np.abs(vec(at, "f")).max() < 1e-11 * a / rv
    rmin = 2 ** (1 / 6) * rv
    at = evaluate(cluster([[0, 0, 0], [rmin, 0, 0]]), p)
    assert at["Ep"].sum() == pytest.approx(-a, rel=1e-13)
    assert np.abs(vec(at, "f")).max() < 1e-12 * a / rv


def _energy(at, pot, box=None, pb=False):
    return evaluate(at, pot, box, pb)["Ep"].sum()


@pytest.mark.parametrize("which", ["ljs", "pair"])
def test_forces_are_minus_energy_gradient(which):
    rng = np.random.default_rng(2)
    pos = rng.uniform(0, 1.2, size=(24, 3))
    if which == "ljs":
        pot = R.LJs(1.0, 0.30, 0.75)
    else:
        t = W.eam_cu_tables()
        pot = R.Pair(t["rc"], t["rphi"], t["dr"])
    base = cluster(pos)
    at = evaluate(base, pot)
    f = vec(at, "f")
    h = 1e-6
    for i in range(0, 24, 5):
        for c in range(3):
            p, m = base.copy(), base.copy()
            p["r"]["xyz"[c]][i] += h
            m["r"]["xyz"[c]][i] -= h
            fd = -(_energy(p, pot) - _energy(m, pot)) / (2 * h)
            assert fd == pytest.approx(f[i, c], rel=1e-5, abs=1e-6 * np.abs(f).max())


def test_pair_table_converges_to_the_analytic_function():
    t = W.eam_cu_tables()
    pot = R.Pair(t["rc"], t["rphi"], t["dr"])
    for r in (0.22, 0.25, 0.31, 0.42):
        at = evaluate(cluster([[0, 0, 0], [r, 0, 0]]), pot)
        ra = r / W.ANG
        phi = W.morse(ra) * W.taper(ra) * EV
        assert at["Ep"].sum() == pytest.approx(phi, rel=2e-6)
        h = 1e-6
        dphi = (W.morse((r + h) / W.ANG) * W.taper((r + h) / W.ANG) - W.morse((r - h) / W.ANG) * W.taper((r - h) / W.ANG)) * EV / (2 * h)
        assert vec(at, "f")[0, 0] == pytest.approx(dphi, rel=2e-5)


def test_pair_equals_eam_without_embedding():
    w = W.eam_cu(reps=3, rattle=0.05)
    t = w["pot"]
    pair = R.Pair(t["rc"], t["rphi"], t["dr"])
    eam0 = R.EAM(rc=t["rc"], F=np.zeros_like(t["F"]), rho=np.zeros_like(t["rho"]), rphi=t["rphi"], dr=t["dr"],
                 drho=t["drho"])
    a = evaluate(w["atoms"], pair, w["box"], True)
    b = evaluate(w["atoms"], eam0, w["box"], True)
    assert max_rel(vec(a, "f"), vec(b, "f")) < 1e-13
    assert max_rel(a["Ep"], b["Ep"]) < 1e-13
