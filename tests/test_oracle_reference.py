"""The oracle (oracle/librefcpu.so) against the reference's own tests and
golden vectors, and against what the mathematics fixes.

Each test names the reference test or property it carries over.  These run
on the CPU only.
"""
import os

import numpy as np
import pytest

from mdcraft_amd.atoms import make_atoms, set_vec, vec, trace_V
from tests.helpers import oracle_mod, pair_set, have_reference, REFERENCE

R = oracle_mod()


def atoms_at(points, rns, rcut=None, m=1.0):
    a = make_atoms(len(points))
    set_vec(a, "r", np.asarray(points, dtype=float))
    a["rns"] = rns
    a["rcut"] = rns if rcut is None else rcut
    a["m"] = m
    return a


# ---------------------------------------------------------------- domain
def test_domain_fit_shortest_belong():
    # test/cxx/domain/test_domain.cxx:9-26
    d = R.Domain([-1, -2, -3], [1, 2, 3], (1, 1, 0))
    p = d.fit_in_period([1.25, -2.25, 0.5])
    assert np.allclose(p, [-0.75, 1.75, 0.5], atol=1e-12)
    s = d.shortest([1.8, -3.5, 0.25])
    assert np.allclose(s, [-0.2, 0.5, 0.25], atol=1e-12)


def test_domain_nearest_distance():
    # test_domain.cxx:28-33
    d = R.Domain([-1, -1, 0], [1, 1, 0], (1, 0, 0))
    assert d.nearest_distance([-0.8, 0, 0], 0) == pytest.approx(0.2, abs=1e-15)
    assert d.nearest_distance([-0.8, 0, 0], 1) == np.finfo(float).max


# ---------------------------------------------------------------- grid
def _grid(atoms, dom):
    import ctypes as C
    n = len(atoms)
    cs = np.zeros(1)
    nc = C.c_int()
    ijk = np.zeros(3 * n, dtype=np.int32)
    cidx = np.zeros(n, dtype=np.int32)
    R.lib.ref_grid_build(atoms.ctypes.data, n, dom.ref(), cs.ctypes.data_as(C.POINTER(C.c_double)), C.byref(nc),
                         ijk.ctypes.data_as(C.POINTER(C.c_int)), cidx.ctypes.data_as(C.POINTER(C.c_int)))
    return cs[0], nc.value, ijk.reshape(n, 3), cidx


def test_grid_local_subgrid_and_indices():
    # test/cxx/neibs/test_grid.cxx:38-70
    a = atoms_at([[0.2, 0.2, 0.2], [1.8, 1.1, 0.7], [2.6, 1.9, 0.8]], 1.0)
    cs, nc, ijk, cidx = _grid(a, R.Domain([0, 0, 0], [4, 4, 4]))
    assert cs == pytest.approx(1.0, abs=1e-12)
    assert nc == 6
    assert ijk.tolist() == [[0, 0, 0], [1, 1, 0], [2, 1, 0]]
    assert cidx.tolist() == [0, 4, 5]


def test_grid_adjacent_cells_wrap():
    # test_grid.cxx:72-94
    import ctypes as C
    a = atoms_at([[-1.8, 0, 0], [-0.2, 0, 0], [0.2, 0, 0], [1.8, 0, 0]], 1.0)
    dom = R.Domain([-2, -2, -2], [2, 2, 2], (1, 0, 0))
    lst = np.zeros(27, dtype=np.int32)
    edge = np.zeros(3, dtype=np.int32)
    q = np.array([-1.8, 0, 0])
    k = R.lib.ref_grid_adjacent(a.ctypes.data, len(a), dom.ref(), q.ctypes.data_as(C.POINTER(C.c_double)),
                                lst.ctypes.data_as(C.POINTER(C.c_int)), edge.ctypes.data_as(C.POINTER(C.c_int)))
    assert edge[0] == 1
    assert sorted(set(lst[:k].tolist())) == [0, 1, 3]


# ---------------------------------------------------------------- Verlet list
def _list(atoms, dom, half=False):
    pool = R.Pool(1)
    nl = R.VerletList(pool, half)
    nl.update(atoms, dom)
    return nl.lists(len(atoms)), nl


def test_verlet_without_periodicity():
    # test/cxx/neibs/test_verlet_list.cxx:25-52
    a = atoms_at([[0.2, 0, 0], [0.85, 0, 0], [2.5, 0, 0]], 0.8)
    L, _ = _list(a, R.Domain([0, -1, -1], [3, 1, 1]))
    assert [len(x) for x in L] == [2, 2, 1]
    assert sorted(L[0].tolist()) == [0, 1]
    assert L[2].tolist() == [2]


def test_verlet_through_periodic_boundary():
    # test_verlet_list.cxx:54-83
    a = atoms_at([[-0.48, 0, 0], [0.48, 0, 0]], 0.1)
    L, _ = _list(a, R.Domain([-0.5] * 3, [0.5] * 3, (1, 0, 0)))
    assert sorted(L[0].tolist()) == [0, 1] and sorted(L[1].tolist()) == [0, 1]


def test_verlet_half_list():
    # test_verlet_list.cxx:144-166
    a = atoms_at([[0.10, 0.1, 0.1], [0.32, 0.1, 0.1], [0.95, 0.1, 0.1]], 0.35)
    L, _ = _list(a, R.Domain([0, 0, 0], [2, 1, 1]), half=True)
    assert sorted(L[0].tolist()) == [0, 1]
    assert L[1].tolist() == [1] and L[2].tolist() == [2]


def test_verlet_kbuf_radius():
    # test_verlet_list.cxx:125-142
    a = atoms_at([[0.10, 0.1, 0.1], [0.35, 0.1, 0.1], [1.80, 0.1, 0.1]], 0.35)
    a["rcut"] = [0.2, 0.6, 0.1]
    dom = R.Domain([0, 0, 0], [2, 1, 1])
    pool = R.Pool(1)
    nl = R.VerletList(pool)
    nl.update(a, dom, -1.0)
    nl.update(a, dom, 2.0)
    assert np.allclose(a["rns"], [1.6, 2.4, 0.4], atol=1e-12)


def test_verlet_decide_half_skin():
    # test_verlet_list.cxx:168-186
    a = atoms_at([[0.10, 0.1, 0.1], [0.50, 0.1, 0.1]], 0.8, rcut=0.6)
    a["uid"] = [1, 2]
    dom = R.Domain([0, 0, 0], [2, 1, 1])
    pool = R.Pool(1)
    nl = R.VerletList(pool, True)
    nl.update(a, dom)
    a["r"]["x"][0] += 0.09
    assert not nl.decide(a, dom)
    a["r"]["x"][0] += 0.02
    assert nl.decide(a, dom)


def test_half_list_rejects_threads():
    # test_verlet_list.cxx:188-196
    with pytest.raises(RuntimeError):
        R.VerletList(R.Pool(2), True)


def brute_pairs(pos, rns, lo, hi, per):
    """O(N^2) pair set of verlet-list.cxx:374-400 (self included)."""
    L = np.asarray(hi) - np.asarray(lo)
    out = set()
    n = len(pos)
    for i in range(n):
        d = pos[i] - pos
        for ax in range(3):
            if per[ax]:
                d[:, ax] = np.where(2 * d[:, ax] < -L[ax], d[:, ax] + L[ax], d[:, ax])
                d[:, ax] = np.where(2 * d[:, ax] > L[ax], d[:, ax] - L[ax], d[:, ax])
        d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        rr = np.maximum(rns[i], rns)
        for j in np.nonzero(d2 <= rr * rr)[0]:
            out.add((i, int(j)))
    return out


@pytest.mark.parametrize("per,n,rns,seed", [((1, 1, 1), 600, 0.31, 1), ((0, 0, 0), 500, 0.35, 2),
                                            ((1, 0, 1), 400, 0.4, 3), ((1, 1, 1), 150, 0.9, 4)])
def test_verlet_pairs_match_brute_force(per, n, rns, seed):
    rng = np.random.default_rng(seed)
    lo, hi = np.array([-1.0, -0.5, 0.0]), np.array([1.0, 1.0, 1.4])
    pos = lo + rng.random((n, 3)) * (hi - lo)
    r = np.full(n, rns) * (1 + 0.2 * rng.random(n) * (seed % 2))
    a = atoms_at(pos, r)
    dom = R.Domain(lo, hi, per)
    pool = R.Pool(1)
    nl = R.VerletList(pool)
    nl.update(a, dom)
    got = pair_set(*nl.export(n))
    assert got == brute_pairs(vec(a, "r"), a["rns"], lo, hi, per)


# ---------------------------------------------------------------- LJsCut
EPS, SIG, RC = 1.03120074782442750, 0.33841043857528683, 0.8125


def lj_pref(r):
    s6 = SIG ** 6
    r2inv = 1 / (r * r)
    r6inv = r2inv ** 3
    return -0.5 * (48 * EPS * s6) * ((2 * s6) * r6inv - 1) * r6inv * r2inv


def lj_half_energy(r):
    sr6 = (SIG / r) ** 6
    return 0.5 * 4 * EPS * sr6 * (sr6 - 1)


@pytest.mark.parametrize("half", [False, True])
def test_lj_closed_form(half):
    # test/cxx/solver/test_ljs_cut.cxx:54-118
    a = atoms_at([[0, 0, 0], [0.4, 0, 0]], RC)
    a["m"] = 40.0
    dom = R.Domain([-2] * 3, [2] * 3)
    pool = R.Pool(1)
    nl = R.VerletList(pool, half)
    nl.update(a, dom)
    pot = R.LJsCut(EPS, SIG, RC)
    s = R.Solver(pot, pool, False)
    s.compute(a, dom, nl)
    fx = lj_pref(0.4) * 0.4
    assert a["f"]["x"][0] == pytest.approx(fx, abs=1e-12)
    assert a["f"]["x"][1] == pytest.approx(-fx, abs=1e-12)
    assert abs(a["f"]["y"][0]) < 1e-12
    assert a["Ep"].sum() == pytest.approx(2 * lj_half_energy(0.4), abs=1e-12)
    assert trace_V(a).sum() == pytest.approx(fx * 0.4, abs=1e-12)


def test_lj_prepare_resets_state():
    # test_ljs_cut.cxx:29-52
    a = atoms_at([[0, 0, 0]], 0.9)
    a["m"] = 40
    set_vec(a, "f", [[1, 2, 3]])
    a["Ep"] = 5
    a["Em"] = 7
    a["nc"] = 9
    a["V"]["xx"] = 1
    dom = R.Domain([-2] * 3, [2] * 3)
    pool = R.Pool(1)
    nl = R.VerletList(pool)
    nl.update(a, dom)
    pot = R.LJsCut(EPS, SIG, RC)
    R.Solver(pot, pool, False).prepare(a, dom, nl)
    assert a["rcut"][0] == RC and a["Ep"][0] == 0 and np.all(vec(a, "f") == 0) and a["V"]["xx"][0] == 0


# ---------------------------------------------------------------- goldens
def fcc(nx, ny, nz, a0, mass):
    """md_test_utils.h make_fcc_bulk"""
    basis = np.array([[0, 0, 0], [.5, .5, 0], [.5, 0, .5], [0, .5, .5]])
    L = np.array([nx, ny, nz]) * a0
    origin = -0.5 * L
    pos = [origin + np.array([ix, iy, iz]) * a0 + a0 * b for iz in range(nz) for iy in range(ny)
           for ix in range(nx) for b in basis]
    at = make_atoms(len(pos))
    set_vec(at, "r", np.array(pos))
    at["uid"] = np.arange(1, len(pos) + 1)
    at["kind"] = 1
    at["m"] = mass
    return at, -0.5 * L, 0.5 * L


def deterministic_velocities(at, scale):
    """md_test_utils.h assign_deterministic_velocities"""
    i = np.arange(len(at))
    v = scale * np.stack([(i % 7) - 3.0, ((i // 7) % 5) - 2.0, ((i // 35) % 3) - 1.0], 1)
    p = (at["m"][:, None] * v).sum(0)
    set_vec(at, "v", v - p / at["m"].sum())


@pytest.mark.reference
@pytest.mark.skipif(not have_reference(), reason="reads the reference's Al2009.eam.alloy fixture")
def test_al_eam_golden():
    # test/cxx/solver/test_aluminum_bulk_eam.cxx:74-87
    pot = R.EAM(path=os.path.join(REFERENCE, "test/python/solver/Al2009.eam.alloy"))
    at, lo, hi = fcc(3, 3, 3, 0.4032, 26.9815384)
    at["rcut"] = pot.rcut
    at["rns"] = 1.25 * pot.rcut
    deterministic_velocities(at, 5e-3)
    dom = R.Domain(lo, hi, (1, 1, 1))
    pool = R.Pool(1)
    h = R.Solver(pot, pool, True).run_md_steps(at, dom, R.VerletList(pool), 4e-3, 2, 1)
    assert h[2, 3] == pytest.approx(-3.4489006959225437e+04, abs=1e-6)
    assert h[2, 0] == pytest.approx(-6.7190981865001777e+03, abs=1e-6)
    assert h[2, 1] == pytest.approx(-2.7770493824864527e+04, abs=1e-6)
    assert h[2, 2] == pytest.approx(5.8505213926282473e-01, abs=1e-9)
    assert h[2, 4] == pytest.approx(7.4272691119381875e+01, abs=1e-9)
    a0 = at[0]
    assert a0["r"]["x"] == pytest.approx(6.0469801558679681e-01, abs=1e-12)
    assert a0["r"]["y"] == pytest.approx(6.0474258946107018e-01, abs=1e-12)
    assert a0["r"]["z"] == pytest.approx(6.0477602485847293e-01, abs=1e-12)
    assert a0["v"]["x"] == pytest.approx(-1.3221213313405629e-02, abs=1e-15)
    assert a0["f"]["x"] == pytest.approx(-1.1358391912725722e+01, abs=1e-12)
    assert a0["Ep"] == pytest.approx(-6.3230321656455359e+01, abs=1e-9)
    assert a0["Em"] == pytest.approx(-2.5721082474679707e+02, abs=1e-9)
    assert trace_V(at)[0] == pytest.approx(1.3557390418206403e+01, abs=1e-9)
    # test_aluminum_bulk_eam.cxx:56-69: energy and momentum over 5 steps
    at2, _, _ = fcc(3, 3, 3, 0.4032, 26.9815384)
    at2["rcut"] = pot.rcut
    at2["rns"] = 1.25 * pot.rcut
    deterministic_velocities(at2, 5e-3)
    h5 = R.Solver(pot, pool, True).run_md_steps(at2, dom, R.VerletList(pool), 4e-3, 5, 1)
    assert abs(h5[-1, 3] - h5[0, 3]) < 40.0
    assert np.linalg.norm(h5[-1, 5:8]) < 1e-12


def load_argon_ball():
    rows = []
    started = False
    with open(os.path.join(REFERENCE, "test/cxx/LMP_BALL_130k_Ar")) as f:
        for ln in f:
            if ln.startswith("Atoms"):
                started = True
                continue
            if started:
                p = ln.split()
                if len(p) >= 5:
                    rows.append([float(x) for x in p[:5]])
    rows = np.array(rows)
    at = make_atoms(len(rows))
    set_vec(at, "r", rows[:, 2:5] * 0.1)
    at["uid"] = rows[:, 0]
    at["kind"] = rows[:, 1]
    at["m"] = 40.0
    lo, hi = np.full(3, -96.934310913) * 0.1, np.full(3, 218.065689087) * 0.1
    return at, lo, hi


@pytest.mark.reference
@pytest.mark.skipif(not have_reference(), reason="reads the reference's LMP_BALL_130k_Ar fixture")
def test_argon_ball_golden():
    # test/cxx/solver/test_argon_ball_md.cxx:48-82
    at, lo, hi = load_argon_ball()
    assert len(at) == 130124
    pot = R.LJsCut(EPS, SIG, RC)
    at["rcut"] = RC
    at["rns"] = 10.125 / 8.125 * RC
    pool = R.Pool(4)
    dom = R.Domain(lo, hi)
    h = R.Solver(pot, pool, False).run_md_steps(at, dom, R.VerletList(pool), 4e-3, 2, 1)
    assert h[2, 0] == pytest.approx(-1.0244383150033894e+06, abs=1e-3)
    assert h[2, 4] == pytest.approx(-1.1164231473675424e+05, abs=1e-3)
    assert abs(h[2, 3] - h[0, 3]) < 2.0
    i1 = int(np.nonzero(at["uid"] == 1)[0][0])
    i2 = int(np.nonzero(at["uid"] == 2)[0][0])
    a = at[i1]
    assert a["r"]["x"] == pytest.approx(-3.9184367989026225, abs=1e-12)
    assert a["r"]["y"] == pytest.approx(2.6440651643797461, abs=1e-12)
    assert a["r"]["z"] == pytest.approx(5.2690712254285517, abs=1e-12)
    assert a["v"]["z"] == pytest.approx(5.79013209455442e-04, abs=1e-15)
    assert a["f"]["z"] == pytest.approx(2.8929546164515996, abs=1e-12)
    assert a["Ep"] == pytest.approx(-4.0629298955784474, abs=1e-12)
    assert trace_V(at)[i1] == pytest.approx(-3.3178872346261723e-01, abs=1e-12)
    a = at[i2]
    assert a["r"]["z"] == pytest.approx(5.7940689066038047, abs=1e-12)
    assert a["v"]["x"] == pytest.approx(-1.4268458571713898e-03, abs=1e-15)
    assert a["f"]["x"] == pytest.approx(-7.1268568973069453, abs=1e-12)
    assert a["Ep"] == pytest.approx(-4.2002585997144832, abs=1e-12)
    assert trace_V(at)[i2] == pytest.approx(4.2842826328841377e-01, abs=1e-12)
