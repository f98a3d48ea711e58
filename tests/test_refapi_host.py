"""Host side of the reference's Python API (mdcraft_amd.refapi and its
``mdcraft`` package names): Domain, Atoms, Threads, the VerletList
constructor checks and the host grid behind VerletList.sort / list_for.

Pinned by the reference's own known answers (test/cxx/domain/
test_domain.cxx:9-33, test/cxx/neibs/test_grid.cxx:38-94) and by the
oracle's grid and domain on random inputs.  Nothing here needs a GPU."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

from mdcraft_amd import refapi as A
from tests.helpers import oracle_mod

COMPAT = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mdcraft_amd", "compat")


def atoms_at(pos, rns):
    a = A.Atoms(len(pos))
    r = np.asarray(pos, dtype=np.float64)
    for k, c in enumerate("xyz"):
        a.data["r"][c] = r[:, k]
    a.data["rns"] = rns
    return a


def test_compat_package_names():
    if COMPAT not in sys.path:
        sys.path.insert(0, COMPAT)
    from mdcraft.data import Atoms
    from mdcraft.lattice import Domain
    from mdcraft.neibs import VerletList
    from mdcraft.solver import Multi, Single, Ternary
    from mdcraft.solver.boundary import Periodic
    from mdcraft.solver.potential import EAM, LJsCut, MaiseBPNNPotential
    from mdcraft.solver.stepper import Verlet
    from mdcraft.solver.thermostat import Langevin
    from mdcraft.tools import Threads
    assert Atoms is A.Atoms and VerletList is A.VerletList and Single is A.Single and Verlet is A.Verlet


def test_domain_known_answers():
    # test_domain.cxx:9-26
    d = A.Domain(-1.0, 1.0, -2.0, 2.0, -3.0, 3.0)
    d.set_periodic(0, True)
    d.set_periodic(1, True)
    a = atoms_at([[1.25, -2.25, 0.5]], 0.1)
    d.fit_in_period(a)
    assert a.data["r"]["x"][0] == pytest.approx(-0.75, abs=1e-12)
    assert a.data["r"]["y"][0] == pytest.approx(1.75, abs=1e-12)
    assert a.data["r"]["z"][0] == pytest.approx(0.5, abs=1e-12)
    assert d.belong(a).tolist() == [1]
    assert d.belong(-0.75, 1.75, 0.5) is True
    assert d.belong(1.5, 0.0, 0.0) is False
    # test_domain.cxx:28-33
    d2 = A.Domain(-1.0, 1.0, -1.0, 1.0, 0.0, 0.0)
    d2.set_periodic(0, True)
    b = atoms_at([[-0.8, 0.0, 0.0]], 0.1)
    assert d2.nearest_distance(b, 0)[0] == pytest.approx(0.2, abs=1e-15)  # EXPECT_DOUBLE_EQ: 4 ulps
    assert d2.nearest_distance(b, 1)[0] == np.finfo(np.float64).max
    assert d2._dim == 2
    assert (d.xmin, d.xmax, d.ysize, d.zsize, d.xperiodic, d.zperiodic) == (-1.0, 1.0, 4.0, 6.0, True, False)


def test_domain_against_oracle():
    R = oracle_mod()
    rng = np.random.default_rng(11)
    lo, hi, per = np.array([-1.0, 0.0, 2.0]), np.array([1.0, 3.0, 2.5]), (True, False, True)
    d = A.Domain(lo[0], hi[0], lo[1], hi[1], lo[2], hi[2])
    for k in range(3):
        d.set_periodic(k, per[k])
    od = R.Domain(lo, hi, per)
    pos = rng.uniform(-3, 5, size=(200, 3))
    a = atoms_at(pos, 0.1)
    d.fit_in_period(a)
    got = np.stack([a.data["r"][c] for c in "xyz"], 1)
    ref = np.stack([od.fit_in_period(p) for p in pos])
    assert np.array_equal(got, ref)
    for axis in range(3):
        want = [od.nearest_distance(p, axis) for p in got]
        assert np.array_equal(d.nearest_distance(a, axis), want)


def test_atoms_container():
    a = A.Atoms(5)
    assert len(a) == 5 and a.data.dtype.itemsize == 208
    assert a.data["tag"].tolist() == [65535] * 5 and not a.data["m"].any()
    a.data["uid"] = np.arange(5)
    a.resize(7)
    assert a.data["uid"].tolist() == [0, 1, 2, 3, 4, 0, 0] and a.data["tag"][6] == 65535
    a.erase(np.array([1, 3]))
    assert a.data["uid"].tolist() == [0, 2, 4, 0, 0]
    a.erase(np.where(a.data["uid"] == 4))
    assert a.data["uid"].tolist() == [0, 2, 0, 0]
    a.erase(1, 3)
    assert a.data["uid"].tolist() == [0, 0]
    b = A.Atoms(2)
    b.data["uid"] = [8, 9]
    a.join(b)
    assert a.data["uid"].tolist() == [0, 0, 8, 9]


def test_threads():
    t = A.Threads(1)
    assert t.count == 1 and not t.active()
    t.on()
    assert t.count == max(1, os.cpu_count() or 1)
    t.off()
    assert t.count == 1
    assert A.Threads(10 ** 6).count == max(1, os.cpu_count() or 1)


def test_verlet_list_constructor_checks():
    a = atoms_at([[0.1, 0.1, 0.1]], 0.5)
    d = A.Domain(0, 1, 0, 1, 0, 1)
    many = A.Threads(4)
    if many.active():
        with pytest.raises(RuntimeError, match="Half Verlet lists"):
            A.VerletList(a, a, d, many, True)
    with pytest.raises(NotImplementedError):
        A.VerletList(a, atoms_at([[0.2, 0.2, 0.2]], 0.5), d)
    nl = A.VerletList(a, a, d, half=True)
    assert nl.half is True and nl.decide() is True  # never updated


def _oracle_cells(atoms, lo, hi, per):
    R = oracle_mod()
    d = atoms.data
    n = len(d)
    cs = np.zeros(1)
    nc = C.c_int()
    ijk = np.zeros(3 * n, dtype=np.int32)
    cidx = np.zeros(n, dtype=np.int32)
    R.lib.ref_grid_build(d.ctypes.data, n, R.Domain(lo, hi, per).ref(), cs.ctypes.data_as(C.POINTER(C.c_double)),
                         C.byref(nc), ijk.ctypes.data_as(C.POINTER(C.c_int)), cidx.ctypes.data_as(C.POINTER(C.c_int)))
    return cs[0], cidx


def test_sort_matches_oracle_cells():
    rng = np.random.default_rng(4)
    lo, hi, per = np.zeros(3), np.array([3.0, 2.5, 2.0]), (True, True, False)
    pos = rng.uniform(-0.4, 3.2, size=(300, 3))
    pos[:, 2] = rng.uniform(0, 2.0, 300)
    a = atoms_at(pos, 0.45)
    a.data["uid"] = np.arange(300)
    d = A.Domain(0, 3.0, 0, 2.5, 0, 2.0)
    d.set_periodic(0, True)
    d.set_periodic(1, True)
    nl = A.VerletList(a, a, d)
    out = nl.sort(a)
    # sort wraps its argument in place, as put_atoms_in_cells does
    _, cidx = _oracle_cells(a, lo, hi, per)
    assert out.data["uid"].tolist() == np.argsort(cidx, kind="stable").tolist()


def test_list_for_matches_oracle_grid():
    """list_for: stencil of the point's cell from the oracle's grid, atoms of
    each cell last-inserted first, kept when within max(cell edge, rns_j)
    with the edge-cell minimum image (verlet-list.cxx:251-276, :361-400)."""
    R = oracle_mod()
    rng = np.random.default_rng(9)
    L = 2.0
    lo, hi, per = np.zeros(3), np.full(3, L), (True, True, True)
    pos = rng.uniform(0, L, size=(250, 3))
    a = atoms_at(pos, 0.4)
    a.data["rns"] = rng.uniform(0.3, 0.45, 250)
    d = A.Domain(0, L, 0, L, 0, L)
    for k in range(3):
        d.set_periodic(k, True)
    g = A._Grid(d, a.data)
    cs, cidx = _oracle_cells(a, lo, hi, per)
    for q in rng.uniform(0, L, size=(20, 3)):
        lst = np.zeros(27, dtype=np.int32)
        edge = np.zeros(3, dtype=np.int32)
        k = R.lib.ref_grid_adjacent(a.data.ctypes.data, 250, R.Domain(lo, hi, per).ref(),
                                    q.ctypes.data_as(C.POINTER(C.c_double)), lst.ctypes.data_as(C.POINTER(C.c_int)),
                                    edge.ctypes.data_as(C.POINTER(C.c_int)))
        want = []
        for c in lst[:k]:
            for j in np.nonzero(cidx == c)[0][::-1]:
                dr = q - pos[j]
                for ax in range(3):
                    if edge[ax] and 2 * dr[ax] < -L:
                        dr[ax] += L
                    if edge[ax] and 2 * dr[ax] > L:
                        dr[ax] -= L
                rr = max(cs, a.data["rns"][j])
                if dr @ dr <= rr * rr:
                    want.append(j)
        assert g.list_for(q).tolist() == want


def _atoms3(xs, rns=0.40):
    a = atoms_at([[x, 0.10, 0.10] for x in xs], rns)
    a.data["rcut"] = 0.40
    a.data["m"] = 1.0
    return a


def test_list_for_known_answer():
    # test/cxx/neibs/test_verlet_list.cxx:85-105
    a = _atoms3([0.15, 0.55, 1.60])
    d = A.Domain(0.0, 2.0, 0.0, 1.0, 0.0, 1.0)
    g = A._Grid(d, a.data)  # the grid update() leaves for list_for
    found = sorted(g.list_for(np.array([0.40, 0.10, 0.10])).tolist())
    assert found == [0, 1]


def test_sort_known_answer():
    # test_verlet_list.cxx:107-123
    a = _atoms3([1.60, 0.15, 0.95])
    d = A.Domain(0.0, 2.0, 0.0, 1.0, 0.0, 1.0)
    out = A.VerletList(a, a, d).sort(a)
    x = out.data["r"]["x"].tolist()
    assert len(x) == 3 and x[0] < x[1] < x[2]
