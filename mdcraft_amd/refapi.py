"""The reference's Python API over the engine.

Classes with the reference's module layout, names, argument names and
defaults (python/mdcraft/*/_*.cxx), so that scripts written against the
reference's ``mdcraft`` package run on the GPU engine.  The package
``mdcraft_amd/compat/mdcraft`` re-exports them under the reference's module
names: put ``mdcraft_amd/compat`` on ``sys.path`` and ``import mdcraft``.

State follows the reference: the host records in ``Atoms.data`` are the
truth between calls.  Every solver, stepper and list call uploads the
records, runs on the device, and writes the fields the reference's call
writes back into ``Atoms.data`` (positions after a list update, forces,
energies, densities and virials after a solver call, positions,
velocities and forces after a step).  The neighbour list lives on the
device; ``get()`` hands out a handle and its ids are copied to the host
only when indexed.

Differences from the reference, each with the reason:

* lists between two different atom sets (``VerletList(atoms, other)``)
  are not supported: the engine builds lists of one set (the reference
  uses them for MPI ghost exchange, which is out of scope);
* ids in a list come in the engine's order (self first, then neighbours
  in cell order); the set of ids is the reference's;
* ``decide()`` compares each atom's displacement with the smallest skin
  (rns - rcut) of the box instead of the atom's own; for the uniform rns
  and rcut every solver sets, the two agree;
* the half list of ``VerletList(half=True)`` is exported as the full
  list's upper half (j >= i); forces are identical either way;
* potentials outside the engine (``DeepModelingPotential``, ``MLIP4``)
  raise NotImplementedError;
* ``Langevin`` draws from a counter-based stream keyed by ``seed`` (an
  added keyword; random when omitted) instead of std::random_device, one
  stream per list's system; the members of a ``Compound`` share the first
  member's seed; atoms with tag == 0 get no thermostat force.
"""
import os

import numpy as np

from . import engine as E
from .atoms import ATOM_DTYPE, make_atoms

_CTX = None


def context():
    """The engine context of this process (device 0), made on first use."""
    global _CTX
    if _CTX is None:
        _CTX = E.Context(0)
    return _CTX


# fields a solver call reads from the host records (rns and uid only
# change through a list update, which re-uploads everything)
_UP = E.F_ALL & ~(E.F_RNS | E.F_UID)


# --------------------------------------------------------------------------
# mdcraft.tools (python/mdcraft/tools/_tools.cxx:14-39, tools/threads.cxx)
# --------------------------------------------------------------------------
class Threads:
    """Thread-pool descriptor.  The engine runs on the GPU; the count is
    kept for the reference's checks (``active()`` when above 1)."""

    def __init__(self, count=100000, comm=None):
        self._max = max(1, os.cpu_count() or 1)
        self._count = 1
        self._resize(count)

    def _resize(self, count):
        self._count = max(1, min(int(count), self._max))

    @property
    def count(self):
        return self._count

    def on(self, count=None):
        self._resize(self._max)

    def off(self):
        self._resize(1)

    def active(self):
        return self._count > 1


_dummy_pool = Threads(1)


# --------------------------------------------------------------------------
# mdcraft.data (python/mdcraft/data/_data.cxx:84-235)
# --------------------------------------------------------------------------
class Atoms:
    """Container of 208-byte atom records; ``data`` is the numpy view."""

    def __init__(self, size=0):
        self._data = make_atoms(int(size))

    @property
    def data(self):
        return self._data

    def resize(self, size):
        size = int(size)
        new = make_atoms(size)
        k = min(size, len(self._data))
        new[:k] = self._data[:k]
        self._data = new

    def __len__(self):
        return len(self._data)

    def erase(self, index, to=None):
        """erase(index array), erase(tuple from numpy.where) or
        erase(from, to); indices sorted and unique, as the reference's
        memmove loop requires."""
        if to is not None:
            self._data = np.concatenate([self._data[:int(index)], self._data[int(to):]])
            return
        if isinstance(index, tuple):
            index = index[0]
        self._data = np.delete(self._data, np.asarray(index, dtype=np.int64))

    def join(self, other):
        self._data = np.concatenate([self._data, other.data])


# --------------------------------------------------------------------------
# mdcraft.lattice (lattice/domain.cxx, python/mdcraft/lattice/_lattice.cxx)
# --------------------------------------------------------------------------
class Domain:
    """Box [min, max] per axis with per-axis periodicity."""

    def __init__(self, xmin=0.0, xmax=0.0, ymin=0.0, ymax=0.0, zmin=0.0, zmax=0.0):
        self._per = [False, False, False]
        self.reshape(xmin, xmax, ymin, ymax, zmin, zmax)
        # domain.cxx:18-19
        self._dim = 3 if abs(zmin - zmax) > 1e-15 else 2 if abs(ymin - ymax) > 1e-15 else 1

    def reshape(self, xmin=0.0, xmax=0.0, ymin=0.0, ymax=0.0, zmin=0.0, zmax=0.0):
        self._lo = np.array([xmin, ymin, zmin], dtype=np.float64)
        self._hi = np.array([xmax, ymax, zmax], dtype=np.float64)
        self._dim = 3 if zmin != zmax else (2 if ymin != ymax else 1)

    def set_periodic(self, axis, periodic=True):
        self._per[int(axis)] = bool(periodic)

    def periodic(self, axis):
        return self._per[axis]

    def box(self):
        """(lo[3], hi[3], periodic[3]) for the engine."""
        return self._lo.copy(), self._hi.copy(), tuple(self._per)

    def belong(self, x, y=None, z=None):
        """belong(x, y, z) -> bool, or belong(atoms) -> int array
        (domain.cxx:33-39, inclusive bounds)."""
        if isinstance(x, Atoms):
            r = _rvec(x.data)
            return np.all((self._lo <= r) & (r <= self._hi), axis=1).astype(np.int32)
        v = np.array([x, y, z], dtype=np.float64)
        return bool(np.all((self._lo <= v) & (v <= self._hi)))

    def fit_in_period(self, atoms):
        """One shift by the box length on periodic axes (domain.cxx:41-47)."""
        d = atoms.data
        for a, name in enumerate("xyz"):
            if not self._per[a]:
                continue
            c = d["r"][name]
            L = self._hi[a] - self._lo[a]
            c[:] = np.where(c < self._lo[a], c + L, np.where(c > self._hi[a], c - L, c))

    def nearest_distance(self, atoms, axis):
        """Distance to the nearer face on a periodic axis, else the largest
        double (domain.cxx:49-57)."""
        c = atoms.data["r"]["xyz"[axis]]
        if not self._per[axis]:
            return np.full(len(c), np.finfo(np.float64).max)
        return np.minimum(np.abs(c - self._lo[axis]), np.abs(c - self._hi[axis]))

    xmin = property(lambda self: float(self._lo[0]))
    ymin = property(lambda self: float(self._lo[1]))
    zmin = property(lambda self: float(self._lo[2]))
    xmax = property(lambda self: float(self._hi[0]))
    ymax = property(lambda self: float(self._hi[1]))
    zmax = property(lambda self: float(self._hi[2]))
    xsize = property(lambda self: float(self._hi[0] - self._lo[0]))
    ysize = property(lambda self: float(self._hi[1] - self._lo[1]))
    zsize = property(lambda self: float(self._hi[2] - self._lo[2]))
    xperiodic = property(lambda self: self._per[0])
    yperiodic = property(lambda self: self._per[1])
    zperiodic = property(lambda self: self._per[2])


def _rvec(d):
    r = d["r"]
    return np.stack([r["x"], r["y"], r["z"]], axis=1)


# --------------------------------------------------------------------------
# mdcraft.neibs (neibs/verlet-list.cxx, python/mdcraft/neibs/_neibs.cxx)
# --------------------------------------------------------------------------
class NeibsList:
    """Handle to a VerletList's lists (``VerletList.get()``); indexing
    copies the ids from the device once per list update."""

    def __init__(self, owner):
        self.owner = owner

    def __len__(self):
        return len(self.owner.atoms)

    def __getitem__(self, i):
        return self.owner[i]

    def __iter__(self):
        for i in range(len(self)):
            yield self.owner[i]


class VerletList:
    """Neighbour list on the device for one set of atoms in one domain."""

    def __init__(self, atoms, neighbors, domain, threads=_dummy_pool, half=False):
        if neighbors is not atoms:
            raise NotImplementedError("the engine builds lists of one atom set: pass the same Atoms twice")
        if half and threads.active():
            # verlet-list.cxx:27-32
            raise RuntimeError(
                "Half Verlet lists are currently incompatible with multithreaded pair-force evaluation. "
                "Disable threads or use a full neighbor list.")
        self.atoms, self.domain, self.threads, self._half = atoms, domain, threads, bool(half)
        self.system = None
        self._key = None
        self._saved = None
        self._ids = None
        self._grid = None

    @property
    def half(self):
        return self._half

    def _ensure_system(self):
        lo, hi, per = self.domain.box()
        key = (len(self.atoms), tuple(lo), tuple(hi), per)
        if self.system is None or self._key != key:
            if self.system is not None:
                self.system.close()
            self.system = E.System(context(), [len(self.atoms)], [(lo, hi, per)], True)
            self.system.set_rebuild(False)
            self._key = key
            return True
        return False

    def update(self, kbuf=-1.0):
        """Optionally rescale rns (kbuf > 0, over the current list), wrap
        positions into the periodic box, rebuild, and save the state
        ``decide()`` compares against (verlet-list.cxx:41-56)."""
        d = self.atoms.data
        self._ids = None
        if len(d) == 0:
            self._saved = {"n": 0}
            self._grid = None
            return
        fresh = self._ensure_system()
        s = self.system
        if kbuf > 0.0:
            if fresh:
                d["rns"] = 0.0  # the reference's lists start empty
            else:
                s.upload(d, E.F_RCUT)
                d["rns"] = s.optimize_rns(kbuf)
        s.upload(d, E.F_ALL)
        s.build()
        out = s.download(fields=E.F_R)
        d["r"] = out["r"]
        self._saved = {"n": len(d), "uid": d["uid"].copy(), "rcut": d["rcut"].copy(), "rns": d["rns"].copy()}
        # the cell grid of this update, made when list_for first asks
        self._grid = d[["r", "rns"]].copy()

    def decide(self):
        """True when a rebuild is needed (verlet-list.cxx:110-158)."""
        d = self.atoms.data
        sv = self._saved
        if sv is None or len(d) != sv["n"]:
            return True
        if len(d) == 0:
            return False
        if (np.any(d["uid"] != sv["uid"]) or np.any(d["rcut"] != sv["rcut"])
                or np.any(d["rns"] != sv["rns"])):
            return True
        skin = sv["rns"] - sv["rcut"]
        if np.any(skin <= 0.0):
            return True
        self.system.upload(d, E.F_R)
        return self.system.decide(float(sv["rns"].min() - skin.min()))

    def get(self):
        return NeibsList(self)

    def _lists(self):
        if self._ids is None:
            if self.system is None:
                n = len(self.atoms)
                self._ids = (np.zeros(n + 1, dtype=np.int64), np.zeros(0, dtype=np.int64))
            else:
                self._ids = self.system.export()
        return self._ids

    def __getitem__(self, i):
        off, ids = self._lists()
        if not 0 <= i < len(off) - 1:
            return np.zeros(0, dtype=np.uint64)
        row = ids[off[i]:off[i + 1]]
        if self._half:
            row = row[row >= i]
        return row.astype(np.uint64)

    def sort(self, elems):
        """Records of ``elems`` wrapped into the box and ordered by cell
        (verlet-list.cxx:237-249, grid.cxx:43-118)."""
        d = elems.data
        self.domain.fit_in_period(elems)
        g = _Grid(self.domain, d)
        order = np.argsort(g.cell, kind="stable")
        out = Atoms(len(d))
        out.data[:] = d[order]
        return out

    def list_for(self, x, y=0.0, z=0.0):
        """Atoms listed for a point, with the cell edge as its search radius
        (verlet-list.cxx:251-276), in stencil order and, within a cell,
        in the reference's linked-list order (last inserted first)."""
        if self._grid is None:
            return np.zeros(0, dtype=np.uint64)
        if not isinstance(self._grid, _Grid):
            self._grid = _Grid(self.domain, self._grid)
        return self._grid.list_for(np.array([x, y, z], dtype=np.float64)).astype(np.uint64)


class _Grid:
    """The reference's cell grid over the wrapped positions (grid.cxx:43-95,
    :107-207), for sort() and list_for()."""

    def __init__(self, domain, d):
        self.dom = domain
        lo, hi = domain._lo, domain._hi
        size = hi - lo
        dim = domain._dim
        r = _rvec(d)
        n = len(d)
        self.dcells = np.ones(3, dtype=np.int64)
        self.csize = size.copy()
        if n == 0:
            self.offset = np.ones(3, dtype=np.int64)
            self.ncells = np.zeros(3, dtype=np.int64)
            self.cell = np.zeros(0, dtype=np.int64)
            return
        rmax = float(d["rns"].max())
        bmin = np.zeros(3)
        bmax = np.zeros(3)
        bmin[:dim] = r[:, :dim].min(axis=0)
        bmax[:dim] = r[:, :dim].max(axis=0)
        for a in range(dim):
            self.dcells[a] = max(1, int(np.floor(size[a] / rmax))) if rmax > 0 else 1
            self.csize[a] = size[a] / self.dcells[a]
        self.offset = self._global(bmin)
        upper = self._global(bmax)
        self.ncells = np.ones(3, dtype=np.int64)
        for a in range(dim):
            upper[a] = min(upper[a] + 1, self.dcells[a])
            self.ncells[a] = upper[a] - self.offset[a]
        loc = self._global(r) - self.offset
        self.cell = loc[:, 0] + self.ncells[0] * (loc[:, 1] + self.ncells[1] * loc[:, 2])
        self.r = r
        self.rns = d["rns"].copy()

    def _global(self, v):
        """global_index3D (grid.cxx:107-112) of one point or of rows."""
        v = np.asarray(v, dtype=np.float64)
        c = np.zeros(v.shape, dtype=np.int64)
        for a in range(self.dom._dim):
            c[..., a] = np.floor((v[..., a] - self.dom._lo[a]) / self.csize[a]).astype(np.int64)
        return c

    def list_for(self, p):
        dom = self.dom
        cell = self._global(p)
        per = dom._per
        cells = []
        for ox in (-1, 0, 1):
            if per[0] and ((self.dcells[0] == 1 and ox != 0) or (self.dcells[0] == 2 and ox == 1)):
                continue
            for oy in (-1, 0, 1):
                if per[1] and ((self.dcells[1] == 1 and oy != 0) or (self.dcells[1] == 2 and oy == 1)):
                    continue
                for oz in (-1, 0, 1):
                    if per[2] and ((self.dcells[2] == 1 and oz != 0) or (self.dcells[2] == 2 and oz == 1)):
                        continue
                    nb = cell + np.array([ox, oy, oz])
                    for a in range(3):
                        if per[a]:
                            nb[a] = (nb[a] + self.dcells[a]) % self.dcells[a]
                    nb -= self.offset
                    if np.all((0 <= nb) & (nb < self.ncells)):
                        cells.append(nb[0] + self.ncells[0] * (nb[1] + self.ncells[1] * nb[2]))
        edge = [False, False, False]
        for a in range(dom._dim):
            edge[a] = cell[a] == 0 or cell[a] == self.ncells[a] - 1
        size = dom._hi - dom._lo
        out = []
        rad = float(self.csize.max())  # Grid::cell_size(), grid.h:116
        for c in cells:
            members = np.nonzero(self.cell == c)[0][::-1]
            for j in members:
                dr = p - self.r[j]
                for a in range(3):
                    if edge[a] and per[a] and 2 * dr[a] < -size[a]:
                        dr[a] += size[a]
                    if edge[a] and per[a] and 2 * dr[a] > size[a]:
                        dr[a] -= size[a]
                rr = max(rad, float(self.rns[j]))
                if dr @ dr <= rr * rr:
                    out.append(j)
        return np.array(out, dtype=np.int64)


# --------------------------------------------------------------------------
# mdcraft.solver.potential (python/mdcraft/solver/potential/_potential.cxx)
# --------------------------------------------------------------------------
class Potential:
    """Base of the engine's potentials; ``rcut`` as the reference's."""

    _engine = None

    @property
    def rcut(self):
        return self._engine.rcut

    def engine_potential(self):
        return self._engine


class LJsCut(Potential):
    """Truncated LJ without shift (LJs-cut.cxx:8-131)."""

    def __init__(self, a, s, rcut):
        self._engine = E.LJsCut(context(), a, s, rcut)


class EAM(Potential):
    """EAM(filename) (single-element setfl) or EAM(Rcutoff, pairPotential,
    chargeDensity, embeddingEnergy, NrPot, drPot, NrEmb, drEmb), tables in
    the reference's units (_potential.cxx:112-146, EAM.cxx:15-87)."""

    def __init__(self, Rcutoff=None, pairPotential=None, chargeDensity=None, embeddingEnergy=None, NrPot=None,
                 drPot=None, NrEmb=None, drEmb=None, filename=None):
        if isinstance(Rcutoff, (str, os.PathLike)) and filename is None:
            filename, Rcutoff = Rcutoff, None
        if filename is not None:
            self._engine = E.EAM(context(), path=os.fspath(filename))
            return
        nr, nrho = int(NrPot), int(NrEmb)
        rphi = np.asarray(pairPotential, dtype=np.float64)[:nr]
        rho = np.asarray(chargeDensity, dtype=np.float64)[:nr]
        F = np.asarray(embeddingEnergy, dtype=np.float64)[:nrho]
        if len(rphi) < nr or len(rho) < nr or len(F) < nrho:
            raise ValueError("tables shorter than NrPot / NrEmb")
        self._engine = E.EAM(context(), rcut=Rcutoff, F=F, rho=rho, rphi=rphi, dr=drPot, drho=drEmb)


class EAMAlloy(Potential):
    """Alloy EAM of 1..4 elements from an N-element setfl ("eam/alloy")
    file, element e acting on atoms of kind e: each atom sums the density
    function of every neighbour's element, embeds with its own element's
    function and adds the pair term of the two elements.  Not in the
    reference, whose EAM(filename) reads one element (EAM.cxx:55-87); it
    plugs into Single like any potential."""

    def __init__(self, filename):
        self._engine = E.Alloy(context(), path=os.fspath(filename))


class PotentialTernary(Potential):
    pass


class MaiseBPNNPotential(PotentialTernary):
    """MAISE BPNN model file (maise_bpnn.cxx, tools/parser/maise_model.cxx);
    images come from the domain of the list the solver is called with."""

    def __init__(self, pot_filename, domain):
        self.domain = domain
        self._engine = E.MaiseBPNNPotential(context(), os.fspath(pot_filename))


def _not_in_engine(name):
    class _Missing(Potential):
        def __init__(self, *args, **kwargs):
            raise NotImplementedError(name + " is not implemented by the GPU engine")

    _Missing.__name__ = name
    return _Missing


class LJs(Potential):
    """LJs(aVr, rVr, Rcutoff): LJ with a polynomial tail taking V and V' to
    zero at the cutoff (LJs.cxx:7-256)."""

    def __init__(self, aVr, rVr, Rcutoff):
        self._engine = E.LJs(context(), aVr, rVr, Rcutoff)


class TabulatedPair(Potential):
    """mdcraft.solver.potential.Pair: Pair(Rcutoff, pairPotential, NrPot,
    drPot) with r*phi in reference units, or Pair(filename)
    (pair.cxx:15-56)."""

    def __init__(self, Rcutoff=None, pairPotential=None, NrPot=None, drPot=None, filename=None):
        if isinstance(Rcutoff, (str, os.PathLike)) and filename is None:
            filename, Rcutoff = Rcutoff, None
        if filename is not None:
            self._engine = E.Pair(context(), path=os.fspath(filename))
            return
        nr = int(NrPot)
        rphi = np.asarray(pairPotential, dtype=np.float64)[:nr]
        if len(rphi) < nr:
            raise ValueError("table shorter than NrPot")
        self._engine = E.Pair(context(), rcut=Rcutoff, rphi=rphi, dr=drPot)


DeepModelingPotential = _not_in_engine("DeepModelingPotential")
MLIP4 = _not_in_engine("MLIP4")


# --------------------------------------------------------------------------
# mdcraft.solver.boundary / thermostat (_boundary.cxx, _thermostat.cxx)
# --------------------------------------------------------------------------
class Boundary:
    """Null boundary."""


class Periodic(Boundary):
    """boundary::Periodic (periodic.cxx:9-102): one image per listed pair."""

    def __init__(self, domain):
        self.domain = domain


class Thermostat:
    """Null thermostat."""


class Langevin(Thermostat):
    """Langevin(beta, temperature, time_step, heat_x, heat_y, heat_z, Ux,
    Uy, Uz, xmin, xmax) (thermostat/langevin.cxx:7-42)."""

    def __init__(self, beta, temperature, time_step, heat_x=1, heat_y=1, heat_z=1, Ux=0.0, Uy=0.0, Uz=0.0,
                 xmin=-1e15, xmax=1e15, seed=None):
        # the reference seeds from std::random_device; ``seed`` (an
        # addition) makes a run reproducible
        self.seed = int.from_bytes(os.urandom(8), "little") if seed is None else int(seed)
        self._beta, self._T, self._dt = float(beta), float(temperature), float(time_step)
        self._heat = [int(heat_x), int(heat_y), int(heat_z)]
        self._U = [float(Ux), float(Uy), float(Uz)]
        self._xmin, self._xmax = float(xmin), float(xmax)

    temperature = property(lambda self: self._T)
    friction = property(lambda self: self._beta)
    Ux = property(lambda self: self._U[0])
    Uy = property(lambda self: self._U[1])
    Uz = property(lambda self: self._U[2])
    xmin = property(lambda self: self._xmin)
    xmax = property(lambda self: self._xmax)

    def set_average_velocity(self, Ux=0.0, Uy=0.0, Uz=0.0):
        self._U = [float(Ux), float(Uy), float(Uz)]

    def set_temperature(self, temperature):
        self._T = float(temperature)

    def set_friction(self, beta):
        self._beta = float(beta)

    def set_heating_axes(self, heat_x=1, heat_y=1, heat_z=1):
        self._heat = [int(heat_x), int(heat_y), int(heat_z)]

    def set_dimensions(self, xmin=-1e15, xmax=1e15):
        self._xmin, self._xmax = float(xmin), float(xmax)

    def row(self):
        """The engine's parameter row (langevin.h:16-28 order)."""
        return (self._beta, self._T, self._dt, *self._heat, *self._U, self._xmin, self._xmax)


class Compound(Thermostat):
    """Thermostats applied one after the other (compound.cxx)."""

    def __init__(self):
        self.members = []

    def add(self, thermostat):
        self.members.append(thermostat)


def _langevins(th):
    if th is None or type(th) is Thermostat:
        return []
    if isinstance(th, Compound):
        out = []
        for m in th.members:
            out += _langevins(m)
        return out
    if isinstance(th, Langevin):
        return [th]
    raise TypeError("unknown thermostat")


# --------------------------------------------------------------------------
# mdcraft.solver (python/mdcraft/solver/_solver.cxx)
# --------------------------------------------------------------------------
class ISolver:
    pass


class _EngineSolver(ISolver):
    """Shared call path: bind the list's device system, upload the host
    records, run one stage, write back what the stage writes."""

    periodic = False
    precision = E.FP64

    def _engine_pot(self):
        raise NotImplementedError

    def _bind(self, atoms, neibs, nlist):
        if not isinstance(nlist, NeibsList):
            raise TypeError("nlist must come from VerletList.get()")
        vl = nlist.owner
        if atoms is not vl.atoms or neibs is not vl.atoms:
            raise ValueError("atoms and neibs must be the Atoms the list was built for")
        if vl.system is None or len(atoms) != vl.system.n:
            raise RuntimeError("the neighbour list is not built for these atoms: call update()")
        s = vl.system
        s.set_boundary(self.periodic)
        th = _langevins(self.thermostat)
        key = (tuple(t.row() for t in th), th[0].seed if th else getattr(s, "_therm_seed", 0))
        if getattr(s, "_therm_key", None) != key:
            # one stream per system; a Compound's members take its first seed
            s.set_thermostat([t.row() for t in th], key[1])
            s._therm_key, s._therm_seed = key, key[1]
        return s

    def _stage(self, atoms, neibs, nlist, stage, fields):
        s = self._bind(atoms, neibs, nlist)
        d = atoms.data
        s.upload(d, _UP)
        s.stage(self._engine_pot(), stage, self.precision)
        out = s.download(fields=fields)
        for f in _names(fields):
            d[f] = out[f]

    def prepare(self, atoms, neibs, nlist):
        self._stage(atoms, neibs, nlist, E.PREPARE,
                    E.F_F | E.F_VIR | E.F_EP | E.F_EM | E.F_NC | E.F_RCUT | E.F_BC)

    def charge(self, atoms, neibs, nlist):
        self._stage(atoms, neibs, nlist, E.CHARGE, E.F_NC)

    def forces(self, atoms, neibs, nlist):
        self._stage(atoms, neibs, nlist, E.FORCES, E.F_F)

    def virials(self, atoms, neibs, nlist):
        self._stage(atoms, neibs, nlist, E.VIRIALS, E.F_VIR | E.F_EP | E.F_EM)

    def _step(self, atoms, nlist, dt):
        s = self._bind(atoms, atoms, nlist)
        d = atoms.data
        s.upload(d, _UP)
        s.step(self._engine_pot(), dt, 1, self.precision)
        out = s.download(fields=_STEP_OUT)
        for f in _names(_STEP_OUT):
            d[f] = out[f]


_FIELD_NAMES = [(E.F_R, "r"), (E.F_V, "v"), (E.F_F, "f"), (E.F_VIR, "V"), (E.F_M, "m"), (E.F_RCUT, "rcut"),
                (E.F_RNS, "rns"), (E.F_EP, "Ep"), (E.F_EM, "Em"), (E.F_NC, "nc"), (E.F_BC, "bc"), (E.F_KIND, "kind"),
                (E.F_UID, "uid"), (E.F_TAG, "tag")]
_STEP_OUT = E.F_R | E.F_V | E.F_F | E.F_VIR | E.F_EP | E.F_EM | E.F_NC | E.F_RCUT | E.F_BC


def _names(fields):
    return [nm for bit, nm in _FIELD_NAMES if fields & bit]


class Single(_EngineSolver):
    """solver::Single(potential, boundary, thermostat, threads)
    (single.cxx:11-83); ``precision`` (engine FP64 or MIXED) is an
    addition, FP64 by default."""

    def __init__(self, potential, boundary=None, thermostat=None, threads=_dummy_pool, precision=E.FP64):
        if isinstance(potential, PotentialTernary):
            raise TypeError("BPNN potentials take the Ternary solver")
        self.potential, self.boundary, self.thermostat, self.threads = potential, boundary, thermostat, threads
        self.periodic = isinstance(boundary, Periodic)
        self.precision = precision

    def _engine_pot(self):
        return self.potential.engine_potential()


class Multi(_EngineSolver):
    """solver::Multi(boundary, thermostat, nkinds, threads) with
    add_potential(i, j, pot) (multi.cxx:11-151), over EAM objects."""

    def __init__(self, boundary=None, thermostat=None, nkinds=1, threads=_dummy_pool, precision=E.FP64):
        self.boundary, self.thermostat, self.nkinds, self.threads = boundary, thermostat, int(nkinds), threads
        self.periodic = isinstance(boundary, Periodic)
        self.precision = precision
        self._pairs = {}
        self._pot = None

    def add_potential(self, i, j, pot):
        self._pairs[(int(i), int(j))] = pot
        self._pot = None

    def _engine_pot(self):
        if self._pot is None:
            pairs = [(i, j, p.engine_potential()) for (i, j), p in self._pairs.items()]
            self._pot = E.Multi(context(), self.nkinds, pairs)
        return self._pot


class Ternary(_EngineSolver):
    """solver::Ternary(potential, thermostat, threads) (ternary.cxx:11-67)
    for MaiseBPNNPotential; images by the minimum image of the list's
    periodic domain, no ghost pass."""

    def __init__(self, potential, thermostat=None, threads=_dummy_pool, precision=E.FP64):
        self.potential, self.thermostat, self.threads = potential, thermostat, threads
        self.boundary = None
        self.precision = precision

    def _engine_pot(self):
        return self.potential.engine_potential()

    def charge(self, atoms, neibs, nlist):
        raise AttributeError("Ternary has no charge stage")

    def prepare(self, atoms, neibs, nlist):
        self._stage(atoms, neibs, nlist, E.PREPARE, E.F_F | E.F_VIR | E.F_EP | E.F_RCUT)


# --------------------------------------------------------------------------
# mdcraft.solver.stepper (python/mdcraft/solver/stepper/_stepper.cxx)
# --------------------------------------------------------------------------
class Stepper:
    pass


class Verlet(Stepper):
    """Velocity Verlet (stepper/verlet.cxx:36-85): half kick, drift,
    prepare + stages of the solver, half kick.  The list is the caller's:
    no rebuild happens inside a step."""

    def __init__(self, solver, threads=_dummy_pool):
        if not isinstance(solver, _EngineSolver):
            raise TypeError("solver must be Single, Multi or Ternary")
        self.solver, self.threads = solver, threads

    def make_step(self, atoms, nlist, dt):
        self.solver._step(atoms, nlist, dt)
