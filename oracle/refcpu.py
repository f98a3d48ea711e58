"""ctypes binding of the FP64 CPU oracle (oracle/librefcpu.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
the cpu_baseline leg of bench.py, always as the checker.  The engine never
imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "librefcpu.so")


class RefDomain(C.Structure):
    _fields_ = [("lo", C.c_double * 3), ("hi", C.c_double * 3), ("periodic", C.c_int32 * 3), ("pad", C.c_int32)]


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def _load():
    if not os.path.exists(_LIB):
        build()
    lib = C.CDLL(_LIB)
    P, D, I, S, Z = C.c_void_p, C.c_double, C.c_int, C.c_size_t, C.POINTER(C.c_double)
    dom = C.POINTER(RefDomain)
    sig = {
        "ref_last_error": (C.c_char_p, []),
        "ref_atom_size": (I, []),
        "ref_pool_create": (P, [I]),
        "ref_pool_free": (None, [P]),
        "ref_lj_create": (P, [D, D, D]),
        "ref_eam_create": (P, [D, Z, Z, Z, I, D, I, D]),
        "ref_eam_create_file": (P, [C.c_char_p]),
        "ref_alloy_create": (P, [I, D, Z, Z, Z, I, D, I, D]),
        "ref_alloy_create_file": (P, [C.c_char_p]),
        "ref_bpnn_create": (P, [C.c_char_p, dom]),
        "ref_ljs_create": (P, [D, D, D]),
        "ref_pair_create": (P, [D, Z, I, D]),
        "ref_multi_create": (P, [I]),
        "ref_multi_add": (I, [P, I, I, P]),
        "ref_pot_free": (None, [P]),
        "ref_pot_rcut": (D, [P]),
        "ref_pot_type": (I, [P]),
        "ref_eam_tables": (I, [P, Z, Z, Z, Z, Z, Z, Z]),
        "ref_bpnn_descriptors": (I, [P, S, Z, Z]),
        "ref_nlist_create": (P, [P, I]),
        "ref_nlist_free": (None, [P]),
        "ref_nlist_update": (I, [P, P, S, dom, D]),
        "ref_nlist_decide": (I, [P, P, S, dom]),
        "ref_nlist_total": (S, [P]),
        "ref_nlist_export": (None, [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
        "ref_grid_build": (I, [P, S, dom, Z, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "ref_grid_adjacent": (I, [P, S, dom, Z, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "ref_domain_op": (I, [dom, I, Z, I, Z]),
        "ref_prepare": (I, [P, P, P, S, dom, P, I]),
        "ref_stages": (I, [P, P, P, S, dom, P, I]),
        "ref_thermo_create": (P, [I, Z, C.c_uint64]),
        "ref_thermo_free": (None, [P]),
        "ref_thermo_use": (None, [P]),
        "ref_thermo_calls": (C.c_uint64, [P]),
        "ref_philox": (None, [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
        "ref_normals": (None, [C.c_uint64, C.c_uint64, C.c_uint32, I, Z]),
        "ref_charge": (I, [P, P, P, S, dom, P, I]),
        "ref_forces": (I, [P, P, P, S, dom, P, I]),
        "ref_virials": (I, [P, P, P, S, dom, P, I]),
        "ref_make_step": (I, [P, P, P, S, dom, P, I, D]),
        "ref_measure_state": (None, [P, S, Z]),
        "ref_run_md_steps": (I, [P, P, P, S, dom, P, I, D, I, I, Z]),
        "ref_run_nve": (I, [P, P, P, S, dom, P, I, D, I, I]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    assert lib.ref_atom_size() == 208
    return lib


lib = _load()


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _err():
    return lib.ref_last_error().decode()


class Domain:
    """Box [lo, hi] with per-axis periodicity (lattice::Domain)."""

    def __init__(self, lo, hi, periodic=(False, False, False)):
        self.lo = np.array(lo, dtype=np.float64)
        self.hi = np.array(hi, dtype=np.float64)
        self.periodic = tuple(bool(p) for p in periodic)
        self.c = RefDomain()
        for k in range(3):
            self.c.lo[k] = self.lo[k]
            self.c.hi[k] = self.hi[k]
            self.c.periodic[k] = int(self.periodic[k])

    def ref(self):
        return C.byref(self.c)

    def fit_in_period(self, v):
        v = np.array(v, dtype=np.float64)
        lib.ref_domain_op(self.ref(), 0, _dp(v), 0, None)
        return v

    def shortest(self, v):
        v = np.array(v, dtype=np.float64)
        lib.ref_domain_op(self.ref(), 1, _dp(v), 0, None)
        return v

    def nearest_distance(self, v, axis):
        v = np.array(v, dtype=np.float64)
        out = np.zeros(1)
        lib.ref_domain_op(self.ref(), 2, _dp(v), axis, _dp(out))
        return float(out[0])


class Pool:
    def __init__(self, threads=1):
        self.h = lib.ref_pool_create(threads)
        self.threads = threads

    def __del__(self):
        if getattr(self, "h", None):
            lib.ref_pool_free(self.h)
            self.h = None


class Potential:
    def __init__(self, h):
        if not h:
            raise RuntimeError(_err())
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib.ref_pot_free(self.h)
            self.h = None

    @property
    def rcut(self):
        return lib.ref_pot_rcut(self.h)


def LJsCut(a, s, rcut):
    return Potential(lib.ref_lj_create(a, s, rcut))


def LJs(aVr, rVr, rcut):
    return Potential(lib.ref_ljs_create(aVr, rVr, rcut))


def Pair(rc, rphi, dr):
    rphi = np.ascontiguousarray(rphi, dtype=np.float64)
    return Potential(lib.ref_pair_create(rc, _dp(rphi), len(rphi), dr))


def EAM(rc=None, F=None, rho=None, rphi=None, dr=None, drho=None, path=None):
    if path is not None:
        return Potential(lib.ref_eam_create_file(path.encode()))
    F, rho, rphi = (np.ascontiguousarray(x, dtype=np.float64) for x in (F, rho, rphi))
    return Potential(lib.ref_eam_create(rc, _dp(F), _dp(rho), _dp(rphi), len(rho), dr, len(F), drho))


def Alloy(rc=None, F=None, rho=None, rphi=None, dr=None, drho=None, path=None):
    """N-element alloy EAM: F [ne][nrho], rho [ne][nr], rphi [ne(ne+1)/2][nr]
    in the setfl pair order, or an N-element setfl file."""
    if path is not None:
        return Potential(lib.ref_alloy_create_file(path.encode()))
    F, rho, rphi = (np.ascontiguousarray(x, dtype=np.float64) for x in (F, rho, rphi))
    ne = F.shape[0]
    if rho.shape[0] != ne or rphi.shape[0] != ne * (ne + 1) // 2 or rphi.shape[1] != rho.shape[1]:
        raise ValueError("alloy tables: F [ne][nrho], rho [ne][nr], rphi [ne(ne+1)/2][nr]")
    return Potential(lib.ref_alloy_create(ne, rc, _dp(F), _dp(rho), _dp(rphi), rho.shape[1], dr, F.shape[1], drho))


class Multi(Potential):
    """solver::Multi(nkinds) with add_potential(i, j, pot) (multi.cxx:11-49);
    keeps its member potentials alive."""

    def __init__(self, nkinds, pairs=()):
        super().__init__(lib.ref_multi_create(nkinds))
        self.nkinds = nkinds
        self.members = []
        for i, j, pot in pairs:
            self.add_potential(i, j, pot)

    def add_potential(self, i, j, pot):
        if lib.ref_multi_add(self.h, i, j, pot.h):
            raise RuntimeError(_err())
        self.members.append(pot)


def MaiseBPNN(path, domain):
    return Potential(lib.ref_bpnn_create(path.encode(), domain.ref()))


def eam_tables(pot):
    meta = np.zeros(5)
    lib.ref_eam_tables(pot.h, _dp(meta), None, None, None, None, None, None)
    nr, nrho = int(meta[3]), int(meta[4])
    t = {k: np.zeros(nrho if k in ("F", "dF") else nr) for k in ("F", "dF", "rho", "drho", "rphi", "drphi")}
    lib.ref_eam_tables(pot.h, _dp(meta), *(_dp(t[k]) for k in ("F", "dF", "rho", "drho", "rphi", "drphi")))
    t.update(rc=meta[0], dr=meta[1], ddens=meta[2])
    return t


def bpnn_descriptors(pot, n):
    nin = lib.ref_bpnn_descriptors(pot.h, n, None, None)
    G = np.zeros((n, nin))
    dE = np.zeros((n, nin))
    if lib.ref_bpnn_descriptors(pot.h, n, _dp(G), _dp(dE)) < 0:
        raise RuntimeError(_err())
    return G, dE


class VerletList:
    """Full (or half) Verlet list including self (neibs::VerletList)."""

    def __init__(self, pool, half=False):
        self.pool = pool
        self.h = lib.ref_nlist_create(pool.h, int(half))
        if not self.h:
            raise RuntimeError(_err())

    def __del__(self):
        if getattr(self, "h", None):
            lib.ref_nlist_free(self.h)
            self.h = None

    def update(self, atoms, domain, kbuf=-1.0):
        if lib.ref_nlist_update(self.h, atoms.ctypes.data, len(atoms), domain.ref(), kbuf):
            raise RuntimeError(_err())

    def decide(self, atoms, domain):
        return bool(lib.ref_nlist_decide(self.h, atoms.ctypes.data, len(atoms), domain.ref()))

    def export(self, n):
        tot = lib.ref_nlist_total(self.h)
        off = np.zeros(n + 1, dtype=np.uint64)
        ids = np.zeros(max(tot, 1), dtype=np.uint64)
        lib.ref_nlist_export(self.h, off.ctypes.data_as(C.POINTER(C.c_uint64)), ids.ctypes.data_as(C.POINTER(C.c_uint64)))
        return off.astype(np.int64), ids[:tot].astype(np.int64)

    def lists(self, n):
        off, ids = self.export(n)
        return [ids[off[i]:off[i + 1]] for i in range(n)]


class Thermostats:
    """Langevin thermostats applied in order (thermostat::Langevin,
    thermostat::Compound); rows of (beta, temperature, time_step, heat_x,
    heat_y, heat_z, Ux, Uy, Uz, xmin, xmax) and the stream seed."""

    def __init__(self, rows, seed):
        q = np.ascontiguousarray(np.asarray(rows, dtype=np.float64).reshape(-1))
        self._q = q
        self.h = lib.ref_thermo_create(len(rows), _dp(q), int(seed) & (2 ** 64 - 1))

    @property
    def calls(self):
        return lib.ref_thermo_calls(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib.ref_thermo_free(self.h)
            self.h = None


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    out = (C.c_uint32 * 4)()
    lib.ref_philox(c, k, out)
    return list(out)


def normals(seed, call, atom, k):
    g = np.zeros(3)
    lib.ref_normals(seed, call, atom, k, _dp(g))
    return g


class Solver:
    """Single / Ternary solver bound to one potential (solver::Single,
    solver::Ternary); ``periodic_boundary`` attaches boundary::Periodic,
    ``thermostats`` a Thermostats object."""

    def __init__(self, pot, pool, periodic_boundary=True, thermostats=None):
        self.pot, self.pool, self.bc = pot, pool, int(periodic_boundary)
        self.thermostats = thermostats

    def _call(self, fn, atoms, domain, nlist, *extra):
        lib.ref_thermo_use(self.thermostats.h if self.thermostats is not None else None)
        try:
            if fn(self.pot.h, self.pool.h, atoms.ctypes.data, len(atoms), domain.ref(), nlist.h, self.bc, *extra):
                raise RuntimeError(_err())
        finally:
            lib.ref_thermo_use(None)

    def prepare(self, atoms, domain, nlist):
        self._call(lib.ref_prepare, atoms, domain, nlist)

    def stages(self, atoms, domain, nlist):
        self._call(lib.ref_stages, atoms, domain, nlist)

    def charge(self, atoms, domain, nlist):
        self._call(lib.ref_charge, atoms, domain, nlist)

    def forces(self, atoms, domain, nlist):
        self._call(lib.ref_forces, atoms, domain, nlist)

    def virials(self, atoms, domain, nlist):
        self._call(lib.ref_virials, atoms, domain, nlist)

    def compute(self, atoms, domain, nlist, virials=True):
        self.prepare(atoms, domain, nlist)
        self.stages(atoms, domain, nlist)
        if virials:
            self.virials(atoms, domain, nlist)

    def make_step(self, atoms, domain, nlist, dt):
        self._call(lib.ref_make_step, atoms, domain, nlist, dt)

    def run_md_steps(self, atoms, domain, nlist, dt, steps, interval=1):
        hist = np.zeros((steps + 1, 8))
        self._call(lib.ref_run_md_steps, atoms, domain, nlist, dt, steps, interval, _dp(hist))
        return hist

    def run_nve(self, atoms, domain, nlist, dt, steps, initial_forces=True):
        lib.ref_thermo_use(self.thermostats.h if self.thermostats is not None else None)
        try:
            r = lib.ref_run_nve(self.pot.h, self.pool.h, atoms.ctypes.data, len(atoms), domain.ref(), nlist.h,
                                self.bc, dt, steps, int(initial_forces))
        finally:
            lib.ref_thermo_use(None)
        if r < 0:
            raise RuntimeError(_err())
        return r


def measure_state(atoms):
    """[Ep, Em, KE, Etot, W, com_x, com_y, com_z] (md_test_utils.h:214-239)."""
    out = np.zeros(8)
    lib.ref_measure_state(atoms.ctypes.data, len(atoms), _dp(out))
    return out
