"""Python host layer over the engine's C ABI (include/mdc.h, libmdc.so).

The classes mirror the reference's names and argument meanings where the
path crosses into the engine: ``LJsCut(a, s, rcut)`` and ``EAM(...)``
(python/mdcraft/solver/potential/_potential.cxx), ``MaiseBPNNPotential``,
a ``VerletList``-style update/decide/export on a resident ``System``, and
``Verlet``-style stepping.  Errors raise ``RuntimeError`` with the engine's
message, as the reference's pybind11 layer turns C++ exceptions into
``RuntimeError``.

The HIP library is the only compute path: importing this module fails when
``libmdc.so`` is missing, and there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

from .atoms import ATOM_DTYPE

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmdc.so")

FP64, MIXED = 0, 1
FORCE, ENERGY, VIRIAL = 1, 2, 4
PREPARE, CHARGE, FORCES, VIRIALS = 0, 1, 2, 3
F_R, F_V, F_F, F_VIR, F_M, F_RCUT, F_RNS, F_EP, F_EM, F_NC, F_BC, F_KIND, F_UID, F_TAG = (1 << k for k in range(14))
F_ALL = (1 << 14) - 1

EXPORTS = [
    "mdc_ctx_create", "mdc_ctx_destroy", "mdc_ctx_last_error", "mdc_last_error", "mdc_atom_record_size",
    "mdc_pot_lj_create", "mdc_pot_ljs_create", "mdc_pot_pair_create", "mdc_pot_pair_create_file", "mdc_pot_eam_create", "mdc_pot_eam_create_file", "mdc_pot_multi_create",
    "mdc_pot_alloy_create", "mdc_pot_alloy_create_file",
    "mdc_pot_bpnn_create", "mdc_pot_bpnn_describe", "mdc_pot_rcut", "mdc_pot_destroy",
    "mdc_sys_create", "mdc_sys_destroy", "mdc_sys_upload", "mdc_sys_download", "mdc_sys_download_range", "mdc_sys_set_active", "mdc_sys_natoms",
    "mdc_sys_set_lanes", "mdc_sys_autotune", "mdc_sys_set_block", "mdc_sys_autotune_tiles", "mdc_sys_set_graphs",
    "mdc_sys_graph_stats", "mdc_ctx_set_spin", "mdc_sys_time_kernels", "mdc_sys_set_rebuild", "mdc_sys_set_boundary", "mdc_sys_set_thermostat",
    "mdc_nlist_build", "mdc_nlist_decide", "mdc_nlist_optimize_rns", "mdc_nlist_export", "mdc_nlist_stats",
    "mdc_compute", "mdc_compute_noalloc", "mdc_stage", "mdc_step", "mdc_thermo", "mdc_sync", "mdc_timing", "mdc_bpnn_inspect",
    "mdc_copy_calibration", "mdc_sys_refill_box", "mdc_sys_step_path",
    "mdc_serve_create", "mdc_serve_add_potential", "mdc_serve_submit", "mdc_serve_wait", "mdc_serve_take",
    "mdc_serve_stats", "mdc_serve_destroy",
]


def load_library(path=LIB_PATH):
    if not os.path.exists(path):
        raise ImportError(
            "libmdc.so is not built (run __graft_entry__.build() or make -C mdcraft_amd/csrc); "
            "the engine has no CPU fallback")
    lib = C.CDLL(path)
    P, I, D, U32, U64 = C.c_void_p, C.c_int, C.c_double, C.c_uint32, C.c_uint64
    PP = C.POINTER(C.c_void_p)
    DP = C.POINTER(C.c_double)
    sig = {
        "mdc_ctx_create": (I, [I, PP]),
        "mdc_ctx_destroy": (I, [P]),
        "mdc_ctx_last_error": (C.c_char_p, [P]),
        "mdc_last_error": (C.c_char_p, []),
        "mdc_atom_record_size": (I, []),
        "mdc_pot_lj_create": (I, [P, D, D, D, PP]),
        "mdc_pot_ljs_create": (I, [P, D, D, D, PP]),
        "mdc_pot_pair_create": (I, [P, D, DP, I, D, PP]),
        "mdc_pot_pair_create_file": (I, [P, C.c_char_p, PP]),
        "mdc_pot_eam_create": (I, [P, D, DP, DP, DP, I, D, I, D, PP]),
        "mdc_pot_eam_create_file": (I, [P, C.c_char_p, PP]),
        "mdc_pot_multi_create": (I, [P, I, I, C.POINTER(I), C.POINTER(I), PP, PP]),
        "mdc_pot_alloy_create": (I, [P, I, D, DP, DP, DP, I, D, I, D, PP]),
        "mdc_pot_alloy_create_file": (I, [P, C.c_char_p, PP]),
        "mdc_pot_bpnn_create": (I, [P, C.c_char_p, PP]),
        "mdc_pot_bpnn_describe": (I, [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double)]),
        "mdc_pot_rcut": (I, [P, DP]),
        "mdc_pot_destroy": (I, [P]),
        "mdc_sys_create": (I, [P, I, C.POINTER(U64), DP, DP, C.POINTER(C.c_uint8), I, PP]),
        "mdc_sys_destroy": (I, [P]),
        "mdc_sys_upload": (I, [P, P, U32]),
        "mdc_sys_download": (I, [P, P, U32]),
        "mdc_sys_download_range": (I, [P, U64, U64, P, U32]),
        "mdc_sys_set_active": (I, [P, I, I]),
        "mdc_sys_natoms": (I, [P, C.POINTER(U64)]),
        "mdc_sys_set_lanes": (I, [P, I]),
        "mdc_sys_autotune": (I, [P, P, I, I, C.POINTER(I), DP]),
        "mdc_sys_set_block": (I, [P, I]),
        "mdc_sys_autotune_tiles": (I, [P, P, I, I, C.POINTER(I), C.POINTER(I), DP]),
        "mdc_sys_set_graphs": (I, [P, I]),
        "mdc_ctx_set_spin": (I, [P, I]),
        "mdc_sys_graph_stats": (I, [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
        "mdc_sys_time_kernels": (I, [P, P, I, D, I, DP]),
        "mdc_sys_set_rebuild": (I, [P, I]),
        "mdc_sys_set_boundary": (I, [P, I]),
        "mdc_sys_set_thermostat": (I, [P, I, DP, U64]),
        "mdc_nlist_build": (I, [P]),
        "mdc_nlist_optimize_rns": (I, [P, D, DP]),
        "mdc_stage": (I, [P, P, I, I]),
        "mdc_nlist_decide": (I, [P, D, C.POINTER(I)]),
        "mdc_nlist_export": (I, [P, C.POINTER(U64), C.POINTER(U64), C.POINTER(U64)]),
        "mdc_nlist_stats": (I, [P, DP, C.POINTER(I), C.POINTER(I), C.POINTER(I)]),
        "mdc_compute": (I, [P, P, U32, I]),
        "mdc_compute_noalloc": (I, [P, P, U32, I]),
        "mdc_copy_calibration": (I, [I, C.c_uint64, I, C.POINTER(C.c_double)]),
        "mdc_step": (I, [P, P, D, I, I, C.POINTER(I)]),
        "mdc_thermo": (I, [P, DP]),
        "mdc_sync": (I, [P]),
        "mdc_timing": (I, [P, I, DP, C.POINTER(C.c_int64)]),
        "mdc_bpnn_inspect": (I, [P, P, DP, DP, C.POINTER(I)]),
        "mdc_sys_refill_box": (I, [P, I, U64, P, DP, DP, C.POINTER(C.c_uint8)]),
        "mdc_sys_step_path": (I, [P, I, C.POINTER(I)]),
        "mdc_serve_create": (I, [I, D, I, I, PP]),
        "mdc_serve_add_potential": (I, [P, P, I, C.POINTER(I)]),
        "mdc_serve_submit": (I, [P, I, P, U64, DP, DP, C.POINTER(C.c_uint8), D, I, C.POINTER(U64)]),
        "mdc_serve_wait": (I, [P, U64, D, C.POINTER(I)]),
        "mdc_serve_take": (I, [P, U64, P, DP, DP]),
        "mdc_serve_stats": (I, [P, DP]),
        "mdc_serve_destroy": (I, [P]),
    }
    for name in EXPORTS:
        fn = getattr(lib, name)
        fn.restype, fn.argtypes = sig[name]
    if lib.mdc_atom_record_size() != ATOM_DTYPE.itemsize:
        raise ImportError("atom record size mismatch")
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = load_library()
    return _lib


def _check(status):
    if status != 0:
        raise RuntimeError(lib().mdc_last_error().decode() or f"mdc error {status}")


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Context:
    """One device and one ordered HIP stream."""

    def __init__(self, device=0):
        h = C.c_void_p()
        _check(lib().mdc_ctx_create(device, C.byref(h)))
        self.h = h
        self.device = device

    def set_spin(self, spin):
        """Wait for the device by polling (True) or by blocking (False)."""
        _check(lib().mdc_ctx_set_spin(self.h, int(bool(spin))))

    def close(self):
        if getattr(self, "h", None):
            lib().mdc_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class Potential:
    def __init__(self, ctx, h, kind):
        self.ctx, self.h, self.kind = ctx, h, kind

    @property
    def rcut(self):
        r = C.c_double()
        _check(lib().mdc_pot_rcut(self.h, C.byref(r)))
        return r.value

    def close(self):
        if getattr(self, "h", None):
            lib().mdc_pot_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def LJsCut(ctx, a, s, rcut):
    """Truncated LJ without energy shift, LJsCut(a, s, rcut)."""
    h = C.c_void_p()
    _check(lib().mdc_pot_lj_create(ctx.h, a, s, rcut, C.byref(h)))
    return Potential(ctx, h, "pair")


def LJs(ctx, aVr, rVr, rcut):
    """LJ with a smooth polynomial tail, LJs(aVr, rVr, Rcutoff)."""
    h = C.c_void_p()
    _check(lib().mdc_pot_ljs_create(ctx.h, aVr, rVr, rcut, C.byref(h)))
    return Potential(ctx, h, "pair")


def Pair(ctx, rcut=None, rphi=None, dr=None, path=None):
    """Tabulated pair potential: Pair(Rcutoff, Vpar, Nr1, dr) or
    Pair(filename)."""
    h = C.c_void_p()
    if path is not None:
        _check(lib().mdc_pot_pair_create_file(ctx.h, path.encode(), C.byref(h)))
    else:
        rphi = np.ascontiguousarray(rphi, dtype=np.float64)
        _check(lib().mdc_pot_pair_create(ctx.h, rcut, _dp(rphi), len(rphi), dr, C.byref(h)))
    return Potential(ctx, h, "pair")


def EAM(ctx, rcut=None, F=None, rho=None, rphi=None, dr=None, drho=None, path=None):
    """Tabulated EAM: EAM(Rcutoff, Femb, Rho, Vpar, Nr1, dr, Nrh1, drh) or
    EAM(setfl filename)."""
    h = C.c_void_p()
    if path is not None:
        _check(lib().mdc_pot_eam_create_file(ctx.h, path.encode(), C.byref(h)))
    else:
        F, rho, rphi = (np.ascontiguousarray(x, dtype=np.float64) for x in (F, rho, rphi))
        _check(lib().mdc_pot_eam_create(ctx.h, rcut, _dp(F), _dp(rho), _dp(rphi), len(rho), dr, len(F), drho,
                                        C.byref(h)))
    return Potential(ctx, h, "eam")


def Multi(ctx, nkinds, pairs):
    """solver::Multi(nkinds) with add_potential(i, j, pot) for every
    (i, j, pot) in pairs (solver/multi.cxx:11-49); members are EAM objects."""
    pairs = list(pairs)
    ki = (C.c_int * len(pairs))(*[p[0] for p in pairs])
    kj = (C.c_int * len(pairs))(*[p[1] for p in pairs])
    hs = (C.c_void_p * len(pairs))(*[p[2].h for p in pairs])
    h = C.c_void_p()
    _check(lib().mdc_pot_multi_create(ctx.h, nkinds, len(pairs), ki, kj, hs, C.byref(h)))
    return Potential(ctx, h, "eam")


def Alloy(ctx, rcut=None, F=None, rho=None, rphi=None, dr=None, drho=None, path=None):
    """Alloy EAM of 1..4 elements (element e is atom kind e) from an
    N-element setfl file or from tables: F [ne][nrho], rho [ne][nr], rphi
    [ne(ne+1)/2][nr] in the setfl pair order (1,1), (2,1), (2,2), ...
    Atom i of element a sums rho_b over neighbours of element b, embeds
    with F_a and adds the pair term of (a, b)."""
    h = C.c_void_p()
    if path is not None:
        _check(lib().mdc_pot_alloy_create_file(ctx.h, path.encode(), C.byref(h)))
    else:
        F, rho, rphi = (np.ascontiguousarray(x, dtype=np.float64) for x in (F, rho, rphi))
        ne = F.shape[0]
        if F.ndim != 2 or rho.shape[0] != ne or rphi.shape != (ne * (ne + 1) // 2, rho.shape[1]):
            raise ValueError("alloy tables: F [ne][nrho], rho [ne][nr], rphi [ne(ne+1)/2][nr]")
        _check(lib().mdc_pot_alloy_create(ctx.h, ne, rcut, _dp(F), _dp(rho), _dp(rphi), rho.shape[1], dr, F.shape[1],
                                          drho, C.byref(h)))
    return Potential(ctx, h, "eam")


def MaiseBPNNPotential(ctx, path):
    h = C.c_void_p()
    _check(lib().mdc_pot_bpnn_create(ctx.h, path.encode(), C.byref(h)))
    return Potential(ctx, h, "ternary")


def copy_calibration(device, nbytes, reps=3):
    """Mean microseconds of a plain device copy of nbytes (the calibration
    kernel of the traffic counters)."""
    us = C.c_double()
    _check(lib().mdc_copy_calibration(device, nbytes, reps, C.byref(us)))
    return us.value


def describe_bpnn(path):
    """Architecture, descriptor counts and cutoff (nm) of a MAISE model as
    the engine reads it, without a device."""
    arch, counts, rc = (C.c_int * 4)(), (C.c_int * 4)(), C.c_double()
    _check(lib().mdc_pot_bpnn_describe(path.encode(), arch, counts, C.byref(rc)))
    return {"arch": list(arch), "radial": counts[0], "angular": counts[1], "g4_eta": counts[2],
            "g4_groups": counts[3], "rcut": rc.value}


class System:
    """One or more independent boxes resident in HBM.

    ``boxes`` is a list of (lo[3], hi[3], periodic[3]); ``natoms`` the atom
    count of each box; atoms of box s are the records
    [sum(natoms[:s]), sum(natoms[:s+1])) of every atoms array passed in.
    ``periodic_boundary`` attaches the reference's boundary::Periodic ghost
    pass to the force evaluation.
    """

    def __init__(self, ctx, natoms, boxes, periodic_boundary=True, lanes=None):
        self.ctx = ctx
        natoms = np.asarray(natoms, dtype=np.uint64).reshape(-1)
        ns = len(natoms)
        lo = np.zeros((ns, 3))
        hi = np.zeros((ns, 3))
        per = np.zeros((ns, 3), dtype=np.uint8)
        for s, (l, h, p) in enumerate(boxes):
            lo[s], hi[s], per[s] = l, h, [1 if x else 0 for x in p]
        self.natoms = natoms
        self.nsys = ns
        self.n = int(natoms.sum())
        h = C.c_void_p()
        _check(lib().mdc_sys_create(ctx.h, ns, natoms.ctypes.data_as(C.POINTER(C.c_uint64)), _dp(lo), _dp(hi),
                                    per.ctypes.data_as(C.POINTER(C.c_uint8)), int(periodic_boundary), C.byref(h)))
        self.h = h
        if lanes is not None:
            self.set_lanes(lanes)

    def close(self):
        if getattr(self, "h", None):
            lib().mdc_sys_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_lanes(self, lanes):
        _check(lib().mdc_sys_set_lanes(self.h, lanes))

    def autotune(self, pot, precision, reps=5):
        """Time one force evaluation for 1, 2, 4, 8, 16 lanes per atom and
        keep the fastest; returns (lanes, {lanes: microseconds})."""
        best = C.c_int(0)
        us = np.zeros(5)
        _check(lib().mdc_sys_autotune(self.h, pot.h, precision, reps, C.byref(best), _dp(us)))
        return best.value, dict(zip((1, 2, 4, 8, 16), us.tolist()))

    def set_block(self, threads):
        """Threads per workgroup (64, 128, 256) of the LJ and EAM kernels."""
        _check(lib().mdc_sys_set_block(self.h, threads))

    def autotune_tiles(self, pot, precision, reps=5):
        """Time the step kernels for every lanes-per-atom x threads-per-
        workgroup tile and keep the fastest; returns (lanes, threads,
        {(lanes, threads): microseconds})."""
        bl, bt = C.c_int(0), C.c_int(0)
        us = np.zeros(15)
        _check(lib().mdc_sys_autotune_tiles(self.h, pot.h, precision, reps, C.byref(bl), C.byref(bt), _dp(us)))
        tiles = [(l, t) for l in (1, 2, 4, 8, 16) for t in (64, 128, 256)]
        return bl.value, bt.value, dict(zip(tiles, us.tolist()))

    def set_graphs(self, on):
        """Graph mode of compute(): each configuration's evaluation captured
        once and replayed."""
        _check(lib().mdc_sys_set_graphs(self.h, int(bool(on))))

    def graph_stats(self):
        cap, lau = C.c_uint64(), C.c_uint64()
        _check(lib().mdc_sys_graph_stats(self.h, C.byref(cap), C.byref(lau)))
        return {"captures": cap.value, "launches": lau.value}

    def time_kernels(self, pot, precision, dt, reps=100):
        """Mean step-mode (density, force) kernel durations in microseconds,
        each from reps back-to-back launches between two HIP events."""
        us = np.zeros(2)
        _check(lib().mdc_sys_time_kernels(self.h, pot.h, precision, dt, reps, _dp(us)))
        return float(us[0]), float(us[1])

    def upload(self, atoms, fields=F_ALL):
        atoms = np.ascontiguousarray(atoms, dtype=ATOM_DTYPE)
        assert len(atoms) == self.n
        _check(lib().mdc_sys_upload(self.h, atoms.ctypes.data, fields))

    def download(self, atoms=None, fields=F_ALL):
        if atoms is None:
            atoms = np.zeros(self.n, dtype=ATOM_DTYPE)
        assert atoms.dtype == ATOM_DTYPE and atoms.flags.c_contiguous and len(atoms) == self.n
        _check(lib().mdc_sys_download(self.h, atoms.ctypes.data, fields))
        return atoms

    def download_range(self, first, count, fields=F_ALL):
        """Records [first, first + count) of the upload order."""
        atoms = np.zeros(count, dtype=ATOM_DTYPE)
        _check(lib().mdc_sys_download_range(self.h, first, count, atoms.ctypes.data, fields))
        return atoms

    def download_box(self, box, fields=F_ALL):
        off = np.concatenate([[0], np.cumsum(self.natoms)]).astype(np.int64)
        return self.download_range(int(off[box]), int(off[box + 1] - off[box]), fields)

    def set_active(self, box, active):
        _check(lib().mdc_sys_set_active(self.h, box, int(bool(active))))

    def refill_box(self, box, atoms, box_geometry):
        """Box `box` takes the records `atoms` (at most its record count) in
        the geometry (lo, hi, periodic); its other records become void."""
        lo, hi, per = box_geometry
        lo = np.ascontiguousarray(lo, dtype=np.float64)
        hi = np.ascontiguousarray(hi, dtype=np.float64)
        pf = np.ascontiguousarray([1 if p else 0 for p in per], dtype=np.uint8)
        a = np.ascontiguousarray(atoms, dtype=ATOM_DTYPE)
        _check(lib().mdc_sys_refill_box(self.h, box, len(a), a.ctypes.data if len(a) else None, _dp(lo), _dp(hi),
                                        pf.ctypes.data_as(C.POINTER(C.c_uint8))))

    def build(self):
        _check(lib().mdc_nlist_build(self.h))

    def optimize_rns(self, kbuf):
        """rns_i = kbuf * max over the current list of rcut_i + rcut_j
        (update(kbuf > 0)); returns the new radii in upload order."""
        out = np.zeros(self.n)
        _check(lib().mdc_nlist_optimize_rns(self.h, kbuf, _dp(out)))
        return out

    def set_rebuild(self, automatic):
        """False: the list is rebuilt only by build(), as the reference's
        callers do; True (default): step() rebuilds on the half-skin trigger."""
        _check(lib().mdc_sys_set_rebuild(self.h, int(bool(automatic))))

    def set_boundary(self, periodic_boundary):
        _check(lib().mdc_sys_set_boundary(self.h, int(bool(periodic_boundary))))

    def set_thermostat(self, langevins, seed=0):
        """langevins: list of (beta, temperature, time_step, heat_x, heat_y,
        heat_z, Ux, Uy, Uz, xmin, xmax); [] detaches."""
        q = np.ascontiguousarray(np.asarray(langevins, dtype=np.float64).reshape(-1))
        _check(lib().mdc_sys_set_thermostat(self.h, len(langevins), _dp(q) if len(langevins) else None,
                                            int(seed) & (2 ** 64 - 1)))

    def stage(self, pot, stage, precision=FP64):
        """One solver call: PREPARE, CHARGE, FORCES or VIRIALS."""
        _check(lib().mdc_stage(self.h, pot.h, stage, precision))

    def decide(self, rcut):
        need = C.c_int()
        _check(lib().mdc_nlist_decide(self.h, rcut, C.byref(need)))
        return bool(need.value)

    def export(self):
        tot = C.c_uint64()
        _check(lib().mdc_nlist_export(self.h, None, None, C.byref(tot)))
        off = np.zeros(self.n + 1, dtype=np.uint64)
        ids = np.zeros(max(tot.value, 1), dtype=np.uint64)
        _check(lib().mdc_nlist_export(self.h, off.ctypes.data_as(C.POINTER(C.c_uint64)),
                                      ids.ctypes.data_as(C.POINTER(C.c_uint64)), None))
        return off.astype(np.int64), ids[:tot.value].astype(np.int64)

    def stats(self):
        mean, mx, cap, reb = C.c_double(), C.c_int(), C.c_int(), C.c_int()
        _check(lib().mdc_nlist_stats(self.h, C.byref(mean), C.byref(mx), C.byref(cap), C.byref(reb)))
        return {"mean": mean.value, "max": mx.value, "capacity": cap.value, "rebuilds": reb.value}

    def compute(self, pot, what=FORCE | ENERGY | VIRIAL, precision=FP64):
        _check(lib().mdc_compute(self.h, pot.h, what, precision))

    def compute_noalloc(self, pot, what=FORCE | ENERGY | VIRIAL, precision=FP64):
        """mdc_compute for a BPNN potential with its work arrays in the
        library's fixed per-device work area (allocates nothing; the list
        must be built)."""
        _check(lib().mdc_compute_noalloc(self.h, pot.h, what, precision))

    def step(self, pot, dt, nsteps=1, precision=FP64):
        reb = C.c_int()
        _check(lib().mdc_step(self.h, pot.h, dt, nsteps, precision, C.byref(reb)))
        return reb.value

    def thermo(self):
        """Per system: [Ep, Em, KE, Etot, W, com_x, com_y, com_z]."""
        out = np.zeros((self.nsys, 8))
        _check(lib().mdc_thermo(self.h, _dp(out)))
        return out

    def sync(self):
        _check(lib().mdc_sync(self.h))

    def bpnn_inspect(self, pot):
        """(G, dE/dG) of the last BPNN evaluation, [n, nin] each."""
        nin = C.c_int()
        _check(lib().mdc_bpnn_inspect(self.h, pot.h, None, None, C.byref(nin)))
        G = np.zeros((self.n, nin.value))
        dE = np.zeros((self.n, nin.value))
        _check(lib().mdc_bpnn_inspect(self.h, pot.h, _dp(G), _dp(dE), None))
        return G, dE

    def step_path(self, set=-1):
        """Step path of single-element EAM in mixed precision: where every
        box has a unique minimum image, step sequences take the fixed-point
        kernels with tables read from global memory (set 0) or staged in
        LDS (set 2); 1 keeps the FP64-difference kernels; -1 leaves the
        choice.  Returns what the last step sequence (or kernel timing) ran:
        0 the FP64-difference kernels, 1 fixed point with global tables, 2
        fixed point with tables in LDS."""
        last = C.c_int(0)
        _check(lib().mdc_sys_step_path(self.h, set, C.byref(last)))
        return last.value

    def timing(self, enable=-1):
        ms = np.zeros(9)
        n = np.zeros(9, dtype=np.int64)
        _check(lib().mdc_timing(self.h, enable, _dp(ms), n.ctypes.data_as(C.POINTER(C.c_int64))))
        return ms, n
