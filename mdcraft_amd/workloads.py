"""Seeded synthetic workloads in the reference's units (nm, ps, amu, kJ/mol).

Random numbers come from a counter-based SplitMix64 stream with Box-Muller
normals, so every structure is reproducible from (base seed, workload id,
index) on any platform, independent of libstdc++ or numpy generators.

Workloads (sizes as benchmarked):
  lj_argon      fcc Ar a = 5.26 A, 10^3 cells = 4000 atoms, LJ eps = 0.0104 eV,
                sigma = 3.405 A, rc = 2.5 sigma, skin 0.3 A, 100 K, dt = 2 fs
  eam_cu        fcc Cu a = 3.615 A, 20^3 cells = 32000 atoms, rattle 0.03 A,
                300 K, analytic Morse/exponential/sqrt EAM tabulated on
                5000-point grids to 5.0 A, dt = 1 fs
  bpnn_cu       fcc Cu 15^3 cells = 13500 atoms, rattle 0.05 A, G2/G4 + a
                16-30-30-1 tanh network with Xavier weights, dt = 1 fs
  nial_batch    512 B2 NiAl cells (a = 2.88 A), atom counts log-uniform in
                [32, 4096], normal strain up to 2 %, rattle 0.05 A, alloy EAM
"""
import math
import os

import numpy as np

from .atoms import EV, KB, make_atoms, set_vec

BASE_SEED = 20261203
ANG = 0.1  # nm per Angstrom
FS = 1e-3  # ps per fs
_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(z):
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def splitmix64(x):
    """One SplitMix64 output for state x (Python int)."""
    with np.errstate(over="ignore"):
        return int(_mix(np.uint64((int(x) + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF)))


class Rng:
    """Counter-based SplitMix64 stream."""

    def __init__(self, seed):
        self.state = np.uint64(int(seed) & 0xFFFFFFFFFFFFFFFF)
        self.counter = 0

    def bits(self, n):
        with np.errstate(over="ignore"):
            k = np.arange(self.counter + 1, self.counter + 1 + n, dtype=np.uint64)
            z = self.state + k * _GAMMA
            self.counter += n
            return _mix(z)

    def uniform(self, n):
        """Uniform doubles in (0, 1)."""
        return ((self.bits(n) >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)

    def normal(self, n):
        m = (n + 1) // 2
        u1 = self.uniform(m)
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log(u1))
        z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
        return z[:n]


def structure_seed(workload_id, index, base=BASE_SEED):
    return splitmix64(base ^ (workload_id * 0x100000001B3) ^ (index * 0x9E3779B1))


# --------------------------------------------------------------------------
# lattices
# --------------------------------------------------------------------------
FCC = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
B2 = np.array([[0, 0, 0], [0.5, 0.5, 0.5]])


def lattice(basis, reps, a):
    """Positions (nm) of a replicated cubic cell with origin at 0; returns
    positions and the basis index of every site."""
    nx, ny, nz = reps
    idx = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1).reshape(-1, 3)
    # z slowest, then y, then x, as md_test_utils.h make_fcc_bulk orders sites
    idx = idx[np.lexsort((idx[:, 0], idx[:, 1], idx[:, 2]))]
    pos = (idx[:, None, :] + basis[None, :, :]).reshape(-1, 3) * a
    site = np.tile(np.arange(len(basis)), len(idx))
    return pos, site


def maxwell_boltzmann(rng, mass, T):
    """Velocities (nm/ps) at temperature T with the net momentum removed."""
    sig = np.sqrt(KB * T / mass)
    v = rng.normal(3 * len(mass)).reshape(3, -1).T * sig[:, None]
    p = (mass[:, None] * v).sum(0)
    return v - p / mass.sum()


def _box(L, periodic=(True, True, True)):
    return (np.zeros(3), np.asarray(L, dtype=np.float64), tuple(periodic))


# --------------------------------------------------------------------------
# analytic EAM tables (Cu-like), reference units
# --------------------------------------------------------------------------
EAM_RC = 5.0  # Angstrom


def taper(r, rc=EAM_RC, h=0.5):
    """Smooth cutoff x^4/(1+x^4), x = (r - rc)/h for r < rc, zero beyond."""
    x = (np.asarray(r, dtype=np.float64) - rc) / h
    s = x ** 4 / (1.0 + x ** 4)
    return np.where(r < rc, s, 0.0)


# The EAM Cu workload's functions: Morse D0 = 0.2 eV, alpha = 1.6 / A,
# r0 = 2.55 A; rho(r) = exp(-3 (r / 2.556 - 1)); F(rho) = -1.2 sqrt(rho) eV.
# With the tapered 5.0 A cutoff their fcc energy minimum lies below
# a = 3.0 A, so the a = 3.615 A crystal is under tension and cavitates after
# a few hundred femtoseconds at 300 K.
EAM_F_A = 1.2  # eV


def morse(r, D0=0.2, alpha=1.6, r0=2.55):
    e = np.exp(-alpha * (r - r0))
    return D0 * (e * e - 2.0 * e)


def density(r):
    return np.exp(-3.0 * (r / 2.556 - 1.0))


def eam_cu_tables(nr=5000, nrho=5000, rho_max=50.0, scale_rho=1.0, scale_F=1.0):
    """Tables on 5000-point grids: r step 5.01/4999 A so (nr-1)*dr exceeds
    the 5.0 A cutoff; rho to 50.  Returns dict in reference units:
    F (kJ/mol), rho (-), rphi (kJ nm/mol), dr (nm), drho, rc (nm)."""
    dr_a = 5.01 / (nr - 1)
    r = np.arange(nr) * dr_a
    t = taper(r)
    phi_ev = morse(r) * t
    rho = density(r) * t * scale_rho
    drho = rho_max / (nrho - 1)
    rh = np.arange(nrho) * drho
    F_ev = -EAM_F_A * np.sqrt(rh) * scale_F
    return {
        "F": F_ev * EV,
        "rho": rho,
        "rphi": (r * ANG) * phi_ev * EV,
        "dr": dr_a * ANG,
        "drho": drho,
        "rc": EAM_RC * ANG,
    }


def eam_analytic(r_nm, rhobar):
    """Analytic phi (kJ/mol), rho, F (kJ/mol) of the tabulated functions."""
    r = np.asarray(r_nm) / ANG
    return morse(r) * taper(r) * EV, density(r) * taper(r), -EAM_F_A * np.sqrt(rhobar) * EV


# --------------------------------------------------------------------------
# workloads
# --------------------------------------------------------------------------
def lj_argon(reps=10, seed_index=0, T=100.0, rattle=0.0):
    """rattle (A) displaces the lattice sites; the benchmark workload uses
    none (forces vanish by symmetry at step 0), parity tests a small one."""
    a = 5.26 * ANG
    pos, _ = lattice(FCC, (reps,) * 3, a)
    n = len(pos)
    rng = Rng(structure_seed(1, seed_index))
    if rattle:
        pos = np.mod(pos + rng.normal(3 * n).reshape(n, 3) * rattle * ANG, reps * a)
    at = make_atoms(n)
    m = np.full(n, 39.948)
    set_vec(at, "r", pos)
    set_vec(at, "v", maxwell_boltzmann(rng, m, T))
    sigma = 3.405 * ANG
    rc = 2.5 * sigma
    at["m"] = m
    at["rcut"] = rc
    at["rns"] = rc + 0.3 * ANG
    at["uid"] = np.arange(1, n + 1)
    pot = {"kind": "lj", "a": 0.0104 * EV, "s": sigma, "rc": rc}
    return {"atoms": at, "box": _box([reps * a] * 3), "pot": pot, "dt": 2 * FS, "steps": 1000, "name": "lj_argon"}


def eam_cu(reps=20, seed_index=0, T=300.0, rattle=0.03):
    a = 3.615 * ANG
    pos, _ = lattice(FCC, (reps,) * 3, a)
    n = len(pos)
    rng = Rng(structure_seed(2, seed_index))
    L = reps * a
    pos = np.mod(pos + rng.normal(3 * n).reshape(n, 3) * rattle * ANG, L)
    at = make_atoms(n)
    m = np.full(n, 63.546)
    set_vec(at, "r", pos)
    set_vec(at, "v", maxwell_boltzmann(rng, m, T))
    tab = eam_cu_tables()
    at["m"] = m
    at["rcut"] = tab["rc"]
    at["rns"] = tab["rc"] + 0.3 * ANG
    at["uid"] = np.arange(1, n + 1)
    pot = dict(kind="eam", **tab)
    return {"atoms": at, "box": _box([L] * 3), "pot": pot, "dt": 1 * FS, "steps": 500, "name": "eam_cu"}


def _write_maise(path, arch, flat, ranges, b2a, scaling, g2_rows, g4_rows, values, model_id="0000BEEF",
                 vmax=100.0, species=(("Cu", 29),)):
    """A MAISE-format model file in the layout the reference parser reads
    (mdcraft/tools/parser/maise_model.cxx:29-129): head fields, the input
    component range block (min, max, range per input, species after
    species), one weight per line (networks species after species), B2A and
    Scaling, the descriptor map (N GN Rc Rs n1 n2 k z l) and the 7 value
    rows (written once: the parser gives every species the first species'
    tables when the file ends, maise_model.cxx:93-98)."""
    bar = "-" * 77

    def head(label, value):
        return "|  " + label.ljust(21) + "|  " + str(value).ljust(49) + "|"

    lines = [bar, "|" + "neural network general information".center(75) + "|", bar,
             head("model unique ID", model_id), head("number of species", len(species)),
             head("species types", " ".join(str(t) for _, t in species)),
             head("species names", " ".join(n for n, _ in species)), head("number of layers", 4),
             head("architecture", " ".join(map(str, arch))),
             head("number of weights", len(flat)), head("reference", "synthetic"), bar,
             "|                          input component range                            |", bar,
             "        minimum                    maximum                     range"]
    for r in ranges:
        lines.append("   0.0000000000000000      %.16f         %.16f" % (vmax, r))
    lines += [bar, "|                        neural network weights                             |", bar]
    lines += ["  %.16f" % w for w in flat]
    lines += [bar, "  B2A          Scaling", "  %s  %s" % (b2a, scaling), "", "  N  GN  Rc  Rs   n1  n2  k   z   l", ""]
    row = 1
    for i in g2_rows:
        lines.append("%3d   2   0   0   %d   0   0   0   0" % (row, i))
        row += 1
    for ie, iz, il in g4_rows:
        lines.append("%3d   4   0   0   0   %d   0   %d   %d" % (row, ie, iz, il))
        row += 1
    lines.append("")
    lines += values
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path


def write_maise_model(path, seed_index=0, hidden=(30, 30)):
    """MAISE-format model file for the BPNN workload: 8 G2 (eta log-spaced
    0.003-0.5 A^-2, Rs = 0) and 8 G4 (eta in {0.003, 0.03}, zeta in {1, 4},
    lambda = +-1), Rc = 6 A, input range 1, Xavier-uniform weights and zero
    biases."""
    rng = Rng(structure_seed(3, 1000 + seed_index))
    nin = 16
    arch = [nin, hidden[0], hidden[1], 1]
    weights = []
    for l in range(3):
        weights.append(np.zeros(arch[l + 1]))
    mats = []
    for l in range(3):
        lim = math.sqrt(6.0 / (arch[l] + arch[l + 1]))
        mats.append((rng.uniform(arch[l] * arch[l + 1]) * 2 - 1) * lim)
    flat = np.concatenate(weights + mats)
    g4 = [(ie, iz, il) for ie in range(2) for iz in range(2) for il in range(2)]
    eta2 = np.exp(np.linspace(np.log(0.003), np.log(0.5), 8))
    values = [" Rc   1   6.0", " Rs   1   0.0", " n1   8   " + "   ".join("%.10f" % e for e in eta2),
              " n2   2   0.0030000000   0.0300000000", " k    1   1.0000", " z    2   1.0000   4.0000",
              " l    2  -1.0000   1.0000"]
    return _write_maise(path, arch, flat, [1.0] * nin, "1.0", "1.0", range(8), g4, values)


def _xavier_net(rng, arch):
    """Zero biases and Xavier-uniform weights, in the file order of one
    species (biases of every layer, then the weights of every layer)."""
    biases = [np.zeros(arch[l + 1]) for l in range(3)]
    mats = []
    for l in range(3):
        lim = math.sqrt(6.0 / (arch[l] + arch[l + 1]))
        mats.append((rng.uniform(arch[l] * arch[l + 1]) * 2 - 1) * lim)
    return biases, mats


# The 16-input Cu descriptor set (8 G2, 8 G4) as blocks of a two-species
# model: nspc radial blocks and nspc (nspc + 1) / 2 angular blocks in the
# order of the reference's shift tables (tools/bpnn/maise_model.cxx:113-131):
# R(d = 0), A(0), R(1), A(1), A(2) for nspc = 2, d the kind difference and
# the angular block that of d_ij^2 + d_ik^2.
BINARY_BLOCKS = (("r", 0), ("a", 0), ("r", 1), ("a", 1), ("a", 2))


def write_maise_model_binary(path, seed_index=0, hidden=(30, 30), tied=False, ranges=False):
    """A two-species MAISE model (Cu and Ag) over the descriptor set of the
    BPNN workload: 2 x 8 G2 and 3 x 8 G4 inputs, one network per species.
    tied: both species take the 16-input workload model's network
    (write_maise_model with the same seed_index) with each W01 column
    repeated in every block of its kind, which must give the single-species
    model's results on any structure.  Otherwise each species has its own
    Xavier network with biases in +-0.1, and ranges gives species 1 input
    ranges of 0.5 and 2 alternately (species 0 keeps 1)."""
    nin = 2 * 8 + 3 * 8
    arch = [nin, hidden[0], hidden[1], 1]
    flat = []
    if tied:
        rng = Rng(structure_seed(3, 1000 + seed_index))
        b, m = _xavier_net(rng, [16, hidden[0], hidden[1], 1])
        w01 = m[0].reshape(hidden[0], 16)
        cols = np.concatenate([w01[:, :8] if kind == "r" else w01[:, 8:] for kind, _ in BINARY_BLOCKS], axis=1)
        for _ in range(2):
            flat += b + [cols.reshape(-1), m[1], m[2]]
    else:
        for s in range(2):
            rng = Rng(structure_seed(3, 2000 + 10 * seed_index + s))
            b, m = _xavier_net(rng, arch)
            b = [(rng.uniform(len(x)) * 2 - 1) * 0.1 for x in b]
            flat += b + m
    flat = np.concatenate(flat)
    rg = [1.0] * nin + ([0.5 if q % 2 else 2.0 for q in range(nin)] if ranges else [1.0] * nin)
    g4 = [(ie, iz, il) for ie in range(2) for iz in range(2) for il in range(2)]
    eta2 = np.exp(np.linspace(np.log(0.003), np.log(0.5), 8))
    values = [" Rc   1   6.0", " Rs   1   0.0", " n1   8   " + "   ".join("%.10f" % e for e in eta2),
              " n2   2   0.0030000000   0.0300000000", " k    1   1.0000", " z    2   1.0000   4.0000",
              " l    2  -1.0000   1.0000"]
    return _write_maise(path, arch, flat, rg, "1.0", "1.0", range(8), g4, values,
                        species=(("Cu", 29), ("Ag", 47)))


# Descriptor set of the wide BPNN model: one term per (eta, zeta, lambda) on
# the G4 grid, the two lowest zeta values for the three smallest eta and all
# four for the others, the last term left out: 43 G4 terms over 7 eta.
# eta in (B2A Scaling)^-2 units, log-spaced, as the 6-decimal values the file holds
WIDE_G2_ETA = np.array([float("%.6f" % e) for e in np.exp(np.linspace(np.log(0.001), np.log(0.4), 8))])
WIDE_G4_ETA = np.array([float("%.6f" % e) for e in np.exp(np.linspace(np.log(0.0001), np.log(0.08), 7))])
WIDE_ZETA = (1.0, 2.0, 4.0, 16.0)
WIDE_B2A, WIDE_SCALING, WIDE_RC = 0.529177249, 1.25, 11.3


def _wide_g4_rows():
    rows = [(ie, iz, il) for ie in range(7) for iz in range(2 if ie < 3 else 4) for il in range(2)]
    return rows[:-1]


def _fcc_descriptor_scale(a_nm, rc_nm, eta2_nm, g4_terms):
    """Magnitudes of the G2 and G4 descriptors of an ideal fcc site (the
    definitions of tools/bpnn/base.h:95-155, angular terms with the factor 2
    of base.cxx:120), to give the wide model input ranges that keep its
    first layer out of saturation."""
    reps = int(math.ceil(rc_nm / a_nm)) + 1
    pts = []
    for i in range(-reps, reps + 1):
        for j in range(-reps, reps + 1):
            for k in range(-reps, reps + 1):
                for b in FCC:
                    pts.append((np.array([i, j, k]) + b) * a_nm)
    pts = np.array(pts)
    d = np.linalg.norm(pts, axis=1)
    pts = pts[(d > 1e-9) & (d < rc_nm)]
    d = np.linalg.norm(pts, axis=1)
    fc = 0.5 * (np.cos(np.pi * d / rc_nm) + 1)
    g2 = [np.sum(np.exp(-e * d * d) * fc) for e in eta2_nm]
    u = pts / d[:, None]
    cs = u @ u.T
    dj = np.linalg.norm(pts[:, None, :] - pts[None, :, :], axis=2)
    ok = (dj > 1e-9) & (dj < rc_nm)
    fcj = np.where(ok, 0.5 * (np.cos(np.pi * np.minimum(dj, rc_nm) / rc_nm) + 1), 0.0)
    F = fc[:, None] * fc[None, :] * fcj
    D2 = d[:, None] ** 2 + d[None, :] ** 2 + dj ** 2
    iu = np.triu_indices(len(d), 1)
    g4 = []
    for eta, z, lam in g4_terms:
        t = 2.0 * 2.0 ** (1 - z) * (1 + lam * cs) ** z * np.exp(-eta * D2) * F
        g4.append(np.sum(t[iu]))
    return np.array(g2), np.array(g4)


def write_maise_model_wide(path, seed_index=0):
    """MAISE-format model of the shape of the reference's Cu model
    (problems/Cu-bpnn/nn_Cu_d3_v0: architecture 51-10-10-1, 8 G2 and 43 G4
    over 7 G4 eta, zeta in {1, 2, 4, 16}, lambda = +-1, B2A Scaling = 0.529177249
    1.25, so a length unit of 0.066 nm) with its own descriptor parameters,
    seeded weights and non-zero biases, and input ranges that differ from 1
    and from each other (each the inverse of the descriptor's ideal-fcc value
    times a factor in [0.6, 1.6])."""
    rng = Rng(structure_seed(3, 2000 + seed_index))
    g4_rows = _wide_g4_rows()
    nin = 8 + len(g4_rows)
    arch = [nin, 10, 10, 1]
    lin = WIDE_B2A * WIDE_SCALING * 0.1
    eta2_nm = WIDE_G2_ETA / lin ** 2
    terms = [(WIDE_G4_ETA[ie] / lin ** 2, WIDE_ZETA[iz], (-1.0, 1.0)[il]) for ie, iz, il in g4_rows]
    g2s, g4s = _fcc_descriptor_scale(0.3615, WIDE_RC * lin, eta2_nm, terms)
    scale = np.concatenate([g2s, g4s])
    ranges = (0.6 + rng.uniform(nin)) / scale
    biases = [(rng.uniform(arch[l + 1]) * 2 - 1) * 0.2 for l in range(3)]
    mats = []
    for l in range(3):
        lim = math.sqrt(6.0 / (arch[l] + arch[l + 1]))
        mats.append((rng.uniform(arch[l] * arch[l + 1]) * 2 - 1) * lim)
    flat = np.concatenate(biases + mats)
    values = [" Rc   1   %.4f" % WIDE_RC, " Rs   1   0.0000",
              " n1   8   " + "   ".join("%.6f" % e for e in WIDE_G2_ETA),
              " n2   7   " + "   ".join("%.6f" % e for e in WIDE_G4_ETA),
              " k    1   1.0000", " z    4   " + "   ".join("%.4f" % z for z in WIDE_ZETA),
              " l    2  -1.0000   1.0000"]
    return _write_maise(path, arch, flat, ranges, str(WIDE_B2A), str(WIDE_SCALING), range(8), g4_rows, values,
                        model_id="0000CAFE", vmax=0.0)


def bpnn_cu(reps=15, seed_index=0, T=300.0, rattle=0.05, model_dir=None):
    a = 3.615 * ANG
    pos, _ = lattice(FCC, (reps,) * 3, a)
    n = len(pos)
    rng = Rng(structure_seed(3, seed_index))
    L = reps * a
    pos = np.mod(pos + rng.normal(3 * n).reshape(n, 3) * rattle * ANG, L)
    at = make_atoms(n)
    m = np.full(n, 63.546)
    set_vec(at, "r", pos)
    set_vec(at, "v", maxwell_boltzmann(rng, m, T))
    at["m"] = m
    at["rcut"] = 0.6
    at["rns"] = 0.6 + 0.3 * ANG
    at["kind"] = 0
    at["uid"] = np.arange(1, n + 1)
    model_dir = model_dir or os.path.join(os.environ.get("TMPDIR", "/tmp"), "mdc_models")
    os.makedirs(model_dir, exist_ok=True)
    path = write_maise_model(os.path.join(model_dir, "bpnn_cu_%d.maise" % seed_index))
    pot = {"kind": "bpnn", "path": path, "rc": 0.6}
    return {"atoms": at, "box": _box([L] * 3), "pot": pot, "dt": 1 * FS, "steps": 200, "name": "bpnn_cu"}


def bpnn_cu_wide(reps=5, seed_index=0, T=300.0, rattle=0.05, model_dir=None, periodic=(True, True, True)):
    """fcc Cu (a = 3.615 A) under the wide BPNN model (51-10-10-1, Rc =
    7.47 A, 43 G4 terms over 7 eta), skin 0.3 A.  The reference's own Cu
    example uses 4^3 cells of a = 3.61 A, a box shorter than twice the
    cutoff; 5^3 cells (500 atoms, 18.1 A) is the smallest cubic cell whose
    minimum images are unique.  reps may be a triple; a non-periodic axis
    gets 5 A of vacuum on both sides."""
    a = 3.615 * ANG
    reps3 = (reps,) * 3 if np.isscalar(reps) else tuple(reps)
    pos, _ = lattice(FCC, reps3, a)
    n = len(pos)
    rng = Rng(structure_seed(6, seed_index))
    L = np.array(reps3, dtype=np.float64) * a
    pos = pos + rng.normal(3 * n).reshape(n, 3) * rattle * ANG
    lo, hi = np.zeros(3), L.copy()
    for c in range(3):
        if periodic[c]:
            pos[:, c] = np.mod(pos[:, c], L[c])
        else:
            lo[c], hi[c] = -0.5, L[c] + 0.5
    at = make_atoms(n)
    m = np.full(n, 63.546)
    set_vec(at, "r", pos)
    set_vec(at, "v", maxwell_boltzmann(rng, m, T))
    rc = WIDE_RC * WIDE_B2A * WIDE_SCALING * 0.1
    at["m"] = m
    at["rcut"] = rc
    at["rns"] = rc + 0.3 * ANG
    at["kind"] = 0
    at["uid"] = np.arange(1, n + 1)
    model_dir = model_dir or os.path.join(os.environ.get("TMPDIR", "/tmp"), "mdc_models")
    os.makedirs(model_dir, exist_ok=True)
    path = write_maise_model_wide(os.path.join(model_dir, "bpnn_cu_wide_%d.maise" % seed_index))
    pot = {"kind": "bpnn", "path": path, "rc": rc}
    box = (lo, hi, tuple(bool(p) for p in periodic))
    return {"atoms": at, "box": box, "pot": pot, "dt": 1 * FS, "steps": 50, "name": "bpnn_cu_wide"}


def cubic_reps(cells, minimum=1):
    """Replicas nx <= ny <= nz (each >= minimum): the most cubic product
    within 10 % of cells, else the closest product."""
    cells = max(cells, minimum ** 3)
    best = None
    for nx in range(minimum, cells + 1):
        if nx ** 3 > 1.1 * cells:
            break
        for ny in range(nx, cells + 1):
            if nx * ny * ny > 1.1 * cells:
                break
            for nz in {max(ny, int(cells // (nx * ny))), max(ny, int(cells // (nx * ny)) + 1)}:
                c = nx * ny * nz
                err = abs(c - cells) / cells
                score = (err > 0.1, nz / nx if err <= 0.1 else err, err)
                if best is None or score < best[0]:
                    best = (score, (nx, ny, nz))
    return best[1]


def nial_cell(index, target=None):
    """One strained, rattled B2 NiAl cell; kind 0 = Ni, 1 = Al."""
    rng = Rng(structure_seed(4, index))
    if target is None:
        target = math.exp(math.log(32) + rng.uniform(1)[0] * (math.log(4096) - math.log(32)))
    reps = cubic_reps(max(1, int(round(target / 2))))
    a = 2.88 * ANG
    pos, site = lattice(B2, reps, a)
    n = len(pos)
    strain = 1.0 + (rng.uniform(3) * 2 - 1) * 0.02
    L = np.array(reps) * a * strain
    pos = pos * strain
    pos = np.mod(pos + rng.normal(3 * n).reshape(n, 3) * 0.05 * ANG, L)
    at = make_atoms(n)
    set_vec(at, "r", pos)
    at["kind"] = site
    m = np.where(site == 0, 58.6934, 26.9815385)
    set_vec(at, "v", maxwell_boltzmann(rng, m, 300.0))
    at["m"] = m
    at["rcut"] = EAM_RC * ANG
    at["rns"] = EAM_RC * ANG + 0.3 * ANG
    at["uid"] = np.arange(1, n + 1)
    return at, _box(L)


def nial_pots():
    """Multi potential over three EAM objects, as the reference's NiAl
    problem builds it (problems/NiAl/test-NiAl.py:120-122): Ni-Ni for kind
    0, Al-Al for kind 1 and Ni-Al for the mixed pairs.  Each object keeps
    the EAM Cu functional forms: Ni scales rho and F by 1.1, Al by 0.9,
    Ni-Al by 1.0; all three share the pair term."""
    ni = dict(kind="eam", **eam_cu_tables(scale_rho=1.1, scale_F=1.1))
    al = dict(kind="eam", **eam_cu_tables(scale_rho=0.9, scale_F=0.9))
    nial = dict(kind="eam", **eam_cu_tables())
    return {"kind": "eam_multi", "nkinds": 2, "pairs": [(0, 0, ni), (1, 1, al), (0, 1, nial)], "rc": ni["rc"]}


# Alloy EAM elements on the EAM Cu forms with constants of their own:
# rho_e(r) = A exp(-beta (r / re - 1)), F_e(rho) = -E sqrt(rho) eV, and a
# Morse pair term (D eV, alpha / A, r0 A) per pair of elements, all tapered
# to the 5.0 A cutoff.  Synthetic functions: they give the elements
# different densities, embeddings and pair terms, not a fitted alloy.
ALLOY_ELEMENTS = {
    "Ni": dict(Z=28, mass=58.6934, a=3.52, lattice="fcc", A=1.1, beta=3.2, re=2.49, E=1.35),
    "Al": dict(Z=13, mass=26.9815385, a=4.05, lattice="fcc", A=0.8, beta=2.6, re=2.86, E=0.95),
    "Cu": dict(Z=29, mass=63.546, a=3.615, lattice="fcc", A=1.0, beta=3.0, re=2.556, E=1.2),
}
ALLOY_PAIRS = {
    ("Ni", "Ni"): (0.22, 1.65, 2.49),
    ("Al", "Ni"): (0.26, 1.55, 2.55),
    ("Al", "Al"): (0.15, 1.45, 2.86),
    ("Cu", "Ni"): (0.21, 1.60, 2.52),
    ("Al", "Cu"): (0.19, 1.50, 2.70),
    ("Cu", "Cu"): (0.20, 1.60, 2.55),
}


def alloy_functions(a, b=None):
    """Analytic rho_a(r), F_a(rho) (eV) or, with b, phi_ab(r) (eV); r in A."""
    if b is None:
        e = ALLOY_ELEMENTS[a]
        rho = lambda r: e["A"] * np.exp(-e["beta"] * (np.asarray(r) / e["re"] - 1.0)) * taper(np.asarray(r))
        F = lambda x: -e["E"] * np.sqrt(x)
        return rho, F
    D, al, r0 = ALLOY_PAIRS[tuple(sorted((a, b)))]
    return lambda r: morse(np.asarray(r), D, al, r0) * taper(np.asarray(r))


def alloy_tables(elements, nr=5000, nrho=5000, rho_max=50.0):
    """Tables of an alloy of the named elements in setfl order and
    reference units: F [ne][nrho] (kJ/mol), rho [ne][nr], rphi
    [ne(ne+1)/2][nr] (kJ nm/mol; pairs (1,1), (2,1), (2,2), ...)."""
    dr_a = 5.01 / (nr - 1)
    r = np.arange(nr) * dr_a
    drho = rho_max / (nrho - 1)
    rh = np.arange(nrho) * drho
    F, rho, rphi = [], [], []
    for e in elements:
        fr, fF = alloy_functions(e)
        rho.append(fr(r))
        F.append(fF(rh) * EV)
    for i in range(len(elements)):
        for j in range(i + 1):
            rphi.append((r * ANG) * alloy_functions(elements[i], elements[j])(r) * EV)
    return {"elements": list(elements), "F": np.array(F), "rho": np.array(rho), "rphi": np.array(rphi),
            "dr": dr_a * ANG, "drho": drho, "rc": EAM_RC * ANG}


def write_setfl(path, elements, nr=5000, nrho=5000, rho_max=50.0):
    """The alloy tables as an N-element setfl file (eV, Angstrom), the
    format of problems/NiAl/*.eam.alloy."""
    t = alloy_tables(elements, nr, nrho, rho_max)
    fmt = lambda v: "\n".join(" ".join("%.16e" % x for x in v[k:k + 5]) for k in range(0, len(v), 5))
    with open(path, "w") as f:
        f.write("synthetic alloy EAM\ntapered Morse pair terms\nexponential densities, square-root embedding\n")
        f.write("%d %s\n" % (len(elements), " ".join(elements)))
        f.write("%d %.16e %d %.16e %.16e\n" % (nrho, t["drho"], nr, t["dr"] / ANG, t["rc"] / ANG))
        for k, e in enumerate(elements):
            p = ALLOY_ELEMENTS[e]
            f.write("%d %.6f %.4f %s\n" % (p["Z"], p["mass"], p["a"], p["lattice"]))
            f.write(fmt(t["F"][k] / EV) + "\n")
            f.write(fmt(t["rho"][k]) + "\n")
        for q in range(len(t["rphi"])):
            f.write(fmt(t["rphi"][q] / (EV * ANG)) + "\n")
    return t


def nial_alloy_batch(count=64):
    """B2 NiAl cells (kind 0 = Ni, 1 = Al) for the alloy EAM."""
    cells = [nial_cell(i) for i in range(count)]
    pot = dict(kind="alloy", **alloy_tables(["Ni", "Al"]))
    return {"cells": cells, "pot": pot, "dt": 1 * FS, "name": "nial_alloy_batch"}


def nial_batch(count=512):
    cells = [nial_cell(i) for i in range(count)]
    return {"cells": cells, "pot": nial_pots(), "dt": 1 * FS, "name": "nial_batch"}


# --------------------------------------------------------------------------
# served mixed stream
# --------------------------------------------------------------------------
SERVED_KINDS = ("lj", "eam", "bpnn")


def served_structure(kind, natoms, seed, vacuum=False, model_path=None):
    """An fcc structure of about natoms atoms built as the LJ argon, EAM Cu
    or BPNN Cu workloads build theirs, rattled so that a single-point
    evaluation has nonzero forces; vacuum adds a 50 A gap along z.
    BPNN boxes keep at least 4 cells per axis (the descriptor path needs
    box lengths of at least twice its 6.3 A search radius)."""
    rng = Rng(seed)
    spec = {"lj": (5.26, 39.948, 100.0, 0.05, 1), "eam": (3.615, 63.546, 300.0, 0.03, 2),
            "bpnn": (3.615, 63.546, 300.0, 0.05, 3)}[kind]
    a_ang, mass, T, rattle, wid = spec
    a = a_ang * ANG
    reps = cubic_reps(max(1, int(round(natoms / 4))), minimum=4 if kind == "bpnn" else 1)
    pos, _ = lattice(FCC, reps, a)
    n = len(pos)
    L = np.array(reps, dtype=np.float64) * a
    if rattle:
        pos = np.mod(pos + rng.normal(3 * n).reshape(n, 3) * rattle * ANG, L)
    if vacuum:
        L[2] += 50.0 * ANG
    at = make_atoms(n)
    m = np.full(n, mass)
    set_vec(at, "r", pos)
    set_vec(at, "v", maxwell_boltzmann(rng, m, T))
    at["m"] = m
    at["uid"] = np.arange(1, n + 1)
    if kind == "lj":
        rc = 2.5 * 3.405 * ANG
        pot = {"kind": "lj", "a": 0.0104 * EV, "s": 3.405 * ANG, "rc": rc}
        dt = 2 * FS
    elif kind == "eam":
        tab = eam_cu_tables()
        rc = tab["rc"]
        pot = dict(kind="eam", **tab)
        dt = 1 * FS
    else:
        rc = 0.6
        pot = {"kind": "bpnn", "path": model_path, "rc": rc}
        dt = 1 * FS
    at["rcut"] = rc
    at["rns"] = rc + 0.3 * ANG
    return at, _box(L), pot, dt


def served_requests(count, seed_index=0, model_path=None):
    """count requests of the served mixed stream: potential uniform over LJ,
    EAM Cu and BPNN Cu; atoms log-uniform over 64-20000; a 50 A vacuum slab
    with probability 0.3; half single-point E/F/virial, half NVE runs of
    10-200 steps (uniform).  Each as (kind, atoms, box, dt, steps)."""
    rng = Rng(structure_seed(5, seed_index))
    out = []
    for q in range(count):
        u = rng.uniform(5)
        kind = SERVED_KINDS[min(2, int(u[0] * 3))]
        natoms = int(round(math.exp(math.log(64) + u[1] * (math.log(20000) - math.log(64)))))
        vacuum = u[2] < 0.3
        steps = 0 if u[3] < 0.5 else 10 + min(190, int(u[4] * 191))
        at, box, pot, dt = served_structure(kind, natoms, structure_seed(5, 100000 + seed_index * 10000 + q), vacuum,
                                            model_path)
        out.append({"kind": kind, "atoms": at, "box": box, "dt": dt, "steps": steps, "pot": pot})
    return out
