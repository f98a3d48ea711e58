"""The oracle's BPNN under the wide model (the shape of the reference's
problems/Cu-bpnn/nn_Cu_d3_v0: 51-10-10-1, 8 G2 and 43 G4 over 7 eta, zeta up
to 16, a length unit of B2A * Scaling = 0.066 nm, input ranges != 1), pinned
by what the mathematics fixes:
  * closed-form dimer and trimer descriptors in the file's length unit;
  * the input range multiplies the columns of the first weight matrix
    (maise_ann.cxx:44-47): a range-r model equals a range-1 model whose
    W01 columns are multiplied by r, and differs from models that divide,
    ignore the range or scale the rows of a transposed read;
  * forces are minus the energy gradient, and energy and forces are
    invariant under rotation, translation and permutation;
  * the diagonal virial of maise_bpnn.cxx:98-120 summed over atoms equals
    dE/d(strain_cc) under homogeneous strain: its radial term moves the
    neighbour's dE/dG onto the atom, which leaves the sum unchanged, and its
    angular term is the strain derivative of G4 (pattern_dG4_s_,
    tools/bpnn/base.h:157-181).
The last two also run on the network of the benchmark workload."""
import math

import numpy as np
import pytest

from mdcraft_amd import workloads as W
from mdcraft_amd.atoms import make_atoms, mat, set_vec, vec
from tests.helpers import have_reference, oracle_mod, REFERENCE

R = oracle_mod()
LIN = W.WIDE_B2A * W.WIDE_SCALING * 0.1  # nm per file length unit
RC = W.WIDE_RC * LIN


@pytest.fixture(scope="module")
def wide(tmp_path_factory):
    return W.write_maise_model_wide(str(tmp_path_factory.mktemp("w") / "wide.maise"))


@pytest.fixture(scope="module")
def small(tmp_path_factory):
    return W.write_maise_model(str(tmp_path_factory.mktemp("s") / "small.maise"))


def compute(pos, path, threads=4):
    n = len(pos)
    at = make_atoms(n)
    set_vec(at, "r", pos)
    at["m"] = 63.546
    at["kind"] = 0
    box = (np.full(3, -5.0), np.full(3, 5.0), (False, False, False))
    dom = R.Domain(*box)
    pool = R.Pool(threads)
    pot = R.MaiseBPNN(path, dom)
    at["rcut"] = pot.rcut
    at["rns"] = pot.rcut + 0.03
    nl = R.VerletList(pool)
    nl.update(at, dom)
    R.Solver(pot, pool, True).compute(at, dom, nl)
    G, dE = R.bpnn_descriptors(pot, n)
    return at, G, dE


def fc(r):
    return 0.5 * (np.cos(np.pi * r / RC) + 1)


def g4_terms():
    out = []
    for ie, iz, il in W._wide_g4_rows():
        out.append((W.WIDE_G4_ETA[ie] / LIN ** 2, W.WIDE_ZETA[iz], (-1.0, 1.0)[il]))
    return out


def test_wide_dimer_in_file_units(wide):
    r = 0.31
    at, G, _ = compute(np.array([[0, 0, 0], [r, 0, 0]]), wide)
    assert G.shape == (2, 51)
    want = np.exp(-(W.WIDE_G2_ETA / LIN ** 2) * r * r) * fc(r)
    assert np.allclose(G[0, :8], want, rtol=1e-13)
    assert np.all(G[:, 8:] == 0)


def test_wide_trimer_angular_closed_form(wide):
    pos = np.array([[0, 0, 0], [0.27, 0, 0], [0.04, 0.29, 0.03]])
    _, G, _ = compute(pos, wide)
    terms = g4_terms()
    assert len(terms) == 43 and {z for _, z, _ in terms} == {1.0, 2.0, 4.0, 16.0}
    for i in range(3):
        j, k = [q for q in range(3) if q != i]
        a, b = pos[j] - pos[i], pos[k] - pos[i]
        da, db, dc = np.linalg.norm(a), np.linalg.norm(b), np.linalg.norm(b - a)
        cs = a @ b / (da * db)
        for q, (eta, z, lam) in enumerate(terms):
            want = 2 * 2 ** (1 - z) * (1 + lam * cs) ** z * math.exp(-eta * (da * da + db * db + dc * dc)) * \
                fc(da) * fc(db) * fc(dc)
            assert G[i, 8 + q] == pytest.approx(want, rel=1e-12, abs=1e-300)


def model_parts(path):
    lines = open(path).read().splitlines()
    k = next(i for i, l in enumerate(lines) if "input component range" in l)
    rows = []
    i = k + 3
    while not lines[i].startswith("---"):
        rows.append(i)
        i += 1
    k = next(i for i, l in enumerate(lines) if "neural network weights" in l)
    wrows = []
    i = k + 2
    while not lines[i].startswith("---"):
        wrows.append(i)
        i += 1
    return lines, rows, wrows


def rewrite(path, out, w01):
    """The model with every input range 1 and W01 (h1 x nin) replaced."""
    lines, rows, wrows = model_parts(path)
    nin, h1, h2 = 51, 10, 10
    for i in rows:
        t = lines[i].split()
        lines[i] = "   %s      %s         %.16f" % (t[0], t[1], 1.0)
    off = h1 + h2 + 1
    for q in range(h1):
        for k in range(nin):
            lines[wrows[off + q * nin + k]] = "  %.17g" % w01[q, k]
    open(out, "w").write("\n".join(lines) + "\n")
    return out


def test_range_scales_first_layer_columns(wide, tmp_path):
    lines, rows, wrows = model_parts(wide)
    rng_vals = np.array([float(lines[i].split()[2]) for i in rows])
    assert np.ptp(rng_vals) > 0 and np.all(np.abs(rng_vals - 1) > 1e-3)
    w = np.array([float(lines[i]) for i in wrows])
    W01 = w[21:21 + 510].reshape(10, 51)
    cols = rewrite(wide, str(tmp_path / "cols.maise"), W01 * rng_vals[None, :])
    div = rewrite(wide, str(tmp_path / "div.maise"), W01 / rng_vals[None, :])
    none = rewrite(wide, str(tmp_path / "none.maise"), W01)
    # a transposed read (nin x h1) with its rows scaled, laid back out
    tr = W01.reshape(51, 10) * rng_vals[:, None]
    trans = rewrite(wide, str(tmp_path / "trans.maise"), tr.reshape(10, 51))
    rng = np.random.default_rng(11)
    pos = rng.random((14, 3)) * 0.6
    ref, _, _ = compute(pos, wide)
    got, _, _ = compute(pos, cols)
    assert np.allclose(got["Ep"], ref["Ep"], rtol=1e-13, atol=0)
    assert np.allclose(vec(got, "f"), vec(ref, "f"), rtol=1e-12, atol=1e-12 * np.abs(vec(ref, "f")).max())
    for other in (div, none, trans):
        e = compute(pos, other)[0]["Ep"]
        assert np.max(np.abs(e - ref["Ep"])) > 1e-3 * np.max(np.abs(ref["Ep"])), other


def cluster(n, spacing, seed, extent):
    rng = np.random.default_rng(seed)
    pts = []
    while len(pts) < n:
        p = rng.random(3) * extent
        if all(np.linalg.norm(p - q) > spacing for q in pts):
            pts.append(p)
    return np.array(pts)


@pytest.mark.parametrize("which", ["wide", "small"])
def test_forces_and_strain_virial(wide, small, which):
    path = wide if which == "wide" else small
    pos = cluster(26, 0.21, 5, 0.75)
    at, _, _ = compute(pos, path)
    f = vec(at, "f")
    assert np.allclose(f.sum(0), 0, atol=1e-9 * np.abs(f).max())
    h = 1e-6
    for i in (0, 7, 19):
        for c in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[i, c] += h
            pm[i, c] -= h
            d = -(compute(pp, path)[0]["Ep"].sum() - compute(pm, path)[0]["Ep"].sum()) / (2 * h)
            assert f[i, c] == pytest.approx(d, rel=1e-6, abs=1e-6 * np.abs(f).max())
    V = mat(at)
    W_diag = V.sum(0)
    assert np.all(W_diag[~np.eye(3, dtype=bool)] == 0)
    c0 = pos.mean(0)
    eps = 1e-6
    for c in range(3):
        sp = np.eye(3)
        sp[c, c] += eps
        sm = np.eye(3)
        sm[c, c] -= eps
        ep = compute((pos - c0) @ sp.T + c0, path)[0]["Ep"].sum()
        em = compute((pos - c0) @ sm.T + c0, path)[0]["Ep"].sum()
        dEde = (ep - em) / (2 * eps)
        assert W_diag[c, c] == pytest.approx(dEde, rel=1e-6, abs=1e-8 * abs(W_diag).max())


def test_wide_invariances(wide):
    pos = cluster(20, 0.22, 9, 0.7)
    at, _, _ = compute(pos, wide)
    th = 0.6
    Rz = np.array([[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0], [0, 0, 1]])
    Ry = np.array([[math.cos(0.4), 0, math.sin(0.4)], [0, 1, 0], [-math.sin(0.4), 0, math.cos(0.4)]])
    Q = Rz @ Ry
    c = pos.mean(0)
    at2, _, _ = compute((pos - c) @ Q.T + c + 0.07, wide)
    assert np.allclose(at2["Ep"], at["Ep"], rtol=1e-11)
    assert np.allclose(vec(at2, "f"), vec(at, "f") @ Q.T, atol=1e-8 * np.abs(vec(at, "f")).max())
    perm = np.random.default_rng(2).permutation(len(pos))
    at3, _, _ = compute(pos[perm], wide)
    assert np.allclose(at3["Ep"], at["Ep"][perm], rtol=1e-11)


REF_MODEL = REFERENCE + "/problems/Cu-bpnn/nn_Cu_d3_v0"


@pytest.mark.reference
@pytest.mark.skipif(not have_reference(), reason="reads the reference's model file in place")
def test_reference_model_on_the_oracle():
    """The reference's own Cu model (read in place, not committed) on a
    20-atom cluster: 51 inputs, forces minus the energy gradient, the
    summed diagonal virial equal to the strain derivative, and rotation
    invariance."""
    pos = cluster(20, 0.22, 3, 0.7)
    at, G, dE = compute(pos, REF_MODEL)
    assert G.shape == (20, 51) and np.all(np.isfinite(at["Ep"]))
    f = vec(at, "f")
    h = 1e-6
    for i in (0, 9):
        for c in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[i, c] += h
            pm[i, c] -= h
            d = -(compute(pp, REF_MODEL)[0]["Ep"].sum() - compute(pm, REF_MODEL)[0]["Ep"].sum()) / (2 * h)
            assert f[i, c] == pytest.approx(d, rel=1e-6, abs=1e-6 * np.abs(f).max())
    W_diag = mat(at).sum(0)
    c0 = pos.mean(0)
    for c in range(3):
        sp, sm = np.eye(3), np.eye(3)
        sp[c, c] += 1e-6
        sm[c, c] -= 1e-6
        d = (compute((pos - c0) @ sp.T + c0, REF_MODEL)[0]["Ep"].sum() -
             compute((pos - c0) @ sm.T + c0, REF_MODEL)[0]["Ep"].sum()) / 2e-6
        assert W_diag[c, c] == pytest.approx(d, rel=1e-6, abs=1e-8 * abs(W_diag).max())
    th = 0.5
    Q = np.array([[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0], [0, 0, 1]])
    at2, _, _ = compute((pos - c0) @ Q.T + c0, REF_MODEL)
    assert np.allclose(at2["Ep"], at["Ep"], rtol=1e-11)


@pytest.mark.reference
@pytest.mark.skipif(not have_reference(), reason="reads the reference's model file in place")
def test_engine_reads_the_reference_model():
    """The engine's host-side parse of the reference's Cu model (no device):
    the shape the oracle evaluates, the cutoff of the file's length unit."""
    from mdcraft_amd import engine as E
    d = E.describe_bpnn(REF_MODEL)
    assert d["arch"] == [51, 10, 10, 1]
    assert (d["radial"], d["angular"], d["g4_eta"], d["g4_groups"]) == (8, 43, 7, 8)
    assert d["rcut"] == pytest.approx(11.338 * 0.529177249 * 1.25 * 0.1, rel=1e-15)
    at, G, _ = compute(np.array([[0, 0, 0], [0.3, 0, 0]]), REF_MODEL)
    assert G.shape[1] == d["arch"][0]


def test_engine_reads_the_wide_model(wide):
    from mdcraft_amd import engine as E
    d = E.describe_bpnn(wide)
    assert d["arch"] == [51, 10, 10, 1]
    assert (d["radial"], d["angular"], d["g4_eta"], d["g4_groups"]) == (8, 43, 7, 8)
    assert d["rcut"] == pytest.approx(RC, rel=1e-15)
