"""BPNN oracle pinned by what the mathematics fixes: closed-form dimer and
trimer descriptors, the network evaluated from an independent parse of the
MAISE file, energy/force finite differences and invariances.  No reference
test exists for this potential (parity against the reference itself is
unpinned); these checks fail for a dropped term, a wrong sign or index, or a
transposed weight matrix."""
import math

import numpy as np
import pytest

from mdcraft_amd import workloads as W
from mdcraft_amd.atoms import make_atoms, set_vec, vec
from tests.helpers import oracle_mod

R = oracle_mod()


@pytest.fixture(scope="module")
def model(tmp_path_factory):
    return W.write_maise_model(str(tmp_path_factory.mktemp("m") / "m.maise"))


def params():
    eta2 = np.exp(np.linspace(np.log(0.003), np.log(0.5), 8)) * 100.0  # A^-2 -> nm^-2
    g4 = [(e * 100.0, z, l) for e in (0.003, 0.03) for z in (1.0, 4.0) for l in (-1.0, 1.0)]
    return eta2, g4, 0.6


def fc(r, rc):
    return 0.5 * (np.cos(np.pi * r / rc) + 1)


def compute(pos, path, box=None):
    n = len(pos)
    at = make_atoms(n)
    set_vec(at, "r", pos)
    at["m"] = 63.546
    at["kind"] = 0
    at["rns"] = 0.63
    if box is None:
        box = (np.full(3, -5.0), np.full(3, 5.0), (False, False, False))
    dom = R.Domain(*box)
    pool = R.Pool(1)
    pot = R.MaiseBPNN(path, dom)
    at["rcut"] = pot.rcut
    nl = R.VerletList(pool)
    nl.update(at, dom)
    R.Solver(pot, pool, True).compute(at, dom, nl)
    G, dE = R.bpnn_descriptors(pot, n)
    return at, G, dE


def test_dimer_radial_closed_form(model):
    eta2, g4, rc = params()
    r = 0.26
    at, G, _ = compute(np.array([[0, 0, 0], [r, 0, 0]]), model)
    want = np.exp(-eta2 * r * r) * fc(r, rc)
    assert np.allclose(G[0, :8], want, rtol=1e-13) and np.allclose(G[1, :8], want, rtol=1e-13)
    assert np.all(G[:, 8:] == 0)


def test_trimer_angular_closed_form(model):
    eta2, g4, rc = params()
    pos = np.array([[0, 0, 0], [0.25, 0, 0], [0.05, 0.27, 0.02]])
    at, G, _ = compute(pos, model)
    for i in range(3):
        j, k = [q for q in range(3) if q != i]
        a, b = pos[j] - pos[i], pos[k] - pos[i]
        da, db, dc = np.linalg.norm(a), np.linalg.norm(b), np.linalg.norm(b - a)
        cs = a @ b / (da * db)
        for q, (eta, z, lam) in enumerate(g4):
            want = 2 * 2 ** (1 - z) * (1 + lam * cs) ** z * math.exp(-eta * (da * da + db * db + dc * dc)) * \
                fc(da, rc) * fc(db, rc) * fc(dc, rc)
            assert G[i, 8 + q] == pytest.approx(want, rel=1e-12)


def read_weights(path):
    lines = open(path).read().splitlines()
    k = next(i for i, l in enumerate(lines) if "neural network weights" in l)
    w = []
    for l in lines[k + 2:]:
        if l.startswith("---"):
            break
        w.append(float(l))
    return np.array(w)


def test_network_from_independent_parse(model):
    rng = np.random.default_rng(3)
    pos = rng.random((12, 3)) * 0.5
    at, G, dE = compute(pos, model)
    w = read_weights(model)
    b1, b2, b3 = w[:30], w[30:60], w[60]
    W1 = w[61:61 + 480].reshape(30, 16)
    W2 = w[541:541 + 900].reshape(30, 30)
    W3 = w[1441:1471]
    h1 = np.tanh(G @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    E = (h2 @ W3 + b3) * 96.48533212331
    assert np.allclose(at["Ep"], E, rtol=1e-12)
    # the input gradient by central differences of the network itself
    eps = 1e-6
    for q in (0, 7, 9, 15):
        Gp, Gm = G.copy(), G.copy()
        Gp[:, q] += eps
        Gm[:, q] -= eps
        f = lambda g: (np.tanh(np.tanh(g @ W1.T + b1) @ W2.T + b2) @ W3 + b3) * 96.48533212331
        assert np.allclose(dE[:, q], (f(Gp) - f(Gm)) / (2 * eps), rtol=1e-6, atol=1e-6)


def cluster(n=24, seed=7):
    rng = np.random.default_rng(seed)
    pts = []
    while len(pts) < n:
        p = rng.random(3) * 0.7
        if all(np.linalg.norm(p - q) > 0.2 for q in pts):
            pts.append(p)
    return np.array(pts)


def test_forces_are_energy_gradient(model):
    pos = cluster()
    at, _, _ = compute(pos, model)
    f = vec(at, "f")
    h = 1e-6
    for i in (0, 5, 11):
        for c in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[i, c] += h
            pm[i, c] -= h
            ep = compute(pp, model)[0]["Ep"].sum()
            em = compute(pm, model)[0]["Ep"].sum()
            assert f[i, c] == pytest.approx(-(ep - em) / (2 * h), rel=1e-6, abs=1e-5)
    assert np.allclose(f.sum(0), 0, atol=1e-9)


def test_invariances(model):
    pos = cluster(seed=9)
    at, _, _ = compute(pos, model)
    th = 0.7
    Rz = np.array([[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0], [0, 0, 1]])
    Rx = np.array([[1, 0, 0], [0, math.cos(0.3), -math.sin(0.3)], [0, math.sin(0.3), math.cos(0.3)]])
    Q = Rz @ Rx
    c = pos.mean(0)
    at2, _, _ = compute((pos - c) @ Q.T + c + 0.05, model)
    assert np.allclose(at2["Ep"], at["Ep"], rtol=1e-11)
    assert np.allclose(vec(at2, "f"), vec(at, "f") @ Q.T, atol=1e-8 * np.abs(vec(at, "f")).max())
    perm = np.random.default_rng(1).permutation(len(pos))
    at3, _, _ = compute(pos[perm], model)
    assert np.allclose(at3["Ep"], at["Ep"][perm], rtol=1e-11)
    assert np.allclose(vec(at3, "f"), vec(at, "f")[perm], atol=1e-9 * np.abs(vec(at, "f")).max())
