"""Two-species BPNN on the engine against the oracle (kind-difference
descriptor blocks and one network per species, maise_bpnn.cxx:141-325):
FP64 descriptors, dE/dG, energies, forces and the diagonal virial within
1e-12 for 1, 8 and 16 lanes per atom; mixed precision force RMS, energy and
virial trace within 1e-5; a two-species model whose networks repeat the
16-input model's first-layer columns in every block gives that model's
results on the engine; 20 FP64 NVE steps through list rebuilds follow the
oracle's run; kinds without a network are refused."""
import numpy as np
import pytest

from mdcraft_amd import workloads as W
from mdcraft_amd.atoms import mat, trace_V, vec
from tests.helpers import force_rms_rel, max_rel, oracle_mod

pytestmark = pytest.mark.gpu


def E():
    from mdcraft_amd import engine
    return engine


@pytest.fixture(scope="module")
def models(tmp_path_factory):
    d = tmp_path_factory.mktemp("bb")
    return {"single": W.write_maise_model(str(d / "s.maise")),
            "tied": W.write_maise_model_binary(str(d / "t.maise"), tied=True),
            "binary": W.write_maise_model_binary(str(d / "b.maise"), ranges=True)}


def alloy(n=500, seed=3, frac=0.4):
    """The BPNN workload's rattled Cu crystal with a random fraction of the
    atoms of the second kind."""
    at, box, _, dt = W.served_structure("bpnn", n, W.structure_seed(21, seed), False, None)
    rng = np.random.default_rng(seed)
    at["kind"] = (rng.random(len(at)) < frac).astype(at["kind"].dtype)
    return at, box, dt


def oracle(at, box, path, steps=0, dt=None):
    R = oracle_mod()
    a = at.copy()
    dom = R.Domain(*box)
    pool = R.Pool(8)
    p = R.MaiseBPNN(path, dom)
    a["rcut"] = p.rcut
    nl = R.VerletList(pool)
    nl.update(a, dom)
    s = R.Solver(p, pool, True)
    if steps:
        s.prepare(a, dom, nl)
        s.stages(a, dom, nl)
        s.run_nve(a, dom, nl, dt, steps, initial_forces=False)
        return a, None, None
    s.compute(a, dom, nl)
    G, dE = R.bpnn_descriptors(p, len(a))
    return a, G, dE


def engine(ctx, at, box, path, precision, lanes=8, steps=0, dt=None):
    eng = E()
    s = eng.System(ctx, [len(at)], [box], lanes=lanes)
    s.upload(at)
    p = eng.MaiseBPNNPotential(ctx, path)
    s.build()
    s.compute(p, 7, precision)
    if steps:
        s.step(p, dt, steps, precision)
        return s.download(), None, None
    G, dE = s.bpnn_inspect(p)
    return s.download(), G, dE


@pytest.mark.parametrize("lanes", [1, 8, 16])
def test_binary_fp64(ctx, models, lanes):
    at, box, _ = alloy()
    ref, G, dE = oracle(at, box, models["binary"])
    got, g, de = engine(ctx, at, box, models["binary"], 0, lanes)
    assert g.shape[1] == 40
    for lo, hi in ((0, 8), (8, 16), (16, 24), (24, 32), (32, 40)):
        assert np.max(np.abs(G[:, lo:hi])) > 0, (lo, hi)
        assert max_rel(g[:, lo:hi], G[:, lo:hi]) < 1e-12, (lo, hi)
    assert max_rel(de, dE) < 1e-11
    assert max_rel(got["Ep"], ref["Ep"]) < 1e-12
    assert max_rel(vec(got, "f"), vec(ref, "f")) < 1e-12
    assert max_rel(mat(got), mat(ref)) < 1e-12


def test_binary_mixed(ctx, models):
    at, box, _ = alloy(seed=5)
    ref, G, _ = oracle(at, box, models["binary"])
    got, g, _ = engine(ctx, at, box, models["binary"], 1)
    assert max_rel(g, G) < 1e-5
    e, er = got["Ep"].sum(), ref["Ep"].sum()
    assert abs(e - er) < 1e-5 * abs(er)
    assert force_rms_rel(vec(got, "f"), vec(ref, "f")) < 1e-5
    w, wr = trace_V(got).sum(), trace_V(ref).sum()
    assert abs(w - wr) < 1e-5 * abs(wr)


def test_tied_model_equals_single_species_on_engine(ctx, models):
    at, box, _ = alloy(seed=7, frac=0.5)
    a, _, _ = engine(ctx, at, box, models["tied"], 0)
    single = at.copy()
    single["kind"] = 0
    b, _, _ = engine(ctx, single, box, models["single"], 0)
    assert max_rel(a["Ep"], b["Ep"]) < 1e-12
    assert max_rel(vec(a, "f"), vec(b, "f")) < 1e-12
    assert max_rel(mat(a), mat(b)) < 1e-12


def test_binary_nve_fp64(ctx, models):
    at, box, dt = alloy(256, seed=9)
    assert np.min(box[1] - box[0]) > 2 * (at["rns"][0] + 0.03)
    ref, _, _ = oracle(at, box, models["binary"], steps=20, dt=dt)
    got, _, _ = engine(ctx, at, box, models["binary"], 0, steps=20, dt=dt)
    L = box[1] - box[0]
    d = vec(got, "r") - vec(ref, "r")
    d -= np.round(d / L) * L
    assert np.max(np.abs(d)) < 1e-9
    assert max_rel(vec(got, "v"), vec(ref, "v")) < 1e-8


def test_kind_without_network_is_refused(ctx, models):
    eng = E()
    at, box, _ = alloy(256, seed=11)
    at["kind"][3] = 2
    s = eng.System(ctx, [len(at)], [box])
    s.upload(at)
    p = eng.MaiseBPNNPotential(ctx, models["binary"])
    s.build()
    with pytest.raises(RuntimeError, match="no network"):
        s.compute(p, 7, 0)
