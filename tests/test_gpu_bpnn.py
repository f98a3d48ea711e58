"""BPNN on the engine against the oracle (FP64: 1e-12; mixed: FP32
descriptors and an FP16-split MFMA network, force RMS within 1e-5 and the
virial trace within 1e-5).  The virial is the reference's diagonal one
(maise_bpnn.cxx:98-120)."""
import numpy as np
import pytest

from mdcraft_amd import workloads as W
from mdcraft_amd.atoms import mat, trace_V, vec
from tests.helpers import force_rms_rel, max_rel, oracle_mod

pytestmark = pytest.mark.gpu


def oracle(w, threads=8):
    R = oracle_mod()
    a = w["atoms"].copy()
    dom = R.Domain(*w["box"])
    pool = R.Pool(threads)
    p = R.MaiseBPNN(w["pot"]["path"], dom)
    a["rcut"] = p.rcut
    nl = R.VerletList(pool)
    nl.update(a, dom)
    s = R.Solver(p, pool, True)
    s.compute(a, dom, nl)
    G, dE = R.bpnn_descriptors(p, len(a))
    return a, G, dE, (s, nl, dom, p)


def engine(ctx, w, precision, steps=0, lanes=16):
    from mdcraft_amd import engine as E
    s = E.System(ctx, [len(w["atoms"])], [w["box"]], lanes=lanes)
    s.upload(w["atoms"])
    p = E.MaiseBPNNPotential(ctx, w["pot"]["path"])
    s.build()
    s.compute(p, 7, precision)
    if steps:
        s.step(p, w["dt"], steps, precision)
        s.compute(p, 7, precision)
    G, dE = s.bpnn_inspect(p)
    return s.download(), G, dE


@pytest.fixture(scope="module")
def small(tmp_path_factory):
    return W.bpnn_cu(reps=4, model_dir=str(tmp_path_factory.mktemp("bp")))


LANES = [1, 2, 4, 8, 16]


@pytest.mark.parametrize("lanes", LANES)
def test_bpnn_fp64_parity(ctx, small, lanes):
    ref, G, dE, _ = oracle(small)
    got, g, de = engine(ctx, small, 0, lanes=lanes)
    assert max_rel(g, G) < 1e-12
    assert max_rel(de, dE) < 1e-11
    assert max_rel(got["Ep"], ref["Ep"]) < 1e-12
    assert max_rel(vec(got, "f"), vec(ref, "f")) < 1e-12
    V, Vr = mat(got), mat(ref)
    assert np.max(np.abs(Vr)) > 0
    assert max_rel(V, Vr) < 1e-12
    # only the diagonal is written, as the reference's virial() does
    off = V.copy()
    for c in range(3):
        off[:, c, c] = 0.0
    assert np.all(off == 0.0)


@pytest.mark.parametrize("lanes", LANES)
def test_bpnn_mixed_accuracy(ctx, small, lanes):
    ref, G, dE, _ = oracle(small)
    got, g, de = engine(ctx, small, 1, lanes=lanes)
    assert max_rel(g, G) < 1e-5
    assert max_rel(de, dE) < 1e-4
    e, er = got["Ep"].sum(), ref["Ep"].sum()
    assert abs(e - er) < 1e-5 * abs(er)
    assert force_rms_rel(vec(got, "f"), vec(ref, "f")) < 1e-5
    w, wr = trace_V(got).sum(), trace_V(ref).sum()
    assert abs(w - wr) < 1e-5 * abs(wr)
    assert max_rel(mat(got), mat(ref)) < 1e-4


def test_bpnn_stage_virials(ctx, small):
    """prepare, forces, then virials() adding the diagonal virial to V, as
    solver::Ternary's calls do one at a time (ternary.cxx:19-67)."""
    from mdcraft_amd import engine as E
    ref, _, _, _ = oracle(small)
    s = E.System(ctx, [len(small["atoms"])], [small["box"]], lanes=4)
    s.upload(small["atoms"])
    p = E.MaiseBPNNPotential(ctx, small["pot"]["path"])
    s.build()
    for st in (E.PREPARE, E.FORCES, E.VIRIALS):
        s.stage(p, st, 0)
    got = s.download()
    assert max_rel(vec(got, "f"), vec(ref, "f")) < 1e-12
    assert max_rel(mat(got), mat(ref)) < 1e-12
    # a second virials() adds the same diagonal again
    s.stage(p, E.VIRIALS, 0)
    got2 = s.download()
    assert max_rel(mat(got2), 2 * mat(ref)) < 1e-12


def test_bpnn_lanes_drop_after_evaluation(ctx, tmp_path):
    """500 atoms (not a multiple of 64) evaluated at 8 lanes per atom, then
    at 1 lane: the list holds more elements at 1 lane, and the work arrays
    made for the first list must be made again for the second."""
    from mdcraft_amd import engine as E
    w = W.bpnn_cu(reps=5, model_dir=str(tmp_path))
    assert len(w["atoms"]) % 64 != 0
    ref, _, _, _ = oracle(w)
    s = E.System(ctx, [len(w["atoms"])], [w["box"]], lanes=8)
    s.upload(w["atoms"])
    p = E.MaiseBPNNPotential(ctx, w["pot"]["path"])
    s.build()
    s.compute(p, 7, 0)
    for lanes in (1, 2, 16):
        s.set_lanes(lanes)
        s.build()
        s.compute(p, 7, 0)
        got = s.download()
        assert max_rel(vec(got, "f"), vec(ref, "f")) < 1e-12, lanes
        assert max_rel(mat(got), mat(ref)) < 1e-12, lanes


def test_bpnn_fp64_short_nve(ctx, small):
    ref, _, _, (s, nl, dom, p) = oracle(small)
    s.run_nve(ref, dom, nl, small["dt"], 5, initial_forces=False)
    s.compute(ref, dom, nl)
    got, _, _ = engine(ctx, small, 0, steps=5)
    L = small["box"][1] - small["box"][0]
    d = vec(got, "r") - vec(ref, "r")
    d -= np.round(d / L) * L
    assert np.max(np.abs(d)) < 1e-11
    assert max_rel(vec(got, "f"), vec(ref, "f")) < 1e-10


def test_bpnn_fp64_nve_through_rebuilds(ctx, small):
    """24 FP64 steps, the list rebuilt on the half-skin trigger inside
    step() (with the reverse-slot index built again), against the oracle's
    run with decide()/update()."""
    from mdcraft_amd import engine as E
    ref, _, _, (s, nl, dom, p) = oracle(small)
    s.run_nve(ref, dom, nl, small["dt"], 24, initial_forces=False)
    es = E.System(ctx, [len(small["atoms"])], [small["box"]], lanes=16)
    es.upload(small["atoms"])
    pot = E.MaiseBPNNPotential(ctx, small["pot"]["path"])
    es.build()
    es.compute(pot, 7, 0)
    rebuilds = es.step(pot, small["dt"], 24, 0)
    got = es.download()
    assert rebuilds >= 1
    L = small["box"][1] - small["box"][0]
    d = vec(got, "r") - vec(ref, "r")
    d -= np.round(d / L) * L
    assert np.max(np.abs(d)) < 1e-9
    assert max_rel(vec(got, "v"), vec(ref, "v")) < 1e-8


def test_bpnn_13500_atoms_against_oracle(ctx, tmp_path):
    """The benchmark's BPNN workload at its full size (13 500 atoms, 15^3
    fcc cells) against the oracle: FP64 forces, energies and virials within
    1e-12, mixed precision force RMS within 1e-5 and energy and virial trace
    within 1e-5."""
    w = W.bpnn_cu(reps=15, model_dir=str(tmp_path))
    print("bpnn 13500: oracle", flush=True)
    ref, _, _, _ = oracle(w, threads=16)
    for prec in (0, 1):
        print("bpnn 13500: engine precision=%d" % prec, flush=True)
        got, _, _ = engine(ctx, w, prec, lanes=8)
        f, fr = vec(got, "f"), vec(ref, "f")
        if prec == 0:
            assert max_rel(f, fr) < 1e-12
            assert max_rel(got["Ep"], ref["Ep"]) < 1e-12
            assert max_rel(mat(got), mat(ref)) < 1e-12
        else:
            assert force_rms_rel(f, fr) < 1e-5
            assert abs(got["Ep"].sum() - ref["Ep"].sum()) < 1e-5 * abs(ref["Ep"].sum())
            w_, wr = trace_V(got).sum(), trace_V(ref).sum()
            assert abs(w_ - wr) < 1e-5 * np.abs(trace_V(ref)).sum()
