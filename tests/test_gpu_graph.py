"""Graph mode of mdc_compute and the launch-tile autotuner on the GPU.

* A batch of NiAl cells evaluated with graph mode on follows a twin system
  evaluated eagerly, bit for bit (the same kernels in the same order), over
  several replays with uploads, downloads and host syncs between them, in
  FP64 and mixed precision; a new configuration (flags, a rebuild) is
  captured anew.
* The same for a BPNN system (the reverse-slot index built before the
  capture).
* Workgroups of 64, 128 and 256 threads give identical FP64 results, and
  the tile the autotuner keeps gives the oracle's within 1e-12.
"""
import numpy as np
import pytest

from mdcraft_amd import workloads as W
from mdcraft_amd.atoms import mat, set_vec, vec
from tests.helpers import engine_potential, max_rel, oracle_compute

pytestmark = pytest.mark.gpu


def E():
    from mdcraft_amd import engine
    return engine


def nial_batch(count=24, seed=0):
    cs = [W.nial_cell(i) for i in range(seed, seed + count)]
    return cs, W.nial_pots()


def twin(ctx, cs, pot):
    eng = E()
    out = []
    for _ in range(2):
        s = eng.System(ctx, [len(a) for a, _ in cs], [b for _, b in cs])
        s.upload(np.concatenate([a for a, _ in cs]))
        s.build()
        out.append(s)
    return out, engine_potential(eng, ctx, pot)


def jiggle(atoms, rng, amp):
    a = atoms.copy()
    set_vec(a, "r", vec(a, "r") + rng.normal(scale=amp, size=(len(a), 3)))
    return a


def same(ga, gb, label):
    for f in ("f", "r"):
        assert np.array_equal(vec(ga, f), vec(gb, f)), (label, f)
    for f in ("Ep", "Em", "nc"):
        assert np.array_equal(ga[f], gb[f]), (label, f)
    assert np.array_equal(mat(ga), mat(gb)), (label, "V")


@pytest.mark.parametrize("precision", [0, 1])
def test_graph_replays_match_eager(ctx, precision):
    eng = E()
    cs, pot = nial_batch()
    (sg, se), p = twin(ctx, cs, pot)
    sg.set_graphs(True)
    base = np.concatenate([a for a, _ in cs])
    rng = np.random.default_rng(7)
    what = eng.FORCE | eng.ENERGY | eng.VIRIAL
    for rep in range(5):
        # positions moved well within the skin: the list stays valid and the
        # replay reads the new positions
        a = jiggle(base, rng, 0.002 * W.ANG)
        sg.upload(a, eng.F_R)
        se.upload(a, eng.F_R)
        sg.compute(p, what, precision)
        se.compute(p, what, precision)
        same(sg.download(), se.download(), rep)
        assert np.array_equal(sg.thermo(), se.thermo()), rep
    st = sg.graph_stats()
    assert st["captures"] == 1 and st["launches"] == 4
    # new flags: a new configuration, run eagerly, then captured
    for _ in range(3):
        sg.compute(p, eng.FORCE, precision)
        se.compute(p, eng.FORCE, precision)
        same(sg.download(), se.download(), "force only")
    st = sg.graph_stats()
    assert st["captures"] == 2 and st["launches"] == 6
    # a rebuild: the list arrays change, the graph is captured again
    sg.build()
    se.build()
    for _ in range(3):
        sg.compute(p, what, precision)
        se.compute(p, what, precision)
        same(sg.download(), se.download(), "after rebuild")
    assert sg.graph_stats()["captures"] == 3


def test_graph_bpnn(ctx, tmp_path):
    eng = E()
    model = W.write_maise_model(str(tmp_path / "cu.maise"))
    at, box, pot, _ = W.served_structure("bpnn", 500, 3, False, model)
    (sg, se), p = twin(ctx, [(at, box)], pot)
    sg.set_graphs(True)
    rng = np.random.default_rng(3)
    for prec in (1, 0):
        for rep in range(4):
            a = jiggle(at, rng, 0.002 * W.ANG)
            sg.upload(a, eng.F_R)
            se.upload(a, eng.F_R)
            sg.compute(p, eng.FORCE | eng.ENERGY | eng.VIRIAL, prec)
            se.compute(p, eng.FORCE | eng.ENERGY | eng.VIRIAL, prec)
            same(sg.download(), se.download(), (prec, rep))
    assert sg.graph_stats()["captures"] == 2


def test_block_sizes_and_tile_autotune(ctx):
    eng = E()
    cs, pot = nial_batch(12, seed=40)
    s = eng.System(ctx, [len(a) for a, _ in cs], [b for _, b in cs])
    s.upload(np.concatenate([a for a, _ in cs]))
    p = engine_potential(eng, ctx, pot)
    s.build()
    outs = []
    for t in (256, 128, 64):
        s.set_block(t)
        s.compute(p, eng.FORCE | eng.ENERGY | eng.VIRIAL, 0)
        outs.append(s.download())
    same(outs[1], outs[0], "128 vs 256")
    same(outs[2], outs[0], "64 vs 256")
    lanes, threads, us = s.autotune_tiles(p, 0, reps=5)
    assert lanes in (1, 2, 4, 8, 16) and threads in (64, 128, 256)
    assert len(us) == 15 and all(v > 0 for v in us.values())
    assert us[(lanes, threads)] == min(us.values())
    s.compute(p, eng.FORCE | eng.ENERGY | eng.VIRIAL, 0)
    got = s.download()
    off = 0
    for q, (a, box) in enumerate(cs[:4]):
        g = got[off:off + len(a)]
        ref, _ = oracle_compute(g, box, pot, threads=8)
        assert max_rel(vec(g, "f"), vec(ref, "f")) <= 1e-12, q
        assert max_rel(mat(g), mat(ref)) <= 1e-12, q
        off += len(a)
