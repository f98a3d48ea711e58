"""The serving loop on the GPU (the engine's mdc_serve_* server).

* FP64: requests served together give the results the same requests give
  alone (1e-9: the extra rebuilds a batch triggers reorder FP64 sums), and
  every request gets exactly its steps.
* The served stream in the benchmark's precisions (LJ FP64, EAM and BPNN
  mixed) with Poisson arrivals, so that requests join batches that are
  already stepping, structures up to 20 000 atoms: each result against its
  alone run at the same precision (FP64 1e-9; mixed force RMS 1e-5,
  positions 1e-6 nm) and against the oracle's evaluation of the returned
  configuration (FP64 1e-12; mixed force RMS 1e-5, energy 1e-5 as in the
  BPNN tests, virial trace 1e-5).
* A memory budget that holds one batch: later batches wait for it to close,
  and a request beyond the whole budget fails.
Boxes of stepped requests exceed 2 (rns + skin), where one image per pair is
fixed between rebuilds, so rebuild timing does not change the physics."""
import numpy as np
import pytest

from mdcraft_amd import serve
from mdcraft_amd import workloads as W
from mdcraft_amd.atoms import trace_V, vec
from tests.helpers import engine_potential, force_rms_rel, max_rel, oracle_compute

pytestmark = pytest.mark.gpu

SERVED_VS_ALONE = 1e-9


def engine():
    from mdcraft_amd import engine as E
    return E


def build(spec, model_path):
    out = []
    for q, sp in enumerate(spec):
        kind, n, steps = sp[:3]
        vac = len(sp) > 3
        at, box, pot, dt = W.served_structure(kind, n, W.structure_seed(9, q), vac, model_path)
        if steps:
            assert np.min(box[1] - box[0]) > 2 * (at["rns"][0] + 0.03)
        out.append((kind, at, box, dt, steps, pot))
    return out


def potentials(E, ctx, specs):
    pots = {}
    for kind, at, box, dt, steps, pot in specs:
        if kind not in pots:
            pots[kind] = engine_potential(E, ctx, pot)
    return pots


def requests(specs):
    return [serve.Request(q, s[0], s[1], s[2], s[3], s[4], 0.0) for q, s in enumerate(specs)]


def test_served_matches_alone(ctx, tmp_path):
    E = engine()
    path = W.write_maise_model(str(tmp_path / "cu.maise"))
    spec = [("lj", 500, 0), ("eam", 864, 30), ("bpnn", 256, 12), ("lj", 400, 45), ("eam", 300, 0),
            ("bpnn", 300, 0), ("eam", 1000, 7), ("lj", 864, 20, True), ("eam", 500, 25, True)]
    specs = build(spec, path)
    pots = potentials(E, ctx, specs)
    fp64 = {k: E.FP64 for k in pots}
    srv = serve.Server(E, ctx, pots, fp64, chunk=8, slots=4)
    served = {r.rid: r for r in srv.run_trace(serve.poisson_trace(requests(specs), None))}
    st = srv.stats()
    srv.close()
    alone = {r.rid: r for r in serve.one_at_a_time(E, ctx, pots, serve.poisson_trace(requests(specs), None))}
    assert sorted(served) == sorted(alone) == list(range(len(specs)))
    assert st["done"] == len(specs) and st["failed"] == 0
    for q in served:
        a, b = served[q], alone[q]
        L = b.box[1] - b.box[0]
        d = vec(a.result, "r") - vec(b.result, "r")
        d -= np.round(d / L) * L
        assert np.max(np.abs(d)) < SERVED_VS_ALONE, q
        fa, fb = vec(a.result, "f"), vec(b.result, "f")
        assert np.max(np.abs(fa - fb)) <= SERVED_VS_ALONE * np.max(np.abs(fb)), q
        for k in (0, 1, 2, 4):  # Ep, Em, KE, W
            assert abs(a.thermo[k] - b.thermo[k]) <= SERVED_VS_ALONE * max(abs(b.thermo[k]), 1.0), (q, k)


def test_served_stream_in_benchmark_precisions(ctx, tmp_path):
    E = engine()
    path = W.write_maise_model(str(tmp_path / "cu.maise"))
    spec = [("lj", 256, 40), ("eam", 256, 50), ("bpnn", 256, 12), ("lj", 864, 0), ("eam", 2048, 0),
            ("bpnn", 500, 0), ("lj", 4000, 30), ("eam", 6912, 25), ("eam", 20000, 10), ("lj", 20000, 0),
            ("bpnn", 2048, 8), ("eam", 13500, 0), ("lj", 500, 20, True), ("eam", 256, 30), ("bpnn", 1372, 0),
            ("eam", 256, 20), ("lj", 256, 12)]
    specs = build(spec, path)
    pots = potentials(E, ctx, specs)
    prec = {"lj": E.FP64, "eam": E.MIXED, "bpnn": E.MIXED}
    srv = serve.Server(E, ctx, pots, prec, chunk=8, slots=4)
    print("served stream: %d requests" % len(specs), flush=True)
    served = {r.rid: r for r in srv.run_trace(serve.poisson_trace(requests(specs), 300.0, seed=5))}
    st = srv.stats()
    srv.close()
    assert st["done"] == len(specs) and st["failed"] == 0
    # arrivals joined batches already open: fewer batches than requests, and
    # a request admitted while another of its kind was stepping
    assert st["batches_opened"] < len(specs)
    stepped = [r for r in served.values() if r.steps]
    assert any(r2 is not r1 and r1.admitted < r2.admitted < r1.done for r1 in stepped for r2 in served.values())
    print("served stream: alone runs", flush=True)
    alone = {r.rid: r for r in serve.one_at_a_time(E, ctx, pots, serve.poisson_trace(requests(specs), None),
                                                  precision=prec)}
    for q, (kind, at, box, dt, steps, pot) in enumerate(specs):
        a, b = served[q], alone[q]
        L = box[1] - box[0]
        d = vec(a.result, "r") - vec(b.result, "r")
        d -= np.round(d / L) * L
        fa, fb, fr = vec(a.result, "f"), vec(b.result, "f"), None
        if prec[kind] == E.FP64:
            assert np.max(np.abs(d)) < SERVED_VS_ALONE, q
            assert np.max(np.abs(fa - fb)) <= SERVED_VS_ALONE * np.max(np.abs(fb)), q
        else:
            assert np.max(np.abs(d)) < 1e-6, q
            assert force_rms_rel(fa, fb) <= 1e-5, q
        # the oracle's evaluation of the configuration the server returned
        print("served stream: oracle request %d (%s, %d atoms)" % (q, kind, len(at)), flush=True)
        ref, _ = oracle_compute(a.result, box, pot, threads=8)
        fr = vec(ref, "f")
        e, er = (a.result["Ep"] + a.result["Em"]).sum(), (ref["Ep"] + ref["Em"]).sum()
        w, wr = trace_V(a.result).sum(), trace_V(ref).sum()
        if prec[kind] == E.FP64:
            assert max_rel(fa, fr) <= 1e-12, q
            assert abs(e - er) <= 1e-12 * abs(er), q
        else:
            assert force_rms_rel(fa, fr) <= 1e-5, q
            assert abs(e - er) <= 1e-5 * abs(er), q
            assert abs(w - wr) <= 1e-5 * np.abs(trace_V(ref)).sum(), q


def test_memory_budget_binds(ctx, tmp_path):
    E = engine()
    spec = [("eam", 256, 20), ("eam", 256, 20), ("eam", 2048, 10), ("eam", 2048, 0), ("eam", 256, 10)]
    specs = build(spec, None)
    pots = potentials(E, ctx, specs)
    prec = {"eam": E.MIXED}
    # the bytes of one batch of each class
    probe = serve.Server(E, ctx, pots, prec, chunk=8, slots=2)
    sizes = []
    for r in requests(specs)[1:3]:
        probe.submit(r)
        probe.wait(r)
        sizes.append(probe.stats()["open_bytes"])
    probe.close()
    small, both = sizes[0], sizes[1]
    big = both - small
    assert big > small > 0
    budget = 1.01 * big  # one batch at a time
    srv = serve.Server(E, ctx, pots, prec, budget_bytes=budget, chunk=8, slots=2)
    done = {r.rid: r for r in srv.run_trace(serve.poisson_trace(requests(specs), None))}
    st = srv.stats()
    assert st["done"] == len(specs) and st["failed"] == 0
    assert 0 < st["peak_bytes"] <= budget
    # the 2048-atom batch opened only after the 256-atom batch closed
    first_small_done = max(done[q].done for q in (0, 1))
    assert min(done[q].admitted for q in (2, 3)) >= first_small_done
    # a request whose batch exceeds the whole budget fails
    at, box, pot, dt = W.served_structure("eam", 20000, 7)
    r = serve.Request(99, "eam", at, box, dt, 0, 0.0)
    srv.submit(r)
    with pytest.raises(RuntimeError, match="budget"):
        srv.wait(r)
    srv.close()
