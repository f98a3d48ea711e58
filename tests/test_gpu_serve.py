"""The serving loop on the GPU: requests served together in batches give
the results the same requests give alone (FP64), and every request gets
exactly its steps.  Boxes exceed 2 (rns + skin), where one image per pair
is fixed between rebuilds, so the extra rebuilds a batch triggers do not
change the result; the tolerance covers the reordered FP64 sums."""
import numpy as np
import pytest

from mdcraft_amd import serve
from mdcraft_amd import workloads as W
from mdcraft_amd.atoms import vec

pytestmark = pytest.mark.gpu

SERVED_VS_ALONE = 1e-9


def engine():
    from mdcraft_amd import engine as E
    return E


def requests(model_path):
    spec = [("lj", 500, 0), ("eam", 864, 30), ("bpnn", 256, 12), ("lj", 400, 45), ("eam", 300, 0),
            ("bpnn", 300, 0), ("eam", 1000, 7), ("lj", 864, 20, True), ("eam", 500, 25, True)]
    out = []
    for q, sp in enumerate(spec):
        kind, n, steps = sp[:3]
        vac = len(sp) > 3
        at, box, pot, dt = W.served_structure(kind, n, W.structure_seed(9, q), vac, model_path)
        assert np.min(box[1] - box[0]) > 2 * (at["rns"][0] + 0.03)
        out.append((kind, at, box, dt, steps, pot))
    return out


def test_served_matches_alone(ctx, tmp_path):
    E = engine()
    path = W.write_maise_model(str(tmp_path / "cu.maise"))
    specs = requests(path)
    pots = {}
    for kind, at, box, dt, steps, pot in specs:
        if kind not in pots:
            from tests.helpers import engine_potential
            pots[kind] = engine_potential(E, ctx, pot)
    fp64 = {k: E.FP64 for k in pots}
    mk = lambda: [serve.Request(q, s[0], s[1], s[2], s[3], s[4], 0.0) for q, s in enumerate(specs)]
    srv = serve.Server(E, ctx, pots, fp64, chunk=8)
    served = {r.rid: r for r in srv.run_trace(serve.poisson_trace(mk(), None))}
    alone = {r.rid: r for r in serve.one_at_a_time(E, ctx, pots, serve.poisson_trace(mk(), None))}
    assert sorted(served) == sorted(alone) == list(range(len(specs)))
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
