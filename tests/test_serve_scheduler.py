"""Host logic of the serving loop (mdcraft_amd/serve.py) against a stand-in
engine that records what it is asked to do: admission in arrival order
within the memory budget, one batch per potential per tick, chunked
stepping that never passes a box's step count, and retirement of each box
at its count."""
import numpy as np
import pytest

from mdcraft_amd import serve
from mdcraft_amd.atoms import make_atoms


class FakeSystem:
    log = []

    def __init__(self, ctx, natoms, boxes, periodic_boundary=True, lanes=None):
        self.natoms = list(natoms)
        self.active = [True] * len(self.natoms)
        self.steps = 0
        FakeSystem.log.append(("create", tuple(self.natoms)))

    def upload(self, atoms):
        assert len(atoms) == sum(self.natoms)

    def build(self):
        pass

    def compute(self, pot, what, precision):
        pass

    def step(self, pot, dt, nsteps, precision):
        self.steps += nsteps
        FakeSystem.log.append(("step", tuple(self.natoms), nsteps))

    def thermo(self):
        return np.tile(np.arange(8.0), (len(self.natoms), 1))

    def download_box(self, box):
        out = make_atoms(self.natoms[box])
        out["Ep"] = self.steps
        return out

    def set_active(self, box, active):
        self.active[box] = bool(active)

    def close(self):
        pass

    def sync(self):
        pass


class FakeE:
    FORCE, ENERGY, VIRIAL, FP64, MIXED = 1, 2, 4, 0, 1
    System = FakeSystem


def req(rid, kind, n, steps, dt=1e-3):
    return serve.Request(rid, kind, make_atoms(n), (np.zeros(3), np.ones(3), (True,) * 3), dt, steps, 0.0)


def make_server(budget, chunk=10):
    FakeSystem.log = []
    clock = iter(np.arange(0, 1e6, 1.0))
    return serve.Server(FakeE, None, {"lj": "LJ", "eam": "EAM", "bpnn": "BP"}, {"lj": 0, "eam": 1, "bpnn": 1},
                        budget_bytes=budget, chunk=chunk, clock=lambda: next(clock))


def test_each_request_gets_exactly_its_steps():
    s = make_server(1e12, chunk=10)
    rs = [req(0, "lj", 100, 0), req(1, "lj", 120, 25), req(2, "eam", 80, 7), req(3, "lj", 90, 40)]
    for r in rs:
        s.submit(r)
    while s.tick():
        pass
    assert sorted(r.rid for r in s.completed) == [0, 1, 2, 3]
    # a box's download happens when the batch has run exactly its steps
    assert {r.rid: int(r.result["Ep"][0]) for r in rs} == {0: 0, 1: 25, 2: 7, 3: 40}
    # one batch per potential for requests admitted together
    creates = [e for e in FakeSystem.log if e[0] == "create"]
    assert sorted(creates) == sorted([("create", (100, 120, 90)), ("create", (80,))])
    # chunks never pass the smallest count left: 10, 10, 5 (box 1 done), then 10, 5 for box 3
    lj_steps = [e[2] for e in FakeSystem.log if e[0] == "step" and e[1] == (100, 120, 90)]
    assert lj_steps == [10, 10, 5, 10, 5]


def test_admission_in_arrival_order_within_budget():
    per = serve.BYTES_PER_ATOM["lj"]
    s = make_server(budget=per * 250, chunk=100)
    big, small = req(0, "lj", 200, 50), req(1, "lj", 100, 0)
    s.submit(big)
    s.submit(small)
    s.tick()
    # the second request does not fit beside the first and waits, in order
    assert big.admitted is not None and small.admitted is None
    while s.tick():
        pass
    assert small.admitted > big.admitted and small.done is not None
    assert s.peak_bytes <= per * 250


def test_request_above_budget_is_refused():
    s = make_server(budget=10.0)
    s.submit(req(0, "eam", 100, 0))
    with pytest.raises(RuntimeError, match="budget"):
        s.tick()


def test_latency_stats():
    rs = [req(i, "lj", 10, 0) for i in range(4)]
    for i, r in enumerate(rs):
        r.arrival, r.done = float(i), float(i) + 0.5 * (i + 1)
    st = serve.latency_stats(rs)
    assert st["requests"] == 4
    assert st["p50_ms"] == pytest.approx(1e3 * np.percentile([0.5, 1.0, 1.5, 2.0], 50))
    assert st["structures_per_s"] == pytest.approx(4 / 5.0)
