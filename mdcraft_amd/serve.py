"""Served mixed stream on one GPU: the engine's continuous-batching server
(mdc_serve_* of include/mdc.h) from Python.

A request is one structure (atoms, box) for one potential, with a task: a
single-point energy/force/virial evaluation (steps = 0) or an NVE run of
`steps` velocity-Verlet steps followed by an energy/force/virial
evaluation.  The server (C++, in the engine) admits requests in arrival
order into free slots of running batches (one system per batch, boxes as
slots of one size class), opens batches within a device-memory budget,
runs every batch on its own stream and host thread, and frees a request's
slot when its count is reached; see the header for the full contract.

This module submits traces of requests at their arrival times, collects the
results, and runs the same traces through the plain FP64 one-at-a-time path
for comparison.
"""
import ctypes as C
import time

import numpy as np

from .atoms import ATOM_DTYPE


class Request:
    __slots__ = ("rid", "kind", "atoms", "box", "dt", "steps", "arrival", "admitted", "done", "result", "thermo",
                 "sid")

    def __init__(self, rid, kind, atoms, box, dt, steps, arrival):
        self.rid, self.kind, self.atoms, self.box, self.dt, self.steps = rid, kind, atoms, box, dt, steps
        self.arrival = arrival
        self.admitted = self.done = None
        self.result = self.thermo = None
        self.sid = None


class Server:
    """pots: {kind: engine potential}; precision: {kind: MDC precision}.
    budget_bytes bounds the estimated device bytes of the open batches;
    chunk is the most steps a batch runs between scheduling points; slots
    is the number of slots a new batch gets."""

    def __init__(self, engine, ctx, pots, precision, budget_bytes=200e9, chunk=25, slots=8):
        self.E = engine
        self._lib = engine.lib()
        h = C.c_void_p()
        engine._check(self._lib.mdc_serve_create(ctx.device, float(budget_bytes), int(chunk), int(slots), C.byref(h)))
        self.h = h
        self.pots = pots  # the potentials must outlive the server
        self.kind_id = {}
        for kind, pot in pots.items():
            k = C.c_int()
            engine._check(self._lib.mdc_serve_add_potential(self.h, pot.h, int(precision[kind]), C.byref(k)))
            self.kind_id[kind] = k.value

    def submit(self, req):
        a = np.ascontiguousarray(req.atoms, dtype=ATOM_DTYPE)
        lo = np.ascontiguousarray(req.box[0], dtype=np.float64)
        hi = np.ascontiguousarray(req.box[1], dtype=np.float64)
        per = np.ascontiguousarray([1 if p else 0 for p in req.box[2]], dtype=np.uint8)
        sid = C.c_uint64()
        self.E._check(self._lib.mdc_serve_submit(
            self.h, self.kind_id[req.kind], a.ctypes.data, len(a), lo.ctypes.data_as(C.POINTER(C.c_double)),
            hi.ctypes.data_as(C.POINTER(C.c_double)), per.ctypes.data_as(C.POINTER(C.c_uint8)), float(req.dt),
            int(req.steps), C.byref(sid)))
        req.sid = sid.value
        return req.sid

    def wait(self, req, timeout=600.0):
        """Blocks until the request is done (or the timeout passes), then
        fills its result, thermo row and server-clock times."""
        state = C.c_int()
        t_end = time.perf_counter() + timeout
        while True:
            self.E._check(self._lib.mdc_serve_wait(self.h, req.sid, 1.0, C.byref(state)))
            if state.value != 0:
                break
            if time.perf_counter() > t_end:
                raise TimeoutError("request %d not done after %.0f s" % (req.rid, timeout))
        out = np.zeros(len(req.atoms), dtype=ATOM_DTYPE)
        th = np.zeros(8)
        tm = np.zeros(3)
        self.E._check(self._lib.mdc_serve_take(self.h, req.sid, out.ctypes.data,
                                               th.ctypes.data_as(C.POINTER(C.c_double)),
                                               tm.ctypes.data_as(C.POINTER(C.c_double))))
        req.result, req.thermo = out, th
        req.arrival, req.admitted, req.done = tm
        return req

    def stats(self):
        """done, failed, queued, running, open batches, batches opened, open
        batch bytes, peak open batch bytes."""
        out = np.zeros(8)
        self.E._check(self._lib.mdc_serve_stats(self.h, out.ctypes.data_as(C.POINTER(C.c_double))))
        keys = ("done", "failed", "queued", "running", "open_batches", "batches_opened", "open_bytes", "peak_bytes")
        return dict(zip(keys, out.tolist()))

    @property
    def peak_bytes(self):
        return self.stats()["peak_bytes"]

    def run_trace(self, trace, timeout=600.0):
        """trace: [(arrival offset s, Request)] sorted by offset.  Submits
        each request at its arrival time, then collects every result.  The
        requests' arrival, admitted and done times are on the server's
        clock (seconds since it started)."""
        t0 = time.perf_counter()
        for off, r in trace:
            wait = t0 + off - time.perf_counter()
            if wait > 0:
                time.sleep(wait)
            self.submit(r)
        return [self.wait(r, timeout) for _, r in trace]

    def close(self):
        if getattr(self, "h", None):
            self._lib.mdc_serve_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def one_at_a_time(engine, ctx, pots, trace, clock=time.perf_counter, precision=None):
    """The plain one-at-a-time path on the same trace: each request alone in
    its own System, in arrival order, evaluated or stepped in FP64 (or in
    precision[kind] when given)."""
    E = engine
    t0 = clock()
    out = []
    for off, r in trace:
        arrival = t0 + off
        now = clock()
        if now < arrival:
            time.sleep(arrival - now)
        r.arrival = arrival
        prec = E.FP64 if precision is None else precision[r.kind]
        s = E.System(ctx, [len(r.atoms)], [r.box], True)
        s.upload(r.atoms)
        s.build()
        p = pots[r.kind]
        if r.steps:
            s.compute(p, E.FORCE, prec)
            s.step(p, r.dt, r.steps, prec)
        s.compute(p, E.FORCE | E.ENERGY | E.VIRIAL, prec)
        r.thermo = s.thermo()[0]
        r.result = s.download()
        r.done = clock()
        s.close()
        out.append(r)
    return out


def latency_stats(reqs):
    lat = np.array([r.done - r.arrival for r in reqs])
    first = min(r.arrival for r in reqs)
    last = max(r.done for r in reqs)
    return {"requests": len(reqs), "structures_per_s": len(reqs) / (last - first),
            "atoms_per_request": float(np.mean([len(r.atoms) for r in reqs])),
            "p50_ms": 1e3 * float(np.percentile(lat, 50)), "p99_ms": 1e3 * float(np.percentile(lat, 99)),
            "mean_ms": 1e3 * float(lat.mean())}


def poisson_trace(requests, rate, seed=0):
    """Arrival offsets of a Poisson process of the given rate (1/s); rate
    None puts every arrival at 0 (saturation)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for r in requests:
        if rate:
            t += rng.exponential(1.0 / rate)
        out.append((t if rate else 0.0, r))
    return out
