"""Served mixed stream on one GPU: continuous batching of MD requests.

A request is one structure (atoms, box) for one potential, with a task: a
single-point energy/force/virial evaluation (steps = 0) or an NVE run of
`steps` velocity-Verlet steps followed by an energy/force/virial
evaluation.  The server

* admits arrived requests in arrival order while the estimated device
  memory of everything admitted stays within the HBM budget;
* at each scheduling tick, puts the newly admitted requests of each
  potential into one engine System (a batch of boxes: one launch per
  kernel for all of them) and evaluates them;
* advances every running batch by up to `chunk` steps, never past the
  smallest step count left among its running boxes;
* retires a request as soon as its count is reached: its box alone is
  downloaded and then switched off on the device, so the rest of the batch
  runs on without it; a batch with no running box is released.

The scheduler is host code; all force and integration work is the
engine's.  One thread drives the GPU: `submit` may be called from other
threads, `tick` / `run_trace` from the driving thread only.
"""
import collections
import threading
import time

import numpy as np

# device bytes per atom used for admission: state and list (about 1 KB for
# LJ / EAM lists of up to ~100 neighbours), BPNN adds its per-neighbour
# descriptor-gradient scratch
BYTES_PER_ATOM = {"lj": 1536, "eam": 1280, "bpnn": 96 * 1024}


class Request:
    __slots__ = ("rid", "kind", "atoms", "box", "dt", "steps", "arrival", "admitted", "done", "result", "thermo",
                 "left", "batch", "box_index")

    def __init__(self, rid, kind, atoms, box, dt, steps, arrival):
        self.rid, self.kind, self.atoms, self.box, self.dt, self.steps = rid, kind, atoms, box, dt, steps
        self.arrival = arrival
        self.admitted = self.done = None
        self.result = self.thermo = None
        self.left = steps
        self.batch = None
        self.box_index = -1

    @property
    def bytes(self):
        return BYTES_PER_ATOM[self.kind] * len(self.atoms)


class _Batch:
    def __init__(self, system, pot, kind, members, precision):
        self.system, self.pot, self.kind, self.members, self.precision = system, pot, kind, members, precision
        self.running = [m for m in members]


class Server:
    """pots: {kind: engine potential}; precision: {kind: MDC precision}."""

    def __init__(self, engine, ctx, pots, precision, budget_bytes=200e9, chunk=25, lanes=None,
                 clock=time.perf_counter):
        self.E, self.ctx, self.pots, self.precision = engine, ctx, pots, precision
        self.budget = float(budget_bytes)
        self.chunk = int(chunk)
        self.lanes = lanes
        self.clock = clock
        self._lock = threading.Lock()
        self._arrived = collections.deque()
        self.waiting = collections.deque()  # arrived, not admitted
        self.admitted_bytes = 0.0
        self.batches = []
        self.completed = []
        self.peak_bytes = 0.0
        self.ticks = 0

    # ---- producer side --------------------------------------------------
    def submit(self, req):
        with self._lock:
            self._arrived.append(req)

    # ---- scheduling -----------------------------------------------------
    def _admit(self):
        with self._lock:
            while self._arrived:
                self.waiting.append(self._arrived.popleft())
        new = collections.defaultdict(list)
        while self.waiting:
            r = self.waiting[0]
            if r.bytes > self.budget:
                raise RuntimeError("request %d needs %.3g bytes, above the whole budget" % (r.rid, r.bytes))
            if self.admitted_bytes + r.bytes > self.budget:
                break  # arrival order: later requests wait behind it
            self.waiting.popleft()
            r.admitted = self.clock()
            self.admitted_bytes += r.bytes
            new[r.kind].append(r)
        self.peak_bytes = max(self.peak_bytes, self.admitted_bytes)
        return new

    def _finish(self, b, reqs):
        """Evaluate E/F/virial of the batch (the finishing boxes take theirs
        from it), download each finishing box and switch it off."""
        s = b.system
        s.compute(b.pot, self.E.FORCE | self.E.ENERGY | self.E.VIRIAL, b.precision)
        th = s.thermo()
        for r in reqs:
            r.result = s.download_box(r.box_index)
            r.thermo = th[r.box_index].copy()
            s.set_active(r.box_index, False)
            r.done = self.clock()
            r.batch = None
            self.admitted_bytes -= r.bytes
            self.completed.append(r)
        b.running = [m for m in b.running if m.done is None]

    def _start(self, kind, reqs):
        E = self.E
        s = E.System(self.ctx, [len(r.atoms) for r in reqs], [r.box for r in reqs], True, lanes=self.lanes)
        s.upload(np.concatenate([r.atoms for r in reqs]))
        s.build()
        b = _Batch(s, self.pots[kind], kind, reqs, self.precision[kind])
        for q, r in enumerate(reqs):
            r.batch, r.box_index = b, q
        singles = [r for r in reqs if r.left == 0]
        if singles:
            self._finish(b, singles)
        else:
            s.compute(b.pot, E.FORCE, b.precision)
        if b.running:
            self.batches.append(b)
        else:
            s.close()

    def _advance(self, b):
        k = min(self.chunk, min(r.left for r in b.running))
        dts = {r.dt for r in b.running}
        if len(dts) != 1:
            raise RuntimeError("one batch steps with one time step")
        b.system.step(b.pot, dts.pop(), k, b.precision)
        for r in b.running:
            r.left -= k
        done = [r for r in b.running if r.left == 0]
        if done:
            self._finish(b, done)

    def tick(self):
        """One scheduling round; returns True while work remains."""
        self.ticks += 1
        for kind, reqs in self._admit().items():
            # one batch per potential and time step
            by_dt = collections.defaultdict(list)
            for r in reqs:
                by_dt[r.dt].append(r)
            for group in by_dt.values():
                self._start(kind, group)
        for b in list(self.batches):
            self._advance(b)
            if not b.running:
                b.system.close()
                self.batches.remove(b)
        with self._lock:
            pending = bool(self._arrived)
        return bool(self.batches or self.waiting or pending)

    def run_trace(self, trace):
        """trace: [(arrival offset s, Request)] sorted by offset.  Submits
        each request when its arrival time has passed and ticks until all
        are done.  Request.arrival is set to its nominal arrival time."""
        t0 = self.clock()
        q = 0
        while True:
            now = self.clock()
            while q < len(trace) and t0 + trace[q][0] <= now:
                r = trace[q][1]
                r.arrival = t0 + trace[q][0]
                self.submit(r)
                q += 1
            busy = self.tick()
            if q == len(trace) and not busy:
                break
            if not busy and q < len(trace):
                # idle until the next arrival
                wait = t0 + trace[q][0] - self.clock()
                if wait > 0:
                    time.sleep(min(wait, 0.01))
        self.ctx_sync()
        return self.completed

    def ctx_sync(self):
        for b in self.batches:
            b.system.sync()


def one_at_a_time(engine, ctx, pots, trace, clock=time.perf_counter):
    """The plain FP64 path on the same trace: each request alone in its own
    System, in arrival order, evaluated or stepped in FP64."""
    E = engine
    t0 = clock()
    out = []
    for off, r in trace:
        arrival = t0 + off
        now = clock()
        if now < arrival:
            time.sleep(arrival - now)
        r.arrival = arrival
        s = E.System(ctx, [len(r.atoms)], [r.box], True)
        s.upload(r.atoms)
        s.build()
        p = pots[r.kind]
        if r.steps:
            s.compute(p, E.FORCE, E.FP64)
            s.step(p, r.dt, r.steps, E.FP64)
        s.compute(p, E.FORCE | E.ENERGY | E.VIRIAL, E.FP64)
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
