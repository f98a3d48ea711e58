"""Host side of serving without a GPU: the arrival traces and latency
statistics of mdcraft_amd/serve.py, and the C ABI server's error path when
no device is present (its scheduling runs on the GPU tests)."""
import ctypes

import numpy as np
import pytest

from mdcraft_amd import engine, serve
from mdcraft_amd.atoms import make_atoms


def req(rid, n=10, steps=0):
    return serve.Request(rid, "lj", make_atoms(n), (np.zeros(3), np.ones(3), (True,) * 3), 1e-3, steps, 0.0)


def test_poisson_trace_is_ordered_at_the_rate():
    reqs = [req(q) for q in range(4000)]
    tr = serve.poisson_trace(reqs, 250.0, seed=3)
    t = np.array([off for off, _ in tr])
    assert np.all(np.diff(t) >= 0) and [r.rid for _, r in tr] == list(range(4000))
    assert abs(len(t) / t[-1] - 250.0) < 0.05 * 250.0
    sat = serve.poisson_trace(reqs[:5], None)
    assert all(off == 0.0 for off, _ in sat)


def test_latency_stats():
    reqs = []
    for q in range(100):
        r = req(q, n=100 + q)
        r.arrival = 0.01 * q
        r.done = r.arrival + 0.002 * (q + 1)
        reqs.append(r)
    st = serve.latency_stats(reqs)
    lat = np.array([0.002 * (q + 1) for q in range(100)])
    assert st["requests"] == 100
    assert st["p50_ms"] == pytest.approx(1e3 * np.percentile(lat, 50))
    assert st["p99_ms"] == pytest.approx(1e3 * np.percentile(lat, 99))
    assert st["structures_per_s"] == pytest.approx(100 / (reqs[-1].done - reqs[0].arrival))
    assert st["atoms_per_request"] == pytest.approx(149.5)


def test_server_without_device_reports_an_error():
    lib = engine.load_library()
    h = ctypes.c_void_p()
    assert lib.mdc_serve_create(-1, 1e9, 10, 4, ctypes.byref(h)) != 0
    assert lib.mdc_last_error()
    assert lib.mdc_serve_create(0, -1.0, 10, 4, ctypes.byref(h)) != 0
