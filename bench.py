"""Benchmark: atom-steps/s of the force-and-integration path on one MI355X.

Headline workload: EAM Cu, fcc a = 3.615 A, 20^3 cells = 32000 atoms,
periodic, tabulated on 5000-point grids to 5.0 A, dt = 1 fs, mixed
precision (FP32 pair arithmetic, FP64 integration / density sum /
energies).  A step is one velocity-Verlet step of the whole system (two
force passes, fused integration, skin-triggered list rebuild), inputs
resident in HBM.  Extra lines in the same JSON object report the FP64 EAM
run and the FP64 LJ argon workload (4000 atoms, dt = 2 fs).

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "atom-steps/s (LJ, EAM, BPNN) at 1 GPU; served structures/s with p50/p99 latency"
HBM_PEAK_GBS = 8000.0
EV_KJMOL = 96.48533212331


def algorithmic_bytes(kind, nbar):
    """Compulsory HBM bytes per atom for one launch of the dominant kernel
    (the pair force kernel): position read (24), list (4 per neighbour),
    velocity read + write (48), force write (12 FP32 / 24 FP64), drifted
    position write (24); EAM adds the F'(rho) read (8)."""
    return {"eam_mixed": 116 + 4 * nbar, "eam_fp64": 128 + 4 * nbar, "lj_fp64": 120 + 4 * nbar,
            "lj_mixed": 108 + 4 * nbar}[kind]


def step_bytes(kind, nbar, k):
    """Per atom-step bytes of the whole step (shared formula): LJ
    120 + 4 n + (64 + 4 n)/k; EAM 160 + 8 n + (64 + 4 n)/k."""
    reb = (64 + 4 * nbar) / max(k, 1)
    return (160 + 8 * nbar + reb) if kind.startswith("eam") else (120 + 4 * nbar + reb)


def read_traffic(kernel_tag):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE,
    WRITE_SIZE), scaled by a copy-kernel calibration, from profiles/."""
    path = os.path.join(ROOT, "profiles", "traffic_%s.txt" % kernel_tag)
    if not os.path.exists(path):
        return None
    vals = {}
    for ln in open(path):
        if ":" in ln:
            k, v = ln.split(":", 1)
            try:
                vals[k.strip()] = float(v.split()[0])
            except (ValueError, IndexError):
                pass
    return vals.get("hbm_bytes_per_launch")


def run_gpu(E, ctx, w, precision, steps, warmup, lanes=None):
    pot = w["pot"]
    at = w["atoms"]
    s = E.System(ctx, [len(at)], [w["box"]], True, lanes=lanes)
    s.upload(at)
    if pot["kind"] == "lj":
        p = E.LJsCut(ctx, pot["a"], pot["s"], pot["rc"])
    else:
        p = E.EAM(ctx, rcut=pot["rc"], F=pot["F"], rho=pot["rho"], rphi=pot["rphi"], dr=pot["dr"], drho=pot["drho"])
    s.build()
    s.compute(p, E.FORCE | E.ENERGY | E.VIRIAL, precision)
    e0 = s.thermo()[0]
    if warmup:
        s.step(p, w["dt"], warmup, precision)
    s.sync()
    t0 = time.perf_counter()
    reb = s.step(p, w["dt"], steps, precision)
    s.sync()
    t1 = time.perf_counter()
    st = s.stats()
    # kernel durations from hipEvents on the launch stream, in a separate
    # run so the events do not perturb the timed region
    s.timing(1)
    s.step(p, w["dt"], min(steps, 200), precision)
    s.sync()
    ms, nl = s.timing(0)
    s.compute(p, E.FORCE | E.ENERGY | E.VIRIAL, precision)
    e1 = s.thermo()[0]
    n = len(at)
    t_ps = (steps + warmup) * w["dt"]
    drift = (e1[3] - e0[3]) / n / EV_KJMOL / (t_ps * 1e-3)  # eV/atom/ns
    kinds = ["force", "density", "predrift", "bp_desc", "bp_mlp", "bp_force", "bp_gather", "rebuild"]
    per_kernel = {kinds[k]: round(ms[k] / nl[k] * 1e3, 2) for k in range(8) if nl[k] > 0}
    return {
        "n": n, "steps": steps, "seconds": t1 - t0, "value": n * steps / (t1 - t0), "kernel_us": per_kernel,
        "ms_per_step": 1e3 * (t1 - t0) / steps, "kernel_ms": ms.tolist(), "kernel_launches": nl.tolist(),
        "nbar": st["mean"], "rebuilds": reb, "drift_eV_atom_ns": drift, "lanes": lanes,
    }


def cpu_baseline(w, max_seconds=20.0, max_steps=None):
    """The FP64 oracle (the reference algorithm: AoS records, full lists
    with self, Periodic ghost pass, decide()/update() rebuilds, std::thread
    pool with the reference chunking rule) on the host cores."""
    from oracle import refcpu as R
    threads = min(16, os.cpu_count() or 1)
    pot = w["pot"]
    a = w["atoms"].copy()
    dom = R.Domain(*w["box"])
    pool = R.Pool(threads)
    p = R.LJsCut(pot["a"], pot["s"], pot["rc"]) if pot["kind"] == "lj" else R.EAM(
        rc=pot["rc"], F=pot["F"], rho=pot["rho"], rphi=pot["rphi"], dr=pot["dr"], drho=pot["drho"])
    a["rcut"] = p.rcut
    nl = R.VerletList(pool)
    nl.update(a, dom)
    s = R.Solver(p, pool, True)
    s.prepare(a, dom, nl)
    s.stages(a, dom, nl)
    t0 = time.perf_counter()
    s.run_nve(a, dom, nl, w["dt"], 2, False)
    per = (time.perf_counter() - t0) / 2
    steps = max(2, int(max_seconds / max(per, 1e-6)))
    if max_steps:
        steps = min(steps, max_steps)
    t0 = time.perf_counter()
    s.run_nve(a, dom, nl, w["dt"], steps, False)
    dt = time.perf_counter() - t0
    return {"value": len(a) * steps / dt, "unit": "atom-steps/s", "cores": threads, "kind": "port",
            "sample": "%s, %d atoms, %d FP64 NVE steps with decide()/update() rebuilds, %d threads, %.1f s" % (
                w["name"], len(a), steps, threads, dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--lanes", type=int, default=None)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU holding one domain; multi-GPU runs are not part of this build "
              "(see the sharding notes in DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    from mdcraft_amd import engine as E
    from mdcraft_amd import workloads as W

    ctx = E.Context(0)
    w = W.eam_cu()
    head = run_gpu(E, ctx, w, E.MIXED, args.steps, args.warmup, args.lanes)
    nbar = head["nbar"]
    avg_ms = head["kernel_ms"][0] / max(head["kernel_launches"][0], 1)
    per_launch = algorithmic_bytes("eam_mixed", nbar) * head["n"]
    achieved = per_launch / (avg_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": read_traffic("eam_force_mixed"),
            "kernel": "k_eam_force<float>", "avg_launch_us": round(avg_ms * 1e3, 2),
            "algorithmic_bytes_per_launch": int(per_launch)}
    k = args.steps / max(head["rebuilds"], 1)
    step_gbs = step_bytes("eam", nbar, k) * head["value"] / 1e9
    keys = ("value", "ms_per_step", "nbar", "rebuilds", "drift_eV_atom_ns", "kernel_us")
    extra = {"eam_cu_mixed": {kk: head[kk] for kk in keys},
             "eam_cu_mixed_step_GBs": round(step_gbs, 1)}
    if not args.no_extra:
        r = run_gpu(E, ctx, w, E.FP64, args.steps, args.warmup, args.lanes)
        extra["eam_cu_fp64"] = {kk: r[kk] for kk in keys}
        wl = W.lj_argon()
        r = run_gpu(E, ctx, wl, E.FP64, 1000, args.warmup, args.lanes)
        extra["lj_argon_fp64"] = {kk: r[kk] for kk in keys}
    cpu = None if args.no_cpu else cpu_baseline(w, max_steps=500)
    line = {
        "metric": METRIC, "value": round(head["value"], 1), "unit": "atom-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(head["ms_per_step"], 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic (seeded fcc Cu, analytic EAM tables)",
        "config": {"workload": "eam_cu: fcc Cu 20^3 cells, 32000 atoms, periodic, 5000-point EAM tables, rc 5.0 A, "
                               "skin 0.3 A, dt 1 fs, mixed precision", "atoms": head["n"],
                   "parallelism": "single GPU, one domain"},
        "roofline": roof, "cpu_baseline": cpu, "workloads": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
