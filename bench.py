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


# mdc_sys_step_path: what a step sequence of single-element EAM ran
STEP_PATHS = {0: "FP64 differences (k_eam_force)", 1: "fixed point, global tables (k_eam_force_q)",
              2: "fixed point, tables in LDS (k_eam_force_ql)"}


def algorithmic_bytes(kind, nbar):
    """Compulsory HBM bytes per atom for one launch of the dominant kernel
    (the pair force kernel): position read (24), list (4 per neighbour),
    velocity read + write (48), force write (12 FP32 / 24 FP64), drifted
    position write (24); EAM adds the F'(rho) read (8)."""
    return {"eam_mixed": 116 + 4 * nbar, "eam_fp64": 128 + 4 * nbar, "lj_fp64": 120 + 4 * nbar,
            "lj_mixed": 108 + 4 * nbar}[kind]


def eval_force_bytes(nbar, nk, precision_bytes):
    """Compulsory HBM bytes per atom of one evaluation-mode launch of the EAM
    force kernel with energy and virial (the batched E/F/virial workload):
    position record read (32), list count (4) and list (4 per neighbour),
    F'(rho) under each of the nk objects of the atom's kind (8 each), force
    write (12 FP32 / 24 FP64), pair energy (8) and virial (48) writes."""
    return 32 + 4 + 4 * nbar + 8 * nk + 3 * precision_bytes + 8 + 48


def step_bytes(kind, nbar, k):
    """Per atom-step bytes of the whole step (shared formula): LJ
    120 + 4 n + (64 + 4 n)/k; EAM 160 + 8 n + (64 + 4 n)/k."""
    reb = (64 + 4 * nbar) / max(k, 1)
    return (160 + 8 * nbar + reb) if kind.startswith("eam") else (120 + 4 * nbar + reb)


def read_traffic(name):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE,
    WRITE_SIZE), scaled by a copy-kernel calibration, from the summary
    profiles/<name>.txt."""
    path = os.path.join(ROOT, "profiles", name + ".txt")
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


def make_pot(E, ctx, pot):
    if pot["kind"] == "lj":
        return E.LJsCut(ctx, pot["a"], pot["s"], pot["rc"])
    if pot["kind"] == "eam":
        return E.EAM(ctx, rcut=pot["rc"], F=pot["F"], rho=pot["rho"], rphi=pot["rphi"], dr=pot["dr"],
                     drho=pot["drho"])
    if pot["kind"] == "eam_multi":
        return E.Multi(ctx, pot["nkinds"], [(i, j, make_pot(E, ctx, q)) for i, j, q in pot["pairs"]])
    if pot["kind"] == "alloy":
        return E.Alloy(ctx, rcut=pot["rc"], F=pot["F"], rho=pot["rho"], rphi=pot["rphi"], dr=pot["dr"],
                       drho=pot["drho"])
    if pot["kind"] == "bpnn":
        return E.MaiseBPNNPotential(ctx, pot["path"])
    raise ValueError(pot["kind"])


KINDS = ["force", "density", "predrift", "bp_desc", "bp_mlp", "bp_force", "bp_gather", "rebuild", "bp_fused"]
MLP_FLOP_PER_ATOM = 2 * (16 * 30 + 30 * 30 + 30) + 2 * (30 + 30 * 30 + 30 * 16)  # forward + input gradient
MFMA_FP16_PEAK_TFLOPS = 2500.0
# FP32 vector peak (published), and the rate without packed FP32 math,
# which these builds leave out
FP32_VALU_PEAK_TFLOPS = 157.3
FP32_VALU_NONPACKED_TFLOPS = 78.65


def bpnn_flops(nr=8, ne=2, lz=3, lz_used=2):
    """FP32 FLOPs of the descriptor and force passes per unit, counted from
    their formulas (a transcendental counts one, an FMA two): per neighbour
    in the cutoff (radial terms) and per neighbour pair in the cutoff of the
    centre and of each other (unordered in the descriptor pass, ordered in
    the force pass).  The ladder kernels: nr radial terms, ne G4 etas, lz
    zeta rungs of which lz_used hold G4 terms (zeta 1 and 4: rungs 1 and 4
    of 1, 2, 4); the powers of every rung below the top are formed."""
    geo = 14  # r_b - r_a, |r_c|^2, 1/|r_c|, the cutoff factor
    desc_pair = geo + 9 + 5 * ne + 2 + 2 * (lz - 1) + 4 * lz_used * ne
    # per used rung and eta the A and Z sums (two FMAs each), then SA, SAe
    # and SZ from them (three FMAs per eta)
    grad_pair = geo + 3 + 9 + 6 * ne + 4 * (lz - 1) + 8 * ne * lz_used + 6 * ne + 1 + 6 + 9 + 2 + 15
    return {"desc_neighbour": 6 * nr, "desc_pair": desc_pair, "grad_neighbour": 10 * nr + 6, "grad_pair": grad_pair}


def bpnn_units(atoms, box, rc):
    """Neighbours within rc of each atom, and unordered pairs of them within
    rc of each other, summed over the atoms (host side, once per workload)."""
    from scipy.spatial import cKDTree
    lo, hi = np.asarray(box[0]), np.asarray(box[1])
    L = hi - lo
    r = np.mod(_vec(atoms, "r") - lo, L)
    nb = cKDTree(r, boxsize=L).query_ball_point(r, rc)
    nrad = npair = 0
    for i, js in enumerate(nb):
        js = np.asarray([j for j in js if j != i], dtype=np.int64)
        d = r[js] - r[i]
        d -= np.round(d / L) * L
        d2 = (d * d).sum(1)
        d = d[(d2 >= 1e-15) & (d2 < rc * rc)]
        nrad += len(d)
        dd = ((d[:, None, :] - d[None, :, :]) ** 2).sum(-1)
        npair += int(np.count_nonzero(np.triu((dd >= 1e-15) & (dd < rc * rc), 1)))
    return nrad, npair


def bpnn_valu(kernel_us, units, atoms):
    """The descriptor and force kernels (or the fused pass, which adds the
    network) against the FP32 vector peak; kernel_us: mean durations of
    full-work launches."""
    nrad, npair = units
    f = bpnn_flops()
    desc = f["desc_neighbour"] * nrad + f["desc_pair"] * npair
    grad = f["grad_neighbour"] * nrad + f["grad_pair"] * 2 * npair
    out = {}
    for kind, flop in (("bp_desc", desc), ("bp_force", grad), ("bp_fused", desc + grad + MLP_FLOP_PER_ATOM * atoms)):
        us = kernel_us.get(kind)
        if not us:
            continue
        tf = flop / (us * 1e-6) / 1e12
        out[kind] = {"bound": "valu", "algorithmic_flop_per_launch": int(flop), "avg_launch_us": us,
                     "achieved": round(tf, 2), "peak": FP32_VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(tf / FP32_VALU_PEAK_TFLOPS, 4),
                     "frac_of_nonpacked": round(tf / FP32_VALU_NONPACKED_TFLOPS, 4)}
    out["units"] = {"neighbours": nrad, "neighbour_pairs": npair, "flop_per_unit": f}
    return out


def run_batch(E, ctx, cells, pot, dt, precision, steps, warmup, lanes=None):
    s = E.System(ctx, [len(a) for a, _ in cells], [b for _, b in cells], True, lanes=lanes)
    s.upload(np.concatenate([a for a, _ in cells]))
    p = make_pot(E, ctx, pot)
    s.build()
    tune = None
    if lanes is None:
        lanes, tune = s.autotune(p, precision, reps=10)
    s.compute(p, E.FORCE, precision)
    if warmup:
        s.step(p, dt, warmup, precision)
    s.sync()
    t0 = time.perf_counter()
    reb = s.step(p, dt, steps, precision)
    s.sync()
    t1 = time.perf_counter()
    s.timing(1)
    s.step(p, dt, min(steps, 100), precision)
    s.sync()
    ms, nl = s.timing(0)
    # full-work launch durations from back-to-back evaluations: in a step
    # sequence the launches queued behind a list rebuild exit early and
    # pull the mean above down
    s.timing(1)
    for _ in range(20):
        s.compute(p, E.FORCE | E.ENERGY, precision)
    s.sync()
    me, ne = s.timing(0)
    n = s.n
    return {"value": n * steps / (t1 - t0), "ms_per_step": 1e3 * (t1 - t0) / steps, "atoms": n,
            "systems": len(cells), "rebuilds": reb, "nbar": s.stats()["mean"], "lanes": lanes,
            "kernel_us": {KINDS[k]: round(ms[k] / nl[k] * 1e3, 2) for k in range(len(KINDS)) if nl[k] > 0},
            "kernel_us_full_work": {KINDS[k]: round(me[k] / ne[k] * 1e3, 2) for k in range(len(KINDS))
                                    if ne[k] > 0}}


def bpnn_mlp_mfma(E, ctx, wb, reps=10):
    """The BPNN network as its own MFMA kernel (k_bp_mlp_mfma): the mixed
    staged evaluation's prepare() runs descriptors and network apart, as
    two-species models always do; a step runs the network inside the fused
    pass instead.  Mean kernel duration against the dense FP16 MFMA peak."""
    s = E.System(ctx, [len(wb["atoms"])], [wb["box"]], True)
    s.upload(wb["atoms"])
    p = make_pot(E, ctx, wb["pot"])
    s.build()
    s.stage(p, E.PREPARE, E.MIXED)
    s.sync()
    s.timing(1)
    for _ in range(reps):
        s.stage(p, E.PREPARE, E.MIXED)
    s.sync()
    ms, nl = s.timing(0)
    k = KINDS.index("bp_mlp")
    if nl[k] == 0:
        return None
    us = ms[k] / nl[k] * 1e3
    tf = MLP_FLOP_PER_ATOM * len(wb["atoms"]) / (us * 1e-6) / 1e12
    return {"kernel": "k_bp_mlp_mfma (staged evaluation, prepare)", "avg_launch_us": round(us, 2),
            "algorithmic_TFLOPs": round(tf, 3), "peak_fp16_dense_TFLOPs": MFMA_FP16_PEAK_TFLOPS,
            "frac": round(tf / MFMA_FP16_PEAK_TFLOPS, 5), "flop_per_atom": MLP_FLOP_PER_ATOM}


def run_batch_eval(E, ctx, cells, pot, precision, reps, lanes=None):
    """Energy, forces and virial of every system of a batch, one density and
    one force launch per evaluation for all systems together (state
    resident; K evaluations timed)."""
    s = E.System(ctx, [len(a) for a, _ in cells], [b for _, b in cells], True, lanes=lanes)
    s.upload(np.concatenate([a for a, _ in cells]))
    p = make_pot(E, ctx, pot)
    s.build()
    threads = 256
    if lanes is None:
        lanes, threads, _ = s.autotune_tiles(p, precision, reps=10)
    what = E.FORCE | E.ENERGY | E.VIRIAL

    def timed():
        s.compute(p, what, precision)
        s.compute(p, what, precision)
        s.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            s.compute(p, what, precision)
        s.sync()
        return time.perf_counter() - t0

    eager = timed()
    # graph mode: the evaluation captured once and replayed
    s.set_graphs(True)
    graph = timed()
    gstats = s.graph_stats()
    s.set_graphs(False)
    t1, t0 = graph, 0.0
    s.timing(1)
    for _ in range(min(reps, 20)):
        s.compute(p, what, precision)
    ms, nl = s.timing(0)
    n = s.n
    nbar = s.stats()["mean"]
    out = {"value": n * reps / (t1 - t0), "unit": "atom-evaluations/s", "ms_per_eval": 1e3 * (t1 - t0) / reps,
           "mode": "graph-captured evaluation, replayed", "graph": gstats,
           "eager": {"value": n * reps / eager, "ms_per_eval": 1e3 * eager / reps},
           "atoms": n, "systems": len(cells), "nbar": nbar, "lanes": lanes, "threads_per_workgroup": threads,
           "kernel_us": {KINDS[k]: round(ms[k] / nl[k] * 1e3, 2) for k in range(len(KINDS)) if nl[k] > 0}}
    if nl[0] > 0 and pot["kind"] in ("eam_multi", "alloy", "eam"):
        nk = pot.get("nkinds", 1) if pot["kind"] == "eam_multi" else (
            len(pot["F"]) if pot["kind"] == "alloy" and np.ndim(pot["F"]) == 2 else 1)
        per_launch = eval_force_bytes(nbar, nk, 4 if precision == E.MIXED else 8) * n
        us = ms[0] / nl[0] * 1e3
        gbs = per_launch / (us * 1e-6) / 1e9
        out["roofline"] = {"bound": "hbm", "kernel": "k_eam_force (evaluation mode, energy and virial)",
                           "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": round(gbs / HBM_PEAK_GBS, 4), "avg_launch_us": round(us, 2),
                           "algorithmic_bytes_per_launch": int(per_launch), "traffic": None}
        if pot["kind"] == "eam_multi" and precision == E.MIXED:
            # measured for this configuration (Multi, mixed precision)
            out["roofline"]["traffic"] = read_traffic("k_eam_force_nial_batch512")
    return out


def one_at_a_time(E, ctx, cells, pot, reps):
    """The plain FP64 GPU path: each system alone in its own engine System,
    evaluated in turn (state resident; creation and upload not timed)."""
    p = make_pot(E, ctx, pot)
    what = E.FORCE | E.ENERGY | E.VIRIAL
    systems = []
    for a, b in cells:
        s = E.System(ctx, [len(a)], [b], True)
        s.upload(a)
        s.build()
        s.compute(p, what, E.FP64)
        systems.append(s)
    t0 = time.perf_counter()
    for _ in range(reps):
        for s in systems:
            s.compute(p, what, E.FP64)
    t1 = time.perf_counter()
    n = sum(len(a) for a, _ in cells)
    return {"value": n * reps / (t1 - t0), "unit": "atom-evaluations/s", "systems": len(cells), "atoms": n}


def served_stream(E, ctx, W, nreq=120, factors=(0.5, 1.0, 2.0)):
    """Served mixed stream: a saturation run (every request at t = 0) gives
    the capacity C; then Poisson arrivals at f x C for each factor, served
    (LJ FP64, EAM and BPNN mixed) and by the plain FP64 one-at-a-time path
    on the same traces."""
    from mdcraft_amd import serve
    model = W.write_maise_model(os.path.join(os.environ.get("TMPDIR", "/tmp"), "mdc_served_cu.maise"))
    specs = W.served_requests(nreq, model_path=model)
    pots = {}
    for sp in specs:
        if sp["kind"] not in pots:
            pots[sp["kind"]] = make_pot(E, ctx, sp["pot"])
    prec = {"lj": E.FP64, "eam": E.MIXED, "bpnn": E.MIXED}

    def mk():
        return [serve.Request(q, sp["kind"], sp["atoms"], sp["box"], sp["dt"], sp["steps"], 0.0)
                for q, sp in enumerate(specs)]

    def served(trace):
        srv = serve.Server(E, ctx, pots, prec)
        t0 = time.perf_counter()
        done = srv.run_trace(trace)
        wall = time.perf_counter() - t0
        st = serve.latency_stats(done)
        info = srv.stats()
        srv.close()
        st["peak_open_batch_GB"] = info["peak_bytes"] / 1e9
        st["batches_opened"] = int(info["batches_opened"])
        return st, wall

    # warm both paths (kernels, allocator) on a few requests
    served(serve.poisson_trace(mk()[:6], None))
    serve.one_at_a_time(E, ctx, pots, serve.poisson_trace(mk()[:6], None))
    sat, sat_wall = served(serve.poisson_trace(mk(), None))
    base_sat = serve.latency_stats(serve.one_at_a_time(E, ctx, pots, serve.poisson_trace(mk(), None)))
    cap = sat["structures_per_s"]
    # The whole stream against the HBM roof: every request's atom-steps
    # (steps plus its final evaluation) at the per-atom-step algorithmic
    # bytes of its potential with the workloads' mean list lengths (LJ
    # 120 + 4 n, n = 79; EAM 160 + 8 n, n = 53; BPNN 120 + 8 n + 128, n = 89),
    # over the saturated run's wall time.
    per = {"lj": 120 + 4 * 79, "eam": 160 + 8 * 53, "bpnn": 120 + 8 * 89 + 128}
    alg = sum(len(sp["atoms"]) * (sp["steps"] + 1) * per[sp["kind"]] for sp in specs)
    gbs = alg / sat_wall / 1e9
    out = {"requests": nreq, "mean_atoms": float(np.mean([len(sp["atoms"]) for sp in specs])),
           "single_point_fraction": float(np.mean([sp["steps"] == 0 for sp in specs])),
           "hbm_budget_GB": 200, "saturation": sat, "one_at_a_time_fp64_saturation": base_sat,
           "saturation_vs_one_at_a_time": round(sat["structures_per_s"] / base_sat["structures_per_s"], 3),
           "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(gbs / HBM_PEAK_GBS, 4), "algorithmic_bytes": int(alg),
                        "wall_s": round(sat_wall, 4), "note": "whole saturated stream, all kernels"},
           "rates": []}
    for f in factors:
        rate = f * cap
        st, _ = served(serve.poisson_trace(mk(), rate, seed=int(100 * f)))
        bt = serve.latency_stats(serve.one_at_a_time(E, ctx, pots, serve.poisson_trace(mk(), rate, seed=int(100 * f))))
        out["rates"].append({"factor": f, "rate_per_s": rate, "served": st, "one_at_a_time_fp64": bt})
    return out


def run_gpu(E, ctx, w, precision, steps, warmup, lanes=None):
    pot = w["pot"]
    at = w["atoms"]
    s = E.System(ctx, [len(at)], [w["box"]], True, lanes=lanes)
    s.upload(at)
    p = make_pot(E, ctx, pot)
    s.build()
    tune = None
    if lanes is None:
        lanes, tune = s.autotune(p, precision, reps=20)
    s.compute(p, E.FORCE | E.ENERGY | E.VIRIAL, precision)
    if warmup:
        s.step(p, w["dt"], warmup, precision)
    # mean durations of the step-mode pair kernels: back-to-back launches on
    # a fresh list of the state the timed steps start from (the hook leaves
    # positions, velocities and forces as they are)
    s.build()
    nbar_b2b = s.stats()["mean"]
    b2b = s.time_kernels(p, precision, w["dt"], reps=200) if pot["kind"] in ("lj", "eam") else (0.0, 0.0)
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
    path = s.step_path(-1) if pot["kind"] == "eam" else None
    n = len(at)
    per_kernel = {KINDS[k]: round(ms[k] / nl[k] * 1e3, 2) for k in range(len(KINDS)) if nl[k] > 0}
    return {
        "n": n, "steps": steps, "seconds": t1 - t0, "value": n * steps / (t1 - t0), "kernel_us": per_kernel,
        "ms_per_step": 1e3 * (t1 - t0) / steps, "kernel_ms": ms.tolist(), "kernel_launches": nl.tolist(),
        "nbar": st["mean"], "rebuilds": reb, "lanes": lanes, "path": STEP_PATHS.get(path),
        "autotune_us": tune, "b2b_us": {"density": round(b2b[0], 2), "force": round(b2b[1], 2)},
        "nbar_b2b": nbar_b2b,
    }


def nve_drift(E, ctx, w, precision, ps=1.0, samples=21, lanes=None):
    """NVE energy drift: the workload from its initial state for `ps`
    picoseconds, the total energy (potential + kinetic, per atom) read at
    `samples` evenly spaced times, and the slope of its least-squares line
    in eV/atom/ns.  Every step between the first and last reading counts."""
    at = w["atoms"]
    s = E.System(ctx, [len(at)], [w["box"]], True, lanes=lanes)
    s.upload(at)
    p = make_pot(E, ctx, w["pot"])
    s.build()
    what = E.FORCE | E.ENERGY | E.VIRIAL
    nsteps = int(round(ps / w["dt"]))
    chunk = max(1, nsteps // (samples - 1))
    t, e = [], []
    s.compute(p, what, precision)
    done = 0
    while True:
        th = s.thermo()[0]
        t.append(done * w["dt"])
        e.append(th[3] / len(at) / EV_KJMOL)
        if done >= nsteps:
            break
        k = min(chunk, nsteps - done)
        s.step(p, w["dt"], k, precision)
        done += k
        s.compute(p, what, precision)  # energies at the new positions (the step sequence leaves them)
    t, e = np.array(t), np.array(e)
    slope = np.polyfit(t, e, 1)[0] * 1e3  # eV/atom/ps -> eV/atom/ns
    return {"drift_eV_atom_ns": float(slope), "fit": "least-squares line of E_total/atom over %d readings" % len(t),
            "span_ps": float(t[-1]), "steps": int(done), "dt_fs": w["dt"] * 1e3,
            "E_total_eV_atom": [round(float(e[0]), 6), round(float(e[-1]), 6)],
            "rms_about_line_eV_atom": float(np.std(e - np.polyval(np.polyfit(t, e, 1), t)))}


COPY_MIB = 256


def copy_calibration_buffers():
    """Two 256 MiB device buffers for the calibration kernel, allocated
    before the engine touches the device (torch initialised after another
    HIP user has crashed under rocprofv3 --pmc)."""
    import torch
    a = torch.ones(COPY_MIB << 18, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    return a, torch.empty_like(a)


def copy_calibration(bufs, E=None):
    """Two kernels of known traffic, each reading 256 MiB and writing 256 MiB:
    the engine's hand-written copy (k_copy_calibration, 16-byte loads and
    stores), whose FETCH_SIZE / WRITE_SIZE readings under rocprofv3 --pmc
    scale the counters of the engine's kernels, and torch's elementwise
    multiply as a second, independent calibration."""
    import torch
    a, b = bufs
    for _ in range(3):
        torch.mul(a, 2.0, out=b)
    torch.cuda.synchronize()
    if E is not None:
        E.copy_calibration(0, COPY_MIB << 20, 3)


def _csv_rows(d, suffix):
    import csv
    import glob
    rows = []
    for f in sorted(glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True)):
        rows += list(csv.DictReader(open(f)))
    if not rows:
        raise SystemExit("no %s under %s" % (suffix, d))
    return rows


def summarise_pmc(fetch_dir, write_dir, kernel):
    """Per-launch HBM bytes of `kernel` from two rocprofv3 --pmc runs
    (FETCH_SIZE, WRITE_SIZE; KiB per dispatch), each scaled by the
    counter's reading for the 256 MiB calibration kernel."""
    out = []
    res = {}
    for tag, d, ctr in (("fetch", fetch_dir, "FETCH_SIZE"), ("write", write_dir, "WRITE_SIZE")):
        rows = [r for r in _csv_rows(d, "counter_collection.csv") if r["Counter_Name"] == ctr]
        mine = sorted(float(r["Counter_Value"]) for r in rows if kernel in r["Kernel_Name"])
        cal = [float(r["Counter_Value"]) for r in rows if "k_copy_calibration" in r["Kernel_Name"]]
        cal_torch = [float(r["Counter_Value"]) for r in rows if "MulFunctor" in r["Kernel_Name"]]
        if not mine or not cal:
            raise SystemExit("missing %s rows for %s or the calibration kernel" % (ctr, kernel))
        # full-work dispatches only (launches queued ahead of a rebuild exit at once)
        med = mine[len(mine) // 2]
        mine = [v for v in mine if v >= 0.5 * med]
        kib, cal_kib = sum(mine) / len(mine), float(np.median(cal))
        factor = COPY_MIB * 1024.0 / cal_kib
        res[tag] = kib * factor * 1024.0
        out += ["%s_launches: %d" % (tag, len(mine)), "%s_kib_per_launch: %.2f" % (tag, kib),
                "%s_calibration_kib: %.1f (expected %d; k_copy_calibration, %d dispatches)" % (
                    tag, cal_kib, COPY_MIB * 1024, len(cal)),
                "%s_factor: %.5f" % (tag, factor)]
        if cal_torch:
            out.append("%s_factor_torch_elementwise: %.5f (cross-check)" % (tag, COPY_MIB * 1024.0 / max(cal_torch)))
    out.append("hbm_bytes_per_launch: %.0f" % (res["fetch"] + res["write"]))
    return "\n".join(out)


def summarise_counters(d, kernel):
    """Median per dispatch of every counter of a rocprofv3 --pmc run for the
    dispatches of `kernel` (full-work dispatches only: launches queued ahead
    of a list rebuild exit at once and read a fraction of the rest)."""
    rows = [r for r in _csv_rows(d, "counter_collection.csv") if kernel in r["Kernel_Name"]]
    if not rows:
        raise SystemExit("no rows for %s under %s" % (kernel, d))
    by = {}
    for r in rows:
        by.setdefault(r["Counter_Name"], []).append(float(r["Counter_Value"]))
    names = sorted({r["Kernel_Name"].split("(")[0] for r in rows})
    out = ["kernel: " + ", ".join(names)]
    for c in sorted(by):
        v = np.sort(np.array(by[c]))
        full = v[v >= 0.5 * np.median(v)] if len(v) else v
        out.append("%-34s dispatches %5d  median %.1f" % (c, len(full), float(np.median(full))))
    return "\n".join(out)


def summarise_stats(d, kernel):
    """Kernel table from rocprofv3 --kernel-trace --stats (CSV output), and
    for `kernel` the mean over full-work dispatches: launches queued ahead
    of a list rebuild exit at once (a few microseconds) and are counted
    apart."""
    import numpy as np
    tr = [r for r in _csv_rows(d, "kernel_trace.csv") if kernel in r["Kernel_Name"]]
    dur = np.array([(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-3 for r in tr])
    full = dur[dur >= 0.5 * np.median(dur)] if len(dur) else dur
    head = ["%s: %d dispatches, mean %.2f us; %d full-work dispatches, mean %.2f us; %d early exits, mean %.2f us"
            % (kernel, len(dur), dur.mean(), len(full), full.mean(), len(dur) - len(full),
               dur[dur < 0.5 * np.median(dur)].mean() if len(dur) > len(full) else 0.0), ""]
    rows = _csv_rows(d, "kernel_stats.csv")
    lines = head + ["%-60s %8s %12s %12s %7s" % ("kernel", "calls", "avg_us", "total_us", "pct")]
    for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"])):
        name = r["Name"].replace("HIP_vector_type<double, 4u>", "double4")
        name = name.split("(")[0] if name.startswith(("mdc::", "void mdc::")) else name[:60]
        lines.append("%-60s %8s %12.2f %12.1f %7.2f" % (name[:60], r["Calls"], float(r["AverageNs"]) / 1e3,
                                                     float(r["TotalDurationNs"]) / 1e3, float(r["Percentage"])))
    return "\n".join(lines)


def engine_check(E, ctx, w):
    """The engine's FP64 and mixed evaluation (forces, energies, virial) of
    the workload's first configuration and its neighbour pairs, for the
    oracle's comparison in cpu_baseline."""
    s = E.System(ctx, [len(w["atoms"])], [w["box"]], True)
    s.upload(w["atoms"])
    p = make_pot(E, ctx, w["pot"])
    s.build()
    out = {}
    for name, prec in (("fp64", E.FP64), ("mixed", E.MIXED)):
        s.compute(p, E.FORCE | E.ENERGY | E.VIRIAL, prec)
        out[name] = s.download()
    out["pairs"] = s.export()
    s.close()
    return out


def _vec(a, f):
    return np.stack([a[f]["x"], a[f]["y"], a[f]["z"]], axis=1)


def _pair_keys(off, ids):
    n = len(off) - 1
    return np.sort(np.repeat(np.arange(n, dtype=np.int64), np.diff(off)) * n + ids)


def accuracy(check, ref, ref_pairs):
    """Errors of the engine's results against the oracle's on the same
    configuration: FP64 max |df| over max |f|, mixed force RMS over the
    force RMS, relative total energy and virial trace; the neighbour pair
    sets compared exactly."""
    fr = _vec(ref, "f")
    e_ref = ref["Ep"].sum() + ref["Em"].sum()
    w_ref = (ref["V"]["xx"] + ref["V"]["yy"] + ref["V"]["zz"]).sum()
    out = {}
    for name in ("fp64", "mixed"):
        a = check[name]
        f = _vec(a, "f")
        e = a["Ep"].sum() + a["Em"].sum()
        wv = (a["V"]["xx"] + a["V"]["yy"] + a["V"]["zz"]).sum()
        out[name] = {"force_max_rel": float(np.abs(f - fr).max() / np.abs(fr).max()),
                     "force_rms_rel": float(np.sqrt(np.mean(np.sum((f - fr) ** 2, 1)) / np.mean(np.sum(fr ** 2, 1)))),
                     "energy_rel": float(abs(e - e_ref) / abs(e_ref)), "virial_trace_rel": float(abs(wv - w_ref) / abs(w_ref))}
    got, want = _pair_keys(*check["pairs"]), _pair_keys(*ref_pairs)
    out["pair_sets_equal"] = bool(len(got) == len(want) and np.array_equal(got, want))
    out["pairs"] = int(len(want))
    out["bars"] = "FP64 <= 1e-12 relative; mixed force RMS <= 1e-5; pair sets bit-exact"
    return out


def cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, max_seconds=3.0, max_steps=None, check=None, runs=3):
    """The FP64 oracle (the reference algorithm: AoS records, full lists
    with self, Periodic ghost pass, decide()/update() rebuilds, std::thread
    pool with the reference chunking rule) on the host's cores and on one
    core, each the median of `runs` timed runs of about max_seconds; with
    check (engine_check's results) the oracle also evaluates the same first
    configuration as the checker of the engine's.  The thread count is the
    cores this process may run on, at most 16: the CPU share of a one-GPU
    host here (nproc is recorded beside it)."""
    from oracle import refcpu as R
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    threads = max(1, min(16, avail))
    pot = w["pot"]
    dom = R.Domain(*w["box"])
    p = R.LJsCut(pot["a"], pot["s"], pot["rc"]) if pot["kind"] == "lj" else R.EAM(
        rc=pot["rc"], F=pot["F"], rho=pot["rho"], rphi=pot["rphi"], dr=pot["dr"], drho=pot["drho"])

    def timed(nthreads, seconds, cap):
        pool = R.Pool(nthreads)
        a = w["atoms"].copy()
        a["rcut"] = p.rcut
        nl = R.VerletList(pool)
        nl.update(a, dom)
        s = R.Solver(p, pool, True)
        s.prepare(a, dom, nl)
        s.stages(a, dom, nl)
        t0 = time.perf_counter()
        s.run_nve(a, dom, nl, w["dt"], 2, False)
        per = (time.perf_counter() - t0) / 2
        steps = max(2, int(seconds / max(per, 1e-6)))
        if cap:
            steps = min(steps, cap)
        t0 = time.perf_counter()
        s.run_nve(a, dom, nl, w["dt"], steps, False)
        dt = time.perf_counter() - t0
        return len(a) * steps / dt, steps, dt

    def median_of(nthreads, seconds):
        rs = sorted((timed(nthreads, seconds, max_steps) for _ in range(runs)), key=lambda r: r[0])
        return rs[len(rs) // 2], [round(r[0], 1) for r in rs]

    (v, steps, dt), all_v = median_of(threads, max_seconds)
    (v1, steps1, dt1), all_v1 = median_of(1, max_seconds)
    out = {"value": v, "unit": "atom-steps/s", "cores": threads, "kind": "port",
           "sample": "%s, %d atoms, %d FP64 NVE steps with decide()/update() rebuilds, %d threads, %.1f s; "
                     "median of %d runs" % (w["name"], len(w["atoms"]), steps, threads, dt, runs),
           "runs": all_v, "cpu_model": cpu_model(), "nproc": os.cpu_count(), "affinity_cpus": avail,
           "single_core": {"value": v1, "runs": all_v1,
                           "sample": "the same, %d steps on 1 thread, %.1f s; median of %d runs" % (steps1, dt1, runs)}}
    if check is not None:
        pool = R.Pool(threads)
        ref = w["atoms"].copy()
        ref["rcut"] = p.rcut
        nl = R.VerletList(pool)
        nl.update(ref, dom)
        R.Solver(p, pool, True).compute(ref, dom, nl)
        out["accuracy"] = accuracy(check, ref, nl.export(len(ref)))
    return out


def eam_sizes(E, ctx, W, lanes=None, reps=(20, 40, 64, 101)):
    """The EAM Cu step (mixed precision) at growing sizes of the same fcc
    crystal: 32 000 atoms up to 4.1 million.  Per size: throughput over
    timed steps with rebuilds, and the force kernel's back-to-back launch
    time and HBM roofline fraction on the same algorithmic-bytes formula
    as the headline line."""
    out = {}
    for r in reps:
        w = W.eam_cu(reps=r)
        steps = 200 if r <= 40 else 100
        g = run_gpu(E, ctx, w, E.MIXED, steps, 10, lanes)
        force_us = g["b2b_us"]["force"]
        per_launch = algorithmic_bytes("eam_mixed", g["nbar_b2b"]) * g["n"]
        gbs = per_launch / (force_us * 1e-6) / 1e9
        out[str(g["n"])] = {
            "value": round(g["value"], 1), "ms_per_step": round(g["ms_per_step"], 5), "steps": steps,
            "rebuilds": g["rebuilds"], "lanes": g["lanes"], "path": g["path"], "nbar": round(g["nbar"], 2),
            "b2b_us": g["b2b_us"], "kernel_us": g["kernel_us"],
            "force_roofline": {"achieved_GBs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4),
                               "algorithmic_bytes_per_launch": int(per_launch)},
        }
        print("eam_sizes: %d atoms done" % g["n"], file=sys.stderr, flush=True)
        del w, g
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--lanes", type=int, default=None)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run the 256 MiB calibration kernel (for rocprofv3 --pmc runs)")
    ap.add_argument("--summarise-pmc", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"))
    ap.add_argument("--summarise-stats", metavar="STATS_DIR")
    ap.add_argument("--summarise-counters", metavar="PMC_DIR",
                    help="median per full-work dispatch of every counter of a rocprofv3 --pmc run for --kernel")
    ap.add_argument("--kernel", default="k_eam_force_q<")
    ap.add_argument("--eam-reps", type=int, nargs="+", default=None,
                    help="--only eam_sizes: fcc cells per edge of the sizes to run (default 20 40 64 101)")
    ap.add_argument("--precision", choices=("mixed", "fp64", "both"), default="both",
                    help="--only bpnn_cu / nial_batch: precisions to run")
    ap.add_argument("--only", choices=("bpnn_cu", "nial_batch", "nial_alloy", "served", "eam_sizes", "lj_argon"),
                    help="run one of the extra workloads alone and print its JSON (for profiling)")
    args = ap.parse_args()
    if args.summarise_pmc:
        print(summarise_pmc(*args.summarise_pmc, args.kernel))
        return
    if args.summarise_stats:
        print(summarise_stats(args.summarise_stats, args.kernel))
        return
    if args.summarise_counters:
        print(summarise_counters(args.summarise_counters, args.kernel))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU holding one domain; runs on several GPUs are not supported "
              "(the engine is one process on one GPU)", file=sys.stderr)
        sys.exit(2)
    cal = copy_calibration_buffers() if args.copy_calibration else None
    from mdcraft_amd import engine as E
    from mdcraft_amd import workloads as W

    ctx = E.Context(0)
    if args.only:
        precs = [(p, pr) for p, pr in (("mixed", E.MIXED), ("fp64", E.FP64)) if args.precision in (p, "both")]
        if args.only == "bpnn_cu":
            wb = W.bpnn_cu()
            out = {p: run_batch(E, ctx, [(wb["atoms"], wb["box"])], wb["pot"], wb["dt"], pr, min(args.steps, 200),
                                10, args.lanes) for p, pr in precs}
            if "mixed" in out:
                out["mixed"]["valu"] = bpnn_valu(out["mixed"]["kernel_us_full_work"],
                                                 bpnn_units(wb["atoms"], wb["box"], wb["pot"]["rc"]), len(wb["atoms"]))
                out["mixed"]["mlp_mfma"] = bpnn_mlp_mfma(E, ctx, wb)
        elif args.only == "lj_argon":
            wl = W.lj_argon()
            r = run_gpu(E, ctx, wl, E.FP64, 1000, args.warmup, args.lanes)
            out = {"fp64": {kk: r[kk] for kk in ("value", "ms_per_step", "nbar", "rebuilds", "kernel_us", "b2b_us",
                                                 "lanes")}}
        elif args.only == "eam_sizes":
            out = eam_sizes(E, ctx, W, args.lanes, **({"reps": tuple(args.eam_reps)} if args.eam_reps else {}))
        elif args.only in ("nial_batch", "nial_alloy"):
            nb = W.nial_batch(512) if args.only == "nial_batch" else W.nial_alloy_batch(512)
            out = {p: run_batch_eval(E, ctx, nb["cells"], nb["pot"], pr, 50, args.lanes) for p, pr in precs}
        else:
            out = served_stream(E, ctx, W)
        print(json.dumps({"workload": args.only, "result": out}), flush=True)
        if cal is not None:
            copy_calibration(cal, E)
        return
    w = W.eam_cu()
    head = run_gpu(E, ctx, w, E.MIXED, args.steps, args.warmup, args.lanes)
    nbar = head["nbar"]
    # mean duration of the step-mode force kernel: 200 back-to-back launches
    # between two HIP events on the engine's stream
    avg_ms = head["b2b_us"]["force"] * 1e-3
    per_launch = algorithmic_bytes("eam_mixed", head["nbar_b2b"]) * head["n"]
    achieved = per_launch / (avg_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": read_traffic("traffic_eam_force_mixed"),
            "kernel": head["path"], "avg_launch_us": round(avg_ms * 1e3, 2),
            "avg_launch_us_in_steps": round(head["kernel_ms"][0] / max(head["kernel_launches"][0], 1) * 1e3, 2),
            "algorithmic_bytes_per_launch": int(per_launch)}
    k = args.steps / max(head["rebuilds"], 1)
    step_gbs = step_bytes("eam", nbar, k) * head["value"] / 1e9
    keys = ("value", "ms_per_step", "nbar", "rebuilds", "kernel_us", "b2b_us", "lanes", "path")
    extra = {"eam_cu_mixed": {kk: head[kk] for kk in keys},
             "eam_cu_mixed_step_GBs": round(step_gbs, 1)}
    # NVE energy drift: 1 ps of each workload from its initial state, slope
    # of a line fitted to the total energy
    drift = {"eam_cu_mixed": nve_drift(E, ctx, w, E.MIXED), "eam_cu_fp64": nve_drift(E, ctx, w, E.FP64)}
    extra["nve_drift"] = drift
    if not args.no_extra:
        drift["lj_argon_fp64"] = nve_drift(E, ctx, W.lj_argon(), E.FP64, ps=2.0)
        wbd = W.bpnn_cu()
        drift["bpnn_cu_mixed"] = nve_drift(E, ctx, wbd, E.MIXED)
        drift["bpnn_cu_fp64"] = nve_drift(E, ctx, wbd, E.FP64)
        r = run_gpu(E, ctx, w, E.FP64, args.steps, args.warmup, args.lanes)
        extra["eam_cu_fp64"] = {kk: r[kk] for kk in keys}
        wl = W.lj_argon()
        r = run_gpu(E, ctx, wl, E.FP64, 1000, args.warmup, args.lanes)
        extra["lj_argon_fp64"] = {kk: r[kk] for kk in keys}
        # 512 NiAl cells (solver::Multi over three EAM objects), one batch
        nb = W.nial_batch(512)
        cells = nb["cells"]
        for name, prec in (("nial_batch512_mixed", E.MIXED), ("nial_batch512_fp64", E.FP64)):
            extra[name] = run_batch_eval(E, ctx, cells, nb["pot"], prec, 50)
        # the same cells under the two-element alloy EAM (per-element
        # densities and embeddings, a pair term per pair of elements)
        na = W.nial_alloy_batch(512)
        for name, prec in (("nial_alloy_batch512_mixed", E.MIXED), ("nial_alloy_batch512_fp64", E.FP64)):
            extra[name] = run_batch_eval(E, ctx, na["cells"], na["pot"], prec, 50)
        seq = one_at_a_time(E, ctx, cells[:64], nb["pot"], 5)
        seq["note"] = "first 64 of the 512 cells, each in its own System, FP64, evaluated in turn"
        extra["nial_one_at_a_time_fp64"] = seq
        # BPNN Cu 13500 atoms
        wb = W.bpnn_cu()
        for name, prec in (("bpnn_cu_mixed", E.MIXED), ("bpnn_cu_fp64", E.FP64)):
            r = run_batch(E, ctx, [(wb["atoms"], wb["box"])], wb["pot"], wb["dt"], prec, 200, 10)
            if prec == E.MIXED:
                r["mlp_mfma"] = bpnn_mlp_mfma(E, ctx, wb)
                if "bp_fused" in r["kernel_us"]:
                    r["mlp"] = ("in the fused pass: FP32 on the vector unit, one wavefront per atom, "
                                f"{MLP_FLOP_PER_ATOM} FLOP per atom")
                r["valu"] = bpnn_valu(r["kernel_us_full_work"], bpnn_units(wb["atoms"], wb["box"], wb["pot"]["rc"]),
                                      r["atoms"])
            extra[name] = r
        extra["served_stream"] = served_stream(E, ctx, W)
        # the headline step at 32 000 to 4.1 million atoms
        extra["eam_cu_sizes_mixed"] = eam_sizes(E, ctx, W)
    cpu = None if args.no_cpu else cpu_baseline(w, max_steps=500, check=engine_check(E, ctx, w))
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
    if cal is not None:
        copy_calibration(cal, E)


if __name__ == "__main__":
    main()
