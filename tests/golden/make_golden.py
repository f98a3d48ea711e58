"""Writes the golden fixtures of this directory from the FP64 oracle.

Each fixture holds a reduced instance of one workload: its inputs (atom
records, box) and what the oracle (the reference algorithm restated in
oracle/refcpu.cpp) gives for them: forces, energies, densities and
virials after prepare + stages + virials, the pair set (as a SHA-256 of
the sorted pairs) and, for BPNN, the descriptors and dE/dG of the first 64
atoms; the NVE fixture adds positions and velocities after 20 steps.  The potential tables are rebuilt from mdcraft_amd/workloads.py,
whose checksum is stored beside them; the BPNN fixtures store the model
file text.

    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from mdcraft_amd import workloads as W  # noqa: E402
from tests.helpers import oracle_mod, oracle_potential  # noqa: E402


def pair_hash(off, ids):
    n = len(off) - 1
    i = np.repeat(np.arange(n, dtype=np.uint64), np.diff(off).astype(np.int64))
    pairs = np.stack([i, ids[:len(i)].astype(np.uint64)], 1)
    pairs = pairs[np.lexsort((pairs[:, 1], pairs[:, 0]))]
    return hashlib.sha256(pairs.tobytes()).hexdigest()


def table_checksum(pot):
    h = hashlib.sha256()
    for k in ("F", "rho", "rphi"):
        h.update(np.ascontiguousarray(pot[k], dtype=np.float64).tobytes())
    return h.hexdigest()


def potential_spec(pot):
    """How the test rebuilds the potential: workload generator + checksum."""
    k = pot["kind"]
    if k == "lj":
        return {"kind": "lj", "a": pot["a"], "s": pot["s"], "rc": pot["rc"]}
    if k == "eam":
        return {"kind": "eam", "tables": "eam_cu_tables", "checksum": table_checksum(pot)}
    if k == "eam_multi":
        return {"kind": "eam_multi", "tables": "nial_pots",
                "checksum": [table_checksum(q) for _, _, q in pot["pairs"]]}
    if k == "bpnn":
        return {"kind": "bpnn", "model": open(pot["path"]).read(), "rc": pot["rc"]}
    raise ValueError(k)


def evaluate(R, atoms, box, pot, nve_steps=0, dt=None):
    a = atoms.copy()
    dom = R.Domain(*box)
    pool = R.Pool(min(8, os.cpu_count() or 4))
    p = oracle_potential(R, pot, box)
    a["rcut"] = p.rcut
    nl = R.VerletList(pool)
    nl.update(a, dom)
    s = R.Solver(p, pool, True)
    out = {}
    if nve_steps:
        s.prepare(a, dom, nl)
        s.stages(a, dom, nl)
        s.run_nve(a, dom, nl, dt, nve_steps, initial_forces=False)
    s.compute(a, dom, nl)
    if pot["kind"] == "bpnn":
        out["G"], out["dEdG"] = R.bpnn_descriptors(p, len(a))
    off, ids = nl.export(len(a))
    return a, pair_hash(off, ids), out


def write(name, atoms, box, pot, nve_steps=0, dt=None, slow=False):
    R = oracle_mod()
    res, ph, extra = evaluate(R, atoms, box, pot, nve_steps, dt)
    meta = {"name": name, "potential": potential_spec(pot), "pair_sha256": ph, "nve_steps": nve_steps,
            "dt": dt, "periodic": [bool(x) for x in box[2]], "slow": slow}
    from mdcraft_amd.atoms import mat, vec
    outs = {"f": vec(res, "f"), "Ep": res["Ep"].copy(), "Em": res["Em"].copy(), "nc": res["nc"].copy(),
            "V": mat(res).reshape(len(res), 9)}
    if nve_steps:
        outs["r"], outs["v"] = vec(res, "r"), vec(res, "v")
    outs.update({k: v[:64] for k, v in extra.items()})
    np.savez_compressed(os.path.join(HERE, name + ".npz"), atoms_in=atoms, lo=box[0], hi=box[1],
                        meta=np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8), **outs)
    print(name, len(atoms), os.path.getsize(os.path.join(HERE, name + ".npz")), "bytes")


def main():
    w = W.lj_argon(reps=4, rattle=0.05)
    write("lj_argon_256", w["atoms"], w["box"], w["pot"])
    w = W.eam_cu(reps=4)
    write("eam_cu_256", w["atoms"], w["box"], w["pot"])
    write("eam_cu_256_nve20", w["atoms"], w["box"], w["pot"], nve_steps=20, dt=w["dt"])
    w = W.eam_cu(reps=3, rattle=0.05)
    write("eam_cu_108_short_box", w["atoms"], w["box"], w["pot"])
    at, box = W.nial_cell(3, 250)
    write("nial_multi_250", at, box, W.nial_pots())
    with tempfile.TemporaryDirectory() as d:
        for reps in (4, 5):
            w = W.bpnn_cu(reps=reps, model_dir=d)
            write("bpnn_cu_%d" % len(w["atoms"]), w["atoms"], w["box"], w["pot"])
    wide()


def wide():
    """The wide BPNN model: a 200-atom slab (periodic in x and y, open in
    z) and the 500-atom periodic crystal (its oracle evaluation takes about
    two minutes on 8 cores, so tests re-check it only on request)."""
    with tempfile.TemporaryDirectory() as d:
        w = W.bpnn_cu_wide(reps=(5, 5, 2), model_dir=d, periodic=(True, True, False))
        write("bpnn_cu_wide_slab_200", w["atoms"], w["box"], w["pot"])
        w = W.bpnn_cu_wide(reps=5, model_dir=d)
        write("bpnn_cu_wide_500", w["atoms"], w["box"], w["pot"], slow=True)


if __name__ == "__main__":
    if sys.argv[1:] == ["wide"]:
        wide()
    else:
        main()
