"""Loading the golden fixtures of tests/golden (written by
tests/golden/make_golden.py from the oracle)."""
import glob
import json
import os

import numpy as np

from mdcraft_amd import workloads as W

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def names():
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*.npz")))


def load(name, tmpdir):
    """(meta, atoms_in, box, pot, outputs); the potential is rebuilt from
    its generator and checked against the stored checksum."""
    from tests.golden.make_golden import table_checksum
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    meta = json.loads(z["meta"].tobytes().decode())
    box = (z["lo"], z["hi"], tuple(meta["periodic"]))
    spec = meta["potential"]
    if spec["kind"] == "lj":
        pot = {k: spec[k] for k in ("kind", "a", "s", "rc")}
    elif spec["kind"] == "eam":
        pot = dict(kind="eam", **W.eam_cu_tables())
        assert table_checksum(pot) == spec["checksum"], "EAM tables changed since the fixture was written"
    elif spec["kind"] == "eam_multi":
        pot = W.nial_pots()
        assert [table_checksum(q) for _, _, q in pot["pairs"]] == spec["checksum"], "NiAl tables changed"
    else:
        path = os.path.join(str(tmpdir), name + ".maise")
        with open(path, "w") as f:
            f.write(spec["model"])
        pot = {"kind": "bpnn", "path": path, "rc": spec["rc"]}
    outs = {k: z[k] for k in z.files if k not in ("atoms_in", "lo", "hi", "meta")}
    return meta, z["atoms_in"], box, pot, outs
