"""Snapshot files (mdcraft_amd.snapshot): the dataset interface of the
reference's checkpoint scripts over npz, records bit-identical after a round
trip, read without pickle, and loadable into Atoms.data."""
import os
import sys

import numpy as np
import pytest

from mdcraft_amd import refapi as A
from mdcraft_amd import snapshot
from mdcraft_amd import workloads as W
from mdcraft_amd.atoms import ATOM_DTYPE


def test_file_round_trip(tmp_path):
    w = W.eam_cu(reps=3, rattle=0.05)
    at = w["atoms"].copy()
    at["tag"][::5] = 0
    at["kind"][::3] = 1
    path = str(tmp_path / "atoms" / "run000100.h5")
    f = snapshot.File(path, "w")
    f["atoms"] = at
    f["physics"] = np.array([1e-3, 0.1, 12.0], dtype=np.float64)
    f["cycles"] = np.array([100, 10, 50], dtype=np.int64)
    f["domain"] = np.array([0, 1.0, 0, 1.0, 0, 1.0])
    f.close()
    with snapshot.File(path, "r") as g:
        assert sorted(g.keys()) == ["atoms", "cycles", "domain", "physics"]
        got = g["atoms"][:]
        # every field equal (the 6 padding bytes of a record carry nothing)
        assert got.dtype == ATOM_DTYPE and bool(np.all(got == at))
        dt, time, tend = g["physics"][:]
        step, ncbuff, ncsave = g["cycles"][:]
        assert (dt, step) == (1e-3, 100)
    # the archive holds plain arrays: numpy reads it with pickle refused
    with np.load(path + ".npz", allow_pickle=False) as z:
        assert z["atoms"].dtype.itemsize == 208
    # into the reference API's container, as continue.py does
    atoms = A.Atoms(len(got))
    atoms.data[:] = got[:]
    assert bool(np.all(atoms.data == at))


def test_modes(tmp_path):
    path = str(tmp_path / "s.npz")
    with snapshot.File(path, "w") as f:
        f["a"] = np.arange(3)
    with snapshot.File(path, "a") as f:
        f["b"] = np.ones(2)
    with snapshot.File(path) as f:
        assert f["a"].tolist() == [0, 1, 2] and f["b"].tolist() == [1.0, 1.0]
        with pytest.raises(OSError):
            f["c"] = 1


def test_compat_io_name():
    compat = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mdcraft_amd", "compat")
    if compat not in sys.path:
        sys.path.insert(0, compat)
    from mdcraft.io import hdf
    assert hdf.File is snapshot.File
