"""Snapshots and restarts in the layout of the reference's scripts.

The reference's production scripts checkpoint with h5py, one file per
snapshot holding the 208-byte ``atoms.data`` records and small arrays
(test/python/solver/test-EAM.py:193-197, problems/ar-mw/continue.py:61-110,
:258-280):

    f = hdf.File("atoms/run%06d.h5" % step, "w")
    f["atoms"] = atoms.data
    f["domain"] = np.array([xmin, xmax, ymin, ymax, zmin, zmax])
    f.close()

``File`` keeps that interface (item assignment, ``f[name][:]``, ``keys()``,
``close()``, use as a context manager) and stores the datasets in one
numpy ``.npz`` archive, read back without pickle.  Records keep the
reference's dtype (field names, offsets, 208-byte stride), so a snapshot
written here loads into ``Atoms.data`` unchanged; when h5py is installed
the scripts can keep using it, with the same dataset names.

``save_system`` / ``load_system`` checkpoint a resident engine System (any
number of boxes) and restore it on a device.
"""
import os

import numpy as np

from .atoms import ATOM_DTYPE


class File:
    """h5py.File-like container of named arrays in a ``.npz`` archive.
    Modes: "r" (read), "w" (create or truncate), "a" (read and add)."""

    def __init__(self, path, mode="r"):
        if mode not in ("r", "w", "a"):
            raise ValueError("mode must be 'r', 'w' or 'a'")
        self.path = os.fspath(path)
        self.mode = mode
        self._data = {}
        self._open = True
        if mode in ("r", "a") and (mode == "r" or os.path.exists(self._file())):
            with np.load(self._file(), allow_pickle=False) as z:
                self._data = {k: z[k] for k in z.files}

    def _file(self):
        # numpy adds .npz to names without it; keep the caller's name
        return self.path if self.path.endswith(".npz") else self.path + ".npz"

    def __setitem__(self, name, value):
        if self.mode == "r":
            raise OSError("file opened read-only")
        self._data[name] = np.array(value, copy=True)

    def __getitem__(self, name):
        return self._data[name]

    def __contains__(self, name):
        return name in self._data

    def keys(self):
        return list(self._data)

    def close(self):
        if self._open and self.mode != "r":
            d = os.path.dirname(self._file())
            if d:
                os.makedirs(d, exist_ok=True)
            np.savez(self._file(), **self._data)
        self._open = False

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False


def save_system(path, system, boxes, atoms=None, **arrays):
    """Write every record of a resident System (downloaded now unless
    ``atoms`` is given), the atom count and box of each system, and any
    extra named arrays."""
    if atoms is None:
        atoms = system.download()
    lo = np.array([b[0] for b in boxes], dtype=np.float64).reshape(-1, 3)
    hi = np.array([b[1] for b in boxes], dtype=np.float64).reshape(-1, 3)
    per = np.array([[bool(x) for x in b[2]] for b in boxes], dtype=np.uint8).reshape(-1, 3)
    with File(path, "w") as f:
        f["atoms"] = np.ascontiguousarray(atoms, dtype=ATOM_DTYPE)
        f["natoms"] = np.asarray(system.natoms, dtype=np.int64)
        f["box_lo"], f["box_hi"], f["periodic"] = lo, hi, per
        for k, v in arrays.items():
            f[k] = v


def load_system(path, engine, ctx, periodic_boundary=True, lanes=None):
    """A new System holding the snapshot's boxes and records (list not yet
    built), the records themselves, and the snapshot's other arrays."""
    with File(path, "r") as f:
        atoms = np.ascontiguousarray(f["atoms"], dtype=ATOM_DTYPE)
        natoms = f["natoms"]
        boxes = [(f["box_lo"][s], f["box_hi"][s], tuple(bool(x) for x in f["periodic"][s])) for s in range(len(natoms))]
        extra = {k: f[k] for k in f.keys() if k not in ("atoms", "natoms", "box_lo", "box_hi", "periodic")}
    s = engine.System(ctx, natoms, boxes, periodic_boundary, lanes=lanes)
    s.upload(atoms)
    return s, atoms, boxes, extra
