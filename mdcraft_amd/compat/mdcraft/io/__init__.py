"""Snapshot files: ``hdf.File`` with the dataset interface the reference's
scripts use (problems/ar-mw/continue.py:9, :258-280), stored as npz."""
import types as _types

from mdcraft_amd.snapshot import File

hdf = _types.SimpleNamespace(File=File)
