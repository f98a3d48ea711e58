"""The 208-byte atom record shared with the reference's Python API.

Field order and sizes follow ``data::Atom`` (mdcraft/data/atom.h:21-38) and
the numpy dtype that the reference's pybind11 module exposes as
``Atoms.data`` (python/mdcraft/data/_data.cxx:29-79).  ``V`` is the Eigen
column-major 3x3 virial; its labels follow memory order, so the label
``xy`` holds element (y, x).  ``tag`` defaults to 65535; ``tag == 0`` marks
an atom that every stage skips.
"""
import numpy as np

VEC3 = np.dtype([("x", "<f8"), ("y", "<f8"), ("z", "<f8")])
MAT3 = np.dtype([(n, "<f8") for n in ("xx", "xy", "xz", "yx", "yy", "yz", "zx", "zy", "zz")])

ATOM_DTYPE = np.dtype(
    {
        "names": ["r", "v", "f", "V", "m", "rcut", "rns", "Ep", "Em", "nc", "bc", "kind", "uid", "tag"],
        "formats": [VEC3, VEC3, VEC3, MAT3, "<f8", "<f8", "<f8", "<f8", "<f8", "<f8", "<u2", "<u2", "<u4", "<u2"],
        "offsets": [0, 24, 48, 72, 144, 152, 160, 168, 176, 184, 192, 194, 196, 200],
        "itemsize": 208,
    }
)

# Physical constants in the reference's unit system (nm, ps, amu, kJ/mol).
KB = 0.008314462175  # kJ/mol/K, test/cxx/support/md_test_utils.h:37
EV = 1.602176634e-19 / (1e-3 / 6.02214076e23 * 1e6)  # kJ/mol per eV (EAM.cxx:8-13)


def make_atoms(n):
    """Zero-initialised records with the reference defaults (tag = 65535)."""
    a = np.zeros(n, dtype=ATOM_DTYPE)
    a["tag"] = 65535
    return a


def vec(a, field):
    """(n, 3) float64 copy of a vector field."""
    f = a[field]
    return np.stack([f["x"], f["y"], f["z"]], axis=1)


def set_vec(a, field, xyz):
    xyz = np.asarray(xyz, dtype=np.float64)
    a[field]["x"] = xyz[:, 0]
    a[field]["y"] = xyz[:, 1]
    a[field]["z"] = xyz[:, 2]


def mat(a):
    """(n, 3, 3) virial with [i, row, col]."""
    V = a["V"]
    out = np.empty((len(a), 3, 3))
    names = ("xx", "xy", "xz", "yx", "yy", "yz", "zx", "zy", "zz")
    for k, nm in enumerate(names):
        col, row = divmod(k, 3)
        out[:, row, col] = V[nm]
    return out


def trace_V(a):
    V = a["V"]
    return V["xx"] + V["yy"] + V["zz"]
