"""BPNN under the wide model (the shape of the reference's Cu model
nn_Cu_d3_v0: 51-10-10-1, 8 G2 and 43 G4 over 7 eta, zeta up to 16, a
length unit of 0.066 nm, input ranges != 1) on the engine against the
oracle's golden fixtures: a 200-atom slab (periodic in x and y) and the
500-atom periodic crystal, for 1 to 16 lanes per atom.  FP64 within 1e-12
(forces, energies, the diagonal virial, descriptors; dE/dG within 1e-11);
mixed precision force RMS within 1e-5, energy and virial trace within
1e-5."""
import numpy as np
import pytest

from mdcraft_amd.atoms import mat, vec
from tests import golden_io
from tests.helpers import force_rms_rel, max_rel

pytestmark = pytest.mark.gpu

NAMES = ["bpnn_cu_wide_slab_200", "bpnn_cu_wide_500"]
LANES = [1, 2, 4, 8, 16]


def run(ctx, name, precision, lanes, tmp_path):
    from mdcraft_amd import engine as E
    meta, atoms, box, pot, outs = golden_io.load(name, tmp_path)
    print("wide %s precision=%d lanes=%d" % (name, precision, lanes), flush=True)
    s = E.System(ctx, [len(atoms)], [box], lanes=lanes)
    s.upload(atoms)
    p = E.MaiseBPNNPotential(ctx, pot["path"])
    s.build()
    s.compute(p, E.FORCE | E.ENERGY | E.VIRIAL, precision)
    G, dE = s.bpnn_inspect(p)
    return s.download(), G, dE, outs


@pytest.mark.parametrize("lanes", LANES)
@pytest.mark.parametrize("name", NAMES)
def test_wide_fp64(ctx, name, lanes, tmp_path):
    got, G, dE, outs = run(ctx, name, 0, lanes, tmp_path)
    assert G.shape[1] == 51
    assert max_rel(G[:64], outs["G"]) < 1e-12
    assert max_rel(dE[:64], outs["dEdG"]) < 1e-11
    assert max_rel(vec(got, "f"), outs["f"]) < 1e-12
    assert max_rel(got["Ep"], outs["Ep"]) < 1e-12
    assert max_rel(mat(got).reshape(len(got), 9), outs["V"]) < 1e-12


@pytest.mark.parametrize("lanes", LANES)
@pytest.mark.parametrize("name", NAMES)
def test_wide_mixed(ctx, name, lanes, tmp_path):
    got, G, dE, outs = run(ctx, name, 1, lanes, tmp_path)
    assert max_rel(G[:64], outs["G"]) < 1e-5
    assert force_rms_rel(vec(got, "f"), outs["f"]) < 1e-5
    e, er = got["Ep"].sum(), outs["Ep"].sum()
    assert abs(e - er) < 1e-5 * abs(er)
    Vr = outs["V"].reshape(-1, 3, 3)
    w, wr = np.trace(mat(got), axis1=1, axis2=2).sum(), np.trace(Vr, axis1=1, axis2=2).sum()
    assert abs(w - wr) < 1e-5 * np.abs(np.trace(Vr, axis1=1, axis2=2)).sum()
