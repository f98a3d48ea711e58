"""Engine (HIP, through the C ABI) against the oracle for the N-element
alloy EAM: B2 NiAl cells, and NiAl cells with copper substituted on
aluminium sites read from a three-element setfl file, alone and batched.

Bars: FP64 within 1e-12 relative (max |difference| over the max magnitude);
mixed precision force RMS within 1e-5 and energy sums within 1e-6;
trajectories of 30 steps within 1e-9 nm on positions and 1e-8 relative on
velocities."""
import numpy as np
import pytest

from mdcraft_amd import workloads as W
from mdcraft_amd.atoms import vec
from tests.helpers import engine_potential, oracle_compute, oracle_mod, oracle_potential
from tests.test_gpu_multi import compare

pytestmark = pytest.mark.gpu


def E():
    from mdcraft_amd import engine
    return engine


def nial_pot():
    return dict(kind="alloy", **W.alloy_tables(["Ni", "Al"]))


def with_copper(at, seed, frac=0.15):
    """Aluminium sites (kind 1) turned to copper (kind 2) at random."""
    a = at.copy()
    rng = np.random.default_rng(seed)
    al = np.flatnonzero(a["kind"] == 1)
    cu = al[rng.random(len(al)) < frac]
    a["kind"][cu] = 2
    a["m"][cu] = W.ALLOY_ELEMENTS["Cu"]["mass"]
    return a


def run_batch(ctx, cs, pot, precision, lanes=None):
    eng = E()
    s = eng.System(ctx, [len(a) for a, _ in cs], [b for _, b in cs], lanes=lanes)
    s.upload(np.concatenate([a for a, _ in cs]))
    p = engine_potential(eng, ctx, pot)
    s.build()
    s.compute(p, 7, precision)
    return s.download(), s, p


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("target", [2000, 32])
def test_alloy_single_cell(ctx, precision, target):
    """A bulk cell (box > 2 rns) and a 32-atom cell (one image per pair)."""
    at, box = W.nial_cell(7, target)
    pot = nial_pot()
    ref, _ = oracle_compute(at, box, pot)
    for lanes in (1, 8):
        got, _, _ = run_batch(ctx, [(at, box)], pot, precision, lanes)
        compare(got, ref, precision, f"alloy{target}/lanes{lanes}")


@pytest.fixture(scope="module")
def ternary_file(tmp_path_factory):
    path = str(tmp_path_factory.mktemp("alloy") / "NiAlCu.eam.alloy")
    W.write_setfl(path, ["Ni", "Al", "Cu"])
    return path


@pytest.mark.parametrize("precision", [0, 1])
def test_three_element_file_batch(ctx, precision, ternary_file):
    cs = []
    for i in range(8):
        at, box = W.nial_cell(i, target=[32, 128, 250, 500, 1000, 54, 432, 686][i])
        cs.append((with_copper(at, i), box))
    assert all((a["kind"] == 2).any() for a, _ in cs[1:])
    pot = {"kind": "alloy", "path": ternary_file}
    got, s, _ = run_batch(ctx, cs, pot, precision)
    off = 0
    for q, (at, box) in enumerate(cs):
        ref, _ = oracle_compute(at, box, pot)
        compare(got[off:off + len(at)], ref, precision, f"cell{q}")
        off += len(at)


def test_alloy_nve_fp64(ctx, ternary_file):
    """A batch stepped together, with skin-triggered rebuilds, follows each
    system's own oracle run (boxes above 2 (rns + skin) on every axis)."""
    R = oracle_mod()
    cs = []
    for i, target in enumerate([250, 500, 1000]):
        at, box = W.nial_cell(20 + i, target)
        cs.append((with_copper(at, 20 + i), box))
        assert np.min(box[1] - box[0]) > 2 * (W.EAM_RC + 0.6) * W.ANG
    pot = {"kind": "alloy", "path": ternary_file}
    dt, steps = 1 * W.FS, 30
    eng = E()
    s = eng.System(ctx, [len(a) for a, _ in cs], [b for _, b in cs])
    s.upload(np.concatenate([a for a, _ in cs]))
    p = engine_potential(eng, ctx, pot)
    s.build()
    s.compute(p, 1, 0)
    s.step(p, dt, steps, 0)
    got = s.download()
    off = 0
    for q, (at, box) in enumerate(cs):
        a = at.copy()
        dom = R.Domain(*box)
        pool = R.Pool(4)
        op = oracle_potential(R, pot)
        nl = R.VerletList(pool)
        nl.update(a, dom)
        sol = R.Solver(op, pool, True)
        sol.prepare(a, dom, nl)
        sol.stages(a, dom, nl)
        sol.run_nve(a, dom, nl, dt, steps, initial_forces=False)
        L = box[1] - box[0]
        d = vec(got[off:off + len(a)], "r") - vec(a, "r")
        d -= np.round(d / L) * L
        assert np.max(np.abs(d)) < 1e-9, q
        g = vec(got[off:off + len(a)], "v")
        assert np.max(np.abs(g - vec(a, "v"))) / np.max(np.abs(vec(a, "v"))) < 1e-8, q
        off += len(a)


def test_one_element_file_is_the_eam_object(ctx, tmp_path):
    path = str(tmp_path / "Cu.eam.alloy")
    W.write_setfl(path, ["Cu"])
    w = W.eam_cu(reps=5, rattle=0.05)
    a, _, _ = run_batch(ctx, [(w["atoms"], w["box"])], {"kind": "alloy", "path": path}, 0)
    b, _, _ = run_batch(ctx, [(w["atoms"], w["box"])], {"kind": "eam", "path": path}, 0)
    for k in ("f", "Ep", "Em", "nc", "V"):
        assert np.array_equal(a[k], b[k]), k


def test_errors(ctx, tmp_path):
    eng = E()
    path = str(tmp_path / "NiAl.eam.alloy")
    W.write_setfl(path, ["Ni", "Al"], nr=500, nrho=500)
    p = eng.Alloy(ctx, path=path)
    at, box = W.nial_cell(1, 32)
    at["kind"][0] = 2
    s = eng.System(ctx, [len(at)], [box])
    s.upload(at)
    s.build()
    with pytest.raises(RuntimeError, match="outside the alloy's elements"):
        s.compute(p, 7, 0)
    with pytest.raises(RuntimeError, match="open"):
        eng.Alloy(ctx, path=str(tmp_path / "none.eam.alloy"))
    t = W.alloy_tables(["Ni", "Al"], nr=100, nrho=100)
    with pytest.raises(ValueError):
        eng.Alloy(ctx, rcut=t["rc"], F=t["F"], rho=t["rho"], rphi=t["rphi"][:2], dr=t["dr"], drho=t["drho"])
    five = W.alloy_tables(["Ni", "Al", "Cu", "Ni", "Al"], nr=100, nrho=100)
    with pytest.raises(RuntimeError, match="1..4"):
        eng.Alloy(ctx, rcut=five["rc"], F=five["F"], rho=five["rho"], rphi=five["rphi"], dr=five["dr"],
                  drho=five["drho"])
