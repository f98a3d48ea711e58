"""The oracle's solver::Multi (multi.cxx:11-151) over EAM objects, pinned by
what the mathematics fixes: the reduction to single-element EAM when every
pair of kinds shares one object, and closed forms of a dimer and a trimer
whose objects differ, so that routing a pair to the wrong object, taking
F'(nc) from the wrong table or setting Em from the wrong sub-list fails."""
import numpy as np
import pytest

from mdcraft_amd import workloads as W
from mdcraft_amd.atoms import EV, make_atoms, mat, set_vec, vec
from tests.helpers import max_rel, oracle_compute

# spline tables against the analytic functions they sample
CLOSED_FORM_RTOL = 2e-6

SCALE = {(0, 0): 1.1, (1, 1): 0.9, (0, 1): 1.0, (1, 0): 1.0}  # rho and F scale of each object


def rho_a(r_nm, s):
    r = np.asarray(r_nm) / W.ANG
    return s * W.density(r) * W.taper(r)


def phi_a(r_nm):
    r = np.asarray(r_nm) / W.ANG
    return W.morse(r) * W.taper(r) * EV


def F_a(x, s):
    return -s * W.EAM_F_A * np.sqrt(x) * EV


def dF_a(x, s):
    return -s * W.EAM_F_A * 0.5 / np.sqrt(x) * EV


def deriv(f, r, h=1e-6):
    return (f(r + h) - f(r - h)) / (2 * h)


def cluster(pos_nm, kinds):
    n = len(kinds)
    at = make_atoms(n)
    set_vec(at, "r", np.asarray(pos_nm, float))
    at["kind"] = kinds
    at["m"] = np.where(np.asarray(kinds) == 0, 58.6934, 26.9815385)
    at["rns"] = (W.EAM_RC + 0.3) * W.ANG
    box = (np.full(3, -2.0), np.full(3, 2.0), (False, False, False))
    return at, box


def test_multi_with_one_object_is_single_eam():
    w = W.eam_cu(reps=4, rattle=0.05)
    at = w["atoms"].copy()
    at["kind"] = np.random.default_rng(3).integers(0, 2, len(at))
    single, _ = oracle_compute(at, w["box"], w["pot"])
    pot = {"kind": "eam_multi", "nkinds": 2, "pairs": [(0, 0, w["pot"]), (1, 1, w["pot"]), (0, 1, w["pot"])]}
    multi, _ = oracle_compute(at, w["box"], pot)
    # the per-kind sub-lists only reorder the sums
    for k in ("nc", "Em", "Ep"):
        assert max_rel(multi[k], single[k]) < 1e-14, k
    assert max_rel(vec(multi, "f"), vec(single, "f")) < 1e-13
    assert max_rel(mat(multi), mat(single)) < 1e-13


def test_multi_dimer_closed_form():
    r = 0.25
    at, box = cluster([[0, 0, 0], [r, 0, 0]], [0, 1])
    got, _ = oracle_compute(at, box, W.nial_pots(), periodic_boundary=False)
    rho = rho_a(r, 1.0)  # both densities come from the Ni-Al object
    assert got["nc"] == pytest.approx([rho, rho], rel=CLOSED_FORM_RTOL)
    FF = 2 * dF_a(rho, 1.0) * deriv(lambda x: rho_a(x, 1.0), r) + deriv(phi_a, r)
    f = vec(got, "f")
    assert f[0, 0] == pytest.approx(FF, rel=CLOSED_FORM_RTOL)
    assert f[1, 0] == pytest.approx(-FF, rel=CLOSED_FORM_RTOL)
    # Em: the last non-empty kind sub-list is kind 1 for both atoms, so Ni
    # takes F of Ni-Al and Al takes F of Al-Al
    assert got["Em"][0] == pytest.approx(F_a(rho, 1.0), rel=CLOSED_FORM_RTOL)
    assert got["Em"][1] == pytest.approx(F_a(rho, 0.9), rel=CLOSED_FORM_RTOL)
    assert got["Ep"] == pytest.approx([0.5 * phi_a(r)] * 2, rel=CLOSED_FORM_RTOL)


def test_multi_trimer_closed_form():
    pos = np.array([[0.0, 0.0, 0.0], [0.25, 0.0, 0.0], [0.0, 0.26, 0.0]])
    kinds = [0, 0, 1]
    at, box = cluster(pos, kinds)
    got, _ = oracle_compute(at, box, W.nial_pots(), periodic_boundary=False)
    d = lambda a, b: float(np.linalg.norm(pos[a] - pos[b]))
    nc = np.zeros(3)
    for a in range(3):
        for b in range(3):
            if a != b:
                nc[a] += rho_a(d(a, b), SCALE[(kinds[a], kinds[b])])
    assert got["nc"] == pytest.approx(nc, rel=CLOSED_FORM_RTOL)
    f = np.zeros((3, 3))
    for a in range(3):
        for b in range(3):
            if a == b:
                continue
            s = SCALE[(kinds[a], kinds[b])]
            r = d(a, b)
            FF = (dF_a(nc[a], s) + dF_a(nc[b], s)) * deriv(lambda x: rho_a(x, s), r) + deriv(phi_a, r)
            f[a] += FF * (pos[b] - pos[a]) / r
    assert np.allclose(vec(got, "f"), f, rtol=CLOSED_FORM_RTOL, atol=CLOSED_FORM_RTOL * np.abs(f).max())
    # Em: atoms 0 and 1 (Ni) end on sub-list 1 (object Ni-Al), atom 2 (Al)
    # on sub-list 1 (object Al-Al)
    assert got["Em"] == pytest.approx([F_a(nc[0], 1.0), F_a(nc[1], 1.0), F_a(nc[2], 0.9)], rel=CLOSED_FORM_RTOL)


def test_multi_rejects_unknown_kind():
    at, box = cluster([[0, 0, 0], [0.25, 0, 0]], [0, 2])
    with pytest.raises(RuntimeError, match="kind"):
        oracle_compute(at, box, W.nial_pots(), periodic_boundary=False)


def pair_member_pots(scale=0.8):
    """The reference's NiAl set-up (problems/NiAl/test-NiAl.py:37-39,
    :120-122): EAM objects for like kinds and a tabulated Pair for the
    cross kinds; here the Pair samples scale * r*phi of the analytic EAM."""
    t = W.eam_cu_tables()
    eam = dict(kind="eam", **t)
    pair = {"kind": "pair", "rc": t["rc"], "rphi": scale * t["rphi"], "dr": t["dr"]}
    return {"kind": "eam_multi", "nkinds": 2, "pairs": [(0, 0, eam), (1, 1, eam), (0, 1, pair)]}


def test_multi_pair_member_trimer_closed_form():
    """Cross-kind pairs under a Pair member give no density and only their
    pair term; like-kind pairs keep the EAM object's terms; Em comes from
    the like-kind EAM object (a Pair's virial leaves Em)."""
    pos = np.array([[0.0, 0.0, 0.0], [0.25, 0.0, 0.0], [0.0, 0.26, 0.0]])
    kinds = [0, 0, 1]
    at, box = cluster(pos, kinds)
    got, _ = oracle_compute(at, box, pair_member_pots(), periodic_boundary=False)
    d = lambda a, b: float(np.linalg.norm(pos[a] - pos[b]))
    nc = np.array([rho_a(d(0, 1), 1.0), rho_a(d(0, 1), 1.0), 0.0])
    assert got["nc"] == pytest.approx(nc, rel=CLOSED_FORM_RTOL, abs=1e-300)
    f = np.zeros((3, 3))
    for a in range(3):
        for b in range(3):
            if a == b:
                continue
            r = d(a, b)
            if kinds[a] == kinds[b]:
                FF = (dF_a(nc[a], 1.0) + dF_a(nc[b], 1.0)) * deriv(lambda x: rho_a(x, 1.0), r) + deriv(phi_a, r)
            else:
                FF = 0.8 * deriv(phi_a, r)
            f[a] += FF * (pos[b] - pos[a]) / r
    assert np.allclose(vec(got, "f"), f, rtol=CLOSED_FORM_RTOL, atol=CLOSED_FORM_RTOL * np.abs(f).max())
    assert got["Em"][:2] == pytest.approx([F_a(nc[0], 1.0), F_a(nc[1], 1.0)], rel=CLOSED_FORM_RTOL)
    assert got["Em"][2] == 0.0  # F(0) of the like-kind object of its self sub-list
    Ep = [0.5 * (phi_a(d(0, 1)) + 0.8 * phi_a(d(0, 2))), 0.5 * (phi_a(d(0, 1)) + 0.8 * phi_a(d(1, 2))),
          0.5 * 0.8 * (phi_a(d(0, 2)) + phi_a(d(1, 2)))]
    assert got["Ep"] == pytest.approx(Ep, rel=CLOSED_FORM_RTOL)
