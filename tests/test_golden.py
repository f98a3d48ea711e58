"""The oracle reproduces the committed golden fixtures (tests/golden), so
the fixtures are what the reference algorithm gives for their inputs and
any drift of the oracle shows here.  Fixtures marked slow (the 500-atom
wide BPNN crystal, about two minutes of oracle time) are re-checked when
MDC_SLOW_TESTS=1."""
import os

import numpy as np
import pytest

from mdcraft_amd.atoms import mat, vec
from tests import golden_io
from tests.golden.make_golden import evaluate
from tests.helpers import oracle_mod

REPRODUCE = 1e-12


@pytest.mark.parametrize("name", golden_io.names())
def test_oracle_reproduces_fixture(name, tmp_path):
    meta, atoms, box, pot, outs = golden_io.load(name, tmp_path)
    if meta.get("slow") and os.environ.get("MDC_SLOW_TESTS") != "1":
        pytest.skip("slow fixture: set MDC_SLOW_TESTS=1 to re-run its oracle evaluation")
    res, ph, extra = evaluate(oracle_mod(), atoms, box, pot, meta["nve_steps"], meta["dt"])
    assert ph == meta["pair_sha256"]
    got = {"f": vec(res, "f"), "Ep": res["Ep"], "Em": res["Em"], "nc": res["nc"], "V": mat(res).reshape(len(res), 9)}
    if meta["nve_steps"]:
        got["r"], got["v"] = vec(res, "r"), vec(res, "v")
    got.update({k: v[:64] for k, v in extra.items()})
    for k, ref in outs.items():
        scale = max(np.max(np.abs(ref)), 1e-300)
        assert np.max(np.abs(got[k] - ref)) <= REPRODUCE * scale, (name, k)


def test_fixtures_present():
    assert {"lj_argon_256", "eam_cu_256", "eam_cu_256_nve20", "eam_cu_108_short_box", "nial_multi_250",
            "bpnn_cu_256", "bpnn_cu_500", "bpnn_cu_wide_slab_200", "bpnn_cu_wide_500"} <= set(golden_io.names())
