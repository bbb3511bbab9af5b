"""Tile-shape autotuner (tmx_contraction_autotune) on the split-precision sector path.

Autotuning executes once and re-tiles the sector GEMMs; the int32 residue sums are exact in
any tile order, so results before and after tuning must agree bit for bit.  The choices are
written to the file named by TAMM_MI355X_TUNE_FILE and reused by later plans.
"""
import os

import numpy as np
import pytest

from tamm_amd import configs
from tamm_amd import engine as E

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("wlf", [lambda: configs.tiny_spin(), lambda: configs.scaled_spin(16, 64, 8, 16)],
                         ids=["tiny", "o16v64"])
def test_autotune_keeps_results(eng, wlf, tmp_path):
    wl = wlf()
    ew = E.EngineWorkload(wl, configs.DEFAULT_SEED)
    R = ew.tensors["R"]
    path = str(tmp_path / "tune.txt")
    os.environ["TAMM_MI355X_TUNE_FILE"] = path
    try:
        for term in wl.terms:
            c = ew.contraction(term, E.EXEC_SPLIT_I8)
            R.fill(0.0)
            c.execute()
            ref = R.download()
            R.fill(0.0)
            c.autotune()  # executes once
            assert np.array_equal(R.download(), ref), term.name
            R.fill(0.0)
            c.execute()
            assert np.array_equal(R.download(), ref), term.name
            c2 = ew.contraction(term, E.EXEC_SPLIT_I8)  # picks the recorded shapes
            R.fill(0.0)
            c2.execute()
            assert np.array_equal(R.download(), ref), term.name
        lines = open(path).read().split("\n")
        assert any(l.startswith("i8 ") for l in lines)
    finally:
        del os.environ["TAMM_MI355X_TUNE_FILE"]
