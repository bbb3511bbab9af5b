"""Scheduler levels with shared products (tmx_level_*).

A CC ring term and its P(ab), P(ij) and P(ab)P(ij) images pack W and T alike and have the same
sector geometry, so a level computes their product once and scatters it four times into R.
Checked on the cfg4 step's forward residual at O=12, V=24 (spin + C2v, drawn tiles) and on a
tiny spin workload:
  * the level finds the shared products (the four ring terms count once);
  * results are bit-identical to executing the terms one by one, in FP64 sector, split and
    AUTO modes, and the whole forward + adjoint program agrees too;
  * after the shared operand changes in place the next level execution follows it.
"""
import numpy as np
import pytest

from tamm_amd import configs

pytestmark = pytest.mark.gpu

SEED = 20261202
RINGS = ("ring", "ring_Pab", "ring_Pij", "ring_PabPij")


@pytest.mark.parametrize("mode", [3, 2, 5])  # SPLIT_I8, SECTOR_F64, AUTO_SPLIT
def test_ring_images_share_one_product(eng, mode):
    wl = configs.config4_mini()
    ew = eng.EngineWorkload(wl, SEED)
    R = ew.tensors["R"]
    terms = [t for t in wl.terms if t.name in RINGS]
    assert len(terms) == 4
    lv = ew.level(terms, mode)
    resolved = [ew.plan(t, mode).mode for t in terms]
    if all(m in (2, 3) for m in resolved):
        assert lv.products() == 1
    R.fill(0.0)
    for t in terms:
        ew.plan(t, mode).execute()
    ref = R.download()
    for _ in range(2):
        R.fill(0.0)
        lv.execute()
        assert np.array_equal(R.download(), ref)
    ew.tensors["T"].scale(0.5)
    R.fill(0.0)
    for t in terms:
        ew.plan(t, mode).execute()
    ref2 = R.download()
    R.fill(0.0)
    lv.execute()
    assert np.array_equal(R.download(), ref2)


@pytest.mark.parametrize("mode", [5, 4])  # AUTO_SPLIT, AUTO
def test_cc_step_program_with_levels(eng, mode):
    wl = configs.config4_mini()
    ew = eng.EngineWorkload(wl, SEED)
    prog = wl.extra["forward"] + wl.extra["adjoint"]
    outs = ["R", "dT", "dV", "dW", "dX", "dU"]
    ew.execute_program(prog, mode)
    ref = {n: ew.tensors[n].download() for n in outs}
    for n in outs:
        ew.tensors[n].fill(0.0)
    ew.execute_program(prog, mode, levels=True)
    for n in outs:
        assert np.array_equal(ew.tensors[n].download(), ref[n]), n


@pytest.mark.parametrize("mode", [3, 2])  # SPLIT_I8, SECTOR_F64
def test_level_graph_replay(eng, mode):
    """A level captured into a hipGraph replays its shared product and four scatters: three
    replays, with eager work between them, equal the eager execution bit for bit; the graph
    keeps working after the level's handle and the terms' plans are dropped."""
    wl = configs.config4_mini()
    ew = eng.EngineWorkload(wl, SEED)
    R = ew.tensors["R"]
    terms = [t for t in wl.terms if t.name in RINGS]
    lv = ew.level(terms, mode)
    R.fill(0.0)
    lv.execute()
    eng.synchronize()
    ref = R.download()
    g = eng.Graph(lambda: (R.fill(0.0), lv.execute()))
    ew.clear_plans()  # the handles go; the graph holds what it captured
    del lv
    for _ in range(3):
        R.fill(float("nan"))
        g.launch()
        eng.synchronize()
        assert np.array_equal(R.download(), ref)
        ew.tensors["T"].scale(1.0)  # eager work between replays
        eng.synchronize()
