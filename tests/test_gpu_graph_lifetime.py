"""Lifetime of plans captured into a hipGraph (tmx_graph_begin / tmx_graph_end).

A graph's kernels point into the plans they were captured from (packed operands, task lists,
row scales).  The engine records every contraction and AddOp executed into a capture; the graph
holds them until tmx_graph_destroy.  So
  * destroying a contraction handle while a graph can replay it leaves its device memory in
    place (pool usage does not drop) and the replays stay bit-identical to the eager result;
  * the autotuner, which rebuilds a plan, refuses a contraction a live graph holds
    (TMX_ERR_STATE) and works again once the graph is destroyed;
  * destroying the graph releases the plan's memory.
"""
import gc
import weakref

import numpy as np
import pytest

from tamm_amd import configs

pytestmark = pytest.mark.gpu

SEED = 20261202
ERR_STATE = 5


@pytest.mark.parametrize("mode", [3, 2, 1])  # SPLIT_I8, SECTOR_F64, GROUPED_F64
def test_graph_holds_destroyed_plan(eng, mode):
    wl = configs.tiny_spin()
    ew = eng.EngineWorkload(wl, SEED)
    term = wl.terms[0]
    R = ew.tensors["R"]
    c = ew.contraction(term, mode)
    fill = eng.AddOp(R, [-1, -2, -3, -4], 0.0, ew.tensors["T"], [-1, -2, -3, -4], is_assign=True)

    def level():
        fill.execute()  # R = 0 * T
        c.execute()

    level()
    eng.synchronize()
    ref = R.download()
    g = eng.Graph(level)
    g.launch()
    eng.synchronize()
    assert np.array_equal(R.download(), ref)
    if mode in (2, 3):
        with pytest.raises(eng.EngineError) as ei:
            c.autotune()
        assert ei.value.code == ERR_STATE
        del ei  # its traceback holds the contraction
    used_before = eng.pool_info()[1]
    alive = weakref.ref(c)
    del c, fill  # the handles go; the graph still holds both plans
    gc.collect()
    assert alive() is None
    assert eng.pool_info()[1] == used_before
    junk = eng.DeviceBuffer(1 << 20)  # memory handed out after the destroy is not the plan's
    eng.check(eng.lib().tmx_fill(junk.ptr, (1 << 20) // 8, 1.0e300, None))
    for _ in range(3):
        R.fill(7.0)
        g.launch()
        eng.synchronize()
        assert np.array_equal(R.download(), ref)
    del junk
    del g
    assert eng.pool_info()[1] < used_before


def test_autotune_after_graph_destroyed(eng):
    wl = configs.tiny_spin()
    ew = eng.EngineWorkload(wl, SEED)
    c = ew.contraction(wl.terms[0], eng.EXEC_SPLIT_I8)
    c.execute()
    g = eng.Graph(lambda: c.execute())
    with pytest.raises(eng.EngineError) as ei:
        c.autotune()
    assert ei.value.code == ERR_STATE
    del ei, g
    c.autotune()  # no graph holds it any more
