"""What holds a contraction against changes: a live graph holds it against every change
(tmx_contraction_set_constant, _set_assign, _autotune); a level only against the autotuner,
which rebuilds the plan its other terms scatter from.  Changing the constant or assign flags
of a level's term is allowed and takes effect in the level's next execution."""
import numpy as np
import pytest

from tamm_amd import configs

pytestmark = pytest.mark.gpu

SEED = 20261202
RINGS = ("ring", "ring_Pab", "ring_Pij", "ring_PabPij")


def test_level_membership_allows_flags_not_autotune(eng):
    wl = configs.config4_mini()
    ew = eng.EngineWorkload(wl, SEED)
    R = ew.tensors["R"]
    terms = [t for t in wl.terms if t.name in RINGS]
    cs = [ew.contraction(t, eng.EXEC_SPLIT_I8) for t in terms]
    R.fill(0.0)
    for c in cs:
        c.execute()
    eng.synchronize()
    ref = R.download()
    lv = eng.Level(cs)
    with pytest.raises(eng.EngineError) as ei:
        cs[0].autotune()  # the level's other terms scatter from this plan's product
    assert ei.value.code == 5
    del ei
    # constant operands and assignment may change under a level
    for c in cs:
        c.set_constant(True, False)
    cs[0].set_assign(True)
    for _ in range(2):
        R.fill(float("nan"))
        lv.execute()
        eng.synchronize()
        assert np.array_equal(R.download(), ref)
    # a graph holds them against every change; once it is gone the flags change again
    g = eng.Graph(lambda: lv.execute())
    for fn in (lambda: cs[1].set_constant(False, False), lambda: cs[0].set_assign(False)):
        with pytest.raises(eng.EngineError) as ei:
            fn()
        assert ei.value.code == 5
        del ei
    R.fill(float("nan"))
    g.launch()
    eng.synchronize()
    assert np.array_equal(R.download(), ref)
    del g
    cs[0].set_assign(False)
    for c in cs:
        c.set_constant(False, False)
    R.fill(0.0)
    lv.execute()
    eng.synchronize()
    assert np.array_equal(R.download(), ref)
    del lv
    cs[0].autotune()  # no level or graph holds it any more
