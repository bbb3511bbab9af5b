"""Assigning contractions, C = alpha*A*B (tmx_contraction_set_assign; TAMM's `C(..) = A*B`, which
TAMM runs as a zero fill of the addressed C blocks followed by `+=`, multop.hpp:255-269).

The engine fuses the fill: the scatter (or the block-pair epilogue) writes instead of adding,
and only the C blocks no block pair contributes to are zeroed.  Checked in every execution mode
against the zero fill followed by accumulation: C starts as NaN, so an element the assignment
misses shows; two executions in a row must not accumulate; the flag can be turned off again;
a level mixes an assigning term with accumulating ones that share its product; a captured graph
replays the assignment.
"""
import numpy as np
import pytest

from tamm_amd import configs

pytestmark = pytest.mark.gpu

SEED = 20261202
MODES = [0, 1, 2, 3, 4, 5]  # BLOCKPAIR, GROUPED, SECTOR, SPLIT, AUTO, AUTO_SPLIT
RINGS = ("ring", "ring_Pab", "ring_Pij", "ring_PabPij")


def _accumulated(eng, c, C):
    C.fill(0.0)
    c.execute()
    eng.synchronize()
    return C.download()


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("term", ["ladder", "ring", "hh"])
def test_assign_matches_fill_then_accumulate(eng, mode, term):
    wl = configs.tiny_spin()
    ew = eng.EngineWorkload(wl, SEED)
    t = next(x for x in wl.terms if x.name == term)
    C = ew.tensors[t.c[0]]
    c = ew.contraction(t, mode)
    ref = _accumulated(eng, c, C)
    c.set_assign(True)
    for _ in range(2):  # the second execution must overwrite, not add
        C.fill(float("nan"))
        c.execute()
        eng.synchronize()
        assert np.array_equal(C.download(), ref), (term, mode)
    c.set_assign(False)
    C.fill(1.0)
    c.execute()
    eng.synchronize()
    assert np.allclose(C.download(), ref + 1.0, rtol=0, atol=1e-12)


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_assign_zeroes_blocks_without_pairs(eng, mode):
    """C(i,j) = A(i,k) B(k,j) with A and B spin-blocked and C dense: the mixed-spin blocks of C
    get no block pair, so the assignment must leave them 0."""
    O = eng.Space(12, 3, spin=[(0, 6, 1), (6, 12, 2)])
    A, B = eng.Tensor([O, O], spin=(1, 1)), eng.Tensor([O, O], spin=(1, 1))
    Cd = eng.Tensor([O, O])
    rng = np.random.default_rng(SEED)
    A.upload(rng.standard_normal(A.size))
    B.upload(rng.standard_normal(B.size))
    c = eng.Contraction(Cd, [1, 2], 0.5, A, [1, 3], B, [3, 2], mode)
    ref = _accumulated(eng, c, Cd)
    assert np.count_nonzero(ref) < ref.size  # some blocks get no pair
    c.set_assign(True)
    Cd.fill(float("nan"))
    c.execute()
    eng.synchronize()
    assert np.array_equal(Cd.download(), ref)


@pytest.mark.parametrize("mode", [3, 2])  # SPLIT_I8, SECTOR_F64
def test_level_with_an_assigning_term(eng, mode):
    """R = ring + its P(ab), P(ij), P(ab)P(ij) images as one level whose first term assigns: the
    three others scatter from its product and accumulate."""
    wl = configs.config4_mini()
    ew = eng.EngineWorkload(wl, SEED)
    R = ew.tensors["R"]
    terms = [t for t in wl.terms if t.name in RINGS]
    R.fill(0.0)
    for t in terms:
        ew.contraction(t, mode).execute()
    eng.synchronize()
    ref = R.download()
    cs = [ew.contraction(t, mode) for t in terms]
    cs[0].set_assign(True)
    lv = eng.Level(cs)
    assert lv.products() == 1
    R.fill(float("nan"))
    lv.execute()
    eng.synchronize()
    assert np.array_equal(R.download(), ref)
    g = eng.Graph(lambda: lv.execute())
    with pytest.raises(eng.EngineError) as ei:
        cs[0].set_assign(False)  # held by the graph
    assert ei.value.code == 5
    del ei
    for _ in range(2):
        R.fill(float("nan"))
        g.launch()
        eng.synchronize()
        assert np.array_equal(R.download(), ref)


@pytest.mark.parametrize("mode", [2, 3])  # SECTOR_F64, SPLIT_I8
@pytest.mark.parametrize("parts", [2, 5])
def test_assign_sums_split_k_parts(eng, mode, parts):
    """cfg0's ladder (K range of 16 blocks of 400) cut into split-K parts: the assigning scatter
    sums every part of a C block before it writes the block, so C starts from NaN and still
    equals the zero fill followed by accumulation."""
    wl = configs.config0()
    ew = eng.EngineWorkload(wl, SEED)
    t = next(x for x in wl.terms if x.name == "ladder")
    C = ew.tensors[t.c[0]]
    eng.check(eng.lib().tmx_set_parallel_split(parts))
    try:
        c = ew.contraction(t, mode)
    finally:
        eng.check(eng.lib().tmx_set_parallel_split(0))
    rounds, _, _ = c.schedule_info()
    assert rounds > 1, rounds
    ref = _accumulated(eng, c, C)
    c.set_assign(True)
    C.fill(float("nan"))
    c.execute()
    eng.synchronize()
    assert np.array_equal(C.download(), ref), (mode, parts, rounds)
