"""Row-scale reuse of split packing (tmx_set_scale_reuse) on the GPU.

Split packing scales every class-matrix row by its maximum |x|.  With reuse on, a plan that
packs the same operand address again scales with the previous execution's maxima, recomputes
them in the same pass, and packs again (gated on the device) when any of them changed.  The
results must be bit-identical to packing with freshly computed maxima (reuse off) whatever the
operands do between executions: stay the same, change every maximum (x2), change one row's
maximum (one element raised), or change values without changing any maximum (negation); and
inside a captured graph replayed over changing operands.  cfg0 (dense O=20/t10, V=80/t20)
covers both packing kernels: the ladder's V(a,b,c,d) has whole contiguous rows (k_pack_kf),
the ring's W(k,b,c,j) and T(a,c,i,k) are tiled (k_pack); an O=20 irrep-split workload adds
blocks of 2-3 rows whose last mode is a row mode (narrow pack tiles, k_rowmax_t).
"""
import numpy as np
import pytest

from tamm_amd import configs

pytestmark = pytest.mark.gpu

SEED = 20261202


def _mutations(a0: np.ndarray, b0: np.ndarray):
    """(name, A data, B data) per execution, including runs of unchanged operands long enough
    for reuse to resume after the pause that follows a change."""
    rng = np.random.default_rng(5)
    a_one = a0.copy()
    k = int(rng.integers(a0.size))
    a_one[k] = 4.0 * np.abs(a0).max()  # raises the maximum of exactly one row
    b_one = b0.copy()
    k = int(rng.integers(b0.size))
    b_one[k] = -3.0 * np.abs(b0).max()
    steps = [("base", a0, b0), ("same", a0, b0), ("same", a0, b0)]
    steps += [("A x2", 2.0 * a0, b0)] + [("A x2 kept", 2.0 * a0, b0)] * 10
    steps += [("one A row", a_one, b0), ("one B row", a_one, b_one)] + [("kept", a_one, b_one)] * 10
    steps += [("negated", -a_one, -b_one), ("negated", -a_one, -b_one)]
    steps += [("base", a0, b0)]
    return steps


def _run(eng, wl, ow, term, mode, reuse, steps):
    eng.check(eng.lib().tmx_set_scale_reuse(1 if reuse else 0))
    try:
        ew = eng.EngineWorkload(wl, SEED, allocate=False)
        for name in wl.tensors:
            ew.tensors[name].upload(ow.data[name])
        an, bn, cn = term.a[0], term.b[0], term.c[0]
        c = ew.contraction(term, mode)
        out = []
        for _, a, b in steps:
            ew.tensors[an].upload(a)
            ew.tensors[bn].upload(b)
            ew.tensors[cn].fill(0.0)
            c.execute()
            out.append(ew.tensors[cn].download())
        del c
        return out
    finally:
        eng.check(eng.lib().tmx_set_scale_reuse(1))


@pytest.fixture(scope="module")
def cfg0(orc):
    wl = configs.config0()
    return wl, orc.OracleWorkload(wl, SEED)


@pytest.mark.parametrize("term_name", ["ladder", "ring"])
def test_reuse_bit_identical(eng, cfg0, term_name):
    wl, ow = cfg0
    term = next(t for t in wl.terms if t.name == term_name)
    steps = _mutations(ow.data[term.a[0]], ow.data[term.b[0]])
    off = _run(eng, wl, ow, term, eng.EXEC_SPLIT_I8, False, steps)
    on = _run(eng, wl, ow, term, eng.EXEC_SPLIT_I8, True, steps)
    for i, ((name, _, _), x, y) in enumerate(zip(steps, off, on)):
        assert np.array_equal(x, y), (term_name, i, name, np.abs(x - y).max())
    # scaling A by 2 scales its row maxima and scales by exactly 2, leaving the residues as
    # they were: the product doubles exactly (C starts at zero), on the execution that sees
    # the change and on the last one of the unchanged run after it
    assert np.array_equal(on[3], 2.0 * on[0]) and np.array_equal(on[13], 2.0 * on[0])
    # negating both operands keeps every maximum and negates every residue: same product
    assert steps[25][0] == "kept" and steps[26][0] == "negated"
    assert np.array_equal(on[26], on[25]) and np.array_equal(on[27], on[25])
    assert np.array_equal(on[-1], on[0])


def test_reuse_bit_identical_few_row_blocks(eng, orc):
    """Blocks with few rows whose last mode is a row mode (L(a,b,i,j) with rows j of 2-3 on an
    O=20 irrep-split space): without reuse their row maxima come from k_rowmax_t's tiles, with
    reuse from the pack's own reduction; both orders must give the same bits."""
    wl = configs._cc_step("tall", 20, 100, 8, 24, SEED, "O=20, V=100: few-row blocks")
    ow = orc.OracleWorkload(wl, SEED)
    term = configs.TermDef("dFo<-LT", ("dFo", "kj"), -1.0, ("L", "abij"), ("T", "abik"))
    a0, b0 = ow.data["L"], ow.data["T"]
    steps = [("base", a0, b0), ("same", a0, b0), ("A x2", 2.0 * a0, b0)]
    steps += [("kept", 2.0 * a0, b0)] * 10 + [("B x3", 2.0 * a0, 3.0 * b0)]
    off = _run(eng, wl, ow, term, eng.EXEC_SPLIT_I8, False, steps)
    on = _run(eng, wl, ow, term, eng.EXEC_SPLIT_I8, True, steps)
    for i, ((name, _, _), x, y) in enumerate(zip(steps, off, on)):
        assert np.array_equal(x, y), (i, name, np.abs(x - y).max())
    assert np.array_equal(on[2], 2.0 * on[0]) and np.array_equal(on[12], 2.0 * on[0])


def test_reuse_matches_oracle(eng, cfg0):
    """Reused scales on unchanged operands: the split result at 1e-10 of the oracle's FP64."""
    wl, ow = cfg0
    for term in wl.terms:
        base = np.zeros(ow.tensors[term.c[0]].size)
        ref = base.copy()
        ow.run_term(term, out=ref)
        steps = [("base", ow.data[term.a[0]], ow.data[term.b[0]])] * 3
        on = _run(eng, wl, ow, term, eng.EXEC_SPLIT_I8, True, steps)
        for i, got in enumerate(on):
            err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
            assert err <= 1e-10, (term.name, i, err)


def test_reuse_in_graph_replays(eng, cfg0):
    """A graph captured after one execution replays the reusing pack with its gated second
    pass; operands change between replays and every replay matches eager reuse-off results."""
    wl, ow = cfg0
    term = next(t for t in wl.terms if t.name == "ladder")
    a0, b0 = ow.data[term.a[0]], ow.data[term.b[0]]
    seq = [("base", a0, b0), ("A x2", 2.0 * a0, b0), ("A x2", 2.0 * a0, b0),
           ("B x-0.5", 2.0 * a0, -0.5 * b0), ("base", a0, b0)]
    ref = _run(eng, wl, ow, term, eng.EXEC_SPLIT_I8, False, seq)
    ew = eng.EngineWorkload(wl, SEED, allocate=False)
    for name in wl.tensors:
        ew.tensors[name].upload(ow.data[name])
    an, bn, cn = term.a[0], term.b[0], term.c[0]
    c = ew.contraction(term, eng.EXEC_SPLIT_I8)
    ew.tensors[cn].fill(0.0)
    c.execute()  # keeps row maxima: the captured pack reuses them
    eng.synchronize()
    g = eng.Graph(lambda: c.execute())
    for i, (name, a, b) in enumerate(seq):
        ew.tensors[an].upload(a)
        ew.tensors[bn].upload(b)
        ew.tensors[cn].fill(0.0)
        g.launch()
        eng.synchronize()
        got = ew.tensors[cn].download()
        assert np.array_equal(got, ref[i]), (i, name, np.abs(got - ref[i]).max())
    del g
    del c


def test_scale_reuse_rejects_other_values(eng):
    assert eng.lib().tmx_set_scale_reuse(2) != 0
    assert eng.lib().tmx_set_scale_reuse(-1) != 0
    eng.check(eng.lib().tmx_set_scale_reuse(1))
