"""Pins the CPU oracle (oracle/) before it is trusted as the checker.

Each oracle function is checked against something it does not compute itself:
closed-form KATs from TAMM's tests, TAMM's documented tile vectors, closed-form
sparsity fractions, brute-force rule evaluation, and numpy.einsum over fully
materialised dense tensors.  A planted-mistake test shows the einsum check has power.
"""
import numpy as np
import pytest

from tamm_amd import configs
from tests.kats import KATS, MULT_OPS, label_ids


# ---- tiling (tiled_index_space.hpp:1135-1260) ----------------------------------------

def test_tiling_cfg1_spin_boundaries(orc):
    O = orc.Space(40, 16, spin=[(0, 20, 1), (20, 40, 2)])
    V = orc.Space(240, 48, spin=[(0, 120, 1), (120, 240, 2)])
    assert O.tile_sizes == [16, 4, 16, 4]
    assert list(O.spin) == [1, 1, 2, 2]
    assert V.tile_sizes == [48, 48, 24, 48, 48, 24]
    assert list(V.spin) == [1, 1, 1, 2, 2, 2]


def test_tiling_cfg3(orc):
    O = orc.Space(30, 16, spin=[(0, 15, 1), (15, 30, 2)])
    V = orc.Space(180, 36, spin=[(0, 90, 1), (90, 180, 2)])
    assert O.tile_sizes == [15, 15]
    assert V.tile_sizes == [36, 36, 18, 36, 36, 18]


def test_tiling_named_subspaces(orc):
    # Test_Ops.cpp:1808-1880: IndexSpace{range(0,10), {nr1:[0,5), nr2:[5,10)}}, tile 3
    s = orc.Space(10, 3, named=[(0, 5), (5, 10)])
    assert s.tile_sizes == [3, 2, 3, 2]
    s1 = orc.Space(10, 1, named=[(0, 5), (5, 10)])
    assert s1.tile_sizes == [1] * 10
    # MO{range(0,14), occ [0,10), virt [10,14)} with tiles {5,5,2,2}
    mo = orc.Space(14, [5, 5, 2, 2], named=[(0, 10), (10, 14)])
    assert mo.tile_sizes == [5, 5, 2, 2]
    with pytest.raises(ValueError):
        orc.Space(14, [5, 6, 3], named=[(0, 10), (10, 14)])  # boundary 10 is not a tile start
    with pytest.raises(ValueError):
        orc.Space(14, [5, 5, 2], named=[(0, 10), (10, 14)])  # does not sum to 14


def test_tiling_irrep_ranges_are_tile_starts(orc):
    wl = configs.config2()
    for sd in wl.spaces.values():
        s = orc.Space(sd.n, sd.tiles, sd.spin, sd.irrep)
        assert sum(s.tile_sizes) == sd.n
        starts = set(int(x) for x in s.offsets)
        for lo, hi, _ in sd.irrep:
            assert lo in starts and hi in starts
        assert all(24 >= 1 for _ in s.tile_sizes)
        # tile sizes are the seeded U{24..96} draws truncated at range ends
        assert max(s.tile_sizes) <= 96


# ---- sparsity (tensor_base.hpp:243-263) ------------------------------------------------

def test_spin_survival_fraction_cfg1(orc):
    V = orc.Space(240, 48, spin=[(0, 120, 1), (120, 240, 2)])
    O = orc.Space(40, 16, spin=[(0, 20, 1), (20, 40, 2)])
    tV = orc.Tensor([V] * 4, (2, 2))
    tT = orc.Tensor([V, V, O, O], (2, 2))
    # closed form: a {2,2} spin tensor keeps 6 of the 16 spin combinations = 3/8
    assert tV.size * 8 == 3 * 240 ** 4
    assert tT.size * 8 == 3 * 240 ** 2 * 40 ** 2
    assert tV.size == 1244160000


def _rule(spins_up, spins_lo, irreps):
    ok = sum(spins_up) == sum(spins_lo)
    if irreps is not None:
        x = 0
        for r in irreps:
            x ^= r
        ok = ok and x == 0
    return ok


@pytest.mark.parametrize("wlf", [configs.tiny_spin, configs.tiny_irrep])
def test_sparsity_brute_force(orc, wlf):
    wl = wlf()
    ow = orc.OracleWorkload(wl, 1)
    for name, td in wl.tensors.items():
        t = ow.tensors[name]
        sp = [ow.spaces[m] for m in td.modes]
        nz = set(int(k) for k in t.keys)
        for key in range(int(np.prod([s.ntiles for s in sp]))):
            bid = t.blockid(key)
            up = [sp[i].spin[bid[i]] for i in range(2)]
            lo = [sp[i].spin[bid[i]] for i in range(2, 4)]
            irr = [sp[i].irrep[bid[i]] for i in range(4)] if td.irrep_rule else None
            assert (key in nz) == _rule(up, lo, irr), (name, bid)


def test_layout_invariants(orc):
    wl = configs.tiny_irrep()
    ow = orc.OracleWorkload(wl, 1)
    for name, t in ow.tensors.items():
        assert np.all(np.diff(t.keys) > 0)
        assert t.offsets[0] == 0
        assert np.all(t.offsets[1:] == np.cumsum(t.sizes)[:-1])
        assert t.size == int(t.sizes.sum())
        for k, sz in zip(t.keys, t.sizes):
            bid = t.blockid(int(k))
            assert sz == np.prod([s.offsets[b + 1] - s.offsets[b] for s, b in zip(t.spaces, bid)])


# ---- contraction KATs (Test_Ops.cpp, Test_Mult_Ops.cpp) ---------------------------------

def _kat_tensors(orc, n, tile, labels):
    s = orc.Space(n, tile, named=[(0, n // 2), (n // 2, n)] if n >= 2 else [])
    return orc.Tensor([s] * len(labels)), s


@pytest.mark.parametrize("kat", KATS, ids=[k[0] for k in KATS])
def test_kat_oracle(orc, kat):
    name, n, tile, cl, cv, alpha, al, av, bl, bv, expected = kat
    L = label_ids(cl, al, bl)
    Ct, _ = _kat_tensors(orc, n, tile, cl)
    At, _ = _kat_tensors(orc, n, tile, al)
    Bt, _ = _kat_tensors(orc, n, tile, bl)
    Cd = np.full(Ct.size, cv)
    orc.contract(Ct, [L[c] for c in cl], Cd, alpha, At, [L[c] for c in al], np.full(At.size, av),
                 Bt, [L[c] for c in bl], np.full(Bt.size, bv))
    assert np.max(np.abs(Cd - expected)) < 1e-10


def test_kat_subspace_labels(orc):
    """Test_Ops.cpp:1855-1968: labels on occ/virt subspaces of MO tensors; tmp1 == tmp2."""
    cv = orc.Space(4, 1)
    mo = orc.Space(14, [5, 5, 2, 2], named=[(0, 10), (10, 14)])
    occ = orc.Space(10, [5, 5])
    vir = orc.Space(4, [2, 2])
    CV3D, CV2D = orc.Tensor([mo, mo, cv]), orc.Tensor([mo, mo])
    res, res1 = orc.Tensor([occ, occ]), orc.Tensor([occ, occ])
    tmp1, tmp2 = orc.Tensor([occ, occ]), orc.Tensor([occ, occ])
    t1, sc1 = orc.Tensor([vir, occ]), orc.Tensor([])
    L = label_ids("hHpcq")  # h=h1, H=h2, p=p1, c=cind, q=h3
    d = {k: v.zeros() for k, v in dict(res=res, res1=res1, tmp1=tmp1, tmp2=tmp2, sc1=sc1).items()}
    d3, d2, dt = np.full(CV3D.size, 42.0), np.full(CV2D.size, 42.0), np.full(t1.size, 2.0)
    # res(h2,h1) += t1(p1,h1) * CV3D(h2,p1,cind): CV3D modes are over MO; h2 -> tiles 0..1,
    # p1 -> tiles 2..3 (virt), cind -> CV
    orc.contract(res, [L["H"], L["h"]], d["res"], 1.0, t1, [L["p"], L["h"]], dt,
                 CV3D, [L["H"], L["p"], L["c"]], d3,
                 toffs=([0, 0], [0, 0], [0, 2, 0]), ntiles=([2, 2], [2, 2], [2, 2, 4]))
    orc.contract(sc1, [], d["sc1"], 1.0, t1, [L["p"], L["h"]], dt, CV3D, [L["h"], L["p"], L["c"]], d3,
                 toffs=([], [0, 0], [0, 2, 0]), ntiles=([], [2, 2], [2, 2, 4]))
    orc.contract(tmp1, [L["H"], L["h"]], d["tmp1"], -1.0, res, [L["q"], L["h"]], d["res"],
                 res, [L["H"], L["q"]], d["res"])
    for _ in range(4):
        orc.contract(res1, [L["H"], L["h"]], d["res1"], 1.0, t1, [L["p"], L["h"]], dt,
                     CV2D, [L["H"], L["p"]], d2, toffs=([0, 0], [0, 0], [0, 2]), ntiles=([2, 2], [2, 2], [2, 2]))
    orc.contract(tmp2, [L["H"], L["h"]], d["tmp2"], -1.0, res1, [L["q"], L["h"]], d["res1"],
                 res1, [L["H"], L["q"]], d["res1"])
    assert np.all(d["res"] == 4 * 4 * 2 * 42)
    assert np.all(d["tmp1"] == d["tmp2"])
    assert np.all(d["tmp1"] == -10 * (4 * 4 * 2 * 42) ** 2)
    assert d["sc1"][0] == 10 * 4 * 4 * 2 * 42


def test_kat_mult_ops_norm_exact(orc):
    k = MULT_OPS
    L = label_ids(k["c"], k["a"], k["b"])
    s = orc.Space(k["n"], k["tile"])
    A, B, Cc = orc.Tensor([s] * 4), orc.Tensor([s] * 4), orc.Tensor([s] * 4)
    Cd = Cc.zeros()
    orc.contract(Cc, [L[c] for c in k["c"]], Cd, 1.0, A, [L[c] for c in k["a"]], np.full(A.size, k["aval"]),
                 B, [L[c] for c in k["b"]], np.full(B.size, k["bval"]))
    assert orc.norm(Cd) == k["norm"]  # exact: every partial sum is an integer


# ---- dense einsum cross-check --------------------------------------------------------------

def _einsum_check(orc, wl, seed, tol=1e-13, corrupt=None):
    ow = orc.OracleWorkload(wl, seed)
    dense = {k: ow.tensors[k].to_dense(ow.data[k]) for k in ow.data}
    errs = {}
    for term in wl.terms:
        cn = term.c[0]
        out = ow.tensors[cn].zeros()
        t = term if corrupt is None else corrupt(term)
        ow.run_term(t, out=out)
        ref = term.alpha * np.einsum(f"{term.a[1]},{term.b[1]}->{term.c[1]}", dense[term.a[0]], dense[term.b[0]])
        ref_b = ow.tensors[cn].from_dense(ref)
        errs[term.name] = np.linalg.norm(out - ref_b) / np.linalg.norm(ref_b)
    return errs


@pytest.mark.parametrize("wlf", [configs.tiny_spin, configs.tiny_irrep, configs.config0])
def test_contract_matches_einsum(orc, wlf):
    errs = _einsum_check(orc, wlf(), 7)
    for name, e in errs.items():
        assert e < 1e-13, (name, e)


def test_planted_label_mistake_is_detected(orc):
    """A transposed operand label in the oracle call must fail the einsum check."""
    from dataclasses import replace

    def corrupt(term):
        bn, bl = term.b
        return replace(term, b=(bn, bl[:2] + bl[3] + bl[2]))  # swap the last two labels of B

    errs = _einsum_check(orc, configs.tiny_spin(), 7, corrupt=corrupt)
    assert min(errs.values()) > 1e-3


# ---- assign (assign.hpp:312-485) -------------------------------------------------------------

@pytest.mark.parametrize("perm", [(1, 0), (2, 0, 1), (3, 1, 0, 2), (0, 1, 2, 3)])
def test_assign_matches_transpose(orc, perm):
    rng = np.random.default_rng(3)
    sdims = [3, 4, 5, 6][:len(perm)]
    src = rng.standard_normal(sdims)
    sl = [-(i + 1) for i in range(len(perm))]
    dl = [sl[p] for p in perm]
    ddims = [sdims[p] for p in perm]
    dst = np.ones(int(np.prod(ddims)))
    orc.assign(dst, ddims, dl, 2.0, src.ravel(), sdims, sl, is_assign=False)
    np.testing.assert_allclose(dst.reshape(ddims), 1.0 + 2.0 * np.transpose(src, perm), rtol=0, atol=1e-15)
    orc.assign(dst, ddims, dl, 1.0, src.ravel(), sdims, sl, is_assign=True)
    np.testing.assert_array_equal(dst.reshape(ddims), np.transpose(src, perm))


# ---- generator ---------------------------------------------------------------------------

def test_generator_symmetry_and_moments(orc):
    wl = configs.tiny_spin()
    ow = orc.OracleWorkload(wl, 20261202)
    V = ow.tensors["V"].to_dense(ow.data["V"])
    assert np.array_equal(V, V.transpose(1, 0, 3, 2))
    big = orc.Tensor([orc.Space(200, 50)] * 3).fill_normal(5, 9, 1.0)
    assert abs(big.mean()) < 5e-3 and abs(big.var() - 1.0) < 5e-3
    again = orc.Tensor([orc.Space(200, 25)] * 3).fill_normal(5, 9, 1.0)
    # layout-independent: same dense values under another tiling
    d1 = orc.Tensor([orc.Space(200, 50)] * 3)
    d2 = orc.Tensor([orc.Space(200, 25)] * 3)
    assert np.array_equal(d1.to_dense(big), d2.to_dense(again))
