"""GPU parity of the engine against the CPU oracle (run on an MI355X).

Tolerances: FP64 paths must agree with the oracle to <= 1e-12 relative Frobenius norm;
closed-form KATs to 1e-10 absolute (TAMM's own tolerance, Test_Ops.cpp:44-46) and the
Test_Mult_Ops norm exactly.  Inputs are generated on the host by the oracle's generator
and uploaded, except where the device generator itself is under test."""
import zlib

import numpy as np
import pytest

from tamm_amd import configs
from tests.kats import KATS, MULT_OPS, label_ids

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
SEED = 20261202


def _rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def test_device_generator_matches_host_bitwise(eng, orc):
    for wlf in (configs.tiny_spin, configs.tiny_irrep, configs.config0):
        wl = wlf()
        ew = eng.EngineWorkload(wl, SEED)
        ow = orc.OracleWorkload(wl, SEED)
        for name in wl.tensors:
            got = ew.tensors[name].download()
            assert np.array_equal(got, ow.data[name]), (wl.name, name)


# ---- block_multiply (kernels::block_multiply) -------------------------------------------

BLOCK_CASES = [
    # (clabels, alabels, blabels) over label dims
    ("abij", "abcd", "cdij"),       # ladder: no permutation
    ("abij", "kbcj", "acik"),       # ring: A, B and C permuted
    ("abij", "klij", "abkl"),       # hole-hole
    ("jikl", "ijmo", "mokl"),       # Test_Mult_Ops order
    ("ijk", "jli", "lik"),          # batch label i
    ("ijkl", "jlkm", "likm"),       # batch labels l, k
    ("ij", "ik", "jk"),             # matrix product with transposed B
    ("ijkl", "ij", "kl"),           # outer product (K = 1)
    ("", "ij", "ij"),               # full contraction to a scalar
    ("i", "ij", "k"),               # A-reduce j, B-reduce k
    ("ab", "ba", ""),               # scalar B
]


@pytest.mark.parametrize("case", BLOCK_CASES, ids=["-".join(c) for c in BLOCK_CASES])
@pytest.mark.parametrize("is_assign", [False, True])
def test_block_multiply_matches_oracle(eng, orc, case, is_assign):
    cl, al, bl = case
    rng = np.random.default_rng(zlib.crc32("-".join(case).encode()))
    dims = {ch: int(rng.integers(1, 24)) for ch in set(cl + al + bl)}
    dims.update({ch: int(rng.integers(40, 90)) for ch in cl[:1]})  # one ragged large mode
    L = label_ids(cl, al, bl)
    ad, bd, cd = [dims[c] for c in al], [dims[c] for c in bl], [dims[c] for c in cl]
    A = rng.standard_normal(int(np.prod(ad)))
    B = rng.standard_normal(int(np.prod(bd)))
    C0 = rng.standard_normal(int(np.prod(cd)))
    ref = C0.copy()
    if is_assign:
        ref[:] = 0.0
    orc.block_multiply(0.7, A, ad, [L[c] for c in al], B, bd, [L[c] for c in bl], ref, cd,
                       [L[c] for c in cl])
    dA, dB, dC = (eng.DeviceBuffer.from_array(x) for x in (A, B, C0))
    eng.block_multiply(0.7, dA, ad, [L[c] for c in al], dB, bd, [L[c] for c in bl], dC, cd,
                       [L[c] for c in cl], is_assign=is_assign)
    got = dC.download(np.float64, C0.size)
    assert _rel(got, ref) <= FP64_TOL


@pytest.mark.parametrize("perm", [(1, 0), (2, 0, 1), (0, 2, 1), (3, 1, 0, 2), (0, 1, 3, 2),
                                  (0, 1, 2, 3), (2, 3, 0, 1)])
@pytest.mark.parametrize("is_assign", [False, True])
def test_block_assign_matches_transpose(eng, orc, perm, is_assign):
    rng = np.random.default_rng(len(perm) * 10 + perm[0])
    sdims = [37, 5, 70, 3][:len(perm)]
    src = rng.standard_normal(sdims)
    sl = [-(i + 1) for i in range(len(perm))]
    dl = [sl[p] for p in perm]
    ddims = [sdims[p] for p in perm]
    dst0 = rng.standard_normal(int(np.prod(ddims)))
    ref = dst0.copy()
    orc.assign(ref, ddims, dl, -1.5, src.ravel(), sdims, sl, is_assign)
    ds, dd = eng.DeviceBuffer.from_array(src.ravel()), eng.DeviceBuffer.from_array(dst0)
    eng.block_assign(dd, ddims, dl, -1.5, ds, sdims, sl, is_assign=is_assign)
    got = dd.download(np.float64, dst0.size)
    assert np.array_equal(got, ref) or _rel(got, ref) < 1e-15


# ---- full contractions --------------------------------------------------------------------

def _run_contractions(eng, orc, wl, modes, tol=FP64_TOL):
    ow = orc.OracleWorkload(wl, SEED)
    ew = eng.EngineWorkload(wl, SEED, allocate=False)
    for name in wl.tensors:
        ew.tensors[name].upload(ow.data[name])
    for term in wl.terms:
        cn = term.c[0]
        base = np.random.default_rng(1).standard_normal(ow.tensors[cn].size)
        ref = base.copy()
        ow.run_term(term, out=ref)
        for mode in modes:
            ew.tensors[cn].upload(base)
            c = ew.contraction(term, mode)
            c.execute()
            got = ew.tensors[cn].download()
            err = _rel(got - base, ref - base)
            assert err <= tol, (wl.name, term.name, eng.EXEC_NAMES[mode], err)


@pytest.mark.parametrize("wlf", [configs.tiny_spin, configs.tiny_irrep, configs.config0],
                         ids=lambda f: f.__name__)
def test_contractions_fp64_match_oracle(eng, orc, wlf):
    _run_contractions(eng, orc, wlf(), [eng.EXEC_BLOCKPAIR_F64, eng.EXEC_GROUPED_F64])


def test_contractions_mid_spin_grouped(eng, orc):
    wl = configs.scaled_spin(12, 48, 8, 16)
    _run_contractions(eng, orc, wl, [eng.EXEC_GROUPED_F64, eng.EXEC_SECTOR_F64])


@pytest.mark.parametrize("wlf", [configs.tiny_spin, configs.tiny_irrep, configs.config0],
                         ids=lambda f: f.__name__)
def test_contractions_sector_fp64_match_oracle(eng, orc, wlf):
    _run_contractions(eng, orc, wlf(), [eng.EXEC_SECTOR_F64])


# ---- split precision (int8 residues, Ozaki scheme II) ---------------------------------------

@pytest.mark.parametrize("wlf", [configs.tiny_spin, configs.tiny_irrep, configs.config0],
                         ids=lambda f: f.__name__)
def test_contractions_split_within_1e10(eng, orc, wlf):
    """Default split precision (35 bits per operand element) meets 1e-10 relative."""
    _run_contractions(eng, orc, wlf(), [eng.EXEC_SPLIT_I8], tol=1e-10)


@pytest.mark.parametrize("bits,lo,hi", [(12, 1e-5, 3e-3), (24, 2e-9, 1e-6), (45, 0.0, 5e-14)])
def test_split_error_follows_bits(eng, orc, bits, lo, hi):
    """Error model: relative error ~ 2^-bits.  A wrong CRT or scaling breaks the high-bit
    case; a silently exact fallback would break the low-bit lower bound."""
    wl = configs.scaled_spin(12, 48, 8, 16)
    ow = orc.OracleWorkload(wl, SEED)
    ew = eng.EngineWorkload(wl, SEED, allocate=False)
    for name in wl.tensors:
        ew.tensors[name].upload(ow.data[name])
    eng.check(eng.lib().tmx_set_split_bits(bits))
    try:
        for term in wl.terms:
            ref = ow.run_term(term, out=ow.tensors["R"].zeros())
            ew.tensors["R"].fill(0.0)
            c = ew.contraction(term, eng.EXEC_SPLIT_I8)
            assert c.split_info()[0] == bits
            c.execute()
            err = _rel(ew.tensors["R"].download(), ref)
            assert lo <= err <= hi, (term.name, bits, err)
    finally:
        eng.check(eng.lib().tmx_set_split_bits(35))


def test_split_exact_on_integer_data(eng):
    """Small-integer operands: the split product carries only the 2^-35 scaling error, so it
    rounds back to the exact integer result."""
    L = label_ids("ijk")
    s = eng.Space(96, 32)
    A, B, Cc = (eng.Tensor([s, s]).allocate() for _ in range(3))
    rng = np.random.default_rng(5)
    a = rng.integers(-7, 8, A.size).astype(np.float64)
    b = rng.integers(-7, 8, B.size).astype(np.float64)
    A.upload(a)
    B.upload(b)
    Cc.fill(0.0)
    eng.Contraction(Cc, [L["i"], L["j"]], 1.0, A, [L["i"], L["k"]], B, [L["k"], L["j"]],
                    eng.EXEC_SPLIT_I8).execute()
    Ad, Bd = a.reshape(3, 3, 32, 32).transpose(0, 2, 1, 3).reshape(96, 96), b.reshape(3, 3, 32, 32).transpose(0, 2, 1, 3).reshape(96, 96)
    ref = (Ad @ Bd).reshape(3, 32, 3, 32).transpose(0, 2, 1, 3).ravel()
    got = Cc.download()
    assert np.array_equal(np.rint(got), ref)
    assert _rel(got, ref) < 1e-10


# ---- KATs through the engine ---------------------------------------------------------------

@pytest.mark.parametrize("kat", KATS, ids=[k[0] for k in KATS])
@pytest.mark.parametrize("mode", [0, 1])
def test_kat_engine(eng, kat, mode):
    name, n, tile, cl, cv, alpha, al, av, bl, bv, expected = kat
    L = label_ids(cl, al, bl)
    s = eng.Space(n, tile, named=[(0, n // 2), (n // 2, n)])
    Ct, At, Bt = (eng.Tensor([s] * len(x)).allocate() for x in (cl, al, bl))
    Ct.fill(cv)
    At.fill(av)
    Bt.fill(bv)
    c = eng.Contraction(Ct, [L[x] for x in cl], alpha, At, [L[x] for x in al], Bt, [L[x] for x in bl], mode)
    c.execute()
    got = Ct.download()
    assert np.max(np.abs(got - expected)) < 1e-10


def test_kat_mult_ops_norm_engine(eng):
    k = MULT_OPS
    L = label_ids(k["c"], k["a"], k["b"])
    s = eng.Space(k["n"], k["tile"])
    A, B, Cc = (eng.Tensor([s] * 4).allocate() for _ in range(3))
    A.fill(k["aval"])
    B.fill(k["bval"])
    Cc.fill(0.0)
    for mode in (eng.EXEC_GROUPED_F64, eng.EXEC_BLOCKPAIR_F64):
        Cc.fill(0.0)
        eng.Contraction(Cc, [L[c] for c in k["c"]], 1.0, A, [L[c] for c in k["a"]], B,
                        [L[c] for c in k["b"]], mode).execute()
        assert Cc.norm() == k["norm"]


def test_addop_matches_oracle(eng, orc):
    wl = configs.tiny_irrep()
    ow = orc.OracleWorkload(wl, SEED)
    ew = eng.EngineWorkload(wl, SEED, allocate=False)
    for n in ("R", "T"):
        ew.tensors[n].upload(ow.data[n])
    L = label_ids("abij")
    # R(a,b,i,j) += -0.5 * T(b,a,j,i)
    op = eng.AddOp(ew.tensors["R"], [L[c] for c in "abij"], -0.5, ew.tensors["T"], [L[c] for c in "baji"])
    op.execute()
    got = ew.tensors["R"].download()
    Td = ow.tensors["T"].to_dense(ow.data["T"])
    ref = ow.tensors["R"].from_dense(-0.5 * np.transpose(Td, (1, 0, 3, 2)))
    assert _rel(got, ref) < 1e-15


def test_norm_and_dot(eng, orc):
    wl = configs.tiny_spin()
    ew = eng.EngineWorkload(wl, SEED)
    ow = orc.OracleWorkload(wl, SEED)
    assert abs(ew.tensors["V"].norm() - orc.norm(ow.data["V"])) <= 1e-14 * orc.norm(ow.data["V"])
    d = eng.dot(ew.tensors["T"].ptr, ew.tensors["T"].ptr, ew.tensors["T"].size)
    assert abs(d - float(ow.data["T"] @ ow.data["T"])) <= 1e-14 * d
