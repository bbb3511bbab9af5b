"""Every execution mode on TAMM's closed-form cases and on labels over named subspaces.

Sources:
  tests/tamm/Test_Ops.cpp:1396-1805   three/four/two-dimensional mult ops (tests/kats.py)
  tests/tamm/Test_Ops.cpp:1808-1880   MultOp with RHS reduction labels
  tests/tamm/Test_Ops.cpp:1882-1968   MO("occ") / MO("virt") labels indexing full-MO tensors

The sector modes (SECTOR_F64, SPLIT_I8) execute contractions with A-outer, B-outer and inner
labels and nothing else; any other label pattern (batch labels, labels reduced on one side
only, outer products, full reductions) is refused with TMX_ERR_UNSUPPORTED at plan creation,
and the AUTO modes then take the grouped block-pair kernel.  Bars: FP64 modes 1e-10 absolute
on the closed forms (TAMM's own tolerance, Test_Ops.cpp:44-46) and 1e-12 relative against the
oracle on random data; split 1e-10 relative.
"""
import numpy as np
import pytest

from tests.kats import KATS, label_ids

pytestmark = pytest.mark.gpu

MODES = [0, 1, 2, 3, 4, 5]  # BLOCKPAIR, GROUPED, SECTOR, SPLIT, AUTO, AUTO_SPLIT
SECTOR_MODES = (2, 3)
# The KATs whose labels the sector modes take (C(i,j) += A(i,k) * B(j,k)); pinned by name so
# a change of the eligibility rule shows up here.
SECTOR_KATS = {"mult_2x2x2", "mult_2x2x2_t4"}
ERR_UNSUPPORTED = 4


def _tol(mode, expected):
    return 1e-10 * max(1.0, abs(expected)) if mode in (3, 5) else 1e-10


@pytest.mark.parametrize("kat", KATS, ids=[k[0] for k in KATS])
@pytest.mark.parametrize("mode", MODES)
def test_kat_every_mode(eng, kat, mode):
    name, n, tile, cl, cv, alpha, al, av, bl, bv, expected = kat
    L = label_ids(cl, al, bl)
    s = eng.Space(n, tile, named=[(0, n // 2), (n // 2, n)])
    Ct, At, Bt = (eng.Tensor([s] * len(x)).allocate() for x in (cl, al, bl))
    Ct.fill(cv)
    At.fill(av)
    Bt.fill(bv)
    args = (Ct, [L[x] for x in cl], alpha, At, [L[x] for x in al], Bt, [L[x] for x in bl], mode)
    if mode in SECTOR_MODES and name not in SECTOR_KATS:
        with pytest.raises(eng.EngineError) as ei:
            eng.Contraction(*args)
        assert ei.value.code == ERR_UNSUPPORTED
        return
    c = eng.Contraction(*args)
    c.execute()
    got = Ct.download()
    assert got.size == Ct.size
    assert np.max(np.abs(got - expected)) < _tol(mode, expected), (name, mode)


# ---- labels on named subspaces (internal::translate_blockid, utils.hpp:263-286) ----------

def _subspace_setup(eng, orc):
    """MO{range(0,14), occ [0,10), virt [10,14)} with tiles {5,5,2,2}; CV{range(0,4)}, tile 1.
    res/tmp tensors over O = MO("occ") (tiles {5,5}), t1 over (V, O) with V = MO("virt")."""
    mk = {"eng": lambda n, t, named=(): eng.Space(n, t, named=named),
          "orc": lambda n, t, named=(): orc.Space(n, t, named=named)}
    out = {}
    for side, new in mk.items():
        T = eng.Tensor if side == "eng" else orc.Tensor
        cv, mo = new(4, 1), new(14, [5, 5, 2, 2], [(0, 10), (10, 14)])
        occ, vir = new(10, [5, 5]), new(4, [2, 2])
        out[side] = dict(CV3D=T([mo, mo, cv]), CV2D=T([mo, mo]), res=T([occ, occ]),
                         res1=T([occ, occ]), tmp1=T([occ, occ]), tmp2=T([occ, occ]),
                         t1=T([vir, occ]), sc1=T([]))
    return out


L_SUB = label_ids("hHpcq")  # h = h1, H = h2, p = p1 (virt), c = cind, q = h3
# (output, out labels, alpha, A, A labels, B, B labels, tile offsets, label tile counts)
SUB_OPS = [
    ("res", "Hh", 1.0, "t1", "ph", "CV3D", "Hpc", ([0, 0], [0, 0], [0, 2, 0]), ([2, 2], [2, 2], [2, 2, 4])),
    ("sc1", "", 1.0, "t1", "ph", "CV3D", "hpc", ([], [0, 0], [0, 2, 0]), ([], [2, 2], [2, 2, 4])),
    ("tmp1", "Hh", -1.0, "res", "qh", "res", "Hq", None, None),
] + [("res1", "Hh", 1.0, "t1", "ph", "CV2D", "Hp", ([0, 0], [0, 0], [0, 2]), ([2, 2], [2, 2], [2, 2]))] * 4 + [
    ("tmp2", "Hh", -1.0, "res1", "qh", "res1", "Hq", None, None),
]
# ops the sector modes take: A-outer, B-outer and inner labels only
SUB_SECTOR_OK = {"tmp1", "res1", "tmp2"}


def _run_subspace(eng, ts, mode):
    """Executes SUB_OPS in order; an op a sector mode refuses (TMX_ERR_UNSUPPORTED, checked)
    runs on the grouped kernel so the chain continues.  Returns {output: (mode used, ...)}."""
    used = {}
    for out, ol, alpha, a, al, b, bl, toffs, ntl in SUB_OPS:
        args = (ts[out], [L_SUB[x] for x in ol], alpha, ts[a], [L_SUB[x] for x in al],
                ts[b], [L_SUB[x] for x in bl])
        m = mode
        if mode in SECTOR_MODES and out not in SUB_SECTOR_OK:
            with pytest.raises(eng.EngineError) as ei:
                eng.Contraction(*args, mode, toffs=toffs, ntiles=ntl)
            assert ei.value.code == ERR_UNSUPPORTED, out
            m = 1
        eng.Contraction(*args, m, toffs=toffs, ntiles=ntl).execute()
        used[out] = m
    return used


@pytest.mark.parametrize("mode", MODES)
def test_subspace_labels_kat_every_mode(eng, orc, mode):
    """Test_Ops.cpp:1882-1968 with its constant fills: res == 4*4*2*42, tmp1 == tmp2 ==
    -10*res^2, sc1 == 10*4*4*2*42."""
    ts = _subspace_setup(eng, orc)["eng"]
    for k, t in ts.items():
        t.allocate()
        t.fill({"CV3D": 42.0, "CV2D": 42.0, "t1": 2.0}.get(k, 0.0))
    _run_subspace(eng, ts, mode)
    r = 4 * 4 * 2 * 42
    tol = 1e-10 if mode not in (3, 5) else 0.0
    got = {k: ts[k].download() for k in ("res", "res1", "tmp1", "tmp2", "sc1")}
    assert np.max(np.abs(got["res"] - r)) <= 1e-10
    assert np.max(np.abs(got["res1"] - r)) <= max(tol, 1e-10 * r)
    assert np.max(np.abs(got["tmp1"] + 10 * r * r)) <= max(tol, 1e-10 * 10 * r * r)
    assert np.max(np.abs(got["tmp2"] + 10 * r * r)) <= max(tol, 1e-10 * 10 * r * r)
    assert abs(got["sc1"][0] - 10 * r) <= 1e-10
    if mode not in (3, 5):
        assert np.array_equal(got["tmp1"], got["tmp2"])


@pytest.mark.parametrize("mode", MODES)
def test_subspace_labels_random_every_mode(eng, orc, mode):
    """The same op chain on seeded random data against the oracle's restatement of
    translate_blockid (tile offsets of the occ/virt labels inside the MO modes)."""
    both = _subspace_setup(eng, orc)
    ts, os_ = both["eng"], both["orc"]
    rng = np.random.default_rng(1968)
    data = {}
    for k, t in os_.items():
        data[k] = rng.standard_normal(t.size) if k in ("CV3D", "CV2D", "t1") else t.zeros()
        ts[k].upload(data[k])
    for out, ol, alpha, a, al, b, bl, toffs, ntl in SUB_OPS:
        kw = {} if toffs is None else dict(toffs=toffs, ntiles=ntl)
        orc.contract(os_[out], [L_SUB[x] for x in ol], data[out], alpha, os_[a], [L_SUB[x] for x in al],
                     data[a], os_[b], [L_SUB[x] for x in bl], data[b], **kw)
    used = _run_subspace(eng, ts, mode)
    for k in ("res", "sc1", "tmp1", "res1", "tmp2"):
        got, ref = ts[k].download(), data[k]
        rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        bar = 1e-10 if used[k] in (3, 5) else 1e-12
        assert rel <= bar, (k, mode, used[k], rel)
