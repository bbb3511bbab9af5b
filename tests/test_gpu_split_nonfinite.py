"""Split precision on rows with non-finite values and on tiny and subnormal rows, on the GPU.

Split packing scales each class-matrix row by its largest |x|, taken over the row's whole K
range.  The behaviour pinned here:
  * Inf or NaN anywhere in a row of A (a column of B) makes every element of the output row
    (column) NaN.  FP64 gives Inf or NaN there: Inf * 0 is NaN, and the sign of an Inf sum
    depends on the data, which residues do not carry.  So the set of non-finite outputs is
    FP64's exactly (checked against the oracle, the FP64 sector path and the grouped path),
    and the finite outputs keep the 1e-10 bar.
  * A row whose largest |x| is tiny (1e-301) or subnormal (4e-317) is scaled without overflow,
    and its inverse scale keeps full precision (held as mantissa and exponent), so blocks of
    R ~ 1e-200 computed from such rows still meet 1e-10 against |V| |T|.
  * With row-scale reuse, an operand that gains a NaN after its maxima were kept is packed
    again with the new maxima, so the second execution shows the same pattern.
"""
import numpy as np
import pytest

from tamm_amd import configs

pytestmark = pytest.mark.gpu

SEED = 20261202


def _nrm(x):
    m = float(np.abs(x).max()) if x.size else 0.0
    return m * float(np.linalg.norm(x / m)) if m > 0 else 0.0


def _setup(eng, orc, V, T):
    wl = configs.tiny_spin()
    ow = orc.OracleWorkload(wl, SEED)
    ow.data["V"], ow.data["T"] = V(ow), T(ow)
    ew = eng.EngineWorkload(wl, SEED, allocate=False)
    for name in wl.tensors:
        ew.tensors[name].upload(ow.data[name])
    term = next(t for t in wl.terms if t.name == "ladder")
    return wl, ow, ew, term


def _poison(ow):
    """NaN in one element of V, +Inf in an element of another V block, -Inf in one T element."""
    V, T = ow.data["V"].copy(), ow.data["T"].copy()
    Vt, Tt = ow.tensors["V"], ow.tensors["T"]
    V[Vt.offsets[1] + 3] = np.nan
    V[Vt.offsets[len(Vt.offsets) // 2] + 1] = np.inf
    T[Tt.offsets[2] + 5] = -np.inf
    return V, T


def test_nonfinite_rows_propagate(eng, orc):
    wl, ow, ew, term = _setup(eng, orc, lambda o: _poison(o)[0], lambda o: _poison(o)[1])
    with np.errstate(invalid="ignore", over="ignore"):
        ref = ow.run_term(term, out=ow.tensors["R"].zeros())
    bad = ~np.isfinite(ref)
    assert 0 < bad.sum() < bad.size
    R = ew.tensors["R"]
    fp64 = {}
    for mode in (eng.EXEC_SECTOR_F64, eng.EXEC_GROUPED_F64, eng.EXEC_SPLIT_I8):
        c = ew.contraction(term, mode)
        for rep in range(2):  # the second split execution reuses its row maxima
            R.fill(0.0)
            c.execute()
            got = R.download()
            name = eng.EXEC_NAMES[mode]
            assert np.array_equal(~np.isfinite(got), bad), (name, rep)
            if mode == eng.EXEC_SPLIT_I8:
                assert np.all(np.isnan(got[bad])), rep
                rel = np.linalg.norm(got[~bad] - ref[~bad]) / np.linalg.norm(ref[~bad])
                assert rel <= 1e-10, (rep, rel)
            else:
                fp64[mode] = got
                rel = np.linalg.norm(got[~bad] - ref[~bad]) / np.linalg.norm(ref[~bad])
                assert rel <= 1e-12, (name, rel)
        del c


def test_nan_after_kept_maxima_repacks(eng, orc):
    wl, ow, ew, term = _setup(eng, orc, lambda o: o.data["V"], lambda o: o.data["T"])
    R = ew.tensors["R"]
    c = ew.contraction(term, eng.EXEC_SPLIT_I8)
    R.fill(0.0)
    c.execute()  # keeps V's and T's row maxima
    clean = R.download()
    V, T = _poison(ow)
    ew.tensors["V"].upload(V)
    ew.tensors["T"].upload(T)
    ow.data["V"], ow.data["T"] = V, T
    with np.errstate(invalid="ignore", over="ignore"):
        ref = ow.run_term(term, out=ow.tensors["R"].zeros())
    bad = ~np.isfinite(ref)
    R.fill(0.0)
    c.execute()  # packs with the kept (finite) maxima, sees them change, packs again
    got = R.download()
    assert np.array_equal(np.isnan(got), bad) and np.all(np.isfinite(got[~bad]))
    rel = np.linalg.norm(got[~bad] - ref[~bad]) / np.linalg.norm(ref[~bad])
    assert rel <= 1e-10
    assert np.all(np.isfinite(clean))


# V blocks scaled per (a,b) tile pair by 1, 1e-300 (row maxima ~4e-301: top / max overflows
# FP64) or 1e-316 (subnormal rows), or zeroed; T blocks per (i,j) tile pair by 1e100, 1e200 or
# 1e250, so every R block is a normal number (1e-217 .. 1e248) and |V||T| does not overflow.
TINY_V = (1.0, 1e-300, 1e-316, 0.0, 1.0)
HUGE_T = (1e100, 1e200, 1e250)


def _scale_blocks(tensor, data, modes, factors, mix):
    out = data.copy()
    fac = {}
    for k, off, sz in zip(tensor.keys, tensor.offsets, tensor.sizes):
        bid = tensor.blockid(int(k))
        t = tuple(bid[m] for m in modes)
        f = factors[(t[0] * mix[0] + t[1] * mix[1]) % len(factors)]
        fac[t] = f
        out[off:off + sz] *= f
    return out, fac


@pytest.mark.parametrize("wlf", [configs.tiny_spin, configs.tiny_irrep], ids=lambda f: f.__name__)
def test_tiny_and_subnormal_rows(eng, orc, wlf):
    wl = wlf()
    ow = orc.OracleWorkload(wl, SEED)
    term = next(t for t in wl.terms if t.name == "ladder")
    V, fv = _scale_blocks(ow.tensors["V"], ow.data["V"], (0, 1), TINY_V, (7, 3))
    T, ft = _scale_blocks(ow.tensors["T"], ow.data["T"], (2, 3), HUGE_T, (2, 1))
    assert {1e-300, 1e-316, 0.0} <= set(fv.values())
    sub = V[np.abs(V) > 0]
    assert np.any(np.abs(sub) < 2.2250738585072014e-308)  # subnormal inputs are present
    Rt = ow.tensors["R"]
    ow.data["V"], ow.data["T"] = np.abs(V), np.abs(T)
    mag = ow.run_term(term, out=Rt.zeros())
    ow.data["V"], ow.data["T"] = V, T
    ref = ow.run_term(term, out=Rt.zeros())
    ew = eng.EngineWorkload(wl, SEED, allocate=False)
    for name in wl.tensors:
        ew.tensors[name].upload(ow.data[name])
    for mode, tol in ((eng.EXEC_SECTOR_F64, 1e-12), (eng.EXEC_SPLIT_I8, 1e-10)):
        c = ew.contraction(term, mode)
        ew.tensors["R"].fill(0.0)
        c.execute()
        got = ew.tensors["R"].download()
        del c
        worst = {}
        for k, off, sz in zip(Rt.keys, Rt.offsets, Rt.sizes):
            a, b, i, j = Rt.blockid(int(k))
            g, r, w = got[off:off + sz], ref[off:off + sz], mag[off:off + sz]
            if fv.get((a, b)) == 0.0:
                assert not np.any(g)
                continue
            nw = _nrm(w)
            assert nw > 0.0
            err = _nrm(g - r) / nw
            f = fv.get((a, b))
            worst[f] = max(worst.get(f, 0.0), err)
            assert err <= tol, (wl.name, eng.EXEC_NAMES[mode], (a, b, i, j), f, ft.get((i, j)), err)
        assert {1e-300, 1e-316} <= set(worst)
        print(f"{wl.name} {eng.EXEC_NAMES[mode]}: worst per-block err / |V||T| by V factor {worst}")
