"""Split precision over operands whose rows span a wide range, including zero rows, on the GPU.

The ladder R(a,b,i,j) += V(a,b,c,d) T(c,d,i,j) packs V's rows (a,b) and T's rows (i,j) with one
scale per row.  V's blocks are multiplied by a factor per (a,b) tile pair (0, 1e-100, 1e100 or
1) and T's by a factor per (i,j) tile pair (1e-100, 1e80 or 1), so every R block is its
unscaled value times one product of factors, from 1e-200 to 1e180.  Errors are therefore taken
per R block, against that block of |V| |T| (the oracle's product of the absolute operands, the
scale of a dot product's rounding error whatever cancels inside it): FP64 sector to 1e-12, split
to 1e-10.  R blocks whose V rows are all zero must be exactly zero in both modes.
"""
import numpy as np
import pytest

from tamm_amd import configs

pytestmark = pytest.mark.gpu

SEED = 20261202
V_FACTORS = (0.0, 1e-100, 1e100, 1.0, 1.0)
T_FACTORS = (1e-100, 1e80, 1.0, 1.0)


def _nrm(x):
    """Frobenius norm without overflow (entries reach 1e180)."""
    m = float(np.abs(x).max()) if x.size else 0.0
    return m * float(np.linalg.norm(x / m)) if m > 0 else 0.0


def _scale_blocks(tensor, data, modes, factors, mix):
    """Multiplies each block by factors[(mix[0]*t0 + mix[1]*t1) % len(factors)], (t0, t1) the
    block's tile ids at `modes`; returns the factor of every tile pair seen."""
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
def test_ladder_rows_over_400_decades(eng, orc, wlf):
    wl = wlf()
    ow = orc.OracleWorkload(wl, SEED)
    term = next(t for t in wl.terms if t.name == "ladder")
    assert term.a[0] == "V" and term.b[0] == "T"
    V, fv = _scale_blocks(ow.tensors["V"], ow.data["V"], (0, 1), V_FACTORS, (7, 3))
    T, ft = _scale_blocks(ow.tensors["T"], ow.data["T"], (2, 3), T_FACTORS, (2, 1))
    assert 0.0 in fv.values() and 1e100 in fv.values() and 1e-100 in fv.values()
    assert 1e80 in ft.values() and 1e-100 in ft.values()
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
        nzero = 0
        worst = 0.0
        for k, off, sz in zip(Rt.keys, Rt.offsets, Rt.sizes):
            a, b, i, j = Rt.blockid(int(k))
            g, r, w = got[off:off + sz], ref[off:off + sz], mag[off:off + sz]
            if fv.get((a, b)) == 0.0:
                nzero += 1
                assert not np.any(g), (wl.name, eng.EXEC_NAMES[mode], a, b, i, j)
                continue
            nw = _nrm(w)
            if nw == 0.0:
                assert not np.any(g)
                continue
            err = _nrm(g - r) / nw
            worst = max(worst, err)
            assert err <= tol, (wl.name, eng.EXEC_NAMES[mode], (a, b, i, j), fv.get((a, b)),
                                ft.get((i, j)), err)
        assert nzero > 0
        print(f"{wl.name} ladder {eng.EXEC_NAMES[mode]}: worst per-block err / |V||T| {worst:.3e}, "
              f"{nzero} all-zero blocks")
