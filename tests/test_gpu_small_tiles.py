"""64x64 GEMM tiles on products with few rows and columns over a long reduction.

Fo(k,i) = 1/2 U(k,l,c,d) T(c,d,i,l) and the adjoint-shaped dFo(k,j) = -L(a,b,i,j) T(a,b,i,k)
have O x O outputs and reductions over (l,c,d) or (a,b,i).  With O = 20 the cost model gives
them 64x64 tiles (a 128-row tile would stream the padding rows of both packed operands over
the whole reduction), and the launch model adds split-K parts.  Each is compared with the
oracle at the mode's tolerance (FP64 1e-12, split 1e-10 relative), in both sector modes,
next to Fv(a,c) = -1/2 U(k,l,c,d) T(a,d,k,l) (V = 100: 128x128 tiles) for contrast.
"""
import numpy as np
import pytest

from tamm_amd import configs
from tamm_amd.configs import TermDef

pytestmark = pytest.mark.gpu

SEED = 20261202

TERMS = {
    "Fo<-UT": TermDef("Fo<-UT", ("Fo", "ki"), 0.5, ("U", "klcd"), ("T", "cdil")),
    "Fv<-UT": TermDef("Fv<-UT", ("Fv", "ac"), -0.5, ("U", "klcd"), ("T", "adkl")),
    "dFo<-LT": TermDef("dFo<-LT", ("dFo", "kj"), -1.0, ("L", "abij"), ("T", "abik")),
    "dFo<-LT'": TermDef("dFo<-LT'", ("dFo", "ki"), -1.0, ("L", "abij"), ("T", "abkj")),
}


@pytest.fixture(scope="module")
def case(orc):
    wl = configs._cc_step("tall", 20, 100, 8, 24, SEED, "O=20, V=100: tall-skinny products")
    ow = orc.OracleWorkload(wl, SEED)
    ew_data = {name: ow.data[name] for name in ("U", "T", "L")}
    return wl, ow, ew_data


@pytest.mark.parametrize("name", sorted(TERMS))
@pytest.mark.parametrize("mode", ["sector_f64", "split_i8"])
def test_tall_skinny_matches_oracle(eng, case, name, mode):
    wl, ow, data = case
    term = TERMS[name]
    cn = term.c[0]
    ref = np.zeros(ow.tensors[cn].size)
    ow.run_term(term, out=ref)
    assert np.linalg.norm(ref) > 0
    m = {"sector_f64": eng.EXEC_SECTOR_F64, "split_i8": eng.EXEC_SPLIT_I8}[mode]
    tol = 1e-12 if mode == "sector_f64" else 1e-10
    ew = eng.EngineWorkload(wl, SEED)
    for t, d in data.items():
        ew.tensors[t].upload(d)
    c = ew.contraction(term, m)
    rounds, launches, tasks = c.schedule_info()
    for rep in range(2):
        ew.tensors[cn].fill(0.0)
        c.execute()
        got = ew.tensors[cn].download()
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        print(f"{name} {mode}: schedule {(rounds, launches, tasks)} rep {rep} err {err:.2e}")
        assert err <= tol, (name, mode, rep, err)
    del c
