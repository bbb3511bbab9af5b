"""The grouped and per-pair block-pair kernels on one-body terms, whose C blocks have one small
side (an O tile of 6-12 rows): the plan takes the 16x256 tile, and for the adjoints whose small
side is N (dT(a,b,k,j) += G(a,b,i,j) Fo(k,i)) builds the plan for C = B*A.

The CC step's terms at O=24, V=48 (spin + C2v, tiles drawn from 6..12) are executed in
GROUPED_F64 and BLOCKPAIR_F64 and compared with the FP64 sector path (itself checked against
the oracle on the mini step) at the FP64 bar, 1e-12 relative Frobenius, and with each other
bit for bit where the sum order is the same (grouped vs per-pair differ in order, so 1e-12).
"""
import numpy as np
import pytest

from tamm_amd import configs

pytestmark = pytest.mark.gpu

SEED = 20261202
NAMES = ["fock_o", "fock_o_Pij", "fock_v", "fock_v_Pab", "dT<-fock_o", "dT<-fock_o_Pij",
         "dT<-fock_v", "dU<-Fo<-UT", "dT<-Fo<-UT", "dU<-Fv<-UT", "dT<-Fv<-UT"]
GROUPED, BLOCKPAIR, SECTOR = 1, 0, 2


@pytest.fixture(scope="module")
def small_cc(eng):
    wl = configs._cc_step("cc_small", 24, 48, 6, 12, SEED, "O=24, V=48, tiles 6..12")
    ew = eng.EngineWorkload(wl, SEED)
    prog = wl.extra["forward"] + wl.extra["adjoint"]
    terms = {op.name: op for op in prog if isinstance(op, configs.TermDef)}
    # fill the operands every term reads (intermediates and G included) with seeded values
    for name, t in wl.tensors.items():
        ew.tensors[name].fill_normal(SEED, 100 + sorted(wl.tensors).index(name), 0.1, False)
    return wl, ew, terms


def _run(eng, ew, term, mode):
    C = ew.tensors[term.c[0]]
    C.fill(0.0)
    c = ew.contraction(term, mode)
    c.execute()
    eng.synchronize()
    return C.download()


@pytest.mark.parametrize("name", NAMES)
def test_one_body_terms_match_sector(eng, small_cc, name):
    wl, ew, terms = small_cc
    term = terms[name]
    ref = _run(eng, ew, term, SECTOR)
    assert np.linalg.norm(ref) > 0
    for mode in (GROUPED, BLOCKPAIR):
        got = _run(eng, ew, term, mode)
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        assert err <= 1e-12, (name, mode, err)
