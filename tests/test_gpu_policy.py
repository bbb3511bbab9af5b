"""Which execution the AUTO modes resolve to, on the cfg4 step's terms (O=60, V=300, spin +
C2v, tiles 32..80), and that the FP64 sector path pads thin classes to 16 rows.

  * AUTO: the sector path for long reductions (ladder, rings, the one-body intermediates
    Fv/Fo <- U T), the grouped block-pair kernel for the Fock terms (reduction over one tile).
  * AUTO_SPLIT: split precision where the GEMMs are MFMA-bound (ladder, rings, the hole-hole
    adjoints, whose (i,j) classes gather many tile pairs); FP64 sector for thin products, whose
    classes are one O tile (Fo: 8 x 8 GEMMs) or one V tile (Fv: 37 x 37) so that the GEMM
    streams its operands, where split would pack 11 bytes per element against 8 for no gain.
"""
import pytest

from tamm_amd import configs

pytestmark = pytest.mark.gpu

GROUPED, SECTOR, SPLIT, AUTO, AUTO_SPLIT = 1, 2, 3, 4, 5


@pytest.fixture(scope="module")
def cfg4(eng):
    wl = configs.config4()
    ew = eng.EngineWorkload(wl, configs.DEFAULT_SEED, allocate=False)
    prog = wl.extra["forward"] + wl.extra["adjoint"]
    terms = {op.name: op for op in prog if isinstance(op, configs.TermDef)}
    return ew, terms


@pytest.mark.parametrize("name,auto,auto_split", [
    ("ladder", SECTOR, SPLIT),
    ("ring_Pab", SECTOR, SPLIT),
    ("dT<-ladder", SECTOR, SPLIT),
    ("Fo<-UT", SECTOR, SECTOR),
    ("Fv<-UT", SECTOR, SECTOR),
    ("dFo<-fock_o", SECTOR, SECTOR),
    ("dFv<-fock_v", SECTOR, SECTOR),
    ("dX<-hh", SECTOR, SPLIT),
    ("hh", SECTOR, SPLIT),
    ("fock_o", GROUPED, GROUPED),
    ("fock_v", GROUPED, GROUPED),
])
def test_auto_policy(cfg4, name, auto, auto_split):
    ew, terms = cfg4
    for mode, want in ((AUTO, auto), (AUTO_SPLIT, auto_split)):
        c = ew.contraction(terms[name], mode)
        assert c.mode == want, (name, mode, c.mode)
        del c


def test_thin_fp64_classes_pad_to_16_rows(cfg4):
    """Fo(k,i) += U(k,l,c,d) T(c,d,i,l): U's and T's classes are one O tile of 7-8 rows; padded
    to 16 their class matrices hold at most twice U's and T's elements (32 rows held four
    times as many)."""
    ew, terms = cfg4
    c = ew.contraction(terms["Fo<-UT"], SECTOR)
    operand_bytes = 8 * (ew.tensors["U"].size + ew.tensors["T"].size)
    assert c.workspace < 2.2 * operand_bytes, (c.workspace, operand_bytes)
