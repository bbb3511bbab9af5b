"""CPU-side checks of the engine library: it loads, exports every symbol its header
declares, returns error codes (no exceptions) at the C boundary, and computes block
structure (tiling, sparsity, Distribution_NW layout) identical bit for bit to the oracle.
No compute call is made here (there is no GPU)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from tamm_amd import configs, engine

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_in_header():
    txt = open(os.path.join(ROOT, "include", "tamm_mi355x.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(tmx_[a-z0-9_]+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    L = engine.lib()
    names = _declared_in_header()
    assert len(names) >= 45
    for n in names:
        assert hasattr(L, n), n
    assert set(engine.declared_symbols()) == set(names)


def test_calls_before_init_return_state_error():
    L = engine.lib()
    if engine.initialised():
        pytest.skip("engine already initialised in this process")
    p = C.c_void_p()
    assert L.tmx_malloc(64, C.byref(p)) == 5
    assert b"tmx_init" in L.tmx_last_error()
    assert L.tmx_fill(None, 0, 0.0, None) == 5


def test_invalid_descriptors_return_invalid():
    with pytest.raises(engine.EngineError) as e:
        engine.Space(14, [5, 6, 3], named=[(0, 10), (10, 14)])
    assert e.value.code == 1
    with pytest.raises(engine.EngineError):
        engine.Space(14, [5, 5, 2])
    s = engine.Space(10, 3)
    with pytest.raises(engine.EngineError):
        engine.Tensor([s] * 9)  # more than 8 modes


WORKLOADS = [configs.tiny_spin, configs.tiny_irrep, configs.config0, configs.config1,
             configs.config2, configs.config3, configs.config4_mini, configs.config4]


@pytest.mark.parametrize("wlf", WORKLOADS, ids=lambda f: f.__name__)
def test_block_structure_matches_oracle_bitwise(orc, wlf):
    wl = wlf()
    for name, sd in wl.spaces.items():
        a = engine.Space(sd.n, sd.tiles, sd.spin, sd.irrep, sd.named)
        b = orc.Space(sd.n, sd.tiles, sd.spin, sd.irrep, sd.named)
        assert np.array_equal(a.offsets, b.offsets)
        assert np.array_equal(a.spin, b.spin) and np.array_equal(a.irrep, b.irrep)
    for name, td in wl.tensors.items():
        sp_e = [engine.Space(wl.spaces[m].n, wl.spaces[m].tiles, wl.spaces[m].spin,
                             wl.spaces[m].irrep) for m in td.modes]
        sp_o = [orc.Space(wl.spaces[m].n, wl.spaces[m].tiles, wl.spaces[m].spin,
                          wl.spaces[m].irrep) for m in td.modes]
        te = engine.Tensor(sp_e, td.spin, td.irrep_rule)
        to = orc.Tensor(sp_o, td.spin, td.irrep_rule)
        assert np.array_equal(te.keys, to.keys), name
        assert np.array_equal(te.offsets, to.offsets), name
        assert np.array_equal(te.sizes, to.sizes), name
        assert te.size == to.size


def test_locate_matches_layout():
    wl = configs.tiny_irrep()
    sp = {k: engine.Space(s.n, s.tiles, s.spin, s.irrep) for k, s in wl.spaces.items()}
    td = wl.tensors["V"]
    t = engine.Tensor([sp[m] for m in td.modes], td.spin, td.irrep_rule)
    nt = [sp[m].ntiles for m in td.modes]
    stored = dict(zip(t.keys.tolist(), t.offsets.tolist()))
    for key in range(int(np.prod(nt))):
        bid, k = [], key
        for i in range(4):
            div = int(np.prod(nt[i + 1:]))
            bid.append(k // div)
            k %= div
        assert t.locate(bid) == stored.get(key, -1)


def test_workload_flop_counts():
    """Algorithmic FLOPs of the configs (closed forms from the config definitions)."""
    # cfg1 ladder: 2 * (sum over spin sectors of M*N*K)
    o, v = 40, 240
    ladder = 2 * (2 * (v // 2) ** 4 * (o // 2) ** 2 + (2 * (v // 2) ** 2) ** 2 * (2 * (o // 2) ** 2))
    assert ladder == 1658880000000
