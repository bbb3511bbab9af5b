"""CCSD-style step (cfg4 program: two intermediates, ten residual terms, denominator, one
adjoint pass) on the GPU.

* cfg4_mini (O=12, V=24): engine against the oracle, per output tensor (FP64 modes <= 1e-12
  relative, split precision <= 1e-10);
* hipGraph capture of the whole step: three replays, with eager engine work between them,
  each bitwise equal to the eager step;
* recomputed intermediates give bitwise the same gradients as stored ones;
* cfg4 (O=60, V=300, spin + C2v, tiles U{32..80}): split precision against plain FP64 on the
  same GPU, per output tensor <= 1e-10, and the captured split step equal to the eager one.
"""
import numpy as np
import pytest

from tamm_amd import configs
from tamm_amd import engine as E
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

OUTS = ["R", "dT", "dV", "dW", "dX", "dU", "dFv", "dFo"]


def _snapshot(wl, ew, names=OUTS):
    out = {}
    for n in names:
        t = ew.tensors[n]
        c = E.Tensor(t.spaces, wl.tensors[n].spin, wl.tensors[n].irrep_rule).allocate()
        c.copy_from(t)
        out[n] = c
    E.synchronize()
    return out


@pytest.fixture(scope="module")
def mini_ref():
    wl = configs.config4_mini()
    ow = orc.OracleWorkload(wl, configs.DEFAULT_SEED)
    ow.execute_program(wl.extra["forward"])
    ow.execute_program(wl.extra["adjoint"])
    return wl, ow


@pytest.mark.parametrize("mode", [E.EXEC_GROUPED_F64, E.EXEC_SECTOR_F64, E.EXEC_AUTO, E.EXEC_SPLIT_I8,
                                  E.EXEC_AUTO_SPLIT])
def test_mini_step_vs_oracle(eng, mini_ref, mode):
    wl, ow = mini_ref
    ew = E.EngineWorkload(wl, configs.DEFAULT_SEED)
    ew.cc_step(mode)
    E.synchronize()
    tol = 1e-10 if mode in (E.EXEC_SPLIT_I8, E.EXEC_AUTO_SPLIT) else 1e-12
    for n in OUTS:
        ref = ow.data[n]
        got = ew.tensors[n].download()
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        assert err <= tol, (n, E.EXEC_NAMES[mode], err)


@pytest.mark.parametrize("mode", [E.EXEC_SECTOR_F64, E.EXEC_SPLIT_I8, E.EXEC_AUTO_SPLIT])
def test_mini_step_graph_replay(eng, mode):
    wl = configs.config4_mini()
    ew = E.EngineWorkload(wl, configs.DEFAULT_SEED)
    ew.cc_step(mode)  # builds and caches every plan; warms the block tables
    E.synchronize()
    eager = _snapshot(wl, ew)
    prog = wl.extra["forward"] + wl.extra["adjoint"]
    g = E.Graph(lambda: ew.execute_program(prog, mode))
    scratch = ew.tensors["G"]
    for rep in range(3):
        for n in OUTS:
            ew.tensors[n].fill(float("nan"))
        g.launch()
        E.synchronize()
        for n in OUTS:
            d = E.rel_diff(ew.tensors[n].ptr, eager[n].ptr, eager[n].size)
            assert d == 0.0, (rep, n, d)
        scratch.fill(1.0)  # eager work between replays
        E.synchronize()


def test_mini_step_recompute_intermediates(eng):
    wl = configs.config4_mini()
    ew = E.EngineWorkload(wl, configs.DEFAULT_SEED)
    ew.cc_step(E.EXEC_SPLIT_I8)
    stored = _snapshot(wl, ew)
    ew.cc_step(E.EXEC_SPLIT_I8, recompute=True)
    E.synchronize()
    for n in OUTS:
        assert E.rel_diff(ew.tensors[n].ptr, stored[n].ptr, stored[n].size) == 0.0, n


@pytest.mark.parametrize("mode", [E.EXEC_SPLIT_I8, E.EXEC_AUTO_SPLIT])
def test_cfg4_split_vs_fp64_and_graph(eng, mode):
    wl = configs.config4()
    ew = E.EngineWorkload(wl, configs.DEFAULT_SEED)
    ew.cc_step(E.EXEC_SECTOR_F64)
    ref = _snapshot(wl, ew)
    ew.clear_plans()
    ew.cc_step(mode)
    E.synchronize()
    for n in OUTS:
        err = E.rel_diff(ew.tensors[n].ptr, ref[n].ptr, ref[n].size)
        print(f"cfg4 split vs FP64 {n}: {err:.3e}")
        assert err <= 1e-10, (n, err)
    eager = _snapshot(wl, ew)
    prog = wl.extra["forward"] + wl.extra["adjoint"]
    g = E.Graph(lambda: ew.execute_program(prog, mode))
    for n in OUTS:
        ew.tensors[n].fill(float("nan"))
    g.launch()
    E.synchronize()
    for n in OUTS:
        assert E.rel_diff(ew.tensors[n].ptr, eager[n].ptr, eager[n].size) == 0.0, n
    del g
    ew.clear_plans()
