"""Adjoint contractions on the GPU.

* tiny workloads: engine gradients against the oracle's (FP64 modes <= 1e-12 relative,
  split precision <= 1e-10);
* cfg3 (O=30/t16, V=180/t36, spin-blocked): the FP64 gradient of L = <Lambda, R> against
  central finite differences (step 1e-6) along 16 seeded random directions, tolerance 1e-6
  relative; split-precision gradients against the FP64 ones (<= 1e-10 per gradient tensor); and
  the split path at 24 bits per operand (FP32's significand: the gradient at FP32-level
  precision) against the same finite differences, along the same 16 directions, at 1e-5.
"""
import numpy as np
import pytest

from tamm_amd import configs
from tamm_amd import engine as E
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

FP64_MODES = [E.EXEC_BLOCKPAIR_F64, E.EXEC_GROUPED_F64, E.EXEC_SECTOR_F64, E.EXEC_AUTO]


@pytest.mark.parametrize("irrep", [False, True])
@pytest.mark.parametrize("mode", FP64_MODES + [E.EXEC_SPLIT_I8])
def test_tiny_gradients_vs_oracle(eng, irrep, mode):
    wl = configs.tiny_grad(irrep)
    ow = orc.OracleWorkload(wl, configs.DEFAULT_SEED)
    ow.gradients()
    ew = E.EngineWorkload(wl, configs.DEFAULT_SEED)
    ew.gradients(mode)
    E.synchronize()
    tol = 1e-10 if mode == E.EXEC_SPLIT_I8 else 1e-12
    for g in wl.extra["grads"].values():
        ref = ow.data[g]
        got = ew.tensors[g].download()
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        assert err <= tol, (g, E.EXEC_NAMES[mode], err)


@pytest.fixture(scope="module")
def cfg3(eng):
    wl = configs.config3()
    ew = E.EngineWorkload(wl, configs.DEFAULT_SEED)
    E.synchronize()
    yield wl, ew
    ew.clear_plans()
    del ew


def _fd_check(wl, ew, grad_mode, tol, ndir=16, h=1e-6):
    """Gradient (in grad_mode) . D against (L(x + hD) - L(x - hD)) / 2h in FP64 (sector)."""
    grads = wl.extra["grads"]
    ew.gradients(grad_mode)
    base, D = {}, {}
    for k in grads:
        src = ew.tensors[k]
        base[k] = E.Tensor(src.spaces, wl.tensors[k].spin, wl.tensors[k].irrep_rule).allocate()
        base[k].copy_from(src)
        D[k] = E.Tensor(src.spaces, wl.tensors[k].spin, wl.tensors[k].irrep_rule).allocate()
    worst = 0.0
    try:
        for d in range(ndir):
            an = 0.0
            for n, k in enumerate(grads):
                D[k].fill_normal(configs.DEFAULT_SEED + d, 100 + n, wl.tensors[k].scale)
                an += ew.tensors[grads[k]].dot(D[k])
            for sign in (1.0, -1.0):
                for k in grads:
                    ew.tensors[k].copy_from(base[k])
                    ew.tensors[k].axpy(sign * h, D[k])
                if sign > 0:
                    lp = ew.loss(E.EXEC_SECTOR_F64)
                else:
                    lm = ew.loss(E.EXEC_SECTOR_F64)
            fd = (lp - lm) / (2 * h)
            rel = abs(fd - an) / abs(an)
            worst = max(worst, rel)
            assert rel <= tol, (d, fd, an, rel)
    finally:
        for k in grads:
            ew.tensors[k].copy_from(base[k])
        E.synchronize()
    print(f"finite differences ({E.EXEC_NAMES[grad_mode]} gradient): worst rel err {worst:.3e}")
    return worst


def test_cfg3_fd_fp64(cfg3):
    wl, ew = cfg3
    _fd_check(wl, ew, E.EXEC_SECTOR_F64, 1e-6)


def test_cfg3_split_gradients_vs_fp64(cfg3):
    wl, ew = cfg3
    grads = wl.extra["grads"]
    ew.gradients(E.EXEC_SECTOR_F64)
    keep = {}
    for k, g in grads.items():
        t = ew.tensors[g]
        keep[g] = E.Tensor(ew.tensors[k].spaces, wl.tensors[k].spin, wl.tensors[k].irrep_rule).allocate()
        keep[g].copy_from(t)
    ew.gradients(E.EXEC_SPLIT_I8)
    E.synchronize()
    for g in grads.values():
        err = E.rel_diff(ew.tensors[g].ptr, keep[g].ptr, keep[g].size)
        print(f"split gradient {g}: rel err {err:.3e}")
        assert err <= 1e-10, (g, err)


def test_cfg3_fd_split24(cfg3):
    wl, ew = cfg3
    old = 35
    E.check(E.lib().tmx_set_split_bits(24))
    try:
        ew.clear_plans()
        _fd_check(wl, ew, E.EXEC_SPLIT_I8, 1e-5, ndir=16)
    finally:
        E.check(E.lib().tmx_set_split_bits(old))
        ew.clear_plans()
