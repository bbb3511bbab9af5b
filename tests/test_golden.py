"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py and
cross-checked there against dense numpy einsum).

CPU: the oracle and the engine's block structures reproduce the fixture layouts bit for bit,
the oracle reproduces the generated inputs bit for bit and the outputs / gradients to
1e-14.  GPU: the engine's generator reproduces the inputs bit for bit and its outputs and
gradients match the fixtures to 1e-12 (FP64 modes) or 1e-10 (split precision).
"""
import os

import numpy as np
import pytest

from tamm_amd import configs
from tamm_amd import engine as E
from oracle import oracle as orc

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = [("tiny_spin", False), ("tiny_irrep", True)]


def _load(name):
    with np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("name,irrep", CASES)
def test_oracle_matches_golden(name, irrep):
    g = _load(name)
    wl = configs.tiny_grad(irrep)
    ow = orc.OracleWorkload(wl, configs.DEFAULT_SEED)
    for k, t in ow.tensors.items():
        assert np.array_equal(t.keys, g[f"{k}.keys"]), k
        assert np.array_equal(t.offsets, g[f"{k}.offsets"]), k
        assert np.array_equal(t.sizes, g[f"{k}.sizes"]), k
    for k in ("V", "W", "X", "T", "L"):
        assert np.array_equal(ow.data[k], g[f"in.{k}"]), k
    for term in wl.terms:
        R = ow.tensors["R"].zeros()
        ow.run_term(term, out=R)
        assert _rel(R, g[f"out.R.{term.name}"]) < 1e-14, term.name
    ow.run(wl.terms, outputs=["R"])
    assert _rel(ow.data["R"], g["out.R.all"]) < 1e-14
    ow.gradients()
    for gname in wl.extra["grads"].values():
        assert _rel(ow.data[gname], g[f"grad.{gname}"]) < 1e-14, gname


@pytest.mark.parametrize("name,irrep", CASES)
def test_engine_structure_matches_golden(name, irrep):
    g = _load(name)
    wl = configs.tiny_grad(irrep)
    ew = E.EngineWorkload(wl, configs.DEFAULT_SEED, allocate=False)
    for k, t in ew.tensors.items():
        assert np.array_equal(t.keys, g[f"{k}.keys"]), k
        assert np.array_equal(t.offsets, g[f"{k}.offsets"]), k
        assert np.array_equal(t.sizes, g[f"{k}.sizes"]), k


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [E.EXEC_GROUPED_F64, E.EXEC_SECTOR_F64, E.EXEC_SPLIT_I8])
@pytest.mark.parametrize("name,irrep", CASES)
def test_engine_matches_golden(eng, name, irrep, mode):
    g = _load(name)
    wl = configs.tiny_grad(irrep)
    ew = E.EngineWorkload(wl, configs.DEFAULT_SEED)
    for k in ("V", "W", "X", "T", "L"):
        assert np.array_equal(ew.tensors[k].download(), g[f"in.{k}"]), k
    tol = 1e-10 if mode == E.EXEC_SPLIT_I8 else 1e-12
    for term in wl.terms:
        ew.run([term], mode, outputs=["R"])
        assert _rel(ew.tensors["R"].download(), g[f"out.R.{term.name}"]) <= tol, term.name
    ew.run(wl.terms, mode, outputs=["R"])
    assert _rel(ew.tensors["R"].download(), g["out.R.all"]) <= tol
    ew.gradients(mode)
    for gname in wl.extra["grads"].values():
        assert _rel(ew.tensors[gname].download(), g[f"grad.{gname}"]) <= tol, gname
