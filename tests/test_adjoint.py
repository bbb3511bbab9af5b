"""Adjoint contractions on the CPU oracle.

The gradient of L = <Lambda, R> with R = sum of C += alpha*A*B terms is pinned three ways,
each independent of how the adjoint terms are written:
  * bilinearity: for one term, <dL/dA, A> = <dL/dB, B> = L (Euler's theorem for a form of
    degree one in each operand) -- a dropped term, wrong label or transposed operand breaks it;
  * dense numpy einsum of the derivative definition on tiny shapes, restricted to the
    operand's nonzero blocks (the engine's variables);
  * central finite differences of L along seeded random directions (L is quadratic in the
    joint perturbation, so the central difference is exact up to rounding).
"""
import numpy as np
import pytest

from tamm_amd import configs
from oracle import oracle as orc


def _workload(irrep):
    return configs.tiny_grad(irrep)


@pytest.mark.parametrize("irrep", [False, True])
def test_adjoint_bilinear_identity(irrep):
    wl = _workload(irrep)
    ow = orc.OracleWorkload(wl, configs.DEFAULT_SEED)
    grads = wl.extra["grads"]
    lam = ow.data["L"]
    for term in wl.terms:
        R = ow.tensors["R"].zeros()
        ow.run_term(term, out=R)
        L = float(np.dot(lam, R))
        assert abs(L) > 0
        for adj in configs.adjoint_terms(term, "L", grads):
            gname = adj.c[0]
            ow.data[gname][:] = 0.0
            ow.run_term(adj)
            src = gname[1:]
            val = float(np.dot(ow.data[gname], ow.data[src]))
            assert abs(val - L) <= 1e-12 * abs(L), (term.name, adj.name, val, L)


@pytest.mark.parametrize("irrep", [False, True])
def test_adjoint_dense_einsum(irrep):
    wl = _workload(irrep)
    ow = orc.OracleWorkload(wl, configs.DEFAULT_SEED)
    ow.gradients()
    dense = {k: ow.tensors[k].to_dense(ow.data[k]) for k in ow.data}
    expect = {g: np.zeros_like(dense[g]) for g in wl.extra["grads"].values()}
    for t in wl.terms:
        for name, labs in (t.a, t.b):
            if name not in wl.extra["grads"]:
                continue
            other = t.b if (name, labs) == t.a else t.a
            spec = f"{t.c[1]},{other[1]}->{labs}"
            expect[wl.extra["grads"][name]] += t.alpha * np.einsum(spec, dense["L"], dense[other[0]])
    for src, g in wl.extra["grads"].items():
        mask = ow.tensors[src].to_dense(np.ones(ow.tensors[src].zeros().shape)) != 0
        e = np.where(mask, expect[g], 0.0)
        err = np.linalg.norm(dense[g] - e) / np.linalg.norm(e)
        assert err < 1e-13, (g, err)


def test_adjoint_finite_differences():
    wl = _workload(False)
    ow = orc.OracleWorkload(wl, configs.DEFAULT_SEED)
    ow.gradients()
    grads = {k: ow.data[g].copy() for k, g in wl.extra["grads"].items()}
    base = {k: ow.data[k].copy() for k in grads}
    rng = np.random.default_rng(7)
    h = 1e-6
    for _ in range(16):
        D = {k: rng.standard_normal(base[k].shape) * np.abs(base[k]).max() for k in grads}
        for k in grads:
            ow.data[k][:] = base[k] + h * D[k]
        lp = ow.loss()
        for k in grads:
            ow.data[k][:] = base[k] - h * D[k]
        lm = ow.loss()
        fd = (lp - lm) / (2 * h)
        an = sum(float(np.dot(grads[k], D[k])) for k in grads)
        assert abs(fd - an) <= 1e-7 * max(abs(an), 1e-300), (fd, an)
    for k in grads:
        ow.data[k][:] = base[k]
