"""Writes the golden fixtures tests/golden/tiny_spin.npz and tiny_irrep.npz.

Each fixture holds, for a tiny spin-blocked (O=4/t2, V=8/t3) or spin + C2v (O=8, V=16,
tiles U{1..3}) workload at seed 20261202: the block lists (keys, offsets, sizes) of every
tensor, the generated inputs (V, W, X, T, Lambda), the output R of each term alone and of all
three terms, and the gradients dL/dV, dW, dX, dT of L = <Lambda, R>.  Values come from the
CPU oracle and are cross-checked against dense numpy einsum before they are written (the
script refuses to write on a mismatch).  Run from the repository root:
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from tamm_amd import configs  # noqa: E402
from oracle import oracle as orc  # noqa: E402


def einsum_check(ow, wl, name, got, spec_terms):
    dense = {k: ow.tensors[k].to_dense(ow.data[k]) for k in ow.data}
    exp = 0
    for alpha, spec, a, b in spec_terms:
        exp = exp + alpha * np.einsum(spec, dense[a], dense[b])
    t = ow.tensors[name]
    mask = t.to_dense(np.ones(t.zeros().shape)) != 0
    exp = np.where(mask, exp, 0.0)
    err = np.linalg.norm(t.to_dense(got) - exp) / np.linalg.norm(exp)
    if not err < 1e-13:
        raise SystemExit(f"einsum cross-check failed for {name}: {err}")


def make(wl, path):
    ow = orc.OracleWorkload(wl, configs.DEFAULT_SEED)
    out = {}
    for k, t in ow.tensors.items():
        out[f"{k}.keys"] = t.keys
        out[f"{k}.offsets"] = t.offsets
        out[f"{k}.sizes"] = t.sizes
    for k in ("V", "W", "X", "T", "L"):
        out[f"in.{k}"] = ow.data[k].copy()
    for term in wl.terms:
        R = ow.tensors["R"].zeros()
        ow.run_term(term, out=R)
        spec = f"{term.a[1]},{term.b[1]}->{term.c[1]}"
        einsum_check(ow, wl, "R", R, [(term.alpha, spec, term.a[0], term.b[0])])
        out[f"out.R.{term.name}"] = R
    ow.run(wl.terms, outputs=["R"])
    out["out.R.all"] = ow.data["R"].copy()
    ow.gradients()
    for src, g in wl.extra["grads"].items():
        specs = []
        for term in wl.terms:
            for (n, labs), other in ((term.a, term.b), (term.b, term.a)):
                if n == src:
                    specs.append((term.alpha, f"{term.c[1]},{other[1]}->{labs}", "L", other[0]))
        einsum_check(ow, wl, g, ow.data[g], specs)
        out[f"grad.{g}"] = ow.data[g].copy()
    np.savez_compressed(path, **out)
    print("wrote", path, sum(v.size for v in out.values()), "values")


if __name__ == "__main__":
    here = os.path.dirname(os.path.abspath(__file__))
    make(configs.tiny_grad(False), os.path.join(here, "tiny_spin.npz"))
    make(configs.tiny_grad(True), os.path.join(here, "tiny_irrep.npz"))
