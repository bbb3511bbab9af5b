"""Workload definitions: index spaces, block-sparse tensors and contraction terms.

These describe the synthetic coupled-cluster-style workloads the engine is
measured and checked on.  A workload is pure data (no device work); the same
definition builds the engine's tensors and the CPU oracle's tensors.

Conventions follow TAMM:
  * an index space is a range of orbitals with Spin attribute ranges
    (alpha = 1, beta = 2, as in tests/tamm/Test_CCSD.cpp:646-648) and, as an
    extension, C2v irrep ranges (irreps 0..3; a block is nonzero only when the
    XOR of its irreps is 0);
  * tiling follows TiledIndexSpace{is, tile} / TiledIndexSpace{is, tiles}
    (src/tamm/tiled_index_space.hpp:1135-1260);
  * 4-index tensors carry the {2,2} spin mask of
    Tensor<T>{{tis...}, {2, 2}} (src/tamm/tensor.hpp:115-134).

Synthetic values come from the counter-based generator in include/tmx_rng.h,
keyed by (seed, tensor id, dense element index).
"""
from __future__ import annotations

from dataclasses import dataclass, field, replace
from typing import Dict, List, Optional, Sequence, Tuple

DEFAULT_SEED = 20261202
MASK64 = (1 << 64) - 1


def _mix64(x: int) -> int:
    """splitmix64 finaliser, identical to tmx_rng_mix64 in include/tmx_rng.h."""
    x = (x + 0x9E3779B97F4A7C15) & MASK64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & MASK64
    return x ^ (x >> 31)


def _key(seed: int, tensor_id: int) -> int:
    return _mix64(seed ^ _mix64((tensor_id * 0xD1B54A32D192ED03 + 0x8CB92BA72F3D8DD7) & MASK64))


def rng_uniform(seed: int, stream: int, counter: int) -> float:
    """Uniform [0,1) deviate, identical to tmx_rng_uniform (53 random bits)."""
    h = _mix64((_key(seed, stream) + _mix64(counter)) & MASK64)
    return (h >> 11) * 2.0 ** -53


@dataclass
class SpaceDef:
    """An IndexSpace plus its tiling."""
    name: str
    n: int
    tiles: object  # int (one tile size) or list of tile sizes
    spin: List[Tuple[int, int, int]] = field(default_factory=list)   # (lo, hi, spin)
    irrep: List[Tuple[int, int, int]] = field(default_factory=list)  # (lo, hi, irrep)
    named: List[Tuple[int, int]] = field(default_factory=list)       # subspace ranges

    @property
    def use_vector(self) -> bool:
        return not isinstance(self.tiles, int)

    @property
    def tile_list(self) -> List[int]:
        return [self.tiles] if isinstance(self.tiles, int) else list(self.tiles)


@dataclass
class TensorDef:
    name: str
    modes: List[str]                 # space names
    spin: Optional[Tuple[int, int]]  # (nupper, nlower) or None (no spin mask)
    irrep_rule: bool = False
    tensor_id: int = 0               # generator stream; 0 = not generated (zero-filled)
    scale: float = 1.0
    symmetrize: bool = False         # v(a,b,c,d) = 0.5*(z(a,b,c,d) + z(b,a,d,c))


@dataclass(frozen=True)
class TermDef:
    """C(clabels) += alpha * A(alabels) * B(blabels); with assign, C(clabels) = alpha * A * B
    (TAMM's `C(..) = A*B`: C's blocks are overwritten, zero where no block pair contributes)."""
    name: str
    c: Tuple[str, str]
    alpha: float
    a: Tuple[str, str]
    b: Tuple[str, str]
    assign: bool = False


@dataclass(frozen=True)
class SetDef:
    """name() = value (SetOp)."""
    name: str
    value: float


@dataclass(frozen=True)
class AddDef:
    """C(clabels) (=|+=) alpha * A(alabels) (AddOp, addop.hpp:242-271)."""
    name: str
    c: Tuple[str, str]
    alpha: float
    a: Tuple[str, str]
    assign: bool


@dataclass(frozen=True)
class DenomDef:
    """X(a,b,i,j) /= e_occ[i] + e_occ[j] - e_vir[a] - e_vir[b] (FP64, element-wise)."""
    name: str


@dataclass
class Workload:
    name: str
    description: str
    spaces: Dict[str, SpaceDef]
    tensors: Dict[str, TensorDef]
    terms: List[TermDef]
    outputs: List[str]
    extra: dict = field(default_factory=dict)

    def label_ids(self) -> Dict[str, int]:
        """Map label characters to int labels (distinct per character)."""
        chars = []
        for t in self.terms:
            for _, labs in (t.c, t.a, t.b):
                for ch in labs:
                    if ch not in chars:
                        chars.append(ch)
        return {ch: -(i + 1) for i, ch in enumerate(chars)}


def _spin_halves(n: int) -> List[Tuple[int, int, int]]:
    h = n // 2
    return [(0, h, 1), (h, n, 2)]


def _irrep_ranges(n: int) -> List[Tuple[int, int, int]]:
    """Round-robin C2v irreps within each spin half, orbitals sorted by irrep.

    Orbital p of a spin block of size s has irrep p mod 4; after sorting by irrep
    the block holds ceil((s - r)/4) orbitals of irrep r, consecutively.
    """
    out = []
    h = n // 2
    for base, size in ((0, h), (h, n - h)):
        lo = base
        for r in range(4):
            cnt = (size - r + 3) // 4 if size > r else 0
            if cnt:
                out.append((lo, lo + cnt, r))
                lo += cnt
    return out


def _drawn_tiles(ranges: Sequence[Tuple[int, int, int]], lo_t: int, hi_t: int, seed: int,
                 stream: int) -> List[int]:
    """Tile sizes drawn uniformly from [lo_t, hi_t] (seeded), truncated at range ends."""
    tiles, c = [], 0
    for lo, hi, _ in ranges:
        rem = hi - lo
        while rem > 0:
            t = lo_t + int(rng_uniform(seed, stream, c) * (hi_t - lo_t + 1))
            c += 1
            t = min(t, rem)
            tiles.append(t)
            rem -= t
    return tiles


def _cc_tensors(O: str, V: str, irrep: bool) -> Dict[str, TensorDef]:
    s = (2, 2)
    return {
        "V": TensorDef("V", [V, V, V, V], s, irrep, tensor_id=1, scale=0.1, symmetrize=True),
        "W": TensorDef("W", [O, V, V, O], s, irrep, tensor_id=2, scale=0.1),
        "T": TensorDef("T", [V, V, O, O], s, irrep, tensor_id=3, scale=0.01),
        "X": TensorDef("X", [O, O, O, O], s, irrep, tensor_id=4, scale=0.1),
        "R": TensorDef("R", [V, V, O, O], s, irrep, tensor_id=0),
    }


def adjoint_terms(term: TermDef, grad_c: str, grads: Dict[str, str]) -> List[TermDef]:
    """Adjoint contractions of C(c) += alpha * A(a) * B(b) for dL/dC = grad_c(c).

    dL/dA(a) += alpha * G(c) * B(b) and dL/dB(b) += alpha * G(c) * A(a): both are labeled
    block-sparse contractions of the same kind, with the output's labels taken from the
    operand being differentiated (its block structure, hence its sparsity, is the operand's).
    `grads` maps an operand tensor name to the name of its gradient tensor; operands without
    an entry are treated as constants.
    """
    out = []
    cn, cl = term.c
    an, al = term.a
    bn, bl = term.b
    if an in grads:
        out.append(TermDef(f"d{an}<-{term.name}", (grads[an], al), term.alpha, (grad_c, cl), (bn, bl)))
    if bn in grads:
        out.append(TermDef(f"d{bn}<-{term.name}", (grads[bn], bl), term.alpha, (grad_c, cl), (an, al)))
    return out


def assign_first(terms: Sequence[TermDef]) -> List[TermDef]:
    """The same terms with the first one into each output tensor assigning (C = ...) and the
    rest accumulating: equal to zeroing every output first, without the zero fill."""
    seen, out = set(), []
    for t in terms:
        out.append(replace(t, assign=t.c[0] not in seen))
        seen.add(t.c[0])
    return out


LADDER = TermDef("ladder", ("R", "abij"), 1.0, ("V", "abcd"), ("T", "cdij"))
RING = TermDef("ring", ("R", "abij"), 1.0, ("W", "kbcj"), ("T", "acik"))
HOLE_HOLE = TermDef("hh", ("R", "abij"), 1.0, ("X", "klij"), ("T", "abkl"))


def config0() -> Workload:
    """Dense O=20 (tile 10), V=80 (tile 20): ladder + permuted ring."""
    spaces = {"O": SpaceDef("O", 20, 10), "V": SpaceDef("V", 80, 20)}
    t = _cc_tensors("O", "V", False)
    for k in t:
        t[k].spin = None
    del t["X"]
    return Workload("cfg0", "dense O=20/t10, V=80/t20; ladder + ring", spaces, t,
                    [LADDER, RING], ["R"])


def spin_blocked(name: str, o: int, v: int, ot: int, vt: int, desc: str) -> Workload:
    spaces = {"O": SpaceDef("O", o, ot, spin=_spin_halves(o)),
              "V": SpaceDef("V", v, vt, spin=_spin_halves(v))}
    return Workload(name, desc, spaces, _cc_tensors("O", "V", False),
                    [LADDER, RING, HOLE_HOLE], ["R"])


def config1() -> Workload:
    """Spin-orbital O=40 (tile 16), V=240 (tile 48): ladder + ring + hole-hole (~10 GB)."""
    return spin_blocked("cfg1", 40, 240, 16, 48, "spin-blocked O=40/t16, V=240/t48; ladder+ring+hh")


def config2(seed: int = DEFAULT_SEED) -> Workload:
    """O=50, V=320 with spin and C2v irreps; tiles drawn from U{24..96}."""
    o, v = 50, 320
    orr, vrr = _irrep_ranges(o), _irrep_ranges(v)
    spaces = {
        "O": SpaceDef("O", o, _drawn_tiles(orr, 24, 96, seed, 101), spin=_spin_halves(o), irrep=orr),
        "V": SpaceDef("V", v, _drawn_tiles(vrr, 24, 96, seed, 102), spin=_spin_halves(v), irrep=vrr),
    }
    return Workload("cfg2", "spin+C2v O=50, V=320, tiles U{24..96}; ladder+ring+hh", spaces,
                    _cc_tensors("O", "V", True), [LADDER, RING, HOLE_HOLE], ["R"])


def with_loss_and_grads(w: Workload) -> Workload:
    """Adds Lambda (loss L = <Lambda, R>) and gradient tensors dV, dW, dX, dT to a
    ladder/ring/hole-hole workload; extra["adjoint_terms"] lists the adjoint contractions."""
    irrep = w.tensors["V"].irrep_rule
    w.tensors["L"] = TensorDef("L", ["V", "V", "O", "O"], (2, 2), irrep, tensor_id=5, scale=0.01)
    grads = {}
    for k in ("V", "W", "X", "T"):
        if k in w.tensors:
            src = w.tensors[k]
            grads[k] = "d" + k
            w.tensors["d" + k] = TensorDef("d" + k, list(src.modes), src.spin, src.irrep_rule)
    w.extra["loss"] = ("L", "R")
    w.extra["grads"] = grads
    w.extra["adjoint_terms"] = [a for t in w.terms for a in adjoint_terms(t, "L", grads)]
    return w


def config3() -> Workload:
    """O=30 (tile 16), V=180 (tile 36) spin-blocked; loss L = <Lambda, R> and its adjoints."""
    return with_loss_and_grads(
        spin_blocked("cfg3", 30, 180, 16, 36, "spin-blocked O=30/t16, V=180/t36; ladder+ring+hh + adjoints"))


def tiny_grad(irrep: bool = False) -> Workload:
    """tiny_spin / tiny_irrep with Lambda and gradient tensors."""
    return with_loss_and_grads(tiny_irrep() if irrep else tiny_spin("tiny_grad"))


def tiny_spin(name: str = "tiny") -> Workload:
    """O=4 (tile 2), V=8 (tile 3), spin-blocked: small enough for dense einsum checks."""
    return spin_blocked(name, 4, 8, 2, 3, "tiny spin-blocked O=4/t2, V=8/t3")


def tiny_irrep(seed: int = DEFAULT_SEED) -> Workload:
    """O=8, V=16 with spin and C2v irreps, tiles drawn from U{1..3}."""
    o, v = 8, 16
    orr, vrr = _irrep_ranges(o), _irrep_ranges(v)
    spaces = {
        "O": SpaceDef("O", o, _drawn_tiles(orr, 1, 3, seed, 101), spin=_spin_halves(o), irrep=orr),
        "V": SpaceDef("V", v, _drawn_tiles(vrr, 1, 3, seed, 102), spin=_spin_halves(v), irrep=vrr),
    }
    return Workload("tiny_irrep", "tiny spin+C2v O=8, V=16", spaces, _cc_tensors("O", "V", True),
                    [LADDER, RING, HOLE_HOLE], ["R"])


def scaled_spin(o: int, v: int, ot: int, vt: int) -> Workload:
    return spin_blocked(f"spin_o{o}_v{v}", o, v, ot, vt, f"spin-blocked O={o}/t{ot}, V={v}/t{vt}")


def _cc_step(name: str, o: int, v: int, tlo: int, thi: int, seed: int, desc: str) -> Workload:
    """CCSD-style doubles-residual step on spin + C2v blocked spaces with drawn tiles.

    Forward: dressed Fock intermediates (each reused by two terms)
        Fv(a,c) = fv(a,c) - 1/2 U(k,l,c,d) T(a,d,k,l),   Fo(k,i) = fo(k,i) + 1/2 U(k,l,c,d) T(c,d,i,l),
    ten residual terms (ladder, hole-hole, ring and its three P(ab)/P(ij) permutations, four
    Fock terms), then R /= D with D(a,b,i,j) = e_i + e_j - e_a - e_b.
    Adjoint pass for L = <Lambda, R>: G = Lambda / D, the adjoints of the ten residual terms
    (into dV, dW, dX, dT, dFv, dFo), then of the two intermediates (into dU, dT).
    """
    orr, vrr = _irrep_ranges(o), _irrep_ranges(v)
    spaces = {
        "O": SpaceDef("O", o, _drawn_tiles(orr, tlo, thi, seed, 101), spin=_spin_halves(o), irrep=orr),
        "V": SpaceDef("V", v, _drawn_tiles(vrr, tlo, thi, seed, 102), spin=_spin_halves(v), irrep=vrr),
    }
    t = _cc_tensors("O", "V", True)
    s2 = (1, 1)
    t["U"] = TensorDef("U", ["O", "O", "V", "V"], (2, 2), True, tensor_id=6, scale=0.1)
    t["fo"] = TensorDef("fo", ["O", "O"], s2, True, tensor_id=7, scale=0.05)
    t["fv"] = TensorDef("fv", ["V", "V"], s2, True, tensor_id=8, scale=0.05)
    t["Fo"] = TensorDef("Fo", ["O", "O"], s2, True)
    t["Fv"] = TensorDef("Fv", ["V", "V"], s2, True)
    t["L"] = TensorDef("L", ["V", "V", "O", "O"], (2, 2), True, tensor_id=5, scale=0.01)
    t["G"] = TensorDef("G", ["V", "V", "O", "O"], (2, 2), True)
    grads = {"V": "dV", "W": "dW", "X": "dX", "T": "dT", "U": "dU", "Fv": "dFv", "Fo": "dFo"}
    for k, g in grads.items():
        t[g] = TensorDef(g, list(t[k].modes), t[k].spin, True)
    fv_t = TermDef("Fv<-UT", ("Fv", "ac"), -0.5, ("U", "klcd"), ("T", "adkl"))
    fo_t = TermDef("Fo<-UT", ("Fo", "ki"), 0.5, ("U", "klcd"), ("T", "cdil"))
    inter = [AddDef("Fv=fv", ("Fv", "ac"), 1.0, ("fv", "ac"), True), fv_t,
             AddDef("Fo=fo", ("Fo", "ki"), 1.0, ("fo", "ki"), True), fo_t]
    resid = [
        LADDER,
        HOLE_HOLE,
        RING,
        TermDef("ring_Pab", ("R", "abij"), -1.0, ("W", "kacj"), ("T", "bcik")),
        TermDef("ring_Pij", ("R", "abij"), -1.0, ("W", "kbci"), ("T", "acjk")),
        TermDef("ring_PabPij", ("R", "abij"), 1.0, ("W", "kaci"), ("T", "bcjk")),
        TermDef("fock_v", ("R", "abij"), 1.0, ("Fv", "ac"), ("T", "cbij")),
        TermDef("fock_v_Pab", ("R", "abij"), 1.0, ("Fv", "bc"), ("T", "acij")),
        TermDef("fock_o", ("R", "abij"), -1.0, ("Fo", "ki"), ("T", "abkj")),
        TermDef("fock_o_Pij", ("R", "abij"), -1.0, ("Fo", "kj"), ("T", "abik")),
    ]
    # the first residual term assigns R (no zero fill), the others accumulate
    forward = inter + assign_first(resid) + [DenomDef("R")]
    adjoint = [AddDef("G=L", ("G", "abij"), 1.0, ("L", "abij"), True), DenomDef("G")]
    adj_terms = []
    for term in resid:
        adj_terms += adjoint_terms(term, "G", grads)
    adj_terms += adjoint_terms(fv_t, "dFv", {"U": "dU", "T": "dT"})
    adj_terms += adjoint_terms(fo_t, "dFo", {"U": "dU", "T": "dT"})
    # the first term into each gradient assigns it (TAMM's `dV(..) = G*T`), so no gradient needs
    # a zero fill except one no term writes
    adj_terms = assign_first(adj_terms)
    written = {t.c[0] for t in adj_terms}
    adjoint += [SetDef(g, 0.0) for g in grads.values() if g not in written]
    adjoint += adj_terms
    w = Workload(name, desc, spaces, t, resid, ["R"])
    w.extra.update(forward=forward, adjoint=adjoint, intermediates=inter, loss=("L", "R"),
                   grads=grads, e_occ=(-2.0, -0.5), e_vir=(0.2, 3.0), seed=seed)
    return w


def config4(seed: int = DEFAULT_SEED) -> Workload:
    """O=60, V=300 spin + C2v, tiles U{32..80}: ten-term CCSD-style step + one adjoint pass."""
    return _cc_step("cfg4", 60, 300, 32, 80, seed,
                    "spin+C2v O=60, V=300, tiles U{32..80}; 10-term CC step, denominator, adjoint")


def config4_mini(seed: int = DEFAULT_SEED) -> Workload:
    """The cfg4 step at O=12, V=24 (tiles U{1..4}), small enough for the CPU oracle."""
    return _cc_step("cfg4_mini", 12, 24, 1, 4, seed, "spin+C2v O=12, V=24, tiles U{1..4}; cfg4 step")


def orbital_energies(wl: Workload, seed: Optional[int] = None):
    """e_occ ~ U(-2,-0.5), e_vir ~ U(0.2,3) per orbital (seeded, streams 201/202)."""
    seed = wl.extra.get("seed", DEFAULT_SEED) if seed is None else seed
    lo, hi = wl.extra["e_occ"]
    eo = [lo + (hi - lo) * rng_uniform(seed, 201, p) for p in range(wl.spaces["O"].n)]
    lo, hi = wl.extra["e_vir"]
    ev = [lo + (hi - lo) * rng_uniform(seed, 202, p) for p in range(wl.spaces["V"].n)]
    return eo, ev


CONFIGS = {"cfg0": config0, "cfg1": config1, "cfg2": config2, "cfg3": config3, "cfg4": config4}
