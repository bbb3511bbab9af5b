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

from dataclasses import dataclass, field
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


@dataclass
class TermDef:
    """C(clabels) += alpha * A(alabels) * B(blabels)."""
    name: str
    c: Tuple[str, str]
    alpha: float
    a: Tuple[str, str]
    b: Tuple[str, str]


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


def config3() -> Workload:
    """O=30 (tile 16), V=180 (tile 36) spin-blocked; loss L = <Lambda, R> and its adjoints."""
    w = spin_blocked("cfg3", 30, 180, 16, 36, "spin-blocked O=30/t16, V=180/t36; ladder+ring+hh + adjoints")
    w.tensors["L"] = TensorDef("L", ["V", "V", "O", "O"], (2, 2), False, tensor_id=5, scale=0.01)
    w.extra["loss"] = ("L", "R")
    return w


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


CONFIGS = {"cfg0": config0, "cfg1": config1, "cfg2": config2, "cfg3": config3}
