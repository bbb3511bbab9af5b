"""ctypes binding of libtamm_mi355x.so and a thin host-side object layer.

The engine itself is the C ABI declared in include/tamm_mi355x.h; this module only
declares prototypes, owns handles and moves numpy arrays to and from HBM.  There is no
compute fallback: if the library is missing or the GPU is absent the calls fail.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, List, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtamm_mi355x.so")

i64p = C.POINTER(C.c_int64)
i32p = C.POINTER(C.c_int)
f64p = C.c_void_p  # device pointers are passed as integers
vp = C.c_void_p

EXEC_BLOCKPAIR_F64, EXEC_GROUPED_F64, EXEC_SECTOR_F64, EXEC_SPLIT_I8, EXEC_AUTO, EXEC_AUTO_SPLIT = range(6)
EXEC_NAMES = {EXEC_BLOCKPAIR_F64: "blockpair_f64", EXEC_GROUPED_F64: "grouped_f64",
              EXEC_SECTOR_F64: "sector_f64", EXEC_SPLIT_I8: "split_i8", EXEC_AUTO: "auto",
              EXEC_AUTO_SPLIT: "auto_split"}
H2D, D2H, D2D = 0, 1, 2

ERRORS = {1: "TMX_ERR_INVALID", 2: "TMX_ERR_HIP", 3: "TMX_ERR_OOM", 4: "TMX_ERR_UNSUPPORTED",
          5: "TMX_ERR_STATE"}


class EngineError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class _Operand(C.Structure):
    _fields_ = [("tensor", vp), ("labels", i32p), ("tile_offset", i64p), ("label_ntiles", i64p)]


# (name, restype, argtypes)
_PROTOS = [
    ("tmx_last_error", C.c_char_p, []),
    ("tmx_version", C.c_char_p, []),
    ("tmx_init", C.c_int, [C.c_int, C.c_double]),
    ("tmx_finalize", C.c_int, []),
    ("tmx_get_stream", C.c_int, [C.POINTER(vp)]),
    ("tmx_synchronize", C.c_int, []),
    ("tmx_malloc", C.c_int, [C.c_size_t, C.POINTER(vp)]),
    ("tmx_free", C.c_int, [vp]),
    ("tmx_pool_info", C.c_int, [C.POINTER(C.c_size_t)] * 3),
    ("tmx_memcpy", C.c_int, [vp, vp, C.c_size_t, C.c_int, vp]),
    ("tmx_memset_zero", C.c_int, [vp, C.c_size_t, vp]),
    ("tmx_space_create", C.c_int, [C.c_int64, C.c_int, i64p, C.c_int64, C.c_int, i64p, i64p, i32p,
                                   C.c_int, i64p, i64p, i32p, C.c_int, i64p, i64p, C.POINTER(vp)]),
    ("tmx_space_destroy", C.c_int, [vp]),
    ("tmx_space_ntiles", C.c_int, [vp, i64p]),
    ("tmx_space_tiles", C.c_int, [vp, i64p, i32p, i32p]),
    ("tmx_tensor_create", C.c_int, [C.c_int, C.POINTER(vp), C.c_int, C.c_int, C.c_int, C.POINTER(vp)]),
    ("tmx_tensor_destroy", C.c_int, [vp]),
    ("tmx_tensor_size", C.c_int, [vp, i64p]),
    ("tmx_tensor_nblocks", C.c_int, [vp, i64p]),
    ("tmx_tensor_blocks", C.c_int, [vp, i64p, i64p, i64p]),
    ("tmx_tensor_locate", C.c_int, [vp, i64p, i64p, i64p]),
    ("tmx_contraction_create", C.c_int, [C.POINTER(_Operand), C.c_double, C.POINTER(_Operand),
                                         C.POINTER(_Operand), C.c_int, C.POINTER(vp)]),
    ("tmx_contraction_execute", C.c_int, [vp, vp, vp, vp, vp]),
    ("tmx_contraction_info", C.c_int, [vp, C.POINTER(C.c_double), i64p, i64p, i32p]),
    ("tmx_contraction_split_info", C.c_int, [vp, i32p, i32p, i32p]),
    ("tmx_contraction_destroy", C.c_int, [vp]),
    ("tmx_contraction_set_constant", C.c_int, [vp, C.c_int, C.c_int]),
    ("tmx_contraction_set_assign", C.c_int, [vp, C.c_int]),
    ("tmx_level_create", C.c_int, [C.POINTER(vp), C.c_int, C.POINTER(vp)]),
    ("tmx_level_products", C.c_int, [vp, i32p]),
    ("tmx_level_execute", C.c_int, [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), vp]),
    ("tmx_level_destroy", C.c_int, [vp]),
    ("tmx_contraction_execute_timed", C.c_int, [vp, vp, vp, vp, vp, C.POINTER(C.c_float)]),
    ("tmx_contraction_execute_stages", C.c_int, [vp, vp, vp, vp, vp, C.POINTER(C.c_float)]),
    ("tmx_contraction_execute_stages_n", C.c_int, [vp, vp, vp, vp, vp, C.POINTER(C.c_float), C.c_int]),
    ("tmx_set_split_bits", C.c_int, [C.c_int]),
    ("tmx_set_parallel_split", C.c_int, [C.c_int]),
    ("tmx_set_scale_reuse", C.c_int, [C.c_int]),
    ("tmx_contraction_schedule_info", C.c_int, [vp, i32p, i32p, i64p]),
    ("tmx_contraction_autotune", C.c_int, [vp, vp, vp, vp, vp, C.POINTER(C.c_int)]),
    ("tmx_tensor_write_hdf5", C.c_int, [vp, vp, C.c_char_p, vp]),
    ("tmx_tensor_read_hdf5", C.c_int, [vp, vp, C.c_char_p, vp]),
    ("tmx_block_multiply", C.c_int, [C.c_double, vp, C.c_int, i64p, i32p, vp, C.c_int, i64p, i32p,
                                     vp, C.c_int, i64p, i32p, C.c_int, vp]),
    ("tmx_block_assign", C.c_int, [vp, C.c_int, i64p, i32p, C.c_double, vp, i64p, i32p, C.c_int, vp]),
    ("tmx_addop_create", C.c_int, [C.POINTER(_Operand), C.c_double, C.POINTER(_Operand), C.c_int,
                                   C.POINTER(vp)]),
    ("tmx_addop_execute", C.c_int, [vp, vp, vp, vp]),
    ("tmx_addop_destroy", C.c_int, [vp]),
    ("tmx_fill", C.c_int, [vp, C.c_int64, C.c_double, vp]),
    ("tmx_scale", C.c_int, [vp, C.c_int64, C.c_double, vp]),
    ("tmx_axpy", C.c_int, [C.c_int64, C.c_double, vp, vp, vp]),
    ("tmx_fill_normal", C.c_int, [vp, vp, C.c_uint64, C.c_uint64, C.c_double, C.c_int, vp]),
    ("tmx_norm", C.c_int, [vp, C.c_int64, C.POINTER(C.c_double), vp]),
    ("tmx_norm_noalloc", C.c_int, [vp, C.c_int64, C.POINTER(C.c_double), vp]),
    ("tmx_dot", C.c_int, [vp, vp, C.c_int64, C.POINTER(C.c_double), vp]),
    ("tmx_rel_diff", C.c_int, [vp, vp, C.c_int64, C.POINTER(C.c_double), vp]),
    ("tmx_energy_denominator", C.c_int, [vp, vp, vp, vp, vp]),
    ("tmx_event_create", C.c_int, [C.POINTER(vp)]),
    ("tmx_event_record", C.c_int, [vp, vp]),
    ("tmx_event_elapsed_ms", C.c_int, [vp, vp, C.POINTER(C.c_float)]),
    ("tmx_event_destroy", C.c_int, [vp]),
    ("tmx_stream_create", C.c_int, [C.POINTER(vp)]),
    ("tmx_stream_destroy", C.c_int, [vp]),
    ("tmx_stream_synchronize", C.c_int, [vp]),
    ("tmx_graph_begin", C.c_int, [vp]),
    ("tmx_graph_end", C.c_int, [vp, C.POINTER(vp)]),
    ("tmx_graph_launch", C.c_int, [vp, vp]),
    ("tmx_graph_destroy", C.c_int, [vp]),
]

_lib = None


def lib():
    """Load the engine library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run tamm_amd/build.py (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, res, args in _PROTOS:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def declared_symbols() -> List[str]:
    return [p[0] for p in _PROTOS]


def check(rc: int) -> None:
    if rc != 0:
        raise EngineError(rc, lib().tmx_last_error().decode())


def _i64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.int64))


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def _p(a, t):
    return a.ctypes.data_as(t)


_initialised = False


def init(device: int = 0, pool_fraction: float = 0.0) -> None:
    global _initialised
    if not _initialised:
        check(lib().tmx_init(device, pool_fraction))
        _initialised = True


def initialised() -> bool:
    return _initialised


def synchronize() -> None:
    check(lib().tmx_synchronize())


def pool_info():
    a, b, c = C.c_size_t(), C.c_size_t(), C.c_size_t()
    check(lib().tmx_pool_info(C.byref(a), C.byref(b), C.byref(c)))
    return a.value, b.value, c.value


class DeviceBuffer:
    """Device memory from the engine pool."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = vp()
        check(lib().tmx_malloc(max(self.nbytes, 1), C.byref(p)))
        self.ptr = p.value

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(lib().tmx_memcpy(self.ptr, a.ctypes.data, a.nbytes, H2D, None))

    def download(self, dtype=np.float64, count: Optional[int] = None) -> np.ndarray:
        n = self.nbytes // np.dtype(dtype).itemsize if count is None else count
        out = np.empty(n, dtype=dtype)
        check(lib().tmx_memcpy(out.ctypes.data, self.ptr, out.nbytes, D2H, None))
        return out

    def free(self) -> None:
        if self.ptr is not None and _lib is not None and _initialised:
            lib().tmx_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Space:
    """TiledIndexSpace (structure only; works without a GPU)."""

    def __init__(self, n: int, tiles, spin=(), irrep=(), named=()):
        use_vector = not isinstance(tiles, (int, np.integer))
        tl = _i64([tiles] if not use_vector else list(tiles))
        arrs = []
        for rng in (list(spin), list(irrep)):
            arrs += [_i64([r[0] for r in rng] or [0]), _i64([r[1] for r in rng] or [0]),
                     _i32([r[2] for r in rng] or [0])]
        nm = list(named)
        nlo, nhi = _i64([r[0] for r in nm] or [0]), _i64([r[1] for r in nm] or [0])
        h = vp()
        check(lib().tmx_space_create(n, int(use_vector), _p(tl, i64p), len(tl), len(list(spin)),
                                     _p(arrs[0], i64p), _p(arrs[1], i64p), _p(arrs[2], i32p),
                                     len(list(irrep)), _p(arrs[3], i64p), _p(arrs[4], i64p),
                                     _p(arrs[5], i32p), len(nm), _p(nlo, i64p), _p(nhi, i64p),
                                     C.byref(h)))
        self.h = h.value
        self.n = n
        nt = C.c_int64()
        check(lib().tmx_space_ntiles(self.h, C.byref(nt)))
        self.offsets = np.zeros(nt.value + 1, dtype=np.int64)
        sp = np.zeros(max(nt.value, 1), dtype=np.int32)
        ir = np.zeros(max(nt.value, 1), dtype=np.int32)
        check(lib().tmx_space_tiles(self.h, _p(self.offsets, i64p), _p(sp, i32p), _p(ir, i32p)))
        self.spin, self.irrep = sp[:nt.value], ir[:nt.value]

    @property
    def ntiles(self) -> int:
        return len(self.offsets) - 1

    @property
    def tile_sizes(self) -> List[int]:
        return [int(x) for x in np.diff(self.offsets)]

    def __del__(self):
        if _lib is not None and getattr(self, "h", None):
            _lib.tmx_space_destroy(self.h)


class Tensor:
    """Block-sparse tensor: structure (layout) plus an optional HBM buffer."""

    def __init__(self, spaces: Sequence[Space], spin=None, irrep_rule: bool = False):
        self.spaces = list(spaces)
        arr = (vp * max(len(spaces), 1))(*[s.h for s in spaces])
        nu, nl = spin if spin is not None else (-1, 0)
        h = vp()
        check(lib().tmx_tensor_create(len(spaces), arr, nu, nl, int(irrep_rule), C.byref(h)))
        self.h = h.value
        n = C.c_int64()
        check(lib().tmx_tensor_size(self.h, C.byref(n)))
        self.size = n.value
        check(lib().tmx_tensor_nblocks(self.h, C.byref(n)))
        nb = n.value
        self.keys = np.zeros(nb, dtype=np.int64)
        self.offsets = np.zeros(nb, dtype=np.int64)
        self.sizes = np.zeros(nb, dtype=np.int64)
        if nb:
            check(lib().tmx_tensor_blocks(self.h, _p(self.keys, i64p), _p(self.offsets, i64p),
                                          _p(self.sizes, i64p)))
        self.buf: Optional[DeviceBuffer] = None

    @property
    def nmodes(self) -> int:
        return len(self.spaces)

    def allocate(self) -> "Tensor":
        if self.buf is None:
            self.buf = DeviceBuffer(self.size * 8)
        return self

    @property
    def ptr(self) -> int:
        assert self.buf is not None, "tensor not allocated"
        return self.buf.ptr

    def upload(self, data: np.ndarray) -> None:
        self.allocate()
        assert data.size == self.size
        self.buf.upload(np.ascontiguousarray(data, dtype=np.float64))

    def download(self) -> np.ndarray:
        return self.buf.download(np.float64, self.size)

    def fill(self, value: float) -> None:
        self.allocate()
        check(lib().tmx_fill(self.ptr, self.size, value, None))

    def fill_normal(self, seed: int, tensor_id: int, scale: float, symmetrize: bool = False) -> None:
        self.allocate()
        check(lib().tmx_fill_normal(self.h, self.ptr, seed, tensor_id, scale, int(symmetrize), None))

    def norm(self) -> float:
        r = C.c_double()
        check(lib().tmx_norm(self.ptr, self.size, C.byref(r), None))
        return r.value

    def write_hdf5(self, path: str) -> None:
        """TAMM's write_to_disk format (dataset "tensor", packed nonzero blocks)."""
        check(lib().tmx_tensor_write_hdf5(self.h, self.ptr, path.encode(), None))

    def read_hdf5(self, path: str) -> None:
        self.allocate()
        check(lib().tmx_tensor_read_hdf5(self.h, self.ptr, path.encode(), None))

    def copy_from(self, other: "Tensor") -> None:
        """self = other (same structure), on the engine stream."""
        assert other.size == self.size
        self.allocate()
        check(lib().tmx_memcpy(self.ptr, other.ptr, self.size * 8, D2D, None))

    def axpy(self, alpha: float, x: "Tensor") -> None:
        """self += alpha * x (same structure)."""
        assert x.size == self.size
        check(lib().tmx_axpy(self.size, alpha, x.ptr, self.ptr, None))

    def scale(self, alpha: float) -> None:
        check(lib().tmx_scale(self.ptr, self.size, alpha, None))

    def dot(self, other: "Tensor") -> float:
        """FP64 inner product over the stored blocks (same structure)."""
        assert other.size == self.size
        return dot(self.ptr, other.ptr, self.size)

    def locate(self, blockid) -> int:
        b = _i64(blockid if len(blockid) else [0])
        off, sz = C.c_int64(), C.c_int64()
        check(lib().tmx_tensor_locate(self.h, _p(b, i64p), C.byref(off), C.byref(sz)))
        return off.value

    def free(self) -> None:
        if self.buf is not None:
            self.buf.free()
            self.buf = None

    def __del__(self):
        try:
            self.free()
            if _lib is not None and getattr(self, "h", None):
                _lib.tmx_tensor_destroy(self.h)
        except Exception:
            pass


class _OperandHolder:
    def __init__(self, t: Tensor, labels, toff=None, ntl=None):
        self.lab = _i32(list(labels) or [0])
        self.toff = None if toff is None else _i64(toff)
        self.ntl = None if ntl is None else _i64(ntl)
        self.op = _Operand(t.h, _p(self.lab, i32p),
                           None if self.toff is None else _p(self.toff, i64p),
                           None if self.ntl is None else _p(self.ntl, i64p))


class Contraction:
    """C(clabels) += alpha * A(alabels) * B(blabels)."""

    def __init__(self, Ct: Tensor, clabels, alpha: float, At: Tensor, alabels, Bt: Tensor, blabels,
                 mode: int = EXEC_AUTO, toffs=None, ntiles=None):
        toffs = toffs or (None, None, None)
        ntiles = ntiles or (None, None, None)
        self._ops = [_OperandHolder(t, l, o, n) for t, l, o, n in
                     zip((Ct, At, Bt), (clabels, alabels, blabels), toffs, ntiles)]
        h = vp()
        check(lib().tmx_contraction_create(C.byref(self._ops[0].op), alpha, C.byref(self._ops[1].op),
                                           C.byref(self._ops[2].op), mode, C.byref(h)))
        self.h = h.value
        fl, npairs, ws, md = C.c_double(), C.c_int64(), C.c_int64(), C.c_int()
        check(lib().tmx_contraction_info(self.h, C.byref(fl), C.byref(npairs), C.byref(ws), C.byref(md)))
        self.flops, self.npairs, self.workspace, self.mode = fl.value, npairs.value, ws.value, md.value
        self.tensors = (Ct, At, Bt)

    def split_info(self):
        a, b, n = C.c_int(), C.c_int(), C.c_int()
        check(lib().tmx_contraction_split_info(self.h, C.byref(a), C.byref(b), C.byref(n)))
        return a.value, b.value, n.value

    def schedule_info(self):
        """(split-K rounds, GEMM launches, GEMM tasks) of a sector plan."""
        r, l, t = C.c_int(), C.c_int(), C.c_int64()
        check(lib().tmx_contraction_schedule_info(self.h, C.byref(r), C.byref(l), C.byref(t)))
        return r.value, l.value, t.value

    def execute(self, stream=None, A: Optional[int] = None, B: Optional[int] = None,
                Cp: Optional[int] = None) -> None:
        Ct, At, Bt = self.tensors
        check(lib().tmx_contraction_execute(self.h, A or At.ptr, B or Bt.ptr, Cp or Ct.ptr, stream))

    def autotune(self, stream=None) -> int:
        """Executes once, then times every tile shape per sector and keeps the fastest;
        returns the number of sectors whose shape changed."""
        Ct, At, Bt = self.tensors
        n = C.c_int()
        check(lib().tmx_contraction_autotune(self.h, At.ptr, Bt.ptr, Ct.ptr, stream, C.byref(n)))
        return n.value

    def execute_timed(self, stream=None):
        """Executes once; returns (pack_ms, gemm_ms, scatter_ms) from HIP events."""
        Ct, At, Bt = self.tensors
        ms = (C.c_float * 3)()
        check(lib().tmx_contraction_execute_timed(self.h, At.ptr, Bt.ptr, Ct.ptr, stream, ms))
        return tuple(ms)

    def set_constant(self, a: bool, b: bool) -> None:
        """Operands whose values never change at their address: packed once and kept
        (calling again drops the kept copies)."""
        check(lib().tmx_contraction_set_constant(self.h, int(a), int(b)))

    def set_assign(self, on: bool) -> None:
        """C = alpha*A*B instead of C += alpha*A*B for later executions (TAMM's `C(..) = A*B`)."""
        check(lib().tmx_contraction_set_assign(self.h, int(on)))

    def execute_stages(self, stream=None, n: int = 5):
        """Executes once; returns (pack A ms, pack B ms, GEMM ms, scatter ms, longest GEMM launch
        ms[, main packing pass of A alone ms]) from HIP events (n = 5 or 6 values)."""
        Ct, At, Bt = self.tensors
        ms = (C.c_float * n)()
        check(lib().tmx_contraction_execute_stages_n(self.h, At.ptr, Bt.ptr, Ct.ptr, stream, ms, n))
        return tuple(ms)

    def __del__(self):
        try:
            if _lib is not None and getattr(self, "h", None):
                _lib.tmx_contraction_destroy(self.h)
        except Exception:
            pass


class Level:
    """Contractions executed in order as one unit; terms whose product (packed operands and
    GEMM) equals an earlier term's run only their scatter (tmx_level_*)."""

    def __init__(self, contractions: Sequence["Contraction"]):
        self.cs = list(contractions)
        arr = (vp * max(len(self.cs), 1))(*[c.h for c in self.cs])
        h = vp()
        check(lib().tmx_level_create(arr, len(self.cs), C.byref(h)))
        self.h = h.value
        n = len(self.cs)
        self._a = (vp * max(n, 1))(*[c.tensors[1].ptr for c in self.cs])
        self._b = (vp * max(n, 1))(*[c.tensors[2].ptr for c in self.cs])
        self._c = (vp * max(n, 1))(*[c.tensors[0].ptr for c in self.cs])

    def products(self) -> int:
        n = C.c_int()
        check(lib().tmx_level_products(self.h, C.byref(n)))
        return n.value

    def execute(self, stream=None) -> None:
        check(lib().tmx_level_execute(self.h, self._a, self._b, self._c, stream))

    def __del__(self):
        try:
            if _lib is not None and getattr(self, "h", None):
                _lib.tmx_level_destroy(self.h)
        except Exception:
            pass


class AddOp:
    """C(clabels) (=|+=) alpha * A(alabels)."""

    def __init__(self, Ct: Tensor, clabels, alpha: float, At: Tensor, alabels, is_assign: bool = False):
        self._ops = [_OperandHolder(Ct, clabels), _OperandHolder(At, alabels)]
        h = vp()
        check(lib().tmx_addop_create(C.byref(self._ops[0].op), alpha, C.byref(self._ops[1].op),
                                     int(is_assign), C.byref(h)))
        self.h = h.value
        self.tensors = (Ct, At)

    def execute(self, stream=None) -> None:
        Ct, At = self.tensors
        check(lib().tmx_addop_execute(self.h, At.ptr, Ct.ptr, stream))

    def __del__(self):
        try:
            if _lib is not None and getattr(self, "h", None):
                _lib.tmx_addop_destroy(self.h)
        except Exception:
            pass


def block_multiply(alpha, A: DeviceBuffer, adims, alabels, B: DeviceBuffer, bdims, blabels,
                   Cb: DeviceBuffer, cdims, clabels, is_assign=False, stream=None) -> None:
    ad, bd, cd = _i64(list(adims) or [0]), _i64(list(bdims) or [0]), _i64(list(cdims) or [0])
    al, bl, cl = _i32(list(alabels) or [0]), _i32(list(blabels) or [0]), _i32(list(clabels) or [0])
    check(lib().tmx_block_multiply(alpha, A.ptr, len(adims), _p(ad, i64p), _p(al, i32p), B.ptr,
                                   len(bdims), _p(bd, i64p), _p(bl, i32p), Cb.ptr, len(cdims),
                                   _p(cd, i64p), _p(cl, i32p), int(is_assign), stream))


def block_assign(dst: DeviceBuffer, ddims, dlabels, scale, src: DeviceBuffer, sdims, slabels,
                 is_assign=True, stream=None) -> None:
    dd, sd = _i64(list(ddims) or [0]), _i64(list(sdims) or [0])
    dl, sl = _i32(list(dlabels) or [0]), _i32(list(slabels) or [0])
    check(lib().tmx_block_assign(dst.ptr, len(ddims), _p(dd, i64p), _p(dl, i32p), scale, src.ptr,
                                 _p(sd, i64p), _p(sl, i32p), int(is_assign), stream))


def rel_diff(x: int, y: int, n: int) -> float:
    r = C.c_double()
    check(lib().tmx_rel_diff(x, y, n, C.byref(r), None))
    return r.value


def dot(x: int, y: int, n: int) -> float:
    r = C.c_double()
    check(lib().tmx_dot(x, y, n, C.byref(r), None))
    return r.value


class Event:
    def __init__(self):
        h = vp()
        check(lib().tmx_event_create(C.byref(h)))
        self.h = h.value

    def record(self, stream=None) -> None:
        check(lib().tmx_event_record(self.h, stream))

    def elapsed_ms(self, stop: "Event") -> float:
        ms = C.c_float()
        check(lib().tmx_event_elapsed_ms(self.h, stop.h, C.byref(ms)))
        return ms.value

    def __del__(self):
        try:
            if _lib is not None and _initialised and getattr(self, "h", None):
                _lib.tmx_event_destroy(self.h)
        except Exception:
            pass


class Graph:
    """hipGraph captured from the engine stream."""

    def __init__(self, fn, stream=None):
        check(lib().tmx_graph_begin(stream))
        try:
            fn()
        finally:
            h = vp()
            rc = lib().tmx_graph_end(stream, C.byref(h))
        check(rc)
        self.h = h.value
        self.stream = stream

    def launch(self) -> None:
        check(lib().tmx_graph_launch(self.h, self.stream))

    def __del__(self):
        try:
            if _lib is not None and getattr(self, "h", None):
                _lib.tmx_graph_destroy(self.h)
        except Exception:
            pass


# --- workloads ----------------------------------------------------------------

class EngineWorkload:
    """Engine objects (spaces, tensors in HBM, contractions) for a configs.Workload."""

    def __init__(self, wl, seed: int, allocate: bool = True):
        self.wl = wl
        self.seed = seed
        self.spaces = {k: Space(s.n, s.tiles, s.spin, s.irrep, s.named) for k, s in wl.spaces.items()}
        self.tensors = {k: Tensor([self.spaces[m] for m in t.modes], t.spin, t.irrep_rule)
                        for k, t in wl.tensors.items()}
        self.labels = wl.label_ids()
        self._plans = {}
        if allocate:
            for k, t in wl.tensors.items():
                self.tensors[k].allocate()
                self.generate(k)

    def plan(self, term, mode: int = EXEC_AUTO) -> Contraction:
        """Cached contraction plan of a term (plans depend on structure only, not data)."""
        key = (term, mode)
        if key not in self._plans:
            self._plans[key] = self.contraction(term, mode)
        return self._plans[key]

    def clear_plans(self) -> None:
        self._plans.clear()

    # --- op programs (configs.SetDef / AddDef / TermDef / DenomDef lists) ---
    def addop(self, op) -> "AddOp":
        if op not in self._plans:
            L = self.labels
            (cn, cl), (an, al) = op.c, op.a
            self._plans[op] = AddOp(self.tensors[cn], [L[c] for c in cl], op.alpha,
                                    self.tensors[an], [L[c] for c in al], op.assign)
        return self._plans[op]

    def denominator(self, name: str) -> None:
        """tensor /= (e_i + e_j - e_a - e_b), orbital energies from configs.orbital_energies."""
        if not hasattr(self, "_eo"):
            from . import configs as _cfg
            eo, ev = _cfg.orbital_energies(self.wl)
            self._eo = DeviceBuffer.from_array(np.asarray(eo, dtype=np.float64))
            self._ev = DeviceBuffer.from_array(np.asarray(ev, dtype=np.float64))
        t = self.tensors[name]
        check(lib().tmx_energy_denominator(t.h, t.ptr, self._eo.ptr, self._ev.ptr, None))

    def cc_step(self, mode: int = EXEC_AUTO, recompute: bool = False) -> None:
        """Forward program, then the adjoint program (workloads with extra["forward"]).

        recompute=True drops the intermediates (extra["intermediates"] outputs) after the
        forward pass and recomputes them for the adjoint pass instead of keeping them in HBM.
        """
        wl = self.wl
        self.execute_program(wl.extra["forward"], mode)
        if recompute:
            inter = [op.c[0] for op in wl.extra["intermediates"]]
            for name in dict.fromkeys(inter):
                self.tensors[name].free()
            for name in dict.fromkeys(inter):
                self.tensors[name].allocate()
                self.tensors[name].fill(float("nan"))  # nothing of the forward pass survives
            self.execute_program(wl.extra["intermediates"], mode)
        self.execute_program(wl.extra["adjoint"], mode)

    def level(self, terms, mode: int = EXEC_AUTO) -> "Level":
        """Cached Level over consecutive contraction terms (their shared products found once)."""
        key = ("level", tuple(terms), mode)
        if key not in self._plans:
            self._plans[key] = Level([self.plan(t, mode) for t in terms])
        return self._plans[key]

    def execute_program(self, ops, mode: int = EXEC_AUTO, levels: bool = False) -> None:
        """Runs a list of ops in stream order (contractions in `mode`; everything else FP64).
        With levels=True each run of consecutive contraction terms executes as one Level."""
        from . import configs as _cfg
        run = []

        def flush():
            if len(run) > 1 and levels:
                self.level(run, mode).execute()
            else:
                for t in run:
                    self.plan(t, mode).execute()
            run.clear()

        for op in ops:
            if isinstance(op, _cfg.TermDef):
                run.append(op)
                continue
            flush()
            if isinstance(op, _cfg.SetDef):
                self.tensors[op.name].fill(op.value)
            elif isinstance(op, _cfg.AddDef):
                self.addop(op).execute()
            elif isinstance(op, _cfg.DenomDef):
                self.denominator(op.name)
        flush()

    def generate(self, name: str) -> None:
        t = self.wl.tensors[name]
        if t.tensor_id:
            self.tensors[name].fill_normal(self.seed, t.tensor_id, t.scale, t.symmetrize)
        else:
            self.tensors[name].fill(0.0)

    def run(self, terms, mode: int = EXEC_AUTO, outputs=()) -> None:
        """Zero `outputs`, then execute every term in order (plans built per call)."""
        for name in outputs:
            self.tensors[name].fill(0.0)
        for t in terms:
            self.plan(t, mode).execute()

    def loss(self, mode: int = EXEC_AUTO) -> float:
        """L = <Lambda, R> after a forward pass (workloads with extra["loss"])."""
        lam, out = self.wl.extra["loss"]
        self.run(self.wl.terms, mode, outputs=[out])
        return self.tensors[lam].dot(self.tensors[out])

    def gradients(self, mode: int = EXEC_AUTO) -> None:
        """dL/dX for every X in extra["grads"] (dL/dR = Lambda), by the adjoint contractions."""
        self.run(self.wl.extra["adjoint_terms"], mode, outputs=list(self.wl.extra["grads"].values()))

    def contraction(self, term, mode: int = EXEC_AUTO) -> Contraction:
        cn, cl = term.c
        an, al = term.a
        bn, bl = term.b
        L = self.labels
        c = Contraction(self.tensors[cn], [L[c] for c in cl], term.alpha,
                        self.tensors[an], [L[c] for c in al],
                        self.tensors[bn], [L[c] for c in bl], mode)
        if getattr(term, "assign", False):
            c.set_assign(True)
        return c
