"""Host-side mirror of the reference LDPC decoding API, backed by the HIP decoder.

Names, argument meaning and error behaviour follow crates/gf2-coding:
  QuasiCyclicLdpc            src/ldpc/core.rs:485-606
  LdpcCode.from_quasi_cyclic src/ldpc/core.rs:168-173 (n, m, k, rate :107-124)
  LdpcDecoder.decode_batch   src/ldpc/core.rs:729-756
  LdpcDecoder.decode_iterative / last_iteration_count / reset
                             src/ldpc/core.rs:888-981 (trait src/traits.rs:314-347)
  LdpcDecoder.decode_soft    src/ldpc/core.rs:877-885
  DecoderResult              src/traits.rs:116-156
  BitVec (result words)      crates/gf2-core/src/bitvec.rs:40-63, 204
  LdpcCode.dvb_t2_short / dvb_t2_normal
                             src/ldpc/core.rs:197-255, dvb_t2/builder.rs:74-118
  LdpcCode.from_gf2data      crates/gf2-core/src/io/sparse.rs:151-200, 397-449
Where the reference panics, these raise ValueError with the reference's
message.  Every decode runs on the GPU through libgf2ldpc; there is no CPU
path.
"""
from __future__ import annotations

import ctypes
import enum
from dataclasses import dataclass
from typing import Iterable, List, Optional, Sequence

import numpy as np

from . import _lib


class BitVec:
    """Read-only bit vector in the reference BitVec layout: bit i is bit
    (i & 63) of little-endian u64 word (i >> 6); tail bits are zero."""

    __slots__ = ("_words", "_len")

    def __init__(self, words: np.ndarray, len_bits: int):
        nw = (len_bits + 63) // 64
        w = np.ascontiguousarray(words, dtype=np.uint64)[:nw].copy()
        if w.size != nw:
            raise ValueError("not enough words for len_bits")
        if len_bits % 64 and nw:
            w[-1] &= np.uint64((1 << (len_bits % 64)) - 1)
        self._words = w
        self._len = len_bits

    @classmethod
    def from_words(cls, words, len_bits: int) -> "BitVec":
        return cls(np.asarray(words, dtype=np.uint64), len_bits)

    @classmethod
    def from_bools(cls, bits: Iterable[bool]) -> "BitVec":
        b = np.asarray(list(bits) if not isinstance(bits, np.ndarray) else bits, dtype=np.uint8)
        n = b.size
        pad = np.zeros(((n + 63) // 64) * 64, dtype=np.uint8)
        pad[:n] = b
        return cls(np.packbits(pad, bitorder="little").view(np.uint64).copy(), n)

    def __len__(self) -> int:
        return self._len

    def len(self) -> int:
        return self._len

    def get(self, i: int) -> bool:
        if not 0 <= i < self._len:
            raise IndexError(i)
        return bool((int(self._words[i >> 6]) >> (i & 63)) & 1)

    def count_ones(self) -> int:
        return int(sum(bin(int(w)).count("1") for w in self._words))

    def words(self) -> np.ndarray:
        return self._words.copy()

    def to_bools(self) -> np.ndarray:
        if self._len == 0:
            return np.zeros(0, dtype=bool)
        bytes_ = self._words.view(np.uint8)
        return np.unpackbits(bytes_, bitorder="little")[: self._len].astype(bool)

    def __eq__(self, other) -> bool:  # PartialEq: len + words (bitvec.rs:77-79)
        return isinstance(other, BitVec) and self._len == other._len and bool(
            np.array_equal(self._words, other._words))

    def __repr__(self) -> str:
        return f"BitVec(len={self._len}, ones={self.count_ones()})"


@dataclass
class DecoderResult:
    decoded_bits: BitVec
    iterations: int
    converged: bool
    syndrome_check_passed: bool


class Context:
    """One HIP device + stream + device buffers (libgf2ldpc gf2l_ctx)."""

    def __init__(self, device: int = 0):
        lib = _lib.load()
        h = ctypes.c_void_p()
        rc = lib.gf2l_create(device, ctypes.byref(h))
        if rc != _lib.GF2L_OK:
            raise _lib.Gf2lError(rc, f"cannot create a context on HIP device {device}")
        self._h = h
        self.device = device

    @property
    def handle(self):
        return self._h

    def set_path(self, path: int) -> None:
        _lib.check(self._h, _lib.load().gf2l_set_path(self._h, path))

    def code_path(self, code: "LdpcCode") -> int:
        """The path (PATH_LDS / _STREAM / _STATE / _CSR) a decode of `code`
        takes under this context's current settings."""
        p = ctypes.c_int()
        _lib.check(self.handle, _lib.load().gf2l_code_path(self.handle, code.handle, ctypes.byref(p)))
        return p.value

    def set_check_rule(self, rule: int) -> None:
        """_lib.RULE_MINSUM (default, the reference decoder's rule) or
        _lib.RULE_SPA (check_node_update_spa, core.rs:762-787)."""
        _lib.check(self._h, _lib.load().gf2l_set_check_rule(self._h, rule))

    def set_tuning(self, key: int, value: int) -> None:
        _lib.check(self._h, _lib.load().gf2l_set_tuning(self._h, key, int(value)))

    def synchronize(self) -> None:
        _lib.check(self._h, _lib.load().gf2l_synchronize(self._h))

    def stream(self) -> int:
        return int(_lib.load().gf2l_stream(self._h) or 0)

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.load().gf2l_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx: dict = {}


def default_context(device: int = 0) -> Context:
    """Per-device context shared by calls that do not pass one (the analogue
    of rayon's global pool)."""
    ctx = _default_ctx.get(device)
    if ctx is None:
        ctx = Context(device)
        _default_ctx[device] = ctx
    return ctx


class QuasiCyclicLdpc:
    """Base matrix of circulant shifts (-1 = zero block) and lift Z."""

    def __init__(self, base_matrix, expansion_factor: int):
        rows = [list(r) for r in base_matrix]
        if len(rows) == 0:
            raise ValueError("Base matrix must have at least one row")
        if expansion_factor <= 0:
            raise ValueError("Expansion factor must be positive")
        cols = len(rows[0])
        if any(len(r) != cols for r in rows):
            raise ValueError("All rows in base matrix must have the same length")
        self._base = np.array(rows, dtype=np.int32).reshape(len(rows), cols)
        self._z = int(expansion_factor)

    @property
    def base_matrix(self) -> np.ndarray:
        return self._base.copy()

    def base_rows(self) -> int:
        return self._base.shape[0]

    def base_cols(self) -> int:
        return self._base.shape[1]

    def expansion_factor(self) -> int:
        return self._z

    def expanded_rows(self) -> int:
        return self.base_rows() * self._z

    def expanded_cols(self) -> int:
        return self.base_cols() * self._z

    def _check_shifts(self) -> None:
        z = self._z
        for (br, bc), s in np.ndenumerate(self._base):
            if s != -1 and not 0 <= s < z:
                raise ValueError(
                    f"Shift value {s} at position ({br},{bc}) must be -1 or in range [0,{z})")

    def to_edges(self) -> np.ndarray:
        """(row, col) pairs in the order of QuasiCyclicLdpc::to_edges
        (core.rs:578-605): block-rows, then block-columns, then circulant
        row i -> column (i + s) mod Z (core.rs:436-447)."""
        self._check_shifts()
        z = self._z
        out = []
        i = np.arange(z, dtype=np.int64)
        for br in range(self.base_rows()):
            for bc in range(self.base_cols()):
                s = int(self._base[br, bc])
                if s == -1:
                    continue
                out.append(np.stack([br * z + i, bc * z + (i + s) % z], axis=1))
        if not out:
            return np.zeros((0, 2), dtype=np.int64)
        return np.concatenate(out)


class CodeRate(enum.IntEnum):
    """DVB-T2 code rates in the order of the reference's CodeRate
    (crates/gf2-coding/src/bch), as the C ABI numbers them."""
    RATE_1_2 = 0
    RATE_3_5 = 1
    RATE_2_3 = 2
    RATE_3_4 = 3
    RATE_4_5 = 4
    RATE_5_6 = 5


class LdpcCode:
    """An LDPC code uploaded to one HIP device: quasi-cyclic (from a base
    matrix) or general sparse (from an edge list)."""

    def __init__(self, qc: Optional[QuasiCyclicLdpc], ctx: Optional[Context] = None,
                 _edges: Optional[tuple] = None, _build=None):
        if _build is not None:
            self._ctx = ctx or default_context()
            self._qc = None
            h = ctypes.c_void_p()
            _lib.check(self._ctx.handle, _build(self._ctx, ctypes.byref(h)))
            self._set_handle(h)
            return
        if qc is not None:
            qc._check_shifts()
        else:
            m, n, edges = _edges
            e = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
            bad = e[(e[:, 0] >= m) | (e[:, 0] < 0)]
            if bad.size:  # SpBitMatrix::from_coo panics (sparse.rs:170-171)
                raise ValueError(f"row index {bad[0, 0]} out of bounds (rows={m})")
            bad = e[(e[:, 1] >= n) | (e[:, 1] < 0)]
            if bad.size:
                raise ValueError(f"col index {bad[0, 1]} out of bounds (cols={n})")
        self._ctx = ctx or default_context()
        self._qc = qc
        lib = _lib.load()
        h = ctypes.c_void_p()
        if qc is not None:
            shifts = np.ascontiguousarray(qc._base, dtype=np.int32)
            mb, nb = shifts.shape
            _lib.check(self._ctx.handle, lib.gf2l_code_qc(
                self._ctx.handle, mb, nb, qc.expansion_factor(), _lib.ptr(shifts, ctypes.c_int32),
                ctypes.byref(h)))
        else:
            e32 = np.ascontiguousarray(e, dtype=np.uint32)
            _lib.check(self._ctx.handle, lib.gf2l_code_edges(
                self._ctx.handle, m, n, e32.ctypes.data if e32.size else None, e32.shape[0],
                ctypes.byref(h)))
        self._set_handle(h)

    def _set_handle(self, h) -> None:
        self._h = h
        n, m, k = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        nnz = ctypes.c_uint64()
        _lib.load().gf2l_code_info(h, ctypes.byref(n), ctypes.byref(m), ctypes.byref(k), ctypes.byref(nnz))
        self._n, self._m, self._k, self._nnz = n.value, m.value, k.value, nnz.value

    @classmethod
    def from_quasi_cyclic(cls, qc: QuasiCyclicLdpc, ctx: Optional[Context] = None) -> "LdpcCode":
        return cls(qc, ctx)

    @classmethod
    def from_edges(cls, m: int, n: int, edges, ctx: Optional[Context] = None) -> "LdpcCode":
        """LdpcCode::from_edges (core.rs:96-104): (check, variable) pairs,
        duplicates cancel in pairs as in SpBitMatrix::from_coo."""
        return cls(None, ctx, _edges=(int(m), int(n), edges))

    @classmethod
    def dvb_t2_short(cls, rate: "CodeRate", ctx: Optional[Context] = None) -> "LdpcCode":
        """LdpcCode::dvb_t2_short (core.rs:197-218): N=16200, ETSI EN 302 755
        Annex B table of `rate`; decoded on the CSR path."""
        return cls(None, ctx, _build=lambda c, h: _lib.load().gf2l_code_dvb_t2(c.handle, 0, int(rate), h))

    @classmethod
    def dvb_t2_normal(cls, rate: "CodeRate", ctx: Optional[Context] = None) -> "LdpcCode":
        """LdpcCode::dvb_t2_normal (core.rs:240-255): N=64800, Annex A table."""
        return cls(None, ctx, _build=lambda c, h: _lib.load().gf2l_code_dvb_t2(c.handle, 1, int(rate), h))

    @classmethod
    def from_dvb_table(cls, n: int, k: int, q: int, table, z: int = 360,
                       ctx: Optional[Context] = None) -> "LdpcCode":
        """build_dvb_edges (dvb_t2/builder.rs:74-118) over a caller's address
        table (one list of parity addresses per z-bit information block)."""
        rows = [list(map(int, r)) for r in table]
        lens = np.array([len(r) for r in rows], dtype=np.uint32)
        addr = np.array([a for r in rows for a in r], dtype=np.uint32)
        return cls(None, ctx, _build=lambda c, h: _lib.load().gf2l_code_dvb_table(
            c.handle, int(n), int(k), int(q), int(z), len(rows), lens.ctypes.data if lens.size else None,
            addr.ctypes.data if addr.size else None, h))

    @classmethod
    def from_gf2data(cls, source, ctx: Optional[Context] = None) -> "LdpcCode":
        """Code from a GF2DATA file (path or bytes) holding the parity-check
        matrix as SpBitMatrix (type 3) or SpBitMatrixDual (type 4)
        (crates/gf2-core/src/io/sparse.rs:151-200, 397-449); rows are checks.
        Equivalent to LdpcCode::from_edges over the loaded matrix's entries."""
        if isinstance(source, (bytes, bytearray, memoryview)):
            data = bytes(source)
        else:
            with open(source, "rb") as f:
                data = f.read()
        return cls(None, ctx, _build=lambda c, h: _lib.load().gf2l_code_gf2data(c.handle, data, len(data), h))

    @property
    def handle(self):
        return self._h

    @property
    def context(self) -> Context:
        return self._ctx

    @property
    def qc(self) -> QuasiCyclicLdpc:
        return self._qc

    def n(self) -> int:
        return self._n

    def m(self) -> int:
        return self._m

    def k(self) -> int:
        return self._k

    def nnz(self) -> int:
        return self._nnz

    def rate(self) -> float:
        return self._k / self._n

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                _lib.load().gf2l_code_destroy(self._h)
                self._h = None
        except Exception:
            pass


def _as_llr_matrix(code: LdpcCode, llr_blocks) -> np.ndarray:
    if isinstance(llr_blocks, np.ndarray) and llr_blocks.ndim == 2:
        arr = llr_blocks
    else:
        blocks = list(llr_blocks)
        for b in blocks:
            if len(b) != code.n():
                raise ValueError("LLR length must equal n")
        arr = np.array(blocks, dtype=np.float32).reshape(len(blocks), code.n())
    if arr.ndim != 2 or (arr.shape[0] and arr.shape[1] != code.n()):
        raise ValueError("LLR length must equal n")
    return np.ascontiguousarray(arr, dtype=np.float32)


def decode_batch_arrays(code: LdpcCode, llr: np.ndarray, max_iterations: int):
    """Array form of decode_batch: llr [F, n] float32 ->
    (msg_words [F, ceil(k/64)] uint64, iterations [F] uint32,
     converged [F] bool, syndrome_ok [F] bool)."""
    llr = _as_llr_matrix(code, llr)
    F = llr.shape[0]
    kw = (code.k() + 63) // 64
    msg = np.zeros((F, kw), dtype=np.uint64)
    it = np.zeros(F, dtype=np.uint32)
    conv = np.zeros(F, dtype=np.uint8)
    synd = np.zeros(F, dtype=np.uint8)
    if F:
        ctx = code.context
        _lib.check(ctx.handle, _lib.load().gf2l_decode_batch(
            ctx.handle, code.handle, llr.ctypes.data, F, int(max_iterations), msg.ctypes.data,
            it.ctypes.data, conv.ctypes.data, synd.ctypes.data))
    return msg, it, conv.astype(bool), synd.astype(bool)


def decode_trace_arrays(code: LdpcCode, llr: np.ndarray, max_iterations: int, trace_iters: int):
    """Per-iteration state: returns (c2v [T, F, E] in block-edge order,
    belief [T, F, n], msg, iterations, converged, syndrome_ok)."""
    llr = _as_llr_matrix(code, llr)
    F = llr.shape[0]
    kw = (code.k() + 63) // 64
    tc = np.zeros((trace_iters, F, code.nnz()), dtype=np.float32)
    tb = np.zeros((trace_iters, F, code.n()), dtype=np.float32)
    msg = np.zeros((F, kw), dtype=np.uint64)
    it = np.zeros(F, dtype=np.uint32)
    conv = np.zeros(F, dtype=np.uint8)
    synd = np.zeros(F, dtype=np.uint8)
    ctx = code.context
    _lib.check(ctx.handle, _lib.load().gf2l_decode_trace(
        ctx.handle, code.handle, llr.ctypes.data, F, int(max_iterations), int(trace_iters),
        tc.ctypes.data, tb.ctypes.data, msg.ctypes.data, it.ctypes.data, conv.ctypes.data,
        synd.ctypes.data))
    return tc, tb, msg, it, conv.astype(bool), synd.astype(bool)


def _results(code: LdpcCode, msg, it, conv, synd) -> List[DecoderResult]:
    k = code.k()
    return [DecoderResult(BitVec(msg[f], k), int(it[f]), bool(conv[f]), bool(synd[f]))
            for f in range(msg.shape[0])]


class LdpcDecoder:
    """Min-sum belief-propagation decoder on the GPU.

    ``decode_batch`` is the drop-in for the reference's rayon batch path; the
    per-frame trait methods run it with one frame.
    """

    def __init__(self, code: LdpcCode):
        self._code = code
        self._last_iterations = 0

    @staticmethod
    def decode_batch(code: LdpcCode, llr_blocks, max_iterations: int) -> List[DecoderResult]:
        msg, it, conv, synd = decode_batch_arrays(code, llr_blocks, max_iterations)
        return _results(code, msg, it, conv, synd)

    def k(self) -> int:
        return self._code.k()

    def n(self) -> int:
        return self._code.n()

    def decode_soft(self, llrs: Sequence[float]) -> BitVec:
        """Hard decisions of the input LLRs (core.rs:877-885)."""
        if len(llrs) != self.n():
            raise ValueError(f"assertion `left == right` failed: {len(llrs)} != {self.n()}")
        return BitVec.from_bools(np.asarray(llrs, dtype=np.float32) < 0)

    def decode_iterative(self, llrs: Sequence[float], max_iterations: int) -> DecoderResult:
        if len(llrs) != self.n():
            raise ValueError("LLR length must equal n")
        res = LdpcDecoder.decode_batch(self._code, [list(llrs)], max_iterations)[0]
        self._last_iterations = res.iterations
        return res

    def last_iteration_count(self) -> int:
        return self._last_iterations

    def reset(self) -> None:
        self._last_iterations = 0


def qc_ira_base_matrix(seed: int, mb: int, nb: int, z: int,
                       info_col_weights: Sequence[int]) -> np.ndarray:
    """Seeded QC-IRA base matrix (libgf2ldpc gf2l_qc_ira_shifts)."""
    w = np.ascontiguousarray(np.asarray(info_col_weights, dtype=np.uint32))
    if w.size != nb - mb:
        raise ValueError("need one weight per info block-column (nb - mb)")
    out = np.zeros((mb, nb), dtype=np.int32)
    rc = _lib.load().gf2l_qc_ira_shifts(int(seed), mb, nb, z, _lib.ptr(w, ctypes.c_uint32),
                                        _lib.ptr(out, ctypes.c_int32))
    if rc != _lib.GF2L_OK:
        raise ValueError("invalid QC-IRA parameters")
    return out
