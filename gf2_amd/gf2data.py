"""GF2DATA files: the reference's versioned binary format for GF(2) data
(crates/gf2-core/src/io/mod.rs:6-98).

Layout: a 32-byte header -- magic "GF2DATA\\0", u16 version (1), u8 type tag
(1 BitVec, 2 BitMatrix, 3 SpBitMatrix, 4 SpBitMatrixDual), u8 flags, 8
reserved zero bytes, u32 metadata length, u64 payload length, all
little-endian (format.rs:3-26, header.rs:8-31) -- then compact JSON metadata
and the payload:
  1 BitVec           u64 words, tail bits zero (bitvec.rs)
  2 BitMatrix        row-major, ceil(cols/64) u64 words per row (matrix.rs)
  3 SpBitMatrix      (row, col) u32 pairs in CSR order (sparse.rs:86-120)
  4 SpBitMatrixDual  row offsets, column indices, column offsets, row
                     indices, all u32 (sparse.rs:315-363)
Readers enforce what the reference's readers enforce (header.rs:33-86,
sparse.rs:151-200, 397-449); compressed payloads and checksums are not
implemented there and are not here (a compressed file is refused, a checksum
trailer is ignored).

    m = SpBitMatrixDual.from_coo(32400, 64800, edges)
    m.save_to_file("h.gf2")
    code = LdpcCode.from_gf2data("h.gf2")     # gf2_amd.ldpc
"""
from __future__ import annotations

import json
import struct
from dataclasses import dataclass
from typing import Iterable, Tuple

import numpy as np

MAGIC = b"GF2DATA\x00"
FORMAT_VERSION = 1
HEADER_SIZE = 32
TYPE_BITVEC, TYPE_BITMATRIX, TYPE_SPBITMATRIX, TYPE_SPBITMATRIX_DUAL = 1, 2, 3, 4
_TYPE_NAMES = {1: "BitVec", 2: "BitMatrix", 3: "SpBitMatrix", 4: "SpBitMatrixDual"}


class Gf2DataError(ValueError):
    """IoError of the reference reader (error.rs:27-40); the message is the
    reference's Display text."""


@dataclass
class Header:
    version: int
    type_tag: int
    flags: int
    metadata_len: int
    data_len: int

    def to_bytes(self) -> bytes:  # Header::write_to (header.rs:8-31)
        return MAGIC + struct.pack("<HBB8xIQ", self.version, self.type_tag, self.flags,
                                   self.metadata_len, self.data_len)

    @classmethod
    def from_bytes(cls, b: bytes) -> "Header":  # Header::read_from (header.rs:33-86)
        if len(b) < HEADER_SIZE:
            raise Gf2DataError("I/O error: failed to fill whole buffer")
        if b[:8] != MAGIC:
            raise Gf2DataError("Invalid magic bytes")
        version, tag, flags, meta_len, data_len = struct.unpack_from("<HBB8xIQ", b, 8)
        if version != FORMAT_VERSION:
            raise Gf2DataError(f"Unsupported format version: {version}")
        if tag not in _TYPE_NAMES:
            raise Gf2DataError(f"Unknown type tag: {tag}")
        return cls(version, tag, flags, meta_len, data_len)


def _dump(tag: int, meta: dict, payload: bytes) -> bytes:
    # serde_json::to_vec: compact, fields in declaration order
    mj = json.dumps(meta, separators=(",", ":")).encode()
    return Header(FORMAT_VERSION, tag, 0, len(mj), len(payload)).to_bytes() + mj + payload


def _load(b: bytes, tag: int) -> Tuple[dict, memoryview]:
    h = Header.from_bytes(b)
    if h.type_tag != tag:
        raise Gf2DataError(f"Invalid data: Expected {_TYPE_NAMES[tag]} type tag, got {_TYPE_NAMES[h.type_tag]}")
    if h.flags & 0x01:
        raise Gf2DataError("Invalid data: compressed payloads are not supported")
    end = HEADER_SIZE + h.metadata_len
    if len(b) < end:
        raise Gf2DataError("I/O error: failed to fill whole buffer")
    try:
        meta = json.loads(bytes(b[HEADER_SIZE:end]))
    except ValueError as e:
        raise Gf2DataError(f"Invalid data: Failed to parse metadata: {e}")
    if meta.get("type") != _TYPE_NAMES[tag]:
        raise Gf2DataError(f"Invalid data: Expected {_TYPE_NAMES[tag]} type, got {meta.get('type')}")
    if len(b) < end + h.data_len:
        raise Gf2DataError("I/O error: failed to fill whole buffer")
    meta["_data_len"] = h.data_len
    return meta, memoryview(b)[end:end + h.data_len]


def _u32(view, count: int, offset: int = 0) -> np.ndarray:
    return np.frombuffer(view, dtype="<u4", count=count, offset=offset).astype(np.int64)


def _check_len(meta: dict, expected: int) -> None:
    if meta["_data_len"] != expected:
        raise Gf2DataError(f"Invalid data: Data length mismatch: header says {meta['_data_len']} bytes, "
                           f"expected {expected}")


# ---- BitVec (type 1) and BitMatrix (type 2) ---------------------------------

def bitvec_to_bytes(words: np.ndarray, len_bits: int) -> bytes:
    w = np.zeros((len_bits + 63) // 64, dtype="<u8")
    src = np.asarray(words, dtype=np.uint64)[:w.size]
    w[:src.size] = src
    if len_bits % 64 and w.size:
        w[-1] &= np.uint64((1 << (len_bits % 64)) - 1)
    return _dump(TYPE_BITVEC, {"type": "BitVec", "len_bits": int(len_bits), "version": 1}, w.tobytes())


def bitvec_from_bytes(b: bytes) -> Tuple[np.ndarray, int]:
    meta, view = _load(b, TYPE_BITVEC)
    nw = (int(meta["len_bits"]) + 63) // 64
    _check_len(meta, nw * 8)
    return np.frombuffer(view, dtype="<u8", count=nw).astype(np.uint64), int(meta["len_bits"])


def bitmatrix_to_bytes(rows_words: np.ndarray, rows: int, cols: int) -> bytes:
    sw = (cols + 63) // 64
    w = np.ascontiguousarray(np.asarray(rows_words, dtype="<u8").reshape(rows, sw))
    return _dump(TYPE_BITMATRIX, {"type": "BitMatrix", "rows": int(rows), "cols": int(cols), "version": 1},
                 w.tobytes())


def bitmatrix_from_bytes(b: bytes) -> Tuple[np.ndarray, int, int]:
    meta, view = _load(b, TYPE_BITMATRIX)
    rows, cols = int(meta["rows"]), int(meta["cols"])
    sw = (cols + 63) // 64
    _check_len(meta, rows * sw * 8)
    return np.frombuffer(view, dtype="<u8", count=rows * sw).astype(np.uint64).reshape(rows, sw), rows, cols


# ---- sparse matrices (types 3 and 4) ------------------------------------------

def _coo_csr(rows: int, cols: int, entries, xor: bool) -> Tuple[np.ndarray, np.ndarray]:
    """CSR of (row, col) entries: rows sorted; repeated entries cancel in
    pairs (xor) or are kept once."""
    e = np.asarray(entries, dtype=np.int64).reshape(-1, 2)
    bad = np.flatnonzero((e[:, 0] < 0) | (e[:, 0] >= rows))
    if bad.size:
        raise ValueError(f"row index {e[bad[0], 0]} out of bounds (rows={rows})")
    bad = np.flatnonzero((e[:, 1] < 0) | (e[:, 1] >= cols))
    if bad.size:
        raise ValueError(f"col index {e[bad[0], 1]} out of bounds (cols={cols})")
    es = e[np.lexsort((e[:, 1], e[:, 0]))]
    if es.shape[0]:
        starts = np.flatnonzero(np.r_[True, np.any(es[1:] != es[:-1], axis=1)])
        counts = np.diff(np.r_[starts, es.shape[0]])
        es = es[starts[counts % 2 == 1]] if xor else es[starts]
    indptr = np.zeros(rows + 1, dtype=np.int64)
    indptr[1:] = np.cumsum(np.bincount(es[:, 0], minlength=rows)) if rows else []
    return indptr, es[:, 1].copy()


@dataclass
class SpBitMatrix:
    """CSR sparse GF(2) matrix (crates/gf2-core/src/sparse.rs)."""
    rows: int
    cols: int
    indptr: np.ndarray
    indices: np.ndarray

    @classmethod
    def from_coo(cls, rows: int, cols: int, entries) -> "SpBitMatrix":
        """Duplicates cancel in pairs (sparse.rs:166-204)."""
        return cls(rows, cols, *_coo_csr(rows, cols, entries, xor=True))

    @classmethod
    def from_coo_deduplicated(cls, rows: int, cols: int, entries) -> "SpBitMatrix":
        """Duplicates kept once (sparse.rs:232-259)."""
        return cls(rows, cols, *_coo_csr(rows, cols, entries, xor=False))

    @property
    def nnz(self) -> int:
        return int(self.indices.size)

    def row_iter(self, r: int) -> Iterable[int]:
        return self.indices[self.indptr[r]:self.indptr[r + 1]].tolist()

    def edges(self) -> np.ndarray:
        rows = np.repeat(np.arange(self.rows, dtype=np.int64), np.diff(self.indptr))
        return np.stack([rows, self.indices], axis=1)

    def transpose(self) -> "SpBitMatrix":
        """Stable transpose (sparse.rs:297-331): each column's rows ascending
        in CSR order of appearance."""
        e = self.edges()
        order = np.argsort(e[:, 1], kind="stable")
        per = np.bincount(e[:, 1], minlength=self.cols) if e.size else np.zeros(self.cols, dtype=np.int64)
        indptr = np.zeros(self.cols + 1, dtype=np.int64)
        indptr[1:] = np.cumsum(per)
        return SpBitMatrix(self.cols, self.rows, indptr, e[order, 0])

    def to_bytes(self) -> bytes:  # write_binary (sparse.rs:86-120)
        meta = {"type": "SpBitMatrix", "rows": self.rows, "cols": self.cols, "nnz": self.nnz,
                "format": "coo", "version": 1}
        return _dump(TYPE_SPBITMATRIX, meta, self.edges().astype("<u4").tobytes())

    @classmethod
    def from_bytes(cls, b: bytes) -> "SpBitMatrix":  # read_binary (sparse.rs:151-200)
        meta, view = _load(b, TYPE_SPBITMATRIX)
        if meta.get("format") != "coo":
            raise Gf2DataError(f"Invalid data: Expected COO format, got {meta.get('format')}")
        nnz = int(meta["nnz"])
        _check_len(meta, nnz * 8)
        e = _u32(view, 2 * nnz).reshape(nnz, 2)
        return cls.from_coo_deduplicated(int(meta["rows"]), int(meta["cols"]), e)

    def save_to_file(self, path) -> None:
        with open(path, "wb") as f:
            f.write(self.to_bytes())

    @classmethod
    def load_from_file(cls, path) -> "SpBitMatrix":
        with open(path, "rb") as f:
            return cls.from_bytes(f.read())


@dataclass
class SpBitMatrixDual:
    """CSR + CSC sparse GF(2) matrix (sparse.rs:441-568): the parity-check
    representation of LdpcCode."""
    csr: SpBitMatrix
    csc: SpBitMatrix

    @classmethod
    def from_coo(cls, rows: int, cols: int, entries) -> "SpBitMatrixDual":
        csr = SpBitMatrix.from_coo(rows, cols, entries)
        return cls(csr, csr.transpose())

    @classmethod
    def from_coo_deduplicated(cls, rows: int, cols: int, entries) -> "SpBitMatrixDual":
        csr = SpBitMatrix.from_coo_deduplicated(rows, cols, entries)
        return cls(csr, csr.transpose())

    @property
    def rows(self) -> int:
        return self.csr.rows

    @property
    def cols(self) -> int:
        return self.csr.cols

    @property
    def nnz(self) -> int:
        return self.csr.nnz

    def row_iter(self, r: int):
        return self.csr.row_iter(r)

    def col_iter(self, c: int):
        return self.csc.row_iter(c)

    def edges(self) -> np.ndarray:
        return self.csr.edges()

    def to_bytes(self) -> bytes:  # write_binary (sparse.rs:315-363)
        meta = {"type": "SpBitMatrixDual", "rows": self.rows, "cols": self.cols, "nnz": self.nnz,
                "version": 1}
        payload = b"".join(np.asarray(a, dtype="<u4").tobytes()
                           for a in (self.csr.indptr, self.csr.indices, self.csc.indptr, self.csc.indices))
        return _dump(TYPE_SPBITMATRIX_DUAL, meta, payload)

    @classmethod
    def from_bytes(cls, b: bytes) -> "SpBitMatrixDual":  # read_binary (sparse.rs:397-449)
        meta, view = _load(b, TYPE_SPBITMATRIX_DUAL)
        rows, cols, nnz = int(meta["rows"]), int(meta["cols"]), int(meta["nnz"])
        _check_len(meta, (rows + 1 + nnz + cols + 1 + nnz) * 4)
        o = 0
        ro = _u32(view, rows + 1, o)
        o += 4 * (rows + 1)
        ri = _u32(view, nnz, o)
        o += 4 * nnz
        co = _u32(view, cols + 1, o)
        o += 4 * (cols + 1)
        ci = _u32(view, nnz, o)
        # from_csr_csc (sparse.rs:547-568) keeps both halves as stored
        return cls(SpBitMatrix(rows, cols, ro, ri), SpBitMatrix(cols, rows, co, ci))

    def save_to_file(self, path) -> None:
        with open(path, "wb") as f:
            f.write(self.to_bytes())

    @classmethod
    def load_from_file(cls, path) -> "SpBitMatrixDual":
        with open(path, "rb") as f:
            return cls.from_bytes(f.read())
