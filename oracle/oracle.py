"""ctypes wrapper of the CPU oracle (liboracle_gf2.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the gf2_amd package.  Build with
``make -C oracle``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_gf2.so")


class OrResult(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_uint32), ("converged", ctypes.c_uint8),
                ("syndrome_ok", ctypes.c_uint8)]


class OrCode(ctypes.Structure):
    _fields_ = [("m", ctypes.c_uint32), ("n", ctypes.c_uint32), ("nnz", ctypes.c_uint64),
                ("row_ptr", ctypes.POINTER(ctypes.c_uint64)),
                ("row_idx", ctypes.POINTER(ctypes.c_uint32)),
                ("col_ptr", ctypes.POINTER(ctypes.c_uint64)),
                ("col_idx", ctypes.POINTER(ctypes.c_uint32)),
                ("csc_to_csr", ctypes.POINTER(ctypes.c_uint64))]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        lib.or_qc_to_edges.restype = ctypes.c_int
        lib.or_qc_to_edges.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, vp, vp,
                                       ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
        lib.or_code_from_edges.restype = ctypes.POINTER(OrCode)
        lib.or_code_from_edges.argtypes = [ctypes.c_uint32, ctypes.c_uint32, vp, ctypes.c_size_t]
        lib.or_code_free.restype = None
        lib.or_code_free.argtypes = [ctypes.POINTER(OrCode)]
        lib.or_minsum_avx2.restype = ctypes.c_float
        lib.or_minsum_avx2.argtypes = [vp, ctypes.c_size_t]
        lib.or_minsum_scalar.restype = ctypes.c_float
        lib.or_minsum_scalar.argtypes = [vp, ctypes.c_size_t]
        lib.or_syndrome.restype = ctypes.c_uint64
        lib.or_syndrome.argtypes = [ctypes.POINTER(OrCode), vp, vp]
        lib.or_decode_iterative_ex.restype = None
        lib.or_decode_iterative_ex.argtypes = [ctypes.POINTER(OrCode), vp, ctypes.c_uint32, vp, vp,
                                               ctypes.POINTER(OrResult), vp, vp, ctypes.c_int]
        lib.or_decode_batch.restype = None
        lib.or_decode_batch.argtypes = [ctypes.POINTER(OrCode), vp, ctypes.c_size_t,
                                        ctypes.c_uint32, vp, vp, vp, ctypes.c_int]
        lib.or_boxplus_n.restype = ctypes.c_float
        lib.or_boxplus_n.argtypes = [vp, ctypes.c_size_t]
        lib.or_decode_spa.restype = None
        lib.or_decode_spa.argtypes = [ctypes.POINTER(OrCode), vp, ctypes.c_uint32, vp,
                                      ctypes.POINTER(OrResult), vp, vp]
        lib.or_decode_batch_spa.restype = None
        lib.or_decode_batch_spa.argtypes = [ctypes.POINTER(OrCode), vp, ctypes.c_size_t, ctypes.c_uint32, vp, vp,
                                            ctypes.c_int]
        lib.or_nms_loss_grad.restype = ctypes.c_double
        lib.or_nms_loss_grad.argtypes = [ctypes.POINTER(OrCode), vp, vp, ctypes.c_uint32, vp, vp]
        lib.or_philox4x32.restype = None
        lib.or_philox4x32.argtypes = [vp, vp, vp]
        lib.or_framegen.restype = None
        lib.or_framegen.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                    ctypes.c_uint64, ctypes.c_uint64, ctypes.c_float,
                                    ctypes.c_float, vp, vp]
        lib.or_framegen_zero.restype = None
        lib.or_framegen_zero.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64,
                                         ctypes.c_float, ctypes.c_float, vp]
        lib.or_layered_decode.restype = None
        lib.or_layered_decode.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, vp,
                                          ctypes.c_uint32, vp, ctypes.c_uint32, vp, ctypes.POINTER(OrResult),
                                          vp, vp]
        lib.or_layered_decode_batch.restype = None
        lib.or_layered_decode_batch.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, vp,
                                                ctypes.c_size_t, ctypes.c_uint32, vp, ctypes.c_uint32, vp, vp,
                                                vp, ctypes.c_int]
        lib.or_layered_loss_grad.restype = ctypes.c_double
        lib.or_layered_loss_grad.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, vp, vp,
                                             ctypes.c_uint32, vp, vp]
        lib.port_code_new.restype = vp
        lib.port_code_new.argtypes = [ctypes.c_uint32, ctypes.c_uint32, vp, ctypes.c_size_t]
        lib.port_code_free.restype = None
        lib.port_code_free.argtypes = [vp]
        lib.port_decode_batch.restype = None
        lib.port_decode_batch.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_uint32, vp, vp,
                                          ctypes.c_int]
        _lib = lib
    return _lib


def _p(a):
    return a.ctypes.data if a is not None else None


def qc_to_edges(base: np.ndarray, z: int) -> np.ndarray:
    base = np.ascontiguousarray(base, dtype=np.int32)
    mb, nb = base.shape
    cap = int((base >= 0).sum()) * z + 1
    out = np.zeros((cap, 2), dtype=np.uint32)
    cnt = ctypes.c_size_t()
    rc = load().or_qc_to_edges(mb, nb, z, _p(base), _p(out), cap, ctypes.byref(cnt))
    if rc != 0:
        raise ValueError(f"or_qc_to_edges failed ({rc})")
    return out[: cnt.value].copy()


def dvb_edges(table, n: int, k: int, q: int, z: int = 360) -> np.ndarray:
    """build_dvb_edges (crates/gf2-coding/src/ldpc/dvb_t2/builder.rs:74-118):
    for table row i, address p and lane j in [0, z), the edge (check
    (p + j*q) mod m, variable i*z + j); then check p -> variable k+p, and for
    p > 0 check p -> variable k+p-1.  (check, variable) pairs in the
    reference's push order."""
    m = n - k
    j = np.arange(z, dtype=np.int64)
    parts = [np.stack([(int(p) + j * q) % m, i * z + j], axis=1) for i, row in enumerate(table) for p in row]
    p = np.arange(m, dtype=np.int64)
    stair = np.empty((2 * m - 1, 2), dtype=np.int64)
    stair[0] = (0, k)
    stair[1::2] = np.stack([p[1:], k + p[1:]], axis=1)
    stair[2::2] = np.stack([p[1:], k + p[1:] - 1], axis=1)
    return np.concatenate(parts + [stair]) if parts else stair


class Code:
    """Oracle code: reference CSR/CSC built by from_coo."""

    def __init__(self, m: int, n: int, edges: np.ndarray):
        e = np.ascontiguousarray(np.asarray(edges, dtype=np.uint32).reshape(-1, 2))
        self._edges = e
        self._p = load().or_code_from_edges(m, n, _p(e), e.shape[0])
        if not self._p:
            raise ValueError("edge index out of bounds")
        c = self._p.contents
        self.m, self.n, self.nnz = c.m, c.n, int(c.nnz)
        self.k = max(self.n - self.m, 0)
        self.row_ptr = np.ctypeslib.as_array(c.row_ptr, (self.m + 1,)).copy()
        self.row_idx = np.ctypeslib.as_array(c.row_idx, (max(self.nnz, 1),))[: self.nnz].copy()
        self.col_ptr = np.ctypeslib.as_array(c.col_ptr, (self.n + 1,)).copy()
        self.col_idx = np.ctypeslib.as_array(c.col_idx, (max(self.nnz, 1),))[: self.nnz].copy()
        self.csc_to_csr = np.ctypeslib.as_array(c.csc_to_csr, (max(self.nnz, 1),))[: self.nnz].copy()
        self._port = None

    @classmethod
    def from_qc(cls, base: np.ndarray, z: int) -> "Code":
        mb, nb = np.asarray(base).shape
        return cls(mb * z, nb * z, qc_to_edges(base, z))

    def __del__(self):
        try:
            if self._p:
                load().or_code_free(self._p)
            if self._port:
                load().port_code_free(self._port)
        except Exception:
            pass

    # -- decoding -------------------------------------------------------
    def decode_batch(self, llr: np.ndarray, max_it: int, edge_w=None, threads: int = 1):
        llr = np.ascontiguousarray(llr, dtype=np.float32).reshape(-1, self.n)
        F = llr.shape[0]
        kw = (self.k + 63) // 64
        msg = np.zeros((F, kw), dtype=np.uint64)
        res = (OrResult * max(F, 1))()
        w = None if edge_w is None else np.ascontiguousarray(edge_w, dtype=np.float32)
        load().or_decode_batch(self._p, _p(llr), F, max_it, _p(w), _p(msg), res, threads)
        it = np.array([res[i].iterations for i in range(F)], dtype=np.uint32)
        conv = np.array([res[i].converged for i in range(F)], dtype=bool)
        synd = np.array([res[i].syndrome_ok for i in range(F)], dtype=bool)
        return msg, it, conv, synd

    def decode_trace(self, llr: np.ndarray, max_it: int, edge_w=None, no_early_stop=False):
        """One frame: (c2v [max_it, nnz] CSR order, belief [max_it, n], msg, result)."""
        llr = np.ascontiguousarray(llr, dtype=np.float32).reshape(self.n)
        tc = np.full((max(max_it, 1), self.nnz), np.nan, dtype=np.float32)
        tb = np.full((max(max_it, 1), self.n), np.nan, dtype=np.float32)
        msg = np.zeros((self.k + 63) // 64, dtype=np.uint64)
        res = OrResult()
        w = None if edge_w is None else np.ascontiguousarray(edge_w, dtype=np.float32)
        load().or_decode_iterative_ex(self._p, _p(llr), max_it, _p(w), _p(msg), ctypes.byref(res),
                                      _p(tc), _p(tb), int(no_early_stop))
        return tc, tb, msg, (res.iterations, bool(res.converged), bool(res.syndrome_ok))

    def decode_batch_spa(self, llr: np.ndarray, max_it: int, threads: int = 1):
        """decode_batch with the sum-product check rule (core.rs:762-787)."""
        llr = np.ascontiguousarray(llr, dtype=np.float32).reshape(-1, self.n)
        F = llr.shape[0]
        kw = (self.k + 63) // 64
        msg = np.zeros((F, kw), dtype=np.uint64)
        res = (OrResult * max(F, 1))()
        load().or_decode_batch_spa(self._p, _p(llr), F, max_it, _p(msg), res, threads)
        it = np.array([res[i].iterations for i in range(F)], dtype=np.uint32)
        conv = np.array([res[i].converged for i in range(F)], dtype=bool)
        synd = np.array([res[i].syndrome_ok for i in range(F)], dtype=bool)
        return msg, it, conv, synd

    def decode_trace_spa(self, llr: np.ndarray, max_it: int):
        """One frame with the sum-product rule: (c2v [max_it, nnz], belief [max_it, n], msg, result)."""
        llr = np.ascontiguousarray(llr, dtype=np.float32).reshape(self.n)
        tc = np.full((max(max_it, 1), self.nnz), np.nan, dtype=np.float32)
        tb = np.full((max(max_it, 1), self.n), np.nan, dtype=np.float32)
        msg = np.zeros((self.k + 63) // 64, dtype=np.uint64)
        res = OrResult()
        load().or_decode_spa(self._p, _p(llr), max_it, _p(msg), ctypes.byref(res), _p(tc), _p(tb))
        return tc, tb, msg, (res.iterations, bool(res.converged), bool(res.syndrome_ok))

    def port_decode_batch(self, llr: np.ndarray, max_it: int, threads: int):
        """The structural restatement of the rayon path (ref_port.cpp)."""
        if self._port is None:
            self._port = load().port_code_new(self.m, self.n, _p(self._edges), self._edges.shape[0])
        llr = np.ascontiguousarray(llr, dtype=np.float32).reshape(-1, self.n)
        F = llr.shape[0]
        kw = (self.k + 63) // 64
        msg = np.zeros((F, kw), dtype=np.uint64)
        res = (OrResult * max(F, 1))()
        load().port_decode_batch(self._port, _p(llr), F, max_it, _p(msg), res, threads)
        it = np.array([res[i].iterations for i in range(F)], dtype=np.uint32)
        conv = np.array([res[i].converged for i in range(F)], dtype=bool)
        synd = np.array([res[i].syndrome_ok for i in range(F)], dtype=bool)
        return msg, it, conv, synd

    def syndrome_weight(self, bits: np.ndarray) -> int:
        words = pack_bits(bits)
        return int(load().or_syndrome(self._p, _p(words), None))

    def nms_loss_grad(self, llr: np.ndarray, bits: np.ndarray, T: int, edge_w: np.ndarray):
        llr = np.ascontiguousarray(llr, dtype=np.float32).reshape(self.n)
        b = np.ascontiguousarray(bits, dtype=np.uint8).reshape(self.n)
        w = np.ascontiguousarray(edge_w, dtype=np.float64).reshape(T * self.nnz)
        g = np.zeros(T * self.nnz, dtype=np.float64)
        loss = load().or_nms_loss_grad(self._p, _p(llr), _p(b), T, _p(w), _p(g))
        return loss, g.reshape(T, self.nnz)


def layered_decode_batch(base: np.ndarray, z: int, llr: np.ndarray, w: np.ndarray, fb_it: int, threads: int = 8):
    """Layered learned decoder (layered_oracle.c): T = w.shape[0] layered
    weighted iterations with early stop, then up to fb_it plain flooding
    iterations continued from the layered state.  w is [T, nbe] block-edge
    weights (block-row-major).  Returns (msg, iterations, converged,
    syndrome_ok, fell_back)."""
    base = np.ascontiguousarray(base, dtype=np.int32)
    mb, nb = base.shape
    n, k = nb * z, (nb - mb) * z
    llr = np.ascontiguousarray(llr, dtype=np.float32).reshape(-1, n)
    w = np.ascontiguousarray(w, dtype=np.float32)
    F = llr.shape[0]
    msg = np.zeros((F, (k + 63) // 64), dtype=np.uint64)
    res = (OrResult * max(F, 1))()
    fb = np.zeros(F, dtype=np.uint8)
    load().or_layered_decode_batch(_p(base), mb, nb, z, _p(llr), F, w.shape[0], _p(w), fb_it, _p(msg), res,
                                   _p(fb), threads)
    it = np.array([res[i].iterations for i in range(F)], dtype=np.uint32)
    conv = np.array([res[i].converged for i in range(F)], dtype=bool)
    synd = np.array([res[i].syndrome_ok for i in range(F)], dtype=bool)
    return msg, it, conv, synd, fb.astype(bool)


def layered_trace(base: np.ndarray, z: int, llr: np.ndarray, w: np.ndarray, fb_it: int = 0):
    """One frame: beliefs after each layered and continuation iteration
    [T + fb_it, n] (rows after an early stop stay NaN), the message words and
    (iterations, converged, syndrome_ok, fell_back)."""
    base = np.ascontiguousarray(base, dtype=np.int32)
    mb, nb = base.shape
    n, k = nb * z, (nb - mb) * z
    w = np.ascontiguousarray(w, dtype=np.float32)
    tb = np.full((w.shape[0] + fb_it, n), np.nan, dtype=np.float32)
    msg = np.zeros(max((k + 63) // 64, 1), dtype=np.uint64)
    res = OrResult()
    fb = np.zeros(1, dtype=np.uint8)
    l1 = np.ascontiguousarray(llr, dtype=np.float32).reshape(n)
    load().or_layered_decode(_p(base), mb, nb, z, _p(l1), w.shape[0], _p(w), fb_it, _p(msg), ctypes.byref(res),
                             _p(fb), _p(tb))
    return tb, msg, (res.iterations, bool(res.converged), bool(res.syndrome_ok), bool(fb[0]))


def layered_loss_grad(base: np.ndarray, z: int, llr: np.ndarray, bits: np.ndarray, w: np.ndarray):
    """Double-precision loss and gradient [T, nbe] of the layered decoder
    (T = w.shape[0] iterations, no early stop)."""
    base = np.ascontiguousarray(base, dtype=np.int32)
    mb, nb = base.shape
    n = nb * z
    w = np.ascontiguousarray(w, dtype=np.float64)
    g = np.zeros_like(w)
    l1 = np.ascontiguousarray(llr, dtype=np.float32).reshape(n)
    b = np.ascontiguousarray(bits, dtype=np.uint8).reshape(n)
    loss = load().or_layered_loss_grad(_p(base), mb, nb, z, _p(l1), _p(b), w.shape[0], _p(w), _p(g))
    return loss, g


def framegen(base: np.ndarray, z: int, seed: int, frame_id: int, ebn0_lo: float,
             ebn0_hi: float):
    """(llr [n] float32, codeword words [ceil(n/64)] uint64) of one frame."""
    base = np.ascontiguousarray(base, dtype=np.int32)
    mb, nb = base.shape
    n = nb * z
    llr = np.zeros(n, dtype=np.float32)
    cw = np.zeros((n + 63) // 64, dtype=np.uint64)
    load().or_framegen(_p(base), mb, nb, z, seed, frame_id, ebn0_lo, ebn0_hi, _p(llr), _p(cw))
    return llr, cw


def framegen_zero(n: int, k: int, seed: int, frame_id: int, ebn0_lo: float, ebn0_hi: float) -> np.ndarray:
    """LLRs [n] float32 of all-zero codeword frame frame_id."""
    llr = np.zeros(n, dtype=np.float32)
    load().or_framegen_zero(n, k, seed, frame_id, ebn0_lo, ebn0_hi, _p(llr))
    return llr


def minsum_avx2(x) -> np.float32:
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    return np.float32(load().or_minsum_avx2(_p(a), a.size))


def boxplus_n(x) -> np.float32:
    """Llr::boxplus_n (llr.rs:210-215) in binary32."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    return np.float32(load().or_boxplus_n(_p(a), a.size))


def minsum_scalar(x) -> np.float32:
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    return np.float32(load().or_minsum_scalar(_p(a), a.size))


def philox4x32(ctr, key) -> np.ndarray:
    c = np.ascontiguousarray(np.asarray(ctr, dtype=np.uint32).reshape(4))
    k = np.ascontiguousarray(np.asarray(key, dtype=np.uint32).reshape(2))
    out = np.zeros(4, dtype=np.uint32)
    load().or_philox4x32(_p(c), _p(k), _p(out))
    return out


def pack_bits(bits: np.ndarray) -> np.ndarray:
    """bool[n] -> LE u64 words (BitVec layout)."""
    b = np.asarray(bits, dtype=np.uint8).reshape(-1)
    n = b.size
    pad = np.zeros(((n + 63) // 64) * 64, dtype=np.uint8)
    pad[:n] = b
    return np.packbits(pad, bitorder="little").view(np.uint64).copy()
