"""Mixed-code batches: frames of several quasi-cyclic codes decoded together.

The reference decodes one code per ``LdpcDecoder::decode_batch`` call
(crates/gf2-coding/src/ldpc/core.rs:729-756); a caller holding frames of
several codes calls it once per code and gets the results in input order.
``decode_mixed`` gives the same per-frame results (message, iteration count,
converged, syndrome flag) for all frames in one device call
(``gf2l_decode_mixed_device``).
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from .device import DeviceBuffer
from .ldpc import BitVec, DecoderResult, LdpcCode


class CodeSet:
    """Up to 16 quasi-cyclic codes on one context (gf2l_codeset_create)."""

    def __init__(self, codes: Sequence[LdpcCode]):
        codes = list(codes)
        if not codes:
            raise ValueError("a code set needs at least one code")
        ctx = codes[0].context
        if any(c.context is not ctx for c in codes):
            raise ValueError("all codes of a set must share one context")
        self.codes = codes
        self.ctx = ctx
        arr = (ctypes.c_void_p * len(codes))(*[c.handle for c in codes])
        h = ctypes.c_void_p()
        _lib.check(ctx.handle, _lib.load().gf2l_codeset_create(ctx.handle, arr, len(codes),
                                                               ctypes.byref(h)))
        self._h = h.value
        self.n = np.array([c.n() for c in codes], dtype=np.int64)
        self.k = np.array([c.k() for c in codes], dtype=np.int64)
        self.kw = (self.k + 63) // 64

    @property
    def handle(self):
        return self._h

    def offsets(self, code_ids: np.ndarray):
        """Element offsets of each frame in the packed LLR and message arrays
        (frames back to back in input order) and their totals."""
        ids = np.asarray(code_ids, dtype=np.int64)
        n = self.n[ids]
        kw = self.kw[ids]
        llr_off = np.concatenate([[0], np.cumsum(n)])
        msg_off = np.concatenate([[0], np.cumsum(kw)])
        return llr_off, msg_off

    def decode_device(self, code_ids: np.ndarray, d_llr: int, max_iterations: int, d_msg: int,
                      d_iters: int, d_conv: int, d_synd: int) -> None:
        """Device-buffer call; code_ids is a host array of uint16."""
        ids = np.ascontiguousarray(code_ids, dtype=np.uint16)
        _lib.check(self.ctx.handle, _lib.load().gf2l_decode_mixed_device(
            self.ctx.handle, self._h, ids.ctypes.data, ids.size, d_llr, int(max_iterations), d_msg,
            d_iters, d_conv, d_synd))

    def decode_nms_device(self, decoders, code_ids: np.ndarray, d_llr: int, fallback_max_it: int, d_msg: int,
                          d_iters: int, d_conv: int, d_synd: int, d_fb: Optional[int] = None) -> None:
        """Learned decode: decoders[c] is an NmsDecoder of code c (same T for all)."""
        if len(decoders) != len(self.codes):
            raise ValueError("one learned decoder per code of the set")
        ids = np.ascontiguousarray(code_ids, dtype=np.uint16)
        arr = (ctypes.c_void_p * len(decoders))(*[d.handle for d in decoders])
        _lib.check(self.ctx.handle, _lib.load().gf2l_decode_mixed_nms_device(
            self.ctx.handle, self._h, arr, ids.ctypes.data, ids.size, d_llr, int(fallback_max_it), d_msg,
            d_iters, d_conv, d_synd, d_fb))

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.load().gf2l_codeset_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def decode_mixed_arrays(cs: CodeSet, code_ids, llr_packed: np.ndarray, max_iterations: int,
                        decoders=None):
    """Host-array form: llr_packed holds each frame's n LLRs back to back.
    Returns (msg_words packed, msg offsets, iterations, converged, syndrome_ok,
    fallback).  With `decoders` (one NmsDecoder per code) the learned decode
    runs and max_iterations is the fallback's cap."""
    ids = np.ascontiguousarray(code_ids, dtype=np.uint16)
    if ids.size and int(ids.max()) >= len(cs.codes):
        raise ValueError(f"code id {int(ids.max())} outside the code set")
    llr_off, msg_off = cs.offsets(ids)
    llr = np.ascontiguousarray(llr_packed, dtype=np.float32).ravel()
    if llr.size != int(llr_off[-1]):
        raise ValueError("packed LLR length must equal the sum of the frames' n")
    F = ids.size
    msg = np.zeros(int(msg_off[-1]), dtype=np.uint64)
    it = np.zeros(F, dtype=np.uint32)
    conv = np.zeros(F, dtype=np.uint8)
    synd = np.zeros(F, dtype=np.uint8)
    fb = np.zeros(F, dtype=np.uint8)
    if F:
        ctx = cs.ctx
        d_llr = DeviceBuffer.from_array(ctx, llr)
        d_msg = DeviceBuffer(ctx, max(msg.nbytes, 8))
        d_it = DeviceBuffer(ctx, it.nbytes)
        d_conv = DeviceBuffer(ctx, F)
        d_synd = DeviceBuffer(ctx, F)
        d_fb = DeviceBuffer(ctx, F)
        if decoders is None:
            cs.decode_device(ids, d_llr.ptr, max_iterations, d_msg.ptr, d_it.ptr, d_conv.ptr, d_synd.ptr)
        else:
            cs.decode_nms_device(decoders, ids, d_llr.ptr, max_iterations, d_msg.ptr, d_it.ptr, d_conv.ptr,
                                 d_synd.ptr, d_fb.ptr)
        ctx.synchronize()
        msg = d_msg.download(np.uint64, msg.shape)
        it = d_it.download(np.uint32, F)
        conv = d_conv.download(np.uint8, F)
        synd = d_synd.download(np.uint8, F)
        if decoders is not None:
            fb = d_fb.download(np.uint8, F)
    return msg, msg_off, it, conv.astype(bool), synd.astype(bool), fb.astype(bool)


def decode_mixed(cs: CodeSet, code_ids, llr_blocks, max_iterations: int) -> List[DecoderResult]:
    """decode_batch over frames of several codes: llr_blocks[f] has
    n(code_ids[f]) LLRs; results in input order."""
    ids = np.ascontiguousarray(code_ids, dtype=np.uint16)
    blocks = list(llr_blocks)
    if len(blocks) != ids.size:
        raise ValueError("one code id per LLR block")
    for f, b in enumerate(blocks):
        if len(b) != int(cs.n[ids[f]]):
            raise ValueError("LLR length must equal n")
    packed = (np.concatenate([np.asarray(b, dtype=np.float32) for b in blocks])
              if blocks else np.zeros(0, dtype=np.float32))
    msg, msg_off, it, conv, synd, _ = decode_mixed_arrays(cs, ids, packed, max_iterations)
    return [DecoderResult(BitVec(msg[msg_off[f]:msg_off[f + 1]], int(cs.k[ids[f]])), int(it[f]),
                          bool(conv[f]), bool(synd[f])) for f in range(ids.size)]
