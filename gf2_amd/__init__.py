"""gf2_amd: MI355X (gfx950) batched LDPC belief-propagation decoding.

A drop-in for the reference's `LdpcDecoder::decode_batch` path
(crates/gf2-coding/src/ldpc/core.rs:729-756).  The decoder runs in
hand-written HIP kernels in libgf2ldpc.so, reached through the C ABI in
include/gf2ldpc.h; this package mirrors the reference API on top of it.
"""
from .ldpc import (BitVec, CodeRate, Context, DecoderResult, LdpcCode, LdpcDecoder, QuasiCyclicLdpc,
                   decode_batch_arrays, decode_trace_arrays, default_context,
                   qc_ira_base_matrix)

__all__ = [
    "BitVec", "CodeRate", "Context", "DecoderResult", "LdpcCode", "LdpcDecoder", "QuasiCyclicLdpc",
    "decode_batch_arrays", "decode_trace_arrays", "default_context", "qc_ira_base_matrix",
]
