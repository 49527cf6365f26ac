"""Test helpers: frame generation (re-exported from gf2_amd.channel, whose
encoder is pinned by test_oracle.test_encoder_produces_codewords) and the
map from device message order to the reference CSR order."""
from __future__ import annotations

import numpy as np

from gf2_amd.channel import (awgn_llr, encode_qc_ira, random_frames, sigma2,  # noqa: F401
                             unpack_words)


def block_to_csr_perm(base: np.ndarray, z: int, row_ptr: np.ndarray) -> np.ndarray:
    """Index map from the device message order (block edge e in block-row-major
    order with block-columns ascending, then lane i) to the reference CSR edge
    order (check, position in its sorted row)."""
    base = np.asarray(base)
    mb, nb = base.shape
    perm = []
    for br in range(mb):
        p = 0
        for bc in range(nb):
            if base[br, bc] >= 0:
                perm.append(row_ptr[br * z + np.arange(z)] + p)
                p += 1
    return np.concatenate(perm).astype(np.int64)


def edge_weights(base: np.ndarray, z: int, w: np.ndarray, oc) -> np.ndarray:
    """[T, nbe] block-edge weights -> oracle [T, nnz] CSR-edge weights."""
    perm = block_to_csr_perm(base, z, oc.row_ptr)
    out = np.zeros((w.shape[0], oc.nnz), dtype=np.float32)
    for t in range(w.shape[0]):
        out[t, perm] = np.repeat(w[t], z)
    return out


def oracle_learned_with_fallback(oc, base: np.ndarray, z: int, w: np.ndarray, llr: np.ndarray, fb_it: int):
    """The learned decoder's results composed from oracle calls: T weighted
    iterations; frames that fail are decoded again from their LLRs by plain
    min-sum for up to fb_it iterations (iterations add up).  Returns
    (msg, iterations, converged, syndrome_ok, fell_back)."""
    T = w.shape[0]
    m1, i1, c1, s1 = oc.decode_batch(llr, T, edge_w=edge_weights(base, z, w, oc), threads=8)
    if fb_it == 0:
        return m1, i1, c1, s1, np.zeros(c1.shape, dtype=bool)
    m2, i2, c2, s2 = oc.decode_batch(llr, fb_it, threads=8)
    fb = ~c1
    return (np.where(fb[:, None], m2, m1), np.where(fb, T + i2, i1), np.where(fb, c2, c1),
            np.where(fb, s2, s1), fb)
