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
