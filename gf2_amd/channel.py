"""Host-side channel model: QC-IRA staircase encoding and BPSK/AWGN LLRs.

Follows the reference channel (crates/gf2-coding/src/channel.rs):
bit 0 -> +1, bit 1 -> -1 (:51-57); sigma^2 = 1 / (2 R 10^(EbN0/10))
(:153-158); LLR = (float)(2 y / sigma^2) computed in double (:100-102).
The noise stream is numpy's PCG64 normal generator (the reference's
thread_rng stream is unseeded and cannot be reproduced).
"""
from __future__ import annotations

import numpy as np


def sigma2(ebn0_db: float, rate: float) -> float:
    if not 0.0 < rate <= 1.0:
        raise ValueError("Code rate must be in (0, 1]")
    return 1.0 / (2.0 * rate * 10.0 ** (ebn0_db / 10.0))


def encode_qc_ira(base: np.ndarray, z: int, info: np.ndarray) -> np.ndarray:
    """info [F, K] uint8 -> codewords [F, N] uint8 for a staircase QC-IRA base
    matrix: info block-columns first; parity block-column j has identity
    blocks at block-rows j and j+1, so parity block p_j = s_j xor p_{j-1}
    where s_j is block-row j's info syndrome (a running XOR)."""
    base = np.asarray(base)
    mb, nb = base.shape
    kb = nb - mb
    info = np.asarray(info, dtype=np.uint8)
    F = info.shape[0]
    u = info.reshape(F, kb, z)
    p = np.zeros((F, mb, z), dtype=np.uint8)
    prev = np.zeros((F, z), dtype=np.uint8)
    for br in range(mb):
        s = prev.copy()
        for bc in range(kb):
            sh = int(base[br, bc])
            if sh >= 0:  # check br*z+i sees info bit bc*z + (i+sh) mod z
                s ^= np.roll(u[:, bc, :], -sh, axis=1)
        prev = s
        p[:, br, :] = s
    return np.concatenate([u.reshape(F, kb * z), p.reshape(F, mb * z)], axis=1)


def awgn_llr(codewords: np.ndarray, ebn0_db: float, rate: float,
             rng: np.random.Generator) -> np.ndarray:
    s2 = sigma2(ebn0_db, rate)
    llr = np.empty(codewords.shape, dtype=np.float32)
    chunk = 1 << 22
    flat_c = codewords.reshape(-1)
    flat_l = llr.reshape(-1)
    for a in range(0, flat_c.size, chunk):
        c = flat_c[a:a + chunk]
        y = (1.0 - 2.0 * c.astype(np.float64)) + np.sqrt(s2) * rng.standard_normal(c.size)
        flat_l[a:a + chunk] = (2.0 * y / s2).astype(np.float32)
    return llr


def random_frames(base: np.ndarray, z: int, frames: int, ebn0_db: float, seed: int):
    """(info bits [F,K], codewords [F,N], llr [F,N] float32)."""
    base = np.asarray(base)
    mb, nb = base.shape
    k = (nb - mb) * z
    rng = np.random.default_rng(seed)
    info = rng.integers(0, 2, size=(frames, k), dtype=np.uint8)
    cw = encode_qc_ira(base, z, info)
    llr = awgn_llr(cw, ebn0_db, k / (nb * z), rng)
    return info, cw, llr


def encode_dvb(table, n: int, k: int, q: int, info: np.ndarray, z: int = 360) -> np.ndarray:
    """Systematic DVB-T2 / DVB-S2 LDPC encoding (ETSI EN 302 755 section
    6.1.2.1): each information bit i*z + j is accumulated into the parity
    bits (p + j*q) mod m for every address p of table row i, then
    p_t ^= p_(t-1) for t = 1..m-1.  info [F, k] -> codewords [F, n]."""
    info = np.asarray(info, dtype=np.uint8).reshape(-1, k)
    m = n - k
    par = np.zeros((info.shape[0], m), dtype=np.uint8)
    lanes = np.arange(z, dtype=np.int64) * q
    for i, row in enumerate(table):
        u = info[:, i * z:(i + 1) * z]
        for p in row:
            par[:, (int(p) + lanes) % m] ^= u
    par = np.bitwise_xor.accumulate(par, axis=1)
    return np.concatenate([info, par], axis=1)


def dvb_random_frames(table, n: int, k: int, q: int, frames: int, ebn0_db: float, seed: int, z: int = 360):
    """(info bits [F,k], codewords [F,n], llr [F,n] float32) of a DVB code."""
    rng = np.random.default_rng(seed)
    info = rng.integers(0, 2, size=(frames, k), dtype=np.uint8)
    cw = encode_dvb(table, n, k, q, info, z)
    return info, cw, awgn_llr(cw, ebn0_db, k / n, rng)


def unpack_words(words: np.ndarray, nbits: int) -> np.ndarray:
    """[F, W] LE u64 words -> [F, nbits] uint8."""
    w = np.ascontiguousarray(words, dtype=np.uint64)
    bits = np.unpackbits(w.view(np.uint8).reshape(w.shape[0], -1), axis=1, bitorder="little")
    return bits[:, :nbits]
