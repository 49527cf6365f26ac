"""Seeded quasi-cyclic IRA codes used by the tests and benchmarks.

All codes have lift Z, info block-columns first and the dual-diagonal
staircase parity part (see gf2l_qc_ira_shifts in include/gf2ldpc.h).

Degree profiles of the Z=360 codes are the info-column weights of the
DVB-T2 LDPC tables (ETSI EN 302 755, as tabulated in the reference's
crates/gf2-coding/src/ldpc/dvb_t2/dvb_t2_matrices.rs: one table row per
group of 360 info bits, its length being that group's column weight),
placed on a block-level staircase so the whole matrix is quasi-cyclic:

  N=64800: K 32400  36 x w8  + 54 x w3    (rate 1/2)
           K 38880  36 x w12 + 72 x w3    (3/5)
           K 43200  12 x w13 + 108 x w3   (2/3)
           K 48600  15 x w12 + 120 x w3   (3/4)
           K 51840  18 x w11 + 126 x w3   (4/5)
           K 54000  15 x w13 + 135 x w3   (5/6)
  N=16200: K 7200   5 x w8  + 15 x w3     (nominal 1/2)
           K 9720   5 x w12 + 22 x w3     (3/5)
           K 10800  3 x w13 + 27 x w3     (2/3)
           K 11880  1 x w12 + 32 x w3     (3/4)
           K 12600  35 x w3               (4/5)
           K 13320  1 x w13 + 36 x w3     (5/6)
A block-level QC graph holds at most one circulant per block, so a column
weight above the number of block-rows is clipped to mb (only the
N=16200, K=13320 code is affected: its w13 column becomes w8).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Dict, List, Tuple

import numpy as np

from .ldpc import QuasiCyclicLdpc, qc_ira_base_matrix

# (N, K) -> [(weight, count), ...] in column order (heavy columns first)
PROFILES: Dict[Tuple[int, int], List[Tuple[int, int]]] = {
    (2304, 1152): [(6, 4), (3, 8)],
    (64800, 32400): [(8, 36), (3, 54)],
    (64800, 38880): [(12, 36), (3, 72)],
    (64800, 43200): [(13, 12), (3, 108)],
    (64800, 48600): [(12, 15), (3, 120)],
    (64800, 51840): [(11, 18), (3, 126)],
    (64800, 54000): [(13, 15), (3, 135)],
    (16200, 7200): [(8, 5), (3, 15)],
    (16200, 9720): [(12, 5), (3, 22)],
    (16200, 10800): [(13, 3), (3, 27)],
    (16200, 11880): [(12, 1), (3, 32)],
    (16200, 12600): [(3, 35)],
    (16200, 13320): [(13, 1), (3, 36)],
}

DEFAULT_SEED = 20241

# Eb/N0 (dB) where plain min-sum BP (50 iterations) has FER about 1e-2 on the
# twelve Z=360 codes of mixed_twelve(): gf2_amd.fer.calibrate, bisection over
# 8 steps of 2048 device-generated frames each (seed 5) on an MI355X.
FER_1E2_EBN0: Dict[Tuple[int, int], float] = {
    (16200, 7200): 2.628, (16200, 9720): 2.362, (16200, 10800): 2.068,
    (16200, 11880): 2.553, (16200, 12600): 2.895, (16200, 13320): 3.182,
    (64800, 32400): 1.579, (64800, 38880): 2.212, (64800, 43200): 1.972,
    (64800, 48600): 2.415, (64800, 51840): 2.779, (64800, 54000): 3.063,
}


@dataclass(frozen=True)
class QcIraSpec:
    n: int
    k: int
    z: int
    seed: int = DEFAULT_SEED

    @property
    def nb(self) -> int:
        return self.n // self.z

    @property
    def mb(self) -> int:
        return (self.n - self.k) // self.z

    def weights(self) -> List[int]:
        w: List[int] = []
        for weight, count in PROFILES[(self.n, self.k)]:
            w.extend([min(weight, self.mb)] * count)
        return w

    def base_matrix(self) -> np.ndarray:
        return qc_ira_base_matrix(self.seed, self.mb, self.nb, self.z, self.weights())

    def qc(self) -> QuasiCyclicLdpc:
        return QuasiCyclicLdpc(self.base_matrix(), self.z)

    @property
    def name(self) -> str:
        return f"qc-ira-n{self.n}-k{self.k}-z{self.z}-s{self.seed}"


def n2304_r12(seed: int = DEFAULT_SEED) -> QcIraSpec:
    """N=2304, K=1152, Z=96 (12x24 base graph, 71 block edges)."""
    return QcIraSpec(2304, 1152, 96, seed)


def n64800_r12(seed: int = DEFAULT_SEED) -> QcIraSpec:
    """N=64800, K=32400, Z=360 (90x180 base graph, 629 block edges)."""
    return QcIraSpec(64800, 32400, 360, seed)


def mixed_twelve(seed: int = DEFAULT_SEED) -> List[QcIraSpec]:
    """The twelve Z=360 codes: N in {16200, 64800} x six rates."""
    return [QcIraSpec(n, k, 360, seed + i)
            for i, (n, k) in enumerate(sorted(p for p in PROFILES if p[0] != 2304))]


@dataclass(frozen=True)
class DvbT2Spec:
    """One DVB-T2 LDPC code: ETSI EN 302 755 parameters and address table, as
    compiled into libgf2ldpc (DvbParams::for_code, crates/gf2-coding/src/ldpc/
    dvb_t2/params.rs:56-86, and dvb_t2/dvb_t2_matrices.rs)."""
    frame: int            # 0 short (16200), 1 normal (64800)
    rate: int             # 0 = 1/2, 1 = 3/5, 2 = 2/3, 3 = 3/4, 4 = 4/5, 5 = 5/6
    n: int
    k: int
    q: int
    table: Tuple[Tuple[int, ...], ...]
    z: int = 360

    @property
    def m(self) -> int:
        return self.n - self.k

    @property
    def name(self) -> str:
        return f"dvb-t2-{'normal' if self.frame else 'short'}-{DVB_RATE_NAMES[self.rate]}"

    def edges(self) -> np.ndarray:
        """(check, variable) pairs of the code as build_dvb_edges lists them
        (dvb_t2/builder.rs:74-118), e.g. for LdpcCode.from_edges (CSR path)
        or a GF2DATA export."""
        j = np.arange(self.z, dtype=np.int64)
        parts = [np.stack([(p + j * self.q) % self.m, i * self.z + j], axis=1)
                 for i, row in enumerate(self.table) for p in row]
        t = np.arange(self.m, dtype=np.int64)
        parts.append(np.stack([t, self.k + t], axis=1))
        parts.append(np.stack([t[1:], self.k + t[1:] - 1], axis=1))
        return np.concatenate(parts)


DVB_RATE_NAMES = ["1/2", "3/5", "2/3", "3/4", "4/5", "5/6"]


def dvb_t2(frame: int, rate: int) -> DvbT2Spec:
    """Parameters and table of a DVB-T2 code, read from the library (no GPU
    needed)."""
    import ctypes

    from . import _lib
    lib = _lib.load()
    n, k, q, rows, entries = (ctypes.c_uint32() for _ in range(5))
    rc = lib.gf2l_dvb_t2_params(int(frame), int(rate), ctypes.byref(n), ctypes.byref(k), ctypes.byref(q),
                                ctypes.byref(rows), ctypes.byref(entries))
    if rc != 0:
        raise ValueError("DVB-T2 frame size must be 0 (short) or 1 (normal) and rate 0..5")
    lens = np.zeros(rows.value, dtype=np.uint32)
    addr = np.zeros(entries.value, dtype=np.uint32)
    lib.gf2l_dvb_t2_table(int(frame), int(rate), lens.ctypes.data, addr.ctypes.data)
    ends = np.cumsum(lens)
    table = tuple(tuple(int(a) for a in addr[e - l:e]) for l, e in zip(lens, ends))
    return DvbT2Spec(int(frame), int(rate), n.value, k.value, q.value, table)


def dvb_t2_all() -> List[DvbT2Spec]:
    return [dvb_t2(f, r) for f in (0, 1) for r in range(6)]
