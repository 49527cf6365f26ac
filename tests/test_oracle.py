"""Pins the CPU oracle before it is trusted as the checker.

Each oracle function is checked against the reference's own known-answer
tests (ported below with their file:line), against the structural port of
the reference's rayon path (ref_port.cpp, real AVX2 intrinsics), and against
what the mathematics fixes (dense GF(2) products, brute-force max-log APP on
cycle-free graphs, finite differences, closed forms).

  or_qc_to_edges        qc_ldpc_tests.rs KATs, dense expansion
  or_code_from_edges    sparse.rs:140-160 doc example, sparse_dual_ldpc.rs KATs
  or_minsum_avx2        llr.rs / llr_simd_integration.rs KATs, closed form,
                        hand-derived -0.0 lane cases, port cross-check
  or_minsum_scalar      llr.rs KATs, closed form
  or_syndrome           sparse_dual_ldpc.rs:95-147 KAT, dense numpy H x
  or_decode_iterative   core.rs decoder tests, backend_integration.rs,
                        brute-force max-log APP on trees, port cross-check
  or_decode_batch       core.rs:1272-1309 batch/empty KATs, port cross-check
  or_nms_loss_grad      finite differences; weights 1 == plain min-sum
  or_philox4x32         Random123 known-answer vectors
"""
import itertools
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from oracle import oracle as orc  # noqa: E402
from tests import frames  # noqa: E402


@pytest.fixture(scope="module", autouse=True)
def _built():
    orc.build()


def dense_h(code: orc.Code) -> np.ndarray:
    h = np.zeros((code.m, code.n), dtype=np.uint8)
    for r in range(code.m):
        h[r, code.row_idx[code.row_ptr[r]:code.row_ptr[r + 1]]] = 1
    return h


# ---------------------------------------------------------------- QC expansion
def test_circulant_identity():  # qc_ldpc_tests.rs:60-72
    e = orc.qc_to_edges(np.array([[0]]), 4)
    assert sorted(map(tuple, e)) == [(0, 0), (1, 1), (2, 2), (3, 3)]


def test_circulant_shift_right():  # qc_ldpc_tests.rs:75-92
    e = orc.qc_to_edges(np.array([[1]]), 4)
    assert set(map(tuple, e)) == {(0, 1), (1, 2), (2, 3), (3, 0)}


def test_circulant_shift_multiple():  # qc_ldpc_tests.rs:95-106
    e = orc.qc_to_edges(np.array([[2]]), 5)
    assert set(map(tuple, e)) == {(0, 2), (1, 3), (2, 4), (3, 0), (4, 1)}


def test_circulant_with_offset():  # qc_ldpc_tests.rs:109-118
    base = np.full((2, 3), -1)
    base[1, 2] = 1
    e = orc.qc_to_edges(base, 3)
    assert set(map(tuple, e)) == {(3, 7), (4, 8), (5, 6)}


def test_qc_zero_blocks_and_dims():  # qc_ldpc_tests.rs:31-58, 139-176
    e = orc.qc_to_edges(np.array([[0, -1], [-1, 1]]), 4)
    assert len(e) == 8
    e = orc.qc_to_edges(np.array([[0, 1, -1], [1, -1, 2]]), 4)
    assert len(e) == 2 * 2 * 4
    code = orc.Code.from_qc(np.array([[0, 1, 2, 3, 4], [3, 0, 1, 2, 4], [2, 3, 0, 1, 4]]), 5)
    assert (code.m, code.n, code.k) == (15, 25, 10)
    assert code.syndrome_weight(np.zeros(25, dtype=np.uint8)) == 0


def test_qc_invalid_shift():  # core.rs:589-596 panics
    with pytest.raises(ValueError):
        orc.qc_to_edges(np.array([[0, 5]]), 5)


def test_qc_expansion_matches_dense_circulants():
    rng = np.random.default_rng(1)
    base = rng.integers(-1, 7, size=(3, 5))
    z = 7
    code = orc.Code.from_qc(base, z)
    h = dense_h(code)
    for br, bc in itertools.product(range(3), range(5)):
        blk = h[br * z:(br + 1) * z, bc * z:(bc + 1) * z]
        s = base[br, bc]
        want = np.zeros((z, z), dtype=np.uint8) if s < 0 else np.roll(np.eye(z, dtype=np.uint8), s, axis=1)
        assert np.array_equal(blk, want)


# ---------------------------------------------------------------- sparse build
def test_from_coo_duplicates_cancel():  # sparse.rs:150-160 doc example
    c = orc.Code(2, 3, np.array([(0, 1), (0, 1), (1, 2)]))
    assert c.nnz == 1
    assert list(c.row_idx) == [2]


def test_regular_36_degrees():  # sparse_dual_ldpc.rs:7-50
    coo = [(0, 0), (1, 0), (2, 0), (0, 1), (1, 1), (3, 1), (0, 2), (2, 2), (3, 2), (1, 3), (2, 3),
           (3, 3), (0, 4), (1, 4), (2, 4), (0, 5), (1, 5), (3, 5), (0, 6), (2, 6), (3, 6), (1, 7),
           (2, 7), (3, 7)]
    c = orc.Code(4, 8, np.array(coo))
    assert c.nnz == 24
    assert all(np.diff(c.row_ptr) == 6)
    assert all(np.diff(c.col_ptr) == 3)


def test_neighbour_order():  # sparse_dual_ldpc.rs:57-93
    coo = [(0, 0), (0, 2), (0, 4), (1, 1), (1, 2), (1, 5), (2, 0), (2, 3), (2, 5)]
    c = orc.Code(3, 6, np.array(coo))
    row = lambda r: list(c.row_idx[c.row_ptr[r]:c.row_ptr[r + 1]])
    col = lambda v: list(c.col_idx[c.col_ptr[v]:c.col_ptr[v + 1]])
    assert row(0) == [0, 2, 4] and row(1) == [1, 2, 5] and row(2) == [0, 3, 5]
    assert col(0) == [0, 2] and col(2) == [0, 1] and col(5) == [1, 2]


def test_syndrome_kat():  # sparse_dual_ldpc.rs:95-125
    coo = [(0, 0), (0, 1), (0, 2), (1, 2), (1, 3), (1, 4), (2, 0), (2, 4), (2, 5)]
    c = orc.Code(3, 6, np.array(coo))
    assert c.syndrome_weight(np.array([1, 1, 0, 0, 0, 1])) == 0
    assert c.syndrome_weight(np.array([1, 0, 0, 0, 0, 0])) == 2


def test_syndrome_dense_and_linear():
    rng = np.random.default_rng(2)
    base = rng.integers(-1, 11, size=(4, 9))
    code = orc.Code.from_qc(base, 11)
    h = dense_h(code)
    for _ in range(20):
        x = rng.integers(0, 2, code.n)
        y = rng.integers(0, 2, code.n)
        assert code.syndrome_weight(x) == int(((h @ x) % 2).sum())
        sx, sy, sxy = (h @ x) % 2, (h @ y) % 2, (h @ (x ^ y)) % 2
        assert np.array_equal(sxy, sx ^ sy)


# ---------------------------------------------------------------- min-sum
@pytest.mark.parametrize("fn", [orc.minsum_avx2, orc.minsum_scalar])
def test_minsum_kats(fn):  # llr.rs:230-244 doc, 552-577, 630-673
    assert fn([3.0, 2.0, 4.0]) == 2.0
    assert fn([3.0, -2.0, 4.0]) == -2.0
    assert fn([-1.0, -2.0, -3.0]) == -1.0
    assert fn([1.5, -2.3, 4.7, -0.8]) == np.float32(0.8)
    assert fn([3.0, -2.0, 1.0]) == -1.0
    assert fn([4.0, 6.0]) == 4.0


def test_minsum_avx2_vs_scalar_long():  # gf2-kernels-simd llr.rs:330-629 (8 and 20 elems)
    rng = np.random.default_rng(3)
    for n in [1, 7, 8, 9, 15, 16, 20, 33]:
        for _ in range(50):
            x = rng.normal(0, 3, n).astype(np.float32)
            a, s = orc.minsum_avx2(x), orc.minsum_scalar(x)
            assert a == s
            # closed form: magnitude = min |x|, sign = parity of negatives
            want = np.float32((-1.0) ** int((x < 0).sum()) * np.abs(x).min())
            assert a == want


def test_minsum_negative_zero_lane_semantics():
    # Inside full 8-lane chunks the sign bit decides (-0.0 counts negative);
    # in the scalar remainder `x < 0` decides (-0.0 counts positive),
    # crates/gf2-kernels-simd/src/llr.rs:146-150 vs :174.
    lane = [-0.0] + [1.0] * 7          # -0.0 inside the chunk
    assert np.signbit(orc.minsum_avx2(lane))            # -> -0.0
    rem = [1.0] * 8 + [-0.0]           # -0.0 in the remainder
    r = orc.minsum_avx2(rem)
    assert r == 0.0 and not np.signbit(r)
    short = [-0.0, 2.0]                # n < 8: all remainder
    assert not np.signbit(orc.minsum_avx2(short))
    # the scalar rule uses >= 0: -0.0 is positive
    assert not np.signbit(orc.minsum_scalar([-0.0, 2.0]))


def test_minsum_empty():
    assert orc.minsum_avx2([]) == 0.0  # gf2-kernels-simd llr.rs:123-125


# ---------------------------------------------------------------- decoder KATs
def spc():
    return orc.Code(1, 3, np.array([(0, 0), (0, 1), (0, 2)]))


def test_trivial_decode_no_errors():  # core.rs:1150-1165
    msg, it, conv, synd = spc().decode_batch(np.array([[10.0, 10.0, 10.0]]), 10)
    assert conv[0] and synd[0] and it[0] <= 2
    assert frames.unpack_words(msg, 2).sum() == 0


def test_decode_with_single_error():  # core.rs:1168-1185
    c = spc()
    tc, tb, msg, (it, conv, synd) = c.decode_trace(np.array([-5.0, -5.0, 2.0]), 20)
    if conv:
        assert synd
        hard = (tb[it - 1] < 0).astype(np.uint8)
        assert hard.sum() == 2


def test_consecutive_and_batch():  # core.rs:1188-1212, 1272-1294
    c = spc()
    llr = np.array([[10, 10, 10], [-10, -10, 10], [10, 10, 10]], dtype=np.float32)
    msg, it, conv, synd = c.decode_batch(llr, 10)
    assert conv.all()
    ones = frames.unpack_words(msg, 2).sum(axis=1)
    # message = first k=2 bits; [1,1,0] has two ones in the message part
    assert list(ones) == [0, 2, 0]


def test_empty_batch():  # core.rs:1297-1309
    msg, it, conv, synd = spc().decode_batch(np.zeros((0, 3), dtype=np.float32), 10)
    assert len(it) == 0


def test_alternating_50_batch():  # tests/backend_integration.rs:44-75
    llr = np.array([[10, 10, 10] if i % 2 == 0 else [-10, -10, 10] for i in range(50)],
                   dtype=np.float32)
    msg, it, conv, synd = spc().decode_batch(llr, 10)
    assert conv.all()
    ones = frames.unpack_words(msg, 2).sum(axis=1)
    assert all(ones[i] == (0 if i % 2 == 0 else 2) for i in range(50))


def test_max_it_zero_semantics():  # core.rs:906-937 with a fresh decoder
    msg, it, conv, synd = spc().decode_batch(np.array([[-3.0, 1.0, 2.0]]), 0)
    assert it[0] == 0 and not conv[0] and synd[0] and msg.sum() == 0


def test_all_zero_codeword_converges_first_iteration():
    spec_base = np.array([[0, 1, 2, -1, 0, -1], [1, -1, 0, 2, 0, 0], [-1, 2, 1, 0, -1, 0]])
    c = orc.Code.from_qc(spec_base, 3)
    msg, it, conv, synd = c.decode_batch(np.full((1, c.n), 10.0, dtype=np.float32), 50)
    assert conv[0] and it[0] == 1


# -------------------------------------------- brute force: max-log APP on trees
def _tree_codes():
    yield orc.Code(1, 3, np.array([(0, 0), (0, 1), (0, 2)]))
    # two checks sharing one variable, plus a pendant check: a tree
    yield orc.Code(3, 7, np.array([(0, 0), (0, 1), (0, 2), (1, 2), (1, 3), (1, 4), (2, 4),
                                   (2, 5), (2, 6)]))
    yield orc.Code(2, 6, np.array([(0, 0), (0, 1), (0, 2), (0, 3), (1, 3), (1, 4), (1, 5)]))


@pytest.mark.parametrize("ci", [0, 1, 2])
def test_minsum_tree_fixed_point_is_maxlog_app(ci):
    code = list(_tree_codes())[ci]
    h = dense_h(code)
    cws = [np.array(c) for c in itertools.product([0, 1], repeat=code.n)
           if not ((h @ np.array(c)) % 2).any()]
    rng = np.random.default_rng(10 + ci)
    for _ in range(20):
        llr = rng.normal(1.0, 2.0, code.n).astype(np.float32)
        tc, tb, msg, _ = code.decode_trace(llr, 12, no_early_stop=True)
        for v in range(code.n):
            m1 = min(float(np.dot(c, llr)) for c in cws if c[v] == 1)
            m0 = min(float(np.dot(c, llr)) for c in cws if c[v] == 0)
            assert abs(tb[-1, v] - (m1 - m0)) <= 1e-4 * (1 + abs(m1 - m0))


# ------------------------------------------ oracle vs structural port (AVX2)
@pytest.mark.parametrize("ebn0", [1.0, 2.0, 3.0])
def test_oracle_matches_rayon_port(ebn0):
    rng = np.random.default_rng(4)
    # rows with degree > 9 exercise full 8-lane chunks in the AVX2 kernel
    base = rng.integers(-1, 12, size=(4, 14))
    base[:, -4:] = -1
    for j in range(4):
        base[j, 10 + j] = 0
        if j + 1 < 4:
            base[j + 1, 10 + j] = 0
    z = 12
    code = orc.Code.from_qc(base, z)
    _, cw, llr = frames.random_frames(base, z, 64, ebn0, seed=int(ebn0 * 10))
    a = code.decode_batch(llr, 30, threads=4)
    b = code.port_decode_batch(llr, 30, threads=4)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_batch_results_independent_of_threads_and_order():
    base = np.array([[0, 3, -1, 5, 0, -1], [2, -1, 4, 1, 0, 0], [-1, 1, 2, 0, -1, 0]])
    code = orc.Code.from_qc(base, 8)
    _, _, llr = frames.random_frames(base, 8, 40, 1.5, seed=5)
    a = code.decode_batch(llr, 25, threads=1)
    b = code.decode_batch(llr[::-1].copy(), 25, threads=3)
    for x, y in zip(a, b):
        assert np.array_equal(x, y[::-1])


# ---------------------------------------------------------------- NMS oracle
def test_nms_unit_weights_equal_plain():
    base = np.array([[0, 3, -1, 5, 0, -1], [2, -1, 4, 1, 0, 0], [-1, 1, 2, 0, -1, 0]])
    code = orc.Code.from_qc(base, 8)
    _, _, llr = frames.random_frames(base, 8, 20, 1.0, seed=6)
    w = np.ones((30, code.nnz), dtype=np.float32)
    a = code.decode_batch(llr, 30)
    b = code.decode_batch(llr, 30, edge_w=w)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_nms_alpha_rule():  # llr.rs:352-355, 676-689: alpha * minsum
    c = spc()
    w = np.full((1, c.nnz), 0.875, dtype=np.float32)
    tc, tb, _, _ = c.decode_trace(np.array([4.0, 6.0, 5.0]), 1, edge_w=w)
    # c2v to var 2 = 0.875 * minsum([4, 6]) = 0.875 * 4
    assert tc[0, 2] == np.float32(0.875) * np.float32(4.0)


def test_nms_gradient_finite_difference():
    base = np.array([[0, 1, -1, 0, -1], [2, -1, 1, 0, 0]])
    z = 3
    code = orc.Code.from_qc(base, z)
    rng = np.random.default_rng(7)
    info, cw, llr = frames.random_frames(base, z, 1, 0.5, seed=8)
    T = 3
    w = rng.uniform(0.6, 1.2, size=(T, code.nnz))
    loss, g = code.nms_loss_grad(llr[0], cw[0], T, w)
    eps = 1e-6
    checked = 0
    for t in range(T):
        for e in range(code.nnz):
            wp, wm = w.copy(), w.copy()
            wp[t, e] += eps
            wm[t, e] -= eps
            lp, _ = code.nms_loss_grad(llr[0], cw[0], T, wp)
            lm, _ = code.nms_loss_grad(llr[0], cw[0], T, wm)
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - g[t, e]) <= 1e-6 + 1e-4 * abs(fd), (t, e, fd, g[t, e])
            checked += 1
    assert checked == T * code.nnz
    assert np.abs(g).sum() > 0


# ---------------------------------------------------------------- Philox KAT
def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32 with 10 rounds
    assert list(orc.philox4x32([0, 0, 0, 0], [0, 0])) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c,
                                                           0x9b00dbd8]
    assert list(orc.philox4x32([0xffffffff] * 4, [0xffffffff] * 2)) == [
        0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert list(orc.philox4x32([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
                               [0xa4093822, 0x299f31d0])) == [0xd16cfe09, 0x94fdcceb,
                                                               0x5001e420, 0x24126ea1]


# ---------------------------------------------------------------- frame helpers
def test_encoder_produces_codewords():
    from gf2_amd import codes
    spec = codes.n2304_r12()
    base = spec.base_matrix()
    code = orc.Code.from_qc(base, spec.z)
    info, cw, llr = frames.random_frames(base, spec.z, 8, 2.0, seed=9)
    assert np.array_equal(cw[:, :spec.k], info)
    for f in range(8):
        assert code.syndrome_weight(cw[f]) == 0


# ---------------------------------------------------------------- frame generator
def test_framegen_restatement_is_a_coded_awgn_channel():
    from gf2_amd import codes
    spec = codes.n2304_r12()
    base = spec.base_matrix()
    code = orc.Code.from_qc(base, spec.z)
    zw = (spec.z + 63) // 64
    s2 = frames.sigma2(2.0, spec.k / spec.n)
    ys = []
    for g in range(6):
        llr, cw = orc.framegen(base, spec.z, 99, g, 2.0, 2.0)
        bits = frames.unpack_words(cw[None, :], spec.n)[0]
        assert code.syndrome_weight(bits) == 0          # a codeword of H
        # message bit 70 of block 0 comes from Philox({0*zw + 1, g, 0, 0})
        o = orc.philox4x32([0 * zw + 70 // 64, g, 0, 0], [99, 0])
        w = int(o[0]) | (int(o[1]) << 32)
        assert bits[70] == (w >> (70 % 64)) & 1
        ys.append(llr * (1 - 2 * bits.astype(np.float32)))
    y = np.concatenate(ys).astype(np.float64)
    # LLR given the sent symbol is N(2/s2, 4/s2)
    assert abs(y.mean() - 2 / s2) < 0.05 * (2 / s2)
    assert abs(y.var() - 4 / s2) < 0.05 * (4 / s2)
    # different frames and seeds give different draws
    a, _ = orc.framegen(base, spec.z, 99, 0, 2.0, 2.0)
    b, _ = orc.framegen(base, spec.z, 99, 1, 2.0, 2.0)
    c, _ = orc.framegen(base, spec.z, 98, 0, 2.0, 2.0)
    assert not np.array_equal(a, b) and not np.array_equal(a, c)
