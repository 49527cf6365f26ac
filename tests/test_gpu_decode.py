"""GPU parity of the HIP decoder against the CPU oracle, through the C ABI.

Bar: bit-exact.  Decoded words, iteration counts, converged and syndrome
flags must be identical to the oracle's, and so must every per-iteration
check message and belief (a 1e-5 relative tolerance on the messages is
tightened to equality because the decoder's arithmetic is exact or uses the
same IEEE operations in the same order).
"""
import sys

import numpy as np
import pytest

from gf2_amd import _lib, codes
from gf2_amd.ldpc import (Context, LdpcCode, LdpcDecoder, QuasiCyclicLdpc, decode_batch_arrays,
                          decode_trace_arrays)
from oracle import oracle as orc
from tests import frames

pytestmark = pytest.mark.gpu

PATHS = {"lds": _lib.PATH_LDS, "stream": _lib.PATH_STREAM, "state": _lib.PATH_STATE}
ALL = ["lds", "stream", "state"]


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    orc.build()
    return Context(0)


def gpu_code(ctx, base, z, path="auto"):
    ctx.set_path(PATHS.get(path, _lib.PATH_AUTO))
    return LdpcCode(QuasiCyclicLdpc(base, z), ctx)


def assert_same(gpu, ref, what=""):
    names = ["msg", "iterations", "converged", "syndrome_ok"]
    for name, a, b in zip(names, gpu, ref):
        a, b = np.asarray(a), np.asarray(b)
        if not np.array_equal(a, b):
            bad = np.nonzero(a != b)[0][:10]
            raise AssertionError(f"{what}: {name} differs at frames {bad.tolist()}")


def z1_base(edges, m, n):
    """Any binary H is quasi-cyclic with Z=1 (shift 0 = one, -1 = zero)."""
    base = -np.ones((m, n), dtype=np.int32)
    for r, c in edges:
        base[r, c] = 0
    return base


# ------------------------------------------------------------ reference KATs
SPC = z1_base([(0, 0), (0, 1), (0, 2)], 1, 3)


@pytest.mark.parametrize("path", ALL)
def test_trivial_decode_no_errors(ctx, path):  # core.rs:1150-1165
    say("spc trivial", path)
    code = gpu_code(ctx, SPC, 1, path)
    dec = LdpcDecoder(code)
    r = dec.decode_iterative([10.0, 10.0, 10.0], 10)
    assert r.converged and r.syndrome_check_passed and r.iterations <= 2
    assert r.decoded_bits.count_ones() == 0
    assert dec.last_iteration_count() == r.iterations
    dec.reset()
    assert dec.last_iteration_count() == 0


@pytest.mark.parametrize("path", ALL)
def test_reference_batch_kats(ctx, path):  # core.rs:1188-1212, 1272-1309; backend_integration.rs:44-75
    say("spc batch kats", path)
    code = gpu_code(ctx, SPC, 1, path)
    res = LdpcDecoder.decode_batch(code, [[10.0] * 3, [-10.0, -10.0, 10.0], [10.0] * 3], 10)
    assert [r.converged for r in res] == [True] * 3
    assert [r.decoded_bits.count_ones() for r in res] == [0, 2, 0]
    assert LdpcDecoder.decode_batch(code, [], 10) == []
    blocks = [[10.0] * 3 if i % 2 == 0 else [-10.0, -10.0, 10.0] for i in range(50)]
    res = LdpcDecoder.decode_batch(code, blocks, 10)
    assert len(res) == 50 and all(r.converged for r in res)
    assert all(r.decoded_bits.count_ones() == (0 if i % 2 == 0 else 2) for i, r in enumerate(res))
    # weak-bit case, core.rs:1168-1185
    r = LdpcDecoder.decode_batch(code, [[-5.0, -5.0, 2.0]], 20)[0]
    if r.converged:
        assert r.syndrome_check_passed


@pytest.mark.parametrize("path", ALL)
def test_max_iterations_zero(ctx, path):  # core.rs:906-937
    say("max_it 0", path)
    code = gpu_code(ctx, SPC, 1, path)
    r = LdpcDecoder.decode_batch(code, [[-3.0, 1.0, 2.0]], 0)[0]
    assert r.iterations == 0 and not r.converged and r.syndrome_check_passed
    assert r.decoded_bits.count_ones() == 0


def test_llr_length_mismatch_raises(ctx):  # core.rs:890
    code = gpu_code(ctx, SPC, 1)
    with pytest.raises(ValueError, match="LLR length must equal n"):
        LdpcDecoder.decode_batch(code, [[1.0, 2.0]], 10)


def test_invalid_shift_rejected_by_c_abi(ctx):  # core.rs:589-596
    import ctypes
    lib = _lib.load()
    shifts = np.array([[0, 7]], dtype=np.int32)
    h = ctypes.c_void_p()
    rc = lib.gf2l_code_qc(ctx.handle, 1, 2, 5, _lib.ptr(shifts, ctypes.c_int32), ctypes.byref(h))
    assert rc == _lib.GF2L_ERR_INVALID
    assert "Shift value 7 at position (0,1) must be -1 or in range [0,5)" in \
        lib.gf2l_last_error(ctx.handle).decode()


# ------------------------------------------------ oracle parity, configs
@pytest.mark.parametrize("path", ALL)
def test_n2304_1024_frames_match_oracle(ctx, path):
    spec = codes.n2304_r12()
    base = spec.base_matrix()
    _, _, llr = frames.random_frames(base, spec.z, 1024, 2.0, seed=11)
    say("n2304 1024 frames 2.0dB", path)
    code = gpu_code(ctx, base, spec.z, path)
    gpu = decode_batch_arrays(code, llr, 50)
    ref = orc.Code.from_qc(base, spec.z).decode_batch(llr, 50, threads=8)
    assert_same(gpu, ref, f"n2304/{path}")
    assert gpu[2].mean() > 0.5  # a real operating point, not all-failing


@pytest.mark.parametrize("ebn0", [0.5, 1.0, 3.0])
def test_n2304_snr_sweep_match_oracle(ctx, ebn0):
    spec = codes.n2304_r12()
    base = spec.base_matrix()
    _, _, llr = frames.random_frames(base, spec.z, 256, ebn0, seed=int(ebn0 * 100))
    say("n2304 sweep", ebn0)
    ref = orc.Code.from_qc(base, spec.z).decode_batch(llr, 50, threads=8)
    for path in ALL:
        code = gpu_code(ctx, base, spec.z, path)
        assert_same(decode_batch_arrays(code, llr, 50), ref, f"n2304@{ebn0}/{path}")


@pytest.mark.parametrize("path", ALL)
def test_per_iteration_messages_bit_exact(ctx, path):
    spec = codes.n2304_r12()
    base = spec.base_matrix()
    _, _, llr = frames.random_frames(base, spec.z, 4, 1.0, seed=12)
    T = 12
    say("trace n2304", path)
    code = gpu_code(ctx, base, spec.z, path)
    tc, tb, msg, it, conv, synd = decode_trace_arrays(code, llr, 50, T)
    oc = orc.Code.from_qc(base, spec.z)
    perm = frames.block_to_csr_perm(base, spec.z, oc.row_ptr)
    for f in range(4):
        rc, rb, rmsg, (rit, rconv, rsynd) = oc.decode_trace(llr[f], 50)
        assert it[f] == rit and conv[f] == rconv
        for t in range(min(T, rit)):
            c2v_csr = np.empty(oc.nnz, dtype=np.float32)
            c2v_csr[perm] = tc[t, f]
            assert np.array_equal(c2v_csr.view(np.uint32), rc[t].view(np.uint32)), (f, t)
            assert np.array_equal(tb[t, f].view(np.uint32), rb[t].view(np.uint32)), (f, t)


@pytest.mark.parametrize("path", ALL)
def test_high_degree_and_degree_one_checks(ctx, path):
    # check degrees from 1 to 27 exercise the 8-lane chunk / remainder sign
    # rule and the degree-1 (+0.0) message; LLRs include exact +-0.0
    rng = np.random.default_rng(13)
    z = 16
    mb, nb = 6, 34
    base = -np.ones((mb, nb), dtype=np.int32)
    degs = [1, 9, 16, 17, 25, 27]
    for r, d in enumerate(degs):
        cols = rng.choice(nb, d, replace=False)
        base[r, cols] = rng.integers(0, z, d)
    llr = rng.normal(0.5, 2.0, size=(64, nb * z)).astype(np.float32)
    llr[rng.random(llr.shape) < 0.02] = 0.0
    llr[rng.random(llr.shape) < 0.02] = -0.0
    say("high degree", path)
    code = gpu_code(ctx, base, z, path)
    gpu = decode_batch_arrays(code, llr, 20)
    ref = orc.Code.from_qc(base, z).decode_batch(llr, 20, threads=8)
    assert_same(gpu, ref, f"highdeg/{path}")
    tc, tb, *_ = decode_trace_arrays(code, llr[:3], 20, 5)
    oc = orc.Code.from_qc(base, z)
    perm = frames.block_to_csr_perm(base, z, oc.row_ptr)
    for f in range(3):
        rc, rb, _, (rit, _, _) = oc.decode_trace(llr[f], 20)
        for t in range(min(5, rit)):
            c2v_csr = np.empty(oc.nnz, dtype=np.float32)
            c2v_csr[perm] = tc[t, f]
            assert np.array_equal(c2v_csr.view(np.uint32), rc[t].view(np.uint32)), (f, t)
            assert np.array_equal(tb[t, f].view(np.uint32), rb[t].view(np.uint32)), (f, t)


@pytest.mark.parametrize("path", ["lds", "stream"])
def test_check_degree_32(ctx, path):
    rng = np.random.default_rng(16)
    z = 64
    base = -np.ones((2, 40), dtype=np.int32)
    base[0, rng.choice(40, 32, replace=False)] = rng.integers(0, z, 32)
    base[1, rng.choice(40, 30, replace=False)] = rng.integers(0, z, 30)
    llr = rng.normal(1.0, 1.5, size=(32, 40 * z)).astype(np.float32)
    say("degree 32", path)
    code = gpu_code(ctx, base, z, path)
    assert_same(decode_batch_arrays(code, llr, 15),
                orc.Code.from_qc(base, z).decode_batch(llr, 15, threads=8), f"deg32/{path}")


@pytest.mark.parametrize("path", ALL)
def test_empty_rows_and_columns(ctx, path):
    # an all-zero block-row (checks with no edges) and block-columns outside
    # every check (degree-0 variables) are valid codes
    rng = np.random.default_rng(18)
    z = 72
    base = rng.integers(-1, z, size=(5, 12)).astype(np.int32)
    base[2, :] = -1
    base[:, 4] = -1
    base[:, 9] = -1
    llr = rng.normal(0.8, 1.5, size=(16, 12 * z)).astype(np.float32)
    say("empty rows/cols", path)
    code = gpu_code(ctx, base, z, path)
    assert_same(decode_batch_arrays(code, llr, 12),
                orc.Code.from_qc(base, z).decode_batch(llr, 12, threads=8), f"empty/{path}")


def test_state_path_rejects_degree_above_27(ctx):
    base = np.zeros((1, 28), dtype=np.int32)
    code = gpu_code(ctx, base, 4, "state")
    with pytest.raises(_lib.Gf2lError) as ei:
        decode_batch_arrays(code, np.ones((2, 28 * 4), dtype=np.float32), 5)
    assert ei.value.code == _lib.GF2L_ERR_UNSUPPORTED
    ctx.set_path(_lib.PATH_AUTO)


def test_infinite_llrs(ctx):
    spec = codes.n2304_r12()
    base = spec.base_matrix()
    llr = np.full((2, spec.n), np.inf, dtype=np.float32)
    for path in ALL:
        say("inf llrs", path)
        code = gpu_code(ctx, base, spec.z, path)
        msg, it, conv, synd = decode_batch_arrays(code, llr, 10)
        assert conv.all() and (it == 1).all() and msg.sum() == 0


def test_n64800_small_batch_matches_oracle(ctx):
    spec = codes.n64800_r12()
    base = spec.base_matrix()
    _, _, llr = frames.random_frames(base, spec.z, 24, 1.55, seed=14)
    ref = orc.Code.from_qc(base, spec.z).decode_batch(llr, 50, threads=8)
    for path in ["stream", "state"]:
        say("n64800 24 frames", path)
        code = gpu_code(ctx, base, spec.z, path)
        assert_same(decode_batch_arrays(code, llr, 50), ref, f"n64800/{path}")


def test_n64800_trace_state_path_bit_exact(ctx):
    spec = codes.n64800_r12()
    base = spec.base_matrix()
    _, _, llr = frames.random_frames(base, spec.z, 2, 1.5, seed=17)
    say("n64800 trace state")
    code = gpu_code(ctx, base, spec.z, "state")
    T = 6
    tc, tb, msg, it, conv, synd = decode_trace_arrays(code, llr, 50, T)
    oc = orc.Code.from_qc(base, spec.z)
    perm = frames.block_to_csr_perm(base, spec.z, oc.row_ptr)
    for f in range(2):
        rc, rb, _, (rit, _, _) = oc.decode_trace(llr[f], 50)
        assert it[f] == rit
        for t in range(min(T, rit)):
            c2v_csr = np.empty(oc.nnz, dtype=np.float32)
            c2v_csr[perm] = tc[t, f]
            assert np.array_equal(c2v_csr.view(np.uint32), rc[t].view(np.uint32)), (f, t)
            assert np.array_equal(tb[t, f].view(np.uint32), rb[t].view(np.uint32)), (f, t)


def test_n64800_full_batch_properties(ctx):
    spec = codes.n64800_r12()
    base = spec.base_matrix()
    info, cw, llr = frames.random_frames(base, spec.z, 4096, 1.75, seed=15)
    say("n64800 4096 frames")
    code = gpu_code(ctx, base, spec.z, "auto")
    msg, it, conv, synd = decode_batch_arrays(code, llr, 50)
    bits = frames.unpack_words(msg, spec.k)
    # converged frames carry a zero syndrome and (almost always) the sent bits
    assert conv.mean() > 0.5
    assert np.array_equal(conv, synd)
    assert (it[~conv] == 50).all() and (it[conv] <= 50).all()
    right = (bits == info).all(axis=1)
    assert right[conv].mean() > 0.999
    # composition independence: a reversed subset decodes identically
    sub = np.arange(4095, 4095 - 256, -1)
    m2, it2, c2, s2 = decode_batch_arrays(code, llr[sub], 50)
    assert np.array_equal(m2, msg[sub]) and np.array_equal(it2, it[sub])
    assert np.array_equal(c2, conv[sub])
    # spot-check against the oracle
    ref = orc.Code.from_qc(base, spec.z).decode_batch(llr[:8], 50, threads=8)
    assert_same((msg[:8], it[:8], conv[:8], synd[:8]), ref, "n64800 spot")


def test_mixed_rate_codes_match_oracle(ctx):
    for spec in codes.mixed_twelve():
        base = spec.base_matrix()
        _, _, llr = frames.random_frames(base, spec.z, 6, 2.5, seed=spec.seed)
        say("mixed code", spec.name)
        code = gpu_code(ctx, base, spec.z, "auto")
        gpu = decode_batch_arrays(code, llr, 30)
        ref = orc.Code.from_qc(base, spec.z).decode_batch(llr, 30, threads=8)
        assert_same(gpu, ref, spec.name)


@pytest.mark.parametrize("chunk", [1, 3, 16, 0])
def test_counter_fed_items_match_oracle(ctx, chunk):
    """The check-state kernels give identical results whether items come from
    the shared counters (any chunk size) or a fixed stride."""
    spec = codes.n2304_r12()
    base = spec.base_matrix()
    _, _, llr = frames.random_frames(base, spec.z, 512, 1.75, seed=61)
    say("counter-fed items, chunk", chunk)
    code = gpu_code(ctx, base, spec.z, "state")
    ctx.set_tuning(_lib.TUNE_CHUNK, chunk)
    try:
        gpu = decode_batch_arrays(code, llr, 30)
    finally:
        ctx.set_tuning(_lib.TUNE_CHUNK, -1)
        ctx.set_path(_lib.PATH_AUTO)
    ref = orc.Code.from_qc(base, spec.z).decode_batch(llr, 30, threads=8)
    assert_same(gpu, ref, f"chunk {chunk}")


# --------------------------------------------------- LDS path thread layouts
@pytest.mark.parametrize("groups", [1, 3, 8])
def test_lds_groups_match_oracle(ctx, groups):
    """The LDS path with a forced thread-group count (group g updates
    block-rows and block-columns g, g + groups, ...): the oracle's results."""
    spec = codes.n2304_r12()
    base = spec.base_matrix()
    _, _, llr = frames.random_frames(base, spec.z, 96, 1.8, seed=40 + groups)
    say("lds groups", groups)
    ctx.set_tuning(_lib.TUNE_LDS_GROUPS, groups)
    try:
        gpu = decode_batch_arrays(gpu_code(ctx, base, spec.z, "lds"), llr, 50)
    finally:
        ctx.set_tuning(_lib.TUNE_LDS_GROUPS, 0)
    assert_same(gpu, orc.Code.from_qc(base, spec.z).decode_batch(llr, 50, threads=8), f"groups {groups}")


@pytest.mark.parametrize("z", [7, 1000])
def test_lds_path_lift_edge_cases(ctx, z):
    """z = 7: an odd number of message and belief floats ahead of the
    8-byte-aligned code tables in LDS; z = 1000: more lifted checks per
    block-row than lanes per group (lane i takes checks i and i + 512)."""
    rng = np.random.default_rng(z)
    base = rng.integers(-1, z, size=(3, 6)).astype(np.int32)
    if z % 2 and (int((base >= 0).sum()) * z + 2 * 6 * z) % 2 == 0:  # make the float count odd
        base[0, 1] = -1 if base[0, 1] >= 0 else 3
    llr = rng.normal(1.0, 1.6, size=(24, 6 * z)).astype(np.float32)
    say("lds lift", z)
    code = gpu_code(ctx, base, z, "lds")
    assert ctx.code_path(code) == _lib.PATH_LDS
    assert_same(decode_batch_arrays(code, llr, 30), orc.Code.from_qc(base, z).decode_batch(llr, 30, threads=8),
                f"lds z={z}")
    ctx.set_path(_lib.PATH_AUTO)


# ------------------------------------------- graph-replayed plain iterations
@pytest.mark.parametrize("path,unroll", [("state", 4), ("state", 6), ("stream", 2)])
def test_graph_replays_match_oracle(ctx, path, unroll):
    """With GF2L_TUNE_HBM_GRAPH, plain iterations after the first run as
    replays of a captured graph (cached on the context while the code, sizes
    and buffers stay the same).
    Several calls on the same device buffers, each replaying the graph, with
    an eager decode (profiling on: iterations launched one by one) between
    them; every result equals the oracle's."""
    from gf2_amd.device import DeviceBuffer
    spec = codes.n2304_r12()
    base = spec.base_matrix()
    F, kw = 64, (spec.k + 63) // 64
    code = gpu_code(ctx, base, spec.z, path)
    oc = orc.Code.from_qc(base, spec.z)
    lib = _lib.load()
    ctx.set_tuning(_lib.TUNE_GRAPH_ITERS, unroll)
    ctx.set_tuning(_lib.TUNE_HBM_GRAPH, 1)
    d_llr = DeviceBuffer(ctx, F * spec.n * 4)
    outs = [DeviceBuffer(ctx, F * kw * 8), DeviceBuffer(ctx, F * 4), DeviceBuffer(ctx, F), DeviceBuffer(ctx, F)]
    try:
        for call, ebn0 in enumerate([1.2, 2.0, 1.6, 2.4]):
            _, _, llr = frames.random_frames(base, spec.z, F, ebn0, seed=70 + call)
            d_llr.upload(llr)
            eager = call % 2 == 1
            say("graph replays", path, unroll, call, "eager" if eager else "graph")
            if eager:
                _lib.check(ctx.handle, lib.gf2l_profile_enable(ctx.handle, 1))
            _lib.check(ctx.handle, lib.gf2l_decode_batch_device(ctx.handle, code.handle, d_llr.ptr, F, 50,
                                                                *[o.ptr for o in outs]))
            ctx.synchronize()
            if eager:
                _lib.check(ctx.handle, lib.gf2l_profile_enable(ctx.handle, 0))
            gpu = (outs[0].download(np.uint64, (F, kw)), outs[1].download(np.uint32, F),
                   outs[2].download(np.uint8, F).astype(bool), outs[3].download(np.uint8, F).astype(bool))
            assert_same(gpu, oc.decode_batch(llr, 50, threads=8), f"{path} call {call}")
    finally:
        ctx.set_tuning(_lib.TUNE_GRAPH_ITERS, 4)
        ctx.set_tuning(_lib.TUNE_HBM_GRAPH, 0)
        ctx.set_path(_lib.PATH_AUTO)
