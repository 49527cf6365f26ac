"""DVB-T2 codes, GF2DATA-loaded codes and the SDR-style API on the GPU,
bit-exact against the oracle."""
import ctypes

import numpy as np
import pytest

from gf2_amd import _lib, codes, gf2data
from gf2_amd.channel import dvb_random_frames, encode_dvb
from gf2_amd.ldpc import CodeRate, Context, LdpcCode, LdpcDecoder, decode_batch_arrays
from oracle import oracle as orc
from tests import frames

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    orc.build()
    return Context(0)


def same(gpu, ref, what):
    for name, a, b in zip(["msg", "iterations", "converged", "syndrome_ok"], gpu, ref):
        a, b = np.asarray(a), np.asarray(b)
        assert np.array_equal(a, b), f"{what}: {name} differs at {np.nonzero(a != b)[0][:8]}"


def oracle_code(spec):
    return orc.Code(spec.m, spec.n, orc.dvb_edges(spec.table, spec.n, spec.k, spec.q))


@pytest.mark.parametrize("rate", range(6))
def test_dvb_short_codes_match_oracle(ctx, rate):
    spec = codes.dvb_t2(0, rate)
    print(f"[dvb] {spec.name}", flush=True)
    code = LdpcCode.dvb_t2_short(CodeRate(rate), ctx)
    assert (code.n(), code.k(), code.m()) == (spec.n, spec.k, spec.m)
    oc = oracle_code(spec)
    assert code.nnz() == oc.nnz
    # Eb/N0 around each code's waterfall: a mix of converging and failing frames
    ebn0 = [1.0, 1.6, 2.0, 2.6, 3.1, 3.6][rate]
    _, _, llr = dvb_random_frames(spec.table, spec.n, spec.k, spec.q, 48, ebn0, seed=40 + rate)
    same(decode_batch_arrays(code, llr, 50), oc.decode_batch(llr, 50, threads=8), spec.name)


@pytest.mark.parametrize("rate", range(6))
def test_dvb_normal_codes_match_oracle(ctx, rate):
    spec = codes.dvb_t2(1, rate)
    print(f"[dvb] {spec.name}", flush=True)
    code = LdpcCode.dvb_t2_normal(CodeRate(rate), ctx)
    oc = oracle_code(spec)
    assert code.nnz() == oc.nnz
    # around each code's min-sum waterfall: converging and failing frames
    ebn0 = [1.55, 2.2, 2.0, 2.45, 2.85, 3.15][rate]
    _, _, llr = dvb_random_frames(spec.table, spec.n, spec.k, spec.q, 24, ebn0, seed=7 + rate)
    same(decode_batch_arrays(code, llr, 30), oc.decode_batch(llr, 30, threads=8), spec.name)


@pytest.mark.parametrize("frame,rate", [(0, 1), (0, 3), (0, 4), (0, 5), (1, 1), (1, 5)])
def test_dvb_signed_zero_and_infinite_llrs(ctx, frame, rate):
    """-0.0 inputs make the min-sum sign rule depend on each input's place in
    the check's variable order, which inside a block of several circulants
    changes from lane to lane; +-inf inputs saturate."""
    spec = codes.dvb_t2(frame, rate)
    print(f"[dvb +-0] {spec.name}", flush=True)
    code = (LdpcCode.dvb_t2_normal if frame else LdpcCode.dvb_t2_short)(CodeRate(rate), ctx)
    rng = np.random.default_rng(100 + 10 * frame + rate)
    _, _, llr = dvb_random_frames(spec.table, spec.n, spec.k, spec.q, 16, 3.5, seed=60 + rate)
    u = rng.random(llr.shape)
    llr[u < 0.10] = -0.0
    llr[(u >= 0.10) & (u < 0.15)] = 0.0
    llr[(u >= 0.15) & (u < 0.16)] = np.inf
    llr[(u >= 0.16) & (u < 0.17)] = -np.inf
    assert np.signbit(llr[llr == 0]).sum() > 0
    same(decode_batch_arrays(code, llr, 20), oracle_code(spec).decode_batch(llr, 20, threads=8), spec.name)


def test_dvb_quasi_cyclic_form_limits(ctx):
    from gf2_amd.ldpc import decode_trace_arrays
    from gf2_amd.mixed import CodeSet
    from gf2_amd.nms import NmsDecoder
    code = LdpcCode.dvb_t2_short(CodeRate.RATE_2_3, ctx)
    llr = np.full((2, code.n()), 4.0, dtype=np.float32)
    with pytest.raises(_lib.Gf2lError, match="not DVB"):
        NmsDecoder(code, 10)
    with pytest.raises(_lib.Gf2lError, match="not DVB"):
        CodeSet([code])
    with pytest.raises(_lib.Gf2lError, match="DVB"):
        decode_trace_arrays(code, llr, 5, 2)
    try:
        ctx.set_path(_lib.PATH_STREAM)
        with pytest.raises(_lib.Gf2lError, match="forced decoder path"):
            decode_batch_arrays(code, llr, 5)
    finally:
        ctx.set_path(_lib.PATH_AUTO)
    assert decode_batch_arrays(code, llr, 5)[2].all()


def test_dvb_reference_kats(ctx):
    # ldpc_simd_validation.rs:26-43: short 1/2, all +5, converges within 10
    code = LdpcCode.dvb_t2_short(CodeRate.RATE_1_2, ctx)
    r = LdpcDecoder(code).decode_iterative([5.0] * code.n(), 10)
    assert r.converged and r.iterations <= 10 and r.decoded_bits.count_ones() == 0
    # the reference benchmark's input (ldpc_throughput.rs:97-124): a codeword
    # as error-free +-10 LLRs decodes in one iteration
    spec = codes.dvb_t2(1, 1)
    info = np.random.default_rng(3).integers(0, 2, size=(4, spec.k), dtype=np.uint8)
    cw = encode_dvb(spec.table, spec.n, spec.k, spec.q, info)
    llr = np.where(cw == 0, 10.0, -10.0).astype(np.float32)
    msg, it, conv, synd = decode_batch_arrays(LdpcCode.dvb_t2_normal(CodeRate.RATE_3_5, ctx), llr, 50)
    assert conv.all() and synd.all() and (it <= 1).all()
    assert np.array_equal(frames.unpack_words(msg, spec.k), info)


def test_dvb_builder_validation(ctx):
    lib = _lib.load()
    h = ctypes.c_void_p()

    def build(n, k, q, table):
        lens = np.array([len(r) for r in table], dtype=np.uint32)
        addr = np.array([a for r in table for a in r], dtype=np.uint32)
        return lib.gf2l_code_dvb_table(ctx.handle, n, k, q, 360, len(table), lens.ctypes.data,
                                       addr.ctypes.data, ctypes.byref(h))

    # builder.rs:123-139
    assert build(16200, 7200, 25, [[0]]) == _lib.GF2L_ERR_INVALID
    assert b"placeholder" in lib.gf2l_last_error(ctx.handle)
    assert build(64800, 32400, 90, [[0, 100]]) == _lib.GF2L_ERR_INVALID
    assert b"must have 90 rows" in lib.gf2l_last_error(ctx.handle)
    bad = [list(r) for r in codes.dvb_t2(0, 0).table]
    bad[3][1] = 9000
    assert build(16200, 7200, 25, bad) == _lib.GF2L_ERR_INVALID
    assert b"Invalid parity index 9000 in table row 3 (must be < 9000)" in lib.gf2l_last_error(ctx.handle)
    with pytest.raises(_lib.Gf2lError):
        LdpcCode(None, ctx, _build=lambda c, hh: lib.gf2l_code_dvb_t2(c.handle, 1, 6, hh))
    # a caller's table gives the same code as the compiled-in one
    s = codes.dvb_t2(0, 2)
    mine = LdpcCode.from_dvb_table(s.n, s.k, s.q, s.table, ctx=ctx)
    _, _, llr = dvb_random_frames(s.table, s.n, s.k, s.q, 16, 2.2, seed=1)
    same(decode_batch_arrays(mine, llr, 40),
         decode_batch_arrays(LdpcCode.dvb_t2_short(CodeRate.RATE_2_3, ctx), llr, 40), "table")


def test_gf2data_codes_match_oracle(ctx, tmp_path):
    spec = codes.n2304_r12()
    base = spec.base_matrix()
    _, _, llr = frames.random_frames(base, spec.z, 64, 1.8, seed=9)
    edges = orc.qc_to_edges(base, spec.z)
    ref = orc.Code.from_qc(base, spec.z).decode_batch(llr, 40)
    dual = gf2data.SpBitMatrixDual.from_coo(base.shape[0] * spec.z, base.shape[1] * spec.z, edges)
    p = tmp_path / "n2304.gf2"
    dual.save_to_file(p)
    same(decode_batch_arrays(LdpcCode.from_gf2data(p, ctx), llr, 40), ref, "type 4 file")
    sp = gf2data.SpBitMatrix.from_coo(dual.rows, dual.cols, edges)
    same(decode_batch_arrays(LdpcCode.from_gf2data(sp.to_bytes(), ctx), llr, 40), ref, "type 3 bytes")
    # a DVB-T2 code through a file
    s = codes.dvb_t2(0, 1)
    e = orc.dvb_edges(s.table, s.n, s.k, s.q)
    b = gf2data.SpBitMatrixDual.from_coo(s.m, s.n, e).to_bytes()
    _, _, llr2 = dvb_random_frames(s.table, s.n, s.k, s.q, 16, 1.7, seed=2)
    same(decode_batch_arrays(LdpcCode.from_gf2data(b, ctx), llr2, 30),
         oracle_code(s).decode_batch(llr2, 30, threads=8), "dvb file")


def test_gf2data_loader_errors(ctx):
    good = gf2data.SpBitMatrixDual.from_coo(2, 3, [(0, 0), (0, 1), (1, 1), (1, 2)]).to_bytes()
    cases = [
        (b"BADMAGIC" + good[8:], "Invalid magic bytes"),
        (good[:8] + b"\x63\x00" + good[10:], "Unsupported format version: 99"),
        (good[:10] + b"\x09" + good[11:], "Unknown type tag: 9"),
        (gf2data.bitvec_to_bytes(np.zeros(1, dtype=np.uint64), 5), "type 3 or 4"),
        (good[:-4], "past the end"),
        (good[:24] + (int.from_bytes(good[24:32], "little") + 4).to_bytes(8, "little") + good[32:] + bytes(4),
         "Data length mismatch"),
        (good[:10] + b"\x03" + good[11:], "Expected SpBitMatrix type"),
    ]
    for data, msg in cases:
        with pytest.raises(_lib.Gf2lError, match=msg):
            LdpcCode.from_gf2data(data, ctx)
    comp = bytearray(good)
    comp[11] = 1
    with pytest.raises(_lib.Gf2lError, match="compressed"):
        LdpcCode.from_gf2data(bytes(comp), ctx)
    chk = bytearray(good)
    chk[11] = 2  # checksum trailer flag: ignored, as by the reference reader
    assert LdpcCode.from_gf2data(bytes(chk) + bytes(32), ctx).nnz() == 4


def test_sdr_api_matches_oracle(ctx):
    """The GNU Radio-style entry points (SDR_INTEGRATION.md:39-57) against the
    oracle decoding the same frames on the DVB-T2 short 1/2 code built from
    build_dvb_edges (dvb_t2/builder.rs:74-118)."""
    lib = _lib.load()
    assert lib.gf2_ldpc_decoder_create(2, 0, 50) is None
    assert lib.gf2_ldpc_decoder_create(0, 6, 50) is None
    dec = lib.gf2_ldpc_decoder_create(0, 0, 50)
    assert dec
    try:
        n, k = ctypes.c_uint32(), ctypes.c_uint32()
        assert lib.gf2_ldpc_decoder_info(dec, ctypes.byref(n), ctypes.byref(k)) == 0
        assert (n.value, k.value) == (16200, 7200)
        s = codes.dvb_t2(0, 0)
        info, _, llr = dvb_random_frames(s.table, s.n, s.k, s.q, 12, 1.2, seed=11)
        msg, it, conv, _ = oracle_code(s).decode_batch(llr, 50, threads=8)
        bits = frames.unpack_words(msg, s.k)
        assert 0 < conv.sum() < 12, "need converging and failing frames"
        out = np.zeros(s.k, dtype=np.uint8)
        for f in range(12):
            l = np.ascontiguousarray(llr[f])
            rc = lib.gf2_ldpc_decode(dec, l.ctypes.data, s.n, out.ctypes.data)
            assert rc == (int(it[f]) if conv[f] else -1)
            assert np.array_equal(out, bits[f])
        outs = np.zeros((12, s.k), dtype=np.uint8)
        its = np.zeros(12, dtype=np.int32)
        rc = lib.gf2_ldpc_decode_frames(dec, llr.ctypes.data, 12, outs.ctypes.data, its.ctypes.data)
        assert rc == conv.sum()
        assert np.array_equal(outs, bits) and np.array_equal(its, np.where(conv, it.astype(np.int32), -1))
        assert lib.gf2_ldpc_decode(dec, llr.ctypes.data, s.n - 1, out.ctypes.data) == -2
        assert b"LLR length must equal n" in lib.gf2_ldpc_last_error(dec)
        assert lib.gf2_ldpc_decode_frames(dec, llr.ctypes.data, 0, outs.ctypes.data, None) == 0
    finally:
        lib.gf2_ldpc_decoder_destroy(dec)


def test_code_path_reports_the_decode_path(ctx):
    """gf2l_code_path: the path a decode takes under the context's settings,
    and the error a decode would return when the forced path or rule cannot
    run the code."""
    from gf2_amd import codes
    from gf2_amd.ldpc import CodeRate
    dvb = LdpcCode.dvb_t2_short(CodeRate.RATE_1_2, ctx)
    qc_small = LdpcCode(codes.n2304_r12().qc(), ctx)
    qc_big = LdpcCode(codes.n64800_r12().qc(), ctx)
    sparse = LdpcCode.from_edges(3, 6, [(0, 0), (0, 1), (0, 3), (1, 1), (1, 2), (1, 4), (2, 0), (2, 2), (2, 5)], ctx)
    try:
        ctx.set_path(_lib.PATH_AUTO)
        assert ctx.code_path(dvb) == _lib.PATH_STATE
        assert ctx.code_path(qc_small) == _lib.PATH_LDS
        assert ctx.code_path(qc_big) == _lib.PATH_STATE
        assert ctx.code_path(sparse) == _lib.PATH_CSR
        ctx.set_path(_lib.PATH_STREAM)
        assert ctx.code_path(qc_big) == _lib.PATH_STREAM
        for code in (dvb, qc_big):  # DVB codes run on the state kernels only; N=64800 does not fit LDS
            ctx.set_path(_lib.PATH_LDS)
            with pytest.raises(_lib.Gf2lError) as e:
                ctx.code_path(code)
            assert e.value.code == _lib.GF2L_ERR_UNSUPPORTED
        ctx.set_path(_lib.PATH_AUTO)
        ctx.set_check_rule(_lib.RULE_SPA)
        assert ctx.code_path(qc_small) == _lib.PATH_STREAM
        for code in (dvb, sparse):
            with pytest.raises(_lib.Gf2lError) as e:
                ctx.code_path(code)
            assert e.value.code == _lib.GF2L_ERR_UNSUPPORTED
    finally:
        ctx.set_check_rule(_lib.RULE_MINSUM)
        ctx.set_path(_lib.PATH_AUTO)
