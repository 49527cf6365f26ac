"""Code ingestion on the CPU: the DVB-T2 tables and parameters compiled into
the library, the oracle's DVB-T2 edge builder, and GF2DATA files."""
import hashlib
import json
import os
import struct

import numpy as np
import pytest

from gf2_amd import codes, gf2data
from gf2_amd.channel import encode_dvb
from oracle import oracle as orc

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "dvb_t2_tables.json")))["tables"]


def digest(table):
    return hashlib.sha256(b"".join(struct.pack("<I", len(r)) for r in table) +
                          b"".join(struct.pack("<I", a) for r in table for a in r)).hexdigest()


# ---- DVB-T2 tables and parameters --------------------------------------------

def test_dvb_parameters_match_reference_kat():
    # tests/ldpc_validation.rs:213-300 (n, k, m) and params.rs:59-73 (q)
    normal = [(64800, 32400, 32400, 90), (64800, 38880, 25920, 72), (64800, 43200, 21600, 60),
              (64800, 48600, 16200, 45), (64800, 51840, 12960, 36), (64800, 54000, 10800, 30)]
    short = [(16200, 7200, 9000, 25), (16200, 9720, 6480, 18), (16200, 10800, 5400, 15),
             (16200, 11880, 4320, 12), (16200, 12600, 3600, 10), (16200, 13320, 2880, 8)]
    for frame, rows in ((1, normal), (0, short)):
        for rate, (n, k, m, q) in enumerate(rows):
            s = codes.dvb_t2(frame, rate)
            assert (s.n, s.k, s.m, s.q) == (n, k, m, q)
            assert len(s.table) == k // 360  # validate_table, builder.rs:22-28
            assert s.m == q * 360
    with pytest.raises(ValueError):
        codes.dvb_t2(2, 0)
    with pytest.raises(ValueError):
        codes.dvb_t2(0, 6)


def test_dvb_tables_match_reference_digests():
    for spec in codes.dvb_t2_all():
        g = GOLDEN[f"{'normal' if spec.frame else 'short'}_{codes.DVB_RATE_NAMES[spec.rate].replace('/', '_')}"]
        assert (g["n"], g["k"], g["q"], g["rows"]) == (spec.n, spec.k, spec.q, len(spec.table))
        assert g["entries"] == sum(map(len, spec.table))
        assert list(spec.table[0]) == g["first_row"] and list(spec.table[-1]) == g["last_row"]
        assert digest(spec.table) == g["sha256"], spec.name


def test_dvb_table_a1_known_rows():
    # dvb_t2/table_interpretation.md:13 and table_a_1_R1_2_N64800.txt:1-2
    t = codes.dvb_t2(1, 0).table
    assert t[0] == (54, 9318, 14392, 27561, 26909, 10219, 2534, 8597)
    assert t[1] == (55, 7263, 4635, 2530, 28130, 3033, 23830, 3651)


def test_dvb_degree_profiles_of_the_standard():
    # ETSI EN 302 755 normal frames: (high degree, blocks with it) + degree-3 blocks
    expect = {0: (8, 36, 54), 1: (12, 36, 72), 2: (13, 12, 108), 3: (12, 15, 120), 4: (11, 18, 126),
              5: (13, 15, 135)}
    for rate, (dh, bh, b3) in expect.items():
        lens = [len(r) for r in codes.dvb_t2(1, rate).table]
        assert lens.count(dh) == bh and lens.count(3) == b3 and len(lens) == bh + b3


@pytest.mark.parametrize("frame,rate", [(f, r) for f in (0, 1) for r in range(6)])
def test_oracle_dvb_code_holds_standard_codewords(frame, rate):
    """The oracle's build_dvb_edges restatement against the standard's own
    encoder (accumulate into parity addresses, then p_t ^= p_(t-1)): every
    codeword has zero syndrome; degrees follow the construction."""
    s = codes.dvb_t2(frame, rate)
    e = orc.dvb_edges(s.table, s.n, s.k, s.q)
    oc = orc.Code(s.m, s.n, e)
    rng = np.random.default_rng(frame * 10 + rate)
    info = rng.integers(0, 2, size=(3, s.k), dtype=np.uint8)
    cw = encode_dvb(s.table, s.n, s.k, s.q, info)
    for f in range(3):
        assert oc.syndrome_weight(cw[f]) == 0
        bad = cw[f].copy()
        bad[rng.integers(0, s.n)] ^= 1
        assert oc.syndrome_weight(bad) > 0
    vdeg = np.diff(oc.col_ptr)
    assert (vdeg[:s.k] == np.repeat([len(r) for r in s.table], 360)).all()
    assert (vdeg[s.k:-1] == 2).all() and vdeg[-1] == 1
    # each table address covers its residue class mod q once per lane
    cdeg = np.diff(oc.row_ptr).astype(np.int64)
    res = np.bincount([a % s.q for r in s.table for a in r], minlength=s.q)
    assert (cdeg - 2 + (np.arange(s.m) == 0) == res[np.arange(s.m) % s.q]).all()


def test_oracle_dvb_edges_wrong_step_is_caught():
    """A planted mistake in the builder (lane step j instead of j*q) breaks
    the codeword check, so the check above has teeth."""
    s = codes.dvb_t2(0, 0)
    m = s.m
    j = np.arange(360)
    parts = [np.stack([(p + j) % m, i * 360 + j], axis=1) for i, row in enumerate(s.table) for p in row]
    stair = orc.dvb_edges(s.table, s.n, s.k, s.q)[sum(map(len, s.table)) * 360:]
    oc = orc.Code(m, s.n, np.concatenate(parts + [stair]))
    info = np.random.default_rng(0).integers(0, 2, size=(1, s.k), dtype=np.uint8)
    assert oc.syndrome_weight(encode_dvb(s.table, s.n, s.k, s.q, info)[0]) > 0


# ---- GF2DATA ------------------------------------------------------------------

def handmade_dual(rows, cols, ro, ri, co, ci):
    """A type-4 file assembled byte by byte from the layout in
    crates/gf2-core/src/io/mod.rs:10-31, 80-98."""
    meta = ('{"type":"SpBitMatrixDual","rows":%d,"cols":%d,"nnz":%d,"version":1}' % (rows, cols, len(ri))).encode()
    payload = b"".join(struct.pack("<%dI" % len(a), *a) for a in (ro, ri, co, ci))
    head = b"GF2DATA\0" + struct.pack("<H", 1) + bytes([4, 0]) + bytes(8) + struct.pack("<I", len(meta)) + \
        struct.pack("<Q", len(payload))
    assert len(head) == 32
    return head + meta + payload


def test_gf2data_dual_bytes_match_the_format():
    # H = [[1,1,0],[0,1,1]]
    m = gf2data.SpBitMatrixDual.from_coo(2, 3, [(0, 0), (0, 1), (1, 1), (1, 2)])
    want = handmade_dual(2, 3, [0, 2, 4], [0, 1, 1, 2], [0, 1, 3, 4], [0, 0, 1, 1])
    assert m.to_bytes() == want
    back = gf2data.SpBitMatrixDual.from_bytes(want)
    assert back.edges().tolist() == [[0, 0], [0, 1], [1, 1], [1, 2]]
    assert back.col_iter(1) == [0, 1]


def test_gf2data_header_kats():
    # header.rs:93-200: 32 bytes, magic first, every type tag round-trips
    for tag in (1, 2, 3, 4):
        b = gf2data.Header(1, tag, 0, 10, 100).to_bytes()
        assert len(b) == 32 and b[:8] == b"GF2DATA\0"
        h = gf2data.Header.from_bytes(b)
        assert (h.type_tag, h.metadata_len, h.data_len) == (tag, 10, 100)
    h = gf2data.Header(1, 3, 0x03, 200, 16000)
    assert gf2data.Header.from_bytes(h.to_bytes()).flags == 0x03
    with pytest.raises(gf2data.Gf2DataError, match="Invalid magic bytes"):
        gf2data.Header.from_bytes(b"BADMAGIC" + bytes(24))
    with pytest.raises(gf2data.Gf2DataError, match="Unsupported format version: 99"):
        gf2data.Header.from_bytes(b"GF2DATA\0" + struct.pack("<H", 99) + bytes(22))
    with pytest.raises(gf2data.Gf2DataError, match="Unknown type tag: 99"):
        gf2data.Header.from_bytes(b"GF2DATA\0" + struct.pack("<H", 1) + bytes([99]) + bytes(21))


def test_gf2data_roundtrips_and_errors(tmp_path):
    rng = np.random.default_rng(5)
    e = np.stack([rng.integers(0, 40, 300), rng.integers(0, 90, 300)], axis=1)
    dual = gf2data.SpBitMatrixDual.from_coo(40, 90, e)
    p = tmp_path / "h.gf2"
    dual.save_to_file(p)
    back = gf2data.SpBitMatrixDual.load_from_file(p)
    for a, b in ((back.csr.indptr, dual.csr.indptr), (back.csr.indices, dual.csr.indices),
                 (back.csc.indptr, dual.csc.indptr), (back.csc.indices, dual.csc.indices)):
        assert np.array_equal(a, b)
    sp = gf2data.SpBitMatrix.from_coo_deduplicated(40, 90, e)
    assert np.array_equal(gf2data.SpBitMatrix.from_bytes(sp.to_bytes()).indices, sp.indices)
    # empty and identity (sparse.rs:563-582)
    for mat in (gf2data.SpBitMatrixDual.from_coo(3, 3, np.zeros((0, 2), dtype=np.int64)),
                gf2data.SpBitMatrixDual.from_coo(3, 3, [(i, i) for i in range(3)])):
        assert gf2data.SpBitMatrixDual.from_bytes(mat.to_bytes()).edges().tolist() == mat.edges().tolist()
    w = np.array([0xDEADBEEF12345678, 0x1F], dtype=np.uint64)
    assert gf2data.bitvec_from_bytes(gf2data.bitvec_to_bytes(w, 69))[0].tolist() == [0xDEADBEEF12345678, 0x1F]
    mw = rng.integers(0, 2**63, size=(3, 2), dtype=np.uint64)
    assert np.array_equal(gf2data.bitmatrix_from_bytes(gf2data.bitmatrix_to_bytes(mw, 3, 100))[0], mw)
    # wrong type tag (sparse.rs:628-635), truncated payload, length mismatch
    with pytest.raises(gf2data.Gf2DataError, match="Expected SpBitMatrixDual"):
        gf2data.SpBitMatrixDual.from_bytes(sp.to_bytes())
    b = bytearray(dual.to_bytes())
    with pytest.raises(gf2data.Gf2DataError):
        gf2data.SpBitMatrixDual.from_bytes(bytes(b[:-4]))
    b[24] += 4
    with pytest.raises(gf2data.Gf2DataError, match="Data length mismatch"):
        gf2data.SpBitMatrixDual.from_bytes(bytes(b) + bytes(4))


def test_host_mirror_exposes_the_constructors():
    from gf2_amd.ldpc import CodeRate, LdpcCode
    for name in ("dvb_t2_short", "dvb_t2_normal", "from_dvb_table", "from_gf2data", "from_edges"):
        assert callable(getattr(LdpcCode, name))
    assert [int(r) for r in CodeRate] == list(range(6))


def test_oracle_zero_codeword_frames_share_the_noise_stream():
    """All-zero codeword frames use framegen's Eb/N0 and noise streams: where
    the encoded frame's bit is 0 the LLRs are identical, where it is 1 they
    differ by exactly the +-1 of BPSK (2*2/sigma^2 before rounding)."""
    spec = codes.n2304_r12()
    base = spec.base_matrix()
    for g in (0, 7, 123456789012):
        llr, cw = orc.framegen(base, spec.z, 99, g, 1.0, 3.0)
        zl = orc.framegen_zero(spec.n, spec.k, 99, g, 1.0, 3.0)
        bits = np.unpackbits(cw.view(np.uint8), bitorder="little")[:spec.n].astype(bool)
        assert np.array_equal(zl[~bits], llr[~bits])
        d = (zl[bits].astype(np.float64) - llr[bits].astype(np.float64))
        assert bits.any() and np.allclose(d, d[0], rtol=0, atol=1e-4 * abs(d[0]))
