"""GPU parity of the mixed-code decoder (gf2l_decode_mixed_device).

Every frame of a mixed batch must give exactly the oracle's result for its
own code (crates/gf2-coding/src/ldpc/core.rs:729-756 decodes one code per
call; results are per frame and in input order), whatever the order of the
frames, the graph unroll and the workspace history.
"""
import sys

import numpy as np
import pytest

from gf2_amd import _lib, codes
from gf2_amd.ldpc import Context, LdpcCode, QuasiCyclicLdpc, decode_batch_arrays, qc_ira_base_matrix
from gf2_amd.mixed import CodeSet, decode_mixed, decode_mixed_arrays
from gf2_amd.nms import NmsDecoder
from oracle import oracle as orc
from tests import frames

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    orc.build()
    c = Context(0)
    yield c
    c.set_tuning(_lib.TUNE_GRAPH_ITERS, 4)


def threshold_db(rate):
    """Rough plain min-sum convergence point of the seeded IRA codes."""
    return 1.0 + 4.0 * (rate - 0.5)


def build_batch(specs, per_code, seed, offsets=(-0.6, 0.2, 1.2)):
    """Frames of every spec at several Eb/N0 offsets, shuffled."""
    ids, blocks = [], []
    for c, spec in enumerate(specs):
        base = spec.base_matrix()
        rate = spec.k / spec.n
        for j, off in enumerate(offsets):
            _, _, llr = frames.random_frames(base, spec.z, per_code, threshold_db(rate) + off,
                                             seed=seed + 97 * c + j)
            for f in range(per_code):
                ids.append(c)
                blocks.append(llr[f])
    order = np.random.default_rng(seed).permutation(len(ids))
    return np.array(ids, dtype=np.uint16)[order], [blocks[i] for i in order]


def oracle_results(specs, ids, blocks, max_it):
    """Per-frame oracle results (msg words, iterations, converged, syndrome)."""
    out = [None] * len(ids)
    for c, spec in enumerate(specs):
        sel = np.nonzero(ids == c)[0]
        if sel.size == 0:
            continue
        llr = np.stack([blocks[i] for i in sel])
        msg, it, conv, synd = orc.Code.from_qc(spec.base_matrix(), spec.z).decode_batch(llr, max_it, threads=8)
        for j, i in enumerate(sel):
            out[i] = (msg[j], int(it[j]), bool(conv[j]), bool(synd[j]))
    return out


def check_against(results, ref, what):
    for f, (r, e) in enumerate(zip(results, ref)):
        got = (r.decoded_bits.words(), r.iterations, r.converged, r.syndrome_check_passed)
        if not (np.array_equal(got[0], e[0]) and got[1:] == e[1:]):
            raise AssertionError(f"{what}: frame {f} differs: iterations {got[1]} vs {e[1]}, "
                                 f"converged {got[2]} vs {e[2]}")


def test_twelve_codes_mixed_match_oracle(ctx):
    specs = codes.mixed_twelve()
    cs = CodeSet([LdpcCode(QuasiCyclicLdpc(s.base_matrix(), s.z), ctx) for s in specs])
    ids, blocks = build_batch(specs, 2, seed=11)
    say("twelve codes", len(ids), "frames")
    res = decode_mixed(cs, ids, blocks, 30)
    ref = oracle_results(specs, ids, blocks, 30)
    check_against(res, ref, "twelve codes")
    its = np.array([r.iterations for r in res])
    assert (its == 30).any() and (its < 30).any()  # both early exits and capped frames


@pytest.mark.parametrize("unroll,max_it", [(1, 7), (3, 7), (4, 1), (7, 12), (64, 9)])
def test_graph_unroll_does_not_change_results(ctx, unroll, max_it):
    specs = [codes.n2304_r12()] + codes.mixed_twelve()[:3]
    cs = CodeSet([LdpcCode(QuasiCyclicLdpc(s.base_matrix(), s.z), ctx) for s in specs])
    ids, blocks = build_batch(specs, 3, seed=23)
    say("unroll", unroll, "max_it", max_it)
    ctx.set_tuning(_lib.TUNE_GRAPH_ITERS, unroll)
    try:
        res = decode_mixed(cs, ids, blocks, max_it)
    finally:
        ctx.set_tuning(_lib.TUNE_GRAPH_ITERS, 4)
    check_against(res, oracle_results(specs, ids, blocks, max_it), f"unroll {unroll}")


def test_small_lift_and_workspace_regrowth(ctx):
    """A z < 64 code beside a z = 96 code; three calls on one set whose sizes
    grow (new workspace and graph) and shrink again, each against the
    single-code decoder path on the same frames."""
    small = qc_ira_base_matrix(5, 4, 10, 40, [3] * 6)
    big = codes.n2304_r12()
    pairs = [(small, 40), (big.base_matrix(), big.z)]
    gcodes = [LdpcCode(QuasiCyclicLdpc(b, z), ctx) for b, z in pairs]
    cs = CodeSet(gcodes)
    rng = np.random.default_rng(3)
    for call, F in enumerate([5, 300, 17]):
        ids = rng.integers(0, 2, size=F).astype(np.uint16)
        blocks = []
        for f, c in enumerate(ids):
            b, z = pairs[c]
            _, _, llr = frames.random_frames(b, z, 1, 2.0, seed=1000 * call + f)
            blocks.append(llr[0])
        say("regrowth call", call, "frames", F)
        res = decode_mixed(cs, ids, blocks, 25)
        for c in (0, 1):
            sel = np.nonzero(ids == c)[0]
            if sel.size == 0:
                continue
            single = decode_batch_arrays(gcodes[c], np.stack([blocks[i] for i in sel]), 25)
            ref = orc.Code.from_qc(*pairs[c]).decode_batch(np.stack([blocks[i] for i in sel]), 25)
            for name, a, b in zip(["msg", "it", "conv", "synd"], single, ref):
                assert np.array_equal(np.asarray(a), np.asarray(b)), f"single-code {name}"
            for j, i in enumerate(sel):
                r = res[i]
                assert np.array_equal(r.decoded_bits.words(), ref[0][j])
                assert (r.iterations, r.converged, r.syndrome_check_passed) == \
                    (int(ref[1][j]), bool(ref[2][j]), bool(ref[3][j])), f"call {call} frame {i}"


def test_zero_iterations_and_empty_batch(ctx):
    specs = codes.mixed_twelve()[:2]
    cs = CodeSet([LdpcCode(QuasiCyclicLdpc(s.base_matrix(), s.z), ctx) for s in specs])
    ids, blocks = build_batch(specs, 1, seed=5)
    res = decode_mixed(cs, ids, blocks, 0)  # fresh decoders (core.rs:906-937)
    for r in res:
        assert r.iterations == 0 and not r.converged and r.syndrome_check_passed
        assert r.decoded_bits.count_ones() == 0
    assert decode_mixed(cs, [], [], 10) == []


def test_errors(ctx):
    specs = codes.mixed_twelve()[:2]
    cs = CodeSet([LdpcCode(QuasiCyclicLdpc(s.base_matrix(), s.z), ctx) for s in specs])
    with pytest.raises(ValueError, match="outside the code set"):
        decode_mixed_arrays(cs, [0, 2], np.zeros(specs[0].n * 2, dtype=np.float32), 5)
    # the C ABI checks ids itself, before touching any buffer (the pointers
    # passed here are never dereferenced)
    buf = np.zeros(16, dtype=np.uint16)
    buf[3] = 9
    rc = _lib.load().gf2l_decode_mixed_device(ctx.handle, cs.handle, buf.ctypes.data, 16, 1, 5, 1, 1, 1, 1)
    assert rc == _lib.GF2L_ERR_INVALID
    assert "code id 9 of frame 3" in _lib.load().gf2l_last_error(ctx.handle).decode()
    csr = LdpcCode.from_edges(2, 4, [(0, 0), (0, 1), (1, 2), (1, 3)], ctx)
    with pytest.raises(_lib.Gf2lError, match="quasi-cyclic"):
        CodeSet([csr])


@pytest.mark.parametrize("fb_it", [0, 20])
def test_learned_mixed_matches_oracle(ctx, fb_it):
    """One learned decoder per code (random weights), frames of four codes at
    low, middle and high Eb/N0 in one call: every frame equals the oracle's
    weighted decode followed, when it fails, by plain min-sum from its LLRs."""
    specs = [codes.n2304_r12()] + [codes.mixed_twelve()[i] for i in (0, 6, 9)]
    gcodes = [LdpcCode(QuasiCyclicLdpc(s.base_matrix(), s.z), ctx) for s in specs]
    cs = CodeSet(gcodes)
    rng = np.random.default_rng(41)
    T = 4
    weights = [rng.uniform(0.6, 0.95, size=(T, int((s.base_matrix() >= 0).sum()))).astype(np.float32)
               for s in specs]
    decs = [NmsDecoder(g, T, weights=w) for g, w in zip(gcodes, weights)]
    ids, blocks = build_batch(specs, 2, seed=43, offsets=(0.2, 1.5, 3.5))
    say("learned mixed", len(ids), "frames, fallback", fb_it)
    n = np.array([s.n for s in specs])
    packed = np.concatenate(blocks)
    msg, msg_off, it, conv, synd, fb = decode_mixed_arrays(cs, ids, packed, fb_it, decoders=decs)
    for c, spec in enumerate(specs):
        sel = np.nonzero(ids == c)[0]
        oc = orc.Code.from_qc(spec.base_matrix(), spec.z)
        ref = frames.oracle_learned_with_fallback(oc, spec.base_matrix(), spec.z, weights[c],
                                                  np.stack([blocks[i] for i in sel]), fb_it)
        for j, f in enumerate(sel):
            got = (msg[msg_off[f]:msg_off[f + 1]], it[f], conv[f], synd[f], fb[f])
            exp = (ref[0][j], ref[1][j], ref[2][j], ref[3][j], ref[4][j])
            assert np.array_equal(got[0], exp[0]) and tuple(got[1:]) == tuple(exp[1:]), \
                f"code {c} frame {f}: got it {got[1]} conv {got[2]} fb {got[4]}, want {exp[1:]}"
    assert (conv & ~fb & (it <= T)).any()  # some frames converge in the weighted phase
    if fb_it:
        assert fb.any() and (~fb).any()
    assert n[ids].sum() == packed.size


@pytest.mark.parametrize("fb_it", [0, 20])
def test_learned_mixed_continue_matches_oracle(ctx, fb_it):
    """Learned decoders with the continuing fallback in one mixed call: every
    frame equals the oracle's flooding decode of T + fb_it iterations with
    the code's weights for the first T and 1.0 after (multiplying by 1.0 is
    exact), as for a single code (gf2l_nms_decode_device)."""
    specs = [codes.n2304_r12()] + [codes.mixed_twelve()[i] for i in (0, 6, 9)]
    gcodes = [LdpcCode(QuasiCyclicLdpc(s.base_matrix(), s.z), ctx) for s in specs]
    cs = CodeSet(gcodes)
    rng = np.random.default_rng(47)
    T = 4
    weights = [rng.uniform(0.6, 0.95, size=(T, int((s.base_matrix() >= 0).sum()))).astype(np.float32)
               for s in specs]
    decs = [NmsDecoder(g, T, weights=w, fallback="continue") for g, w in zip(gcodes, weights)]
    ids, blocks = build_batch(specs, 2, seed=49, offsets=(0.2, 1.5, 3.5))
    say("learned mixed continue", len(ids), "frames, fallback", fb_it)
    packed = np.concatenate(blocks)
    msg, msg_off, it, conv, synd, fb = decode_mixed_arrays(cs, ids, packed, fb_it, decoders=decs)
    for c, spec in enumerate(specs):
        sel = np.nonzero(ids == c)[0]
        base = spec.base_matrix()
        oc = orc.Code.from_qc(base, spec.z)
        nbe = int((base >= 0).sum())
        wfull = np.concatenate([weights[c], np.ones((fb_it, nbe), np.float32)])
        rm, ri, rc, rs = oc.decode_batch(np.stack([blocks[i] for i in sel]), T + fb_it,
                                         edge_w=frames.edge_weights(base, spec.z, wfull, oc), threads=8)
        rfb = ((ri > T) | ~rc) if fb_it else np.zeros(rc.shape, dtype=bool)
        for j, f in enumerate(sel):
            got = (msg[msg_off[f]:msg_off[f + 1]], it[f], conv[f], synd[f], fb[f])
            assert np.array_equal(got[0], rm[j]) and tuple(got[1:]) == (ri[j], rc[j], rs[j], rfb[j]), \
                f"code {c} frame {f}: got it {got[1]} conv {got[2]} fb {got[4]}, want {ri[j]} {rc[j]} {rfb[j]}"
    assert (conv & ~fb & (it <= T)).any()  # some frames converge in the weighted phase
    if fb_it:
        assert fb.any() and (conv & fb).any()  # and some converge in the continuation


def test_learned_mixed_rejects_mismatched_decoders(ctx):
    specs = codes.mixed_twelve()[:2]
    gcodes = [LdpcCode(QuasiCyclicLdpc(s.base_matrix(), s.z), ctx) for s in specs]
    cs = CodeSet(gcodes)
    ids, blocks = build_batch(specs, 1, seed=5)
    packed = np.concatenate(blocks)
    swapped = [NmsDecoder(gcodes[1], 4, init=0.8), NmsDecoder(gcodes[0], 4, init=0.8)]
    with pytest.raises(_lib.Gf2lError, match="another code"):
        decode_mixed_arrays(cs, ids, packed, 10, decoders=swapped)
    uneven = [NmsDecoder(gcodes[0], 4, init=0.8), NmsDecoder(gcodes[1], 5, init=0.8)]
    with pytest.raises(_lib.Gf2lError, match="iteration count"):
        decode_mixed_arrays(cs, ids, packed, 10, decoders=uneven)
    modes = [NmsDecoder(gcodes[0], 4, init=0.8), NmsDecoder(gcodes[1], 4, init=0.8, fallback="continue")]
    with pytest.raises(_lib.Gf2lError, match="fallback mode"):
        decode_mixed_arrays(cs, ids, packed, 10, decoders=modes)
    layered = [NmsDecoder(g, 4, init=0.8, schedule="layered") for g in gcodes]
    with pytest.raises(_lib.Gf2lError) as e:
        decode_mixed_arrays(cs, ids, packed, 10, decoders=layered)
    assert e.value.code == _lib.GF2L_ERR_UNSUPPORTED


@pytest.mark.parametrize("chunk", [1, 4])
def test_counter_fed_mixed_items_match_oracle(ctx, chunk):
    specs = [codes.n2304_r12()] + codes.mixed_twelve()[:2]
    cs = CodeSet([LdpcCode(QuasiCyclicLdpc(s.base_matrix(), s.z), ctx) for s in specs])
    ids, blocks = build_batch(specs, 3, seed=71)
    say("mixed counter-fed items, chunk", chunk)
    ctx.set_tuning(_lib.TUNE_CHUNK, chunk)
    try:
        res = decode_mixed(cs, ids, blocks, 20)
    finally:
        ctx.set_tuning(_lib.TUNE_CHUNK, -1)
    check_against(res, oracle_results(specs, ids, blocks, 20), f"chunk {chunk}")
