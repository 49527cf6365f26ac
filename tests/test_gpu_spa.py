"""The sum-product check rule (GF2L_RULE_SPA: check_node_update_spa,
core.rs:762-787, with Llr::boxplus_n, llr.rs:210-215) on the GPU against the
oracle's or_decode_spa (pinned in tests/test_oracle_spa.py).

Tolerance parity (the device's tanhf/atanhf are not the host C library's):
- one check update, from the same inputs, agrees within 1e-5 relative
  (1e-6 absolute near zero) where |m| <= 8; above, atanh's slope
  1/(1 - y^2) grows like e^|m| and a last-bit difference of the product y
  becomes a large one of m, so there the messages agree within 1e-5 in the
  tanh domain: |tanh(m_gpu / 2) - tanh(m_ref / 2)| <= 1e-5. The first
  iteration's inputs are the channel LLRs, so its messages are compared with
  the oracle's trace directly; every later iteration's messages are compared
  with the oracle's boxplus_n applied to the GPU's own inputs of that
  iteration (belief - message of the previous one), because the variable
  messages x = b - m cancel: a 1e-7 relative difference of b becomes a
  |b| / |x| times larger one of x, and so of every message it feeds;
- beliefs within 1e-5 of (|llr| + sum |m|) of the llr plus the column's
  messages of the same iteration (binary32 summation in any order);
- end to end, the beliefs of the first traced iterations keep the oracle's
  signs wherever the oracle's belief exceeds 0.05 in magnitude;
- decoded words identical on at least 99.99% of frames (every frame at
  these batch sizes), iteration counts and flags on at least 99.5%.
"""
import sys

import numpy as np
import pytest

from gf2_amd import _lib, codes
from gf2_amd.ldpc import Context, LdpcCode, decode_batch_arrays, decode_trace_arrays
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
    c.set_check_rule(_lib.RULE_SPA)
    yield c
    c.set_check_rule(_lib.RULE_MINSUM)
    c.set_path(_lib.PATH_AUTO)


def messages_close(gpu, ref):
    gpu, ref = gpu.astype(np.float64), ref.astype(np.float64)
    small = np.abs(ref) <= 8.0
    ok_small = np.abs(gpu - ref) <= 1e-5 * np.abs(ref) + 1e-6
    ok_large = np.abs(np.tanh(gpu / 2) - np.tanh(ref / 2)) <= 1e-5
    return np.where(small, ok_small, ok_large)


def test_spa_messages_per_iteration(ctx):
    spec = codes.n2304_r12()
    base = spec.base_matrix()
    _, _, llr = frames.random_frames(base, spec.z, 8, 1.2, seed=61)
    T = 6
    say("spa trace n2304")
    code = LdpcCode(spec.qc(), ctx)
    tc, tb, msg, it, conv, synd = decode_trace_arrays(code, llr, 50, T)
    oc = orc.Code.from_qc(base, spec.z)
    perm = frames.block_to_csr_perm(base, spec.z, oc.row_ptr)
    checked = 0
    for f in range(8):
        rc, rb, rmsg, (rit, rconv, rsynd) = oc.decode_trace_spa(llr[f], 50)
        prev_b, prev_m = llr[f].astype(np.float32), np.zeros(oc.nnz, np.float32)
        for t in range(min(T, rit, int(it[f]))):
            c2v = np.empty(oc.nnz, dtype=np.float32)
            c2v[perm] = tc[t, f]
            if t == 0:
                ok = messages_close(c2v, rc[0])
                assert ok.all(), (f, np.nonzero(~ok)[0][:5], c2v[~ok][:5], rc[0][~ok][:5])
            if f < 3:
                # one check update of the oracle's rule on the GPU's inputs
                x = prev_b[oc.row_idx] - prev_m
                want = np.empty(oc.nnz, np.float32)
                for r in range(oc.m):
                    a, b = int(oc.row_ptr[r]), int(oc.row_ptr[r + 1])
                    for e in range(a, b):
                        want[e] = orc.boxplus_n(np.delete(x[a:b], e - a))
                ok = messages_close(c2v, want)
                assert ok.all(), (f, t, np.nonzero(~ok)[0][:5], c2v[~ok][:5], want[~ok][:5])
            # beliefs: the llr plus this iteration's messages of the column
            bel = tb[t, f].astype(np.float64)
            tot = llr[f].astype(np.float64) + np.bincount(oc.row_idx, c2v.astype(np.float64), oc.n)
            mag = np.abs(llr[f]).astype(np.float64) + np.bincount(oc.row_idx, np.abs(c2v).astype(np.float64), oc.n)
            assert (np.abs(bel - tot) <= 1e-5 * mag + 1e-6).all(), (f, t)
            # end to end: the oracle's signs where its beliefs are clear
            clear = np.abs(rb[t]) > 0.05
            assert (np.signbit(tb[t, f][clear]) == np.signbit(rb[t][clear])).all(), (f, t)
            prev_b, prev_m = tb[t, f], c2v
            checked += 1
    assert checked >= 20, checked


@pytest.mark.parametrize("ebn0", [1.0, 1.5, 2.0])
def test_spa_batch_matches_oracle(ctx, ebn0):
    spec = codes.n2304_r12()
    base = spec.base_matrix()
    F = 1024
    _, _, llr = frames.random_frames(base, spec.z, F, ebn0, seed=62 + int(ebn0 * 10))
    say("spa batch n2304", ebn0)
    gpu = decode_batch_arrays(LdpcCode(spec.qc(), ctx), llr, 50)
    ref = orc.Code.from_qc(base, spec.z).decode_batch_spa(llr, 50, threads=8)
    same_words = (gpu[0] == ref[0]).all(axis=1)
    assert same_words.mean() >= 0.9999, (ebn0, np.nonzero(~same_words)[0][:8])
    for name, a, b in zip(["iterations", "converged", "syndrome_ok"], gpu[1:], ref[1:]):
        assert (np.asarray(a) == np.asarray(b)).mean() >= 0.995, (ebn0, name)


def test_spa_n64800_subset(ctx):
    spec = codes.n64800_r12()
    base = spec.base_matrix()
    _, _, llr = frames.random_frames(base, spec.z, 16, 1.3, seed=63)
    say("spa n64800")
    gpu = decode_batch_arrays(LdpcCode(spec.qc(), ctx), llr, 50)
    ref = orc.Code.from_qc(base, spec.z).decode_batch_spa(llr, 50, threads=8)
    assert (gpu[0] == ref[0]).all()
    assert (gpu[1] == ref[1]).mean() >= 0.9


@pytest.mark.parametrize("ebn0", [1.0, 1.6])
def test_spa_gain_and_saturation(ctx, ebn0):
    """What the exact rule gains, and where binary32 ends it.

    Llr::boxplus_n (llr.rs:210-215) in binary32: once every other input
    exceeds about 18 in magnitude, tanhf(x / 2) rounds to 1.0f, atanhf(1) is
    inf, the belief becomes inf and the next x = b - m is NaN; NaN < 0 is
    false, so the hard word is all zero, whose syndrome is zero. The rule is
    unused by the reference decoder, and this is its behaviour as written:
    - every frame that ends with syndrome_ok and a wrong word returned the
      all-zero word (the saturation, not a wrong codeword);
    - at 1.0 dB, where most frames end before their LLRs grow that large,
      it fails fewer frames than the reference's min-sum (the oracle: 328
      against 446 of 512 frames, seed 64)."""
    spec = codes.n2304_r12()
    base = spec.base_matrix()
    F = 1024
    info, cw, llr = frames.random_frames(base, spec.z, F, ebn0, seed=64)
    code = LdpcCode(spec.qc(), ctx)
    say("spa vs minsum", ebn0)
    spa = decode_batch_arrays(code, llr, 50)
    ctx.set_check_rule(_lib.RULE_MINSUM)
    try:
        ms = decode_batch_arrays(code, llr, 50)
    finally:
        ctx.set_check_rule(_lib.RULE_SPA)
    fe = {}
    for name, r in [("spa", spa), ("minsum", ms)]:
        wrong_word = (frames.unpack_words(r[0], spec.k) != info).any(axis=1)
        synd = np.asarray(r[3], dtype=bool)
        fe[name] = int((~synd | wrong_word).sum())
        if name == "spa":
            zero = (np.asarray(r[0]) == 0).all(axis=1)
            assert zero[wrong_word & synd].all()
            fe["spa_saturated"] = int((wrong_word & synd).sum())
    say("frame errors", fe)
    if ebn0 == 1.0:
        assert fe["spa"] < fe["minsum"], fe


def test_spa_unsupported_combinations(ctx):
    spec = codes.n2304_r12()
    base = spec.base_matrix()
    _, _, llr = frames.random_frames(base, spec.z, 4, 2.0, seed=65)
    code = LdpcCode(spec.qc(), ctx)
    for path in (_lib.PATH_LDS, _lib.PATH_STATE):
        ctx.set_path(path)
        with pytest.raises(_lib.Gf2lError) as e:
            decode_batch_arrays(code, llr, 10)
        assert e.value.code == _lib.GF2L_ERR_UNSUPPORTED
    ctx.set_path(_lib.PATH_STREAM)
    decode_batch_arrays(code, llr, 10)  # the STREAM path runs it
    ctx.set_path(_lib.PATH_AUTO)
    nbe = int((base >= 0).sum())
    dec = NmsDecoder(code, 2, weights=np.full((2, nbe), 0.8, np.float32))
    with pytest.raises(_lib.Gf2lError) as e:
        dec.decode(llr, fallback_max_it=5)
    assert e.value.code == _lib.GF2L_ERR_UNSUPPORTED
    with pytest.raises(_lib.Gf2lError):
        ctx.set_check_rule(7)
