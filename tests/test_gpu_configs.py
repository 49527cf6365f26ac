"""The full-size workloads, decoded on the GPU and checked against the oracle.

N=64800 plain BP: the rate-1/2 QC-IRA code (Z=360, 90x180 base graph,
E=226,440), 4096 frames per batch at Eb/N0 1.0, 1.25 and 1.5 dB, 50
iterations.  At 1.0 dB almost every frame runs all 50 iterations and retires
through the last-iteration path; at 1.5 dB most converge early.  Full-batch
properties at every point, and a 32-frame subset (16 from each end of the
batch) compared frame by frame with the oracle's decode_batch
(crates/gf2-coding/src/ldpc/core.rs:729-756, 889-938).

Twelve learned decoders in one call: the shipped 10-iteration weight sets of
the twelve Z=360 codes (gf2_amd/weights/nms_n*-k*_t10.npz), 16384 frames
spread over the codes in shuffled order, one gf2l_decode_mixed_nms_device
call with a 50-iteration plain fallback, at 1.5 dB.  Full-size properties,
and four frames per code compared with the oracle's weighted decode composed
with a plain fallback decode from the LLRs.
"""
import os
import sys

import numpy as np
import pytest

from gf2_amd import _lib, codes
from gf2_amd.channel import unpack_words
from gf2_amd.device import DeviceBuffer, generate_frames
from gf2_amd.ldpc import Context, LdpcCode, decode_batch_arrays
from gf2_amd.mixed import CodeSet
from gf2_amd.nms import NmsDecoder
from gf2_amd.train_nms import default_path, mixed_name
from oracle import oracle as orc
from tests import frames

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    orc.build()
    return Context(0)


def assert_frames_equal(got, ref, what):
    for name, a, b in zip(["msg", "iterations", "converged", "syndrome_ok"], got, ref):
        a, b = np.asarray(a), np.asarray(b)
        if not np.array_equal(a, b):
            bad = np.nonzero((a != b).reshape(a.shape[0], -1).any(axis=1))[0][:10]
            raise AssertionError(f"{what}: {name} differs at subset frames {bad.tolist()}")


@pytest.mark.parametrize("ebn0", [1.0, 1.25, 1.5])
def test_n64800_4096_frames(ctx, ebn0):
    spec = codes.n64800_r12()
    base = spec.base_matrix()
    F, max_it, kw, nw = 4096, 50, (spec.k + 63) // 64, (spec.n + 63) // 64
    code = LdpcCode(spec.qc(), ctx)
    say("n64800 4096 frames at", ebn0, "dB")
    d_llr, d_cw = generate_frames(code, 2024, 0, F, ebn0, ebn0)
    outs = [DeviceBuffer(ctx, F * kw * 8), DeviceBuffer(ctx, F * 4), DeviceBuffer(ctx, F),
            DeviceBuffer(ctx, F)]
    _lib.check(ctx.handle, _lib.load().gf2l_decode_batch_device(
        ctx.handle, code.handle, d_llr.ptr, F, max_it, outs[0].ptr, outs[1].ptr, outs[2].ptr, outs[3].ptr))
    ctx.synchronize()
    msg = outs[0].download(np.uint64, (F, kw))
    it = outs[1].download(np.uint32, F)
    conv = outs[2].download(np.uint8, F).astype(bool)
    synd = outs[3].download(np.uint8, F).astype(bool)
    info = unpack_words(d_cw.download(np.uint64, (F, nw)), spec.n)[:, :spec.k]
    bits = unpack_words(msg, spec.k)
    # full batch: flags consistent, iteration counts in range, frames that
    # stopped early carry the sent message (an undetected error is a codeword
    # other than the sent one: allowed but rare)
    assert np.array_equal(conv, synd)
    assert (it[~conv] == max_it).all() and (it >= 1).all() and (it <= max_it).all()
    right = (bits == info).all(axis=1)
    if conv.any():
        assert right[conv].mean() > 0.999
    fer = float(((~conv) | ~right).mean())
    say(f"  FER {fer:.4f}, mean iterations {it.mean():.2f}, converged {int(conv.sum())}")
    if ebn0 == 1.0:
        assert (it == max_it).mean() > 0.9  # the last-iteration retirement path
    if ebn0 == 1.5:
        assert conv.mean() > 0.5
    # subset against the oracle: 16 frames from each end of the batch
    sub = np.r_[0:16, F - 16:F]
    llr = np.concatenate([d_llr.download(np.float32, (16, spec.n)),
                          d_llr.download(np.float32, (16, spec.n), offset=(F - 16) * spec.n * 4)])
    ref = orc.Code.from_qc(base, spec.z).decode_batch(llr, max_it, threads=min(16, os.cpu_count() or 1))
    assert_frames_equal((msg[sub], it[sub], conv[sub], synd[sub]), ref, f"n64800 {ebn0} dB")
    # the same 32 frames decoded as their own batch give the same results
    again = decode_batch_arrays(code, llr, max_it)
    assert_frames_equal(again, ref, f"n64800 {ebn0} dB, subset batch")


def test_twelve_learned_sets_16384_frames(ctx):
    specs = codes.mixed_twelve()
    ebn0, F, T, fb_it = 1.5, 16384, 10, 50
    gcodes = [LdpcCode(s.qc(), ctx) for s in specs]
    decs = [NmsDecoder.load(g, default_path(mixed_name(s))) for g, s in zip(gcodes, specs)]
    assert all(d.iterations == T for d in decs)
    cs = CodeSet(gcodes)
    rng = np.random.default_rng(31)
    ids = rng.integers(0, len(specs), size=F).astype(np.uint16)
    n = np.array([s.n for s in specs])
    k = np.array([s.k for s in specs])
    kw, nw = (k + 63) // 64, (n + 63) // 64
    llr_off = np.concatenate([[0], np.cumsum(n[ids])])
    msg_off = np.concatenate([[0], np.cumsum(kw[ids])])
    say("twelve learned sets,", F, "frames at", ebn0, "dB")
    # frames generated per code on the device, scattered into their slots of
    # the packed (shuffled) batch
    packed = np.empty(int(llr_off[-1]), dtype=np.float32)
    cw = [None] * F
    for c, g in enumerate(gcodes):
        sel = np.nonzero(ids == c)[0]
        d_one, d_cw = generate_frames(g, 900 + c, 0, len(sel), ebn0, ebn0)
        words = d_cw.download(np.uint64, (len(sel), nw[c]))
        host = d_one.download(np.float32, (len(sel), n[c]))
        for j, f in enumerate(sel):
            cw[f] = words[j]
            packed[llr_off[f]:llr_off[f + 1]] = host[j]
        del d_one, d_cw
    d_llr = DeviceBuffer.from_array(ctx, packed)
    outs = [DeviceBuffer(ctx, int(msg_off[-1]) * 8), DeviceBuffer(ctx, F * 4), DeviceBuffer(ctx, F),
            DeviceBuffer(ctx, F), DeviceBuffer(ctx, F)]
    cs.decode_nms_device(decs, ids, d_llr.ptr, fb_it, *[o.ptr for o in outs])
    ctx.synchronize()
    msg = outs[0].download(np.uint64, int(msg_off[-1]))
    it = outs[1].download(np.uint32, F)
    conv = outs[2].download(np.uint8, F).astype(bool)
    synd = outs[3].download(np.uint8, F).astype(bool)
    fb = outs[4].download(np.uint8, F).astype(bool)
    # full size: flags consistent; weighted-phase frames stopped within T
    # iterations, fallback frames ran T plus 1..fb_it
    assert np.array_equal(conv, synd)
    assert (it[~fb] <= T).all() and (it[~fb] >= 1).all()
    assert (it[fb] > T).all() and (it[fb] <= T + fb_it).all()
    assert (it[~conv] == T + fb_it).all() and fb[~conv].all()
    right = np.array([np.array_equal(msg[msg_off[f]:msg_off[f + 1]], cw[f][:kw[ids[f]]] &
                                     mask_words(k[ids[f]], kw[ids[f]])) for f in range(F)])
    for c, spec in enumerate(specs):
        sel = ids == c
        say(f"  {mixed_name(spec)}: converged {conv[sel].mean():.3f}, fallback {fb[sel].mean():.3f}, "
            f"converged to another codeword {int((conv & ~right & sel).sum())}")
    # converged frames carry the sent message on the rate-1/2 N=64800 code,
    # which is at its waterfall here; the short low-rate codes are far below
    # their thresholds and some of their frames converge to another codeword
    # (the oracle shows the same undetected errors on these codes)
    r12 = ids == [mixed_name(s) for s in specs].index("n64800-k32400")
    assert conv[r12].any() and right[r12 & conv].mean() > 0.999
    # four frames per code against the oracle composition
    for c, spec in enumerate(specs):
        sel = np.nonzero(ids == c)[0]
        pick = np.r_[sel[:2], sel[-2:]]
        llr = np.stack([packed[llr_off[f]:llr_off[f + 1]] for f in pick])
        oc = orc.Code.from_qc(spec.base_matrix(), spec.z)
        ref = frames.oracle_learned_with_fallback(oc, spec.base_matrix(), spec.z, decs[c].weights, llr, fb_it)
        for j, f in enumerate(pick):
            got = (msg[msg_off[f]:msg_off[f + 1]], it[f], conv[f], synd[f], fb[f])
            exp = tuple(r[j] for r in ref)
            assert np.array_equal(got[0], exp[0]) and tuple(got[1:]) == tuple(exp[1:]), \
                f"{mixed_name(spec)} frame {f}: got it {got[1]} conv {got[2]} fb {got[4]}, want {exp[1:]}"


def mask_words(k: int, kw: int) -> np.ndarray:
    """Mask of the first k bits in kw little-endian words."""
    m = np.full(kw, np.uint64(0xFFFFFFFFFFFFFFFF), dtype=np.uint64)
    if k % 64:
        m[-1] = np.uint64((1 << (k % 64)) - 1)
    return m
