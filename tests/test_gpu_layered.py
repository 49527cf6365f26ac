"""The layered learned decoder and the continuing fallback on the GPU,
against the CPU oracle.

Decoding is bit-exact: the layered kernel (gf2l_nms_create_layered) and the
plain flooding continuation must give exactly oracle/layered_oracle.c's
messages, iteration counts, flags and fallback marks, with the beliefs of
every block-column in LDS (N=2304), with most in LDS and the rest in global
memory (N=64800), and with a forced small LDS share (GF2L_TUNE_LAYERED_COLUMNS).
The flooding decoder's continue mode must equal the flooding oracle run T +
fb_it iterations with unit weights after the learned ones (llr.rs:352-355;
core.rs:889-938).  The training gradient is float32 against the oracle's
double-precision reverse mode (pinned by finite differences in
tests/test_layered_oracle.py); the tolerance is written in the test.
"""
import sys

import numpy as np
import pytest

from gf2_amd import _lib, codes
from gf2_amd.device import generate_frames
from gf2_amd.ldpc import Context, LdpcCode
from gf2_amd.nms import NmsDecoder, train
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
    c.set_tuning(_lib.TUNE_LAYERED_COLUMNS, 0)


def same(got, ref, what):
    for name, a, b in zip(["msg", "iterations", "converged", "syndrome_ok", "fallback"], got, ref):
        a, b = np.asarray(a), np.asarray(b)
        assert np.array_equal(a, b), f"{what}: {name} differs at {np.nonzero((a != b).reshape(len(a), -1).any(1))[0][:8]}"


# kernel: the automatic choice (the staircase kernel on both codes: QC-IRA
# staircases of unshifted circulants, mb = 12 and 90) or the general kernel,
# with the automatic column count or a forced small one
@pytest.mark.parametrize("which,cols,kernel", [("n2304", 0, _lib.LAYERED_AUTO), ("n2304", 0, _lib.LAYERED_GENERAL),
                                               ("n2304", 5, _lib.LAYERED_AUTO), ("n64800", 0, _lib.LAYERED_AUTO),
                                               ("n64800", 0, _lib.LAYERED_GENERAL),
                                               ("n64800", 37, _lib.LAYERED_AUTO)])
@pytest.mark.parametrize("fb_it", [0, 30])
def test_layered_decode_matches_oracle(ctx, which, cols, kernel, fb_it):
    spec = codes.n2304_r12() if which == "n2304" else codes.n64800_r12()
    base = spec.base_matrix()
    nbe = int((base >= 0).sum())
    T = 10 if which == "n2304" else 8
    rng = np.random.default_rng(51)
    w = rng.uniform(0.6, 1.0, size=(T, nbe)).astype(np.float32) if which == "n2304" else \
        rng.uniform(0.8, 0.95, size=(T, nbe)).astype(np.float32)
    F = 256 if which == "n2304" else 32
    _, _, llr = frames.random_frames(base, spec.z, F, 1.7 if which == "n2304" else 1.75, seed=52)
    ctx.set_tuning(_lib.TUNE_LAYERED_COLUMNS, cols)
    ctx.set_tuning(_lib.TUNE_LAYERED_KERNEL, kernel)
    dec = NmsDecoder(LdpcCode(spec.qc(), ctx), T, weights=w, schedule="layered")
    x = dec.layered_columns
    kb = spec.n // spec.z - (spec.n - spec.k) // spec.z
    stair = kernel == _lib.LAYERED_AUTO and cols == 0
    assert (x == kb) if stair else x == (cols or x) > 0, x
    say("layered", which, "lds columns", x, "kernel", "staircase" if stair else "general", "fallback", fb_it)
    got = dec.decode(llr, fallback_max_it=fb_it)
    ctx.set_tuning(_lib.TUNE_LAYERED_COLUMNS, 0)
    ctx.set_tuning(_lib.TUNE_LAYERED_KERNEL, _lib.LAYERED_AUTO)
    ref = orc.layered_decode_batch(base, spec.z, llr, w, fb_it, threads=8)
    same(got, ref, f"{which} cols {cols} fb {fb_it}")
    conv, fb = ref[2], ref[4]
    assert (conv & ~fb).any(), "some frames converge in the layered phase"
    if fb_it:
        assert fb.any(), "some frames run the continuation"


def test_staircase_kernel_where_it_does_not_apply(ctx):
    """Asking for the staircase kernel on codes without its structure fails
    with GF2L_ERR_UNSUPPORTED; the automatic choice runs the general kernel
    there (bit-exact against the oracle)."""
    two = {(s.n, s.k): s for s in codes.mixed_twelve()}
    # 8 block-rows (no load-ahead distance divides them with two groups);
    # 108 information columns (more than one workgroup's LDS)
    for spec in (two[(16200, 13320)], two[(64800, 38880)]):
        base = spec.base_matrix()
        nbe = int((base >= 0).sum())
        w = np.full((4, nbe), 0.8, dtype=np.float32)
        _, _, llr = frames.random_frames(base, spec.z, 4, 3.5, seed=57)
        dec = NmsDecoder(LdpcCode(spec.qc(), ctx), 4, weights=w, schedule="layered")
        ctx.set_tuning(_lib.TUNE_LAYERED_KERNEL, _lib.LAYERED_STAIRCASE)
        say("staircase kernel forced on", spec.name)
        try:
            with pytest.raises(_lib.Gf2lError) as e:
                dec.decode(llr, fallback_max_it=5)
            assert e.value.code == _lib.GF2L_ERR_UNSUPPORTED
        finally:
            ctx.set_tuning(_lib.TUNE_LAYERED_KERNEL, _lib.LAYERED_AUTO)
        say("automatic kernel on", spec.name, "lds columns", dec.layered_columns)
        same(dec.decode(llr, fallback_max_it=5), orc.layered_decode_batch(base, spec.z, llr, w, 5), spec.name)


def test_layered_signed_zero_and_infinite_llrs(ctx):
    spec = codes.n2304_r12()
    base = spec.base_matrix()
    nbe = int((base >= 0).sum())
    _, _, llr = frames.random_frames(base, spec.z, 16, 2.0, seed=53)
    llr[0, ::7] = 0.0
    llr[1, ::5] = -0.0
    llr[2, ::11] = np.inf
    llr[3, ::13] = -np.inf
    llr[4] = -0.0
    w = np.full((5, nbe), 0.75, dtype=np.float32)
    say("layered +-0 / inf")
    dec = NmsDecoder(LdpcCode(spec.qc(), ctx), 5, weights=w, schedule="layered")
    same(dec.decode(llr, fallback_max_it=10), orc.layered_decode_batch(base, spec.z, llr, w, 10), "signed zero")


@pytest.mark.parametrize("which", ["n2304", "n64800"])
def test_flooding_continue_mode_matches_oracle(ctx, which):
    """Flooding learned decoder with the continuing fallback: exactly the
    flooding oracle run T + fb_it iterations, weights w for the first T and 1
    after (multiplying by 1.0 is exact)."""
    spec = codes.n2304_r12() if which == "n2304" else codes.n64800_r12()
    base = spec.base_matrix()
    nbe = int((base >= 0).sum())
    T, fb_it = 10, 20
    rng = np.random.default_rng(54)
    w = rng.uniform(0.7, 0.95, size=(T, nbe)).astype(np.float32)
    F = 256 if which == "n2304" else 16
    _, _, llr = frames.random_frames(base, spec.z, F, 1.6 if which == "n2304" else 1.45, seed=55)
    paths = ["lds", "state", "stream"] if which == "n2304" else ["state", "stream"]
    oc = orc.Code.from_qc(base, spec.z)
    wfull = np.concatenate([w, np.ones((fb_it, nbe), np.float32)])
    msg, it, conv, synd = oc.decode_batch(llr, T + fb_it, edge_w=frames.edge_weights(base, spec.z, wfull, oc),
                                          threads=8)
    ref = (msg, it, conv, synd, (it > T) | ~conv)
    for path in paths:
        ctx.set_path({"lds": _lib.PATH_LDS, "state": _lib.PATH_STATE, "stream": _lib.PATH_STREAM}[path])
        say("flooding continue", which, path)
        dec = NmsDecoder(LdpcCode(spec.qc(), ctx), T, weights=w, fallback="continue")
        same(dec.decode(llr, fallback_max_it=fb_it), ref, f"{which}/{path}")
    ctx.set_path(_lib.PATH_AUTO)
    assert ref[4].any()  # frames that ran the continuation
    if which == "n2304":
        assert (~ref[4]).any()  # and frames that converged in the weighted phase


@pytest.mark.parametrize("which", ["n2304", "n64800"])
def test_layered_gradient_matches_double_oracle(ctx, which):
    spec = codes.n2304_r12() if which == "n2304" else codes.n64800_r12()
    base = spec.base_matrix()
    nbe = int((base >= 0).sum())
    T = 5 if which == "n2304" else 2
    F = 3 if which == "n2304" else 2
    rng = np.random.default_rng(56)
    w = rng.uniform(0.6, 1.0, size=(T, nbe)).astype(np.float32)
    code = LdpcCode(spec.qc(), ctx)
    d_llr, d_cw = generate_frames(code, 78, 0, F, 1.0, 2.0)
    say("layered gradient", which)
    dec = NmsDecoder(code, T, weights=w, schedule="layered")
    g, loss = dec.gradient(d_llr, d_cw, F)
    llr = d_llr.download(np.float32, (F, spec.n))
    bits = frames.unpack_words(d_cw.download(np.uint64, (F, (spec.n + 63) // 64)), spec.n)
    ref = np.zeros((T, nbe))
    ref_loss = 0.0
    for f in range(F):
        lf, gf = orc.layered_loss_grad(base, spec.z, llr[f], bits[f], w.astype(np.float64))
        ref_loss += lf / F
        ref += gf / F
    # float32 forward/backward against double precision
    assert abs(loss - ref_loss) <= 1e-4 * abs(ref_loss), (loss, ref_loss)
    err = np.linalg.norm(g - ref) / np.linalg.norm(ref)
    assert err < 2e-3, err


def test_layered_training_gain(ctx):
    """Training the layered network lowers its 10-iteration FER on fresh frames
    below the same decoder with every weight 0.85 (the best constant found for
    this schedule on the N=2304 code)."""
    spec = codes.n2304_r12()
    code = LdpcCode(spec.qc(), ctx)
    dec = NmsDecoder(code, 10, init=0.85, schedule="layered")
    say("layered training")
    losses = train(dec, steps=150, batch=1024, lr=3e-3, ebn0_lo=1.0, ebn0_hi=3.0, seed=13)
    assert losses[-20:].mean() < losses[:20].mean()
    F = 8192
    d_llr, d_cw = generate_frames(code, 998, 0, F, 2.0, 2.0)
    llr = d_llr.download(np.float32, (F, spec.n))
    info = frames.unpack_words(d_cw.download(np.uint64, (F, (spec.n + 63) // 64)), spec.n)[:, :spec.k]
    fer = {}
    for name, d in [("trained", dec), ("constant", NmsDecoder(code, 10, init=0.85, schedule="layered"))]:
        msg, it, conv, synd, fb = d.decode(llr, fallback_max_it=0)
        bits = frames.unpack_words(msg, spec.k)
        fer[name] = float(((~conv) | (bits != info).any(axis=1)).mean())
    say("fer", fer)
    assert fer["trained"] < fer["constant"], fer
