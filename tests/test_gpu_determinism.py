"""Determinism: the same input run twice gives identical bytes.

The kernels hand out work through atomic counters (items per XCD, frames of
the layered decoder, syndrome workgroups), so the order in which workgroups
take work changes from run to run; the results must not.  Each case decodes
(or differentiates, or generates) the same input twice on one context and
compares every output byte.
"""
import ctypes
import os
import sys

import numpy as np
import pytest

from gf2_amd import _lib, codes
from gf2_amd.device import DeviceBuffer, generate_frames
from gf2_amd.ldpc import Context, LdpcCode, QuasiCyclicLdpc
from gf2_amd.mixed import CodeSet, decode_mixed_arrays
from gf2_amd.nms import NmsDecoder

pytestmark = pytest.mark.gpu
WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gf2_amd", "weights")


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    return Context(0)


def outputs(ctx, F, k, extra=0):
    kw = (k + 63) // 64
    return [DeviceBuffer(ctx, F * kw * 8), DeviceBuffer(ctx, F * 4), DeviceBuffer(ctx, F), DeviceBuffer(ctx, F)] + \
        [DeviceBuffer(ctx, F) for _ in range(extra)]


def snapshot(bufs):
    return [b.download(np.uint8, b.nbytes) for b in bufs]


def twice(run, bufs):
    run()
    a = snapshot(bufs)
    run()
    b = snapshot(bufs)
    return all(np.array_equal(x, y) for x, y in zip(a, b))


@pytest.mark.parametrize("chunk", [-1, 0])
def test_plain_decode_twice(ctx, chunk):
    spec = codes.n64800_r12()
    code = LdpcCode(spec.qc(), ctx)
    F = 512
    d_llr, _ = generate_frames(code, 31, 0, F, 1.5, 1.5, with_codewords=False)
    outs = outputs(ctx, F, spec.k)
    lib = _lib.load()
    ctx.set_tuning(_lib.TUNE_CHUNK, chunk)
    say("plain twice, chunk", chunk)

    def run():
        _lib.check(ctx.handle, lib.gf2l_decode_batch_device(ctx.handle, code.handle, d_llr.ptr, F, 50,
                                                            *[o.ptr for o in outs]))
        ctx.synchronize()

    try:
        assert twice(run, outs)
    finally:
        ctx.set_tuning(_lib.TUNE_CHUNK, -1)


@pytest.mark.parametrize("path", ["lds", "state", "stream"])
def test_small_code_paths_twice(ctx, path):
    spec = codes.n2304_r12()
    code = LdpcCode(spec.qc(), ctx)
    F = 2048
    d_llr, _ = generate_frames(code, 35, 0, F, 1.0, 2.5, with_codewords=False)
    outs = outputs(ctx, F, spec.k)
    lib = _lib.load()
    ctx.set_path({"lds": _lib.PATH_LDS, "state": _lib.PATH_STATE, "stream": _lib.PATH_STREAM}[path])
    say("N=2304 twice, path", path)

    def run():
        _lib.check(ctx.handle, lib.gf2l_decode_batch_device(ctx.handle, code.handle, d_llr.ptr, F, 50,
                                                            *[o.ptr for o in outs]))
        ctx.synchronize()

    try:
        assert twice(run, outs)
    finally:
        ctx.set_path(_lib.PATH_AUTO)


def test_layered_decode_twice(ctx):
    spec = codes.n64800_r12()
    code = LdpcCode(spec.qc(), ctx)
    dec = NmsDecoder.load(code, os.path.join(WEIGHTS, "nms_n64800_layered_t10.npz"))
    F = 512
    d_llr, _ = generate_frames(code, 32, 0, F, 1.45, 1.45, with_codewords=False)
    outs = outputs(ctx, F, spec.k, extra=1)
    say("layered twice")

    def run():
        dec.decode_device(d_llr, F, 50, *outs)
        ctx.synchronize()

    assert twice(run, outs)
    assert outs[4].download(np.uint8, F).any()  # some frames ran the continuation


def test_mixed_decode_twice(ctx):
    specs = codes.mixed_twelve()[:4]
    gcodes = [LdpcCode(QuasiCyclicLdpc(s.base_matrix(), s.z), ctx) for s in specs]
    cs = CodeSet(gcodes)
    rng = np.random.default_rng(33)
    ids = rng.permutation(np.repeat(np.arange(4, dtype=np.uint16), 24))
    blocks = [rng.normal(1.2, 1.7, size=specs[c].n).astype(np.float32) for c in ids]
    packed = np.concatenate(blocks)
    say("mixed twice")
    a = decode_mixed_arrays(cs, ids, packed, 30)
    b = decode_mixed_arrays(cs, ids, packed, 30)
    assert all(np.array_equal(np.asarray(x), np.asarray(y)) for x, y in zip(a, b))


def test_gradient_and_frames_twice(ctx):
    spec = codes.n2304_r12()
    code = LdpcCode(spec.qc(), ctx)
    F = 1024
    l1, c1 = generate_frames(code, 34, 0, F, 1.0, 3.0)
    l2, c2 = generate_frames(code, 34, 0, F, 1.0, 3.0)
    assert np.array_equal(l1.download(np.uint8, l1.nbytes), l2.download(np.uint8, l2.nbytes))
    assert np.array_equal(c1.download(np.uint8, c1.nbytes), c2.download(np.uint8, c2.nbytes))
    say("gradient twice")
    for schedule in ("flooding", "layered"):
        dec = NmsDecoder(code, 10, init=0.85, schedule=schedule)
        g1, loss1 = dec.gradient(l1, c1, F)
        g2, loss2 = dec.gradient(l1, c1, F)
        assert np.array_equal(g1, g2) and loss1 == loss2, schedule


@pytest.mark.parametrize("level", [0, 1, 2])
def test_profiling_levels_leave_results_unchanged(ctx, level):
    """Profiling records events (and, for every launch, per-iteration
    readbacks) but decodes the same bytes; the layered-only level times the
    layered kernel and none of the continuation's launches."""
    spec = codes.n64800_r12()
    code = LdpcCode(spec.qc(), ctx)
    dec = NmsDecoder.load(code, os.path.join(WEIGHTS, "nms_n64800_layered_t10.npz"))
    F = 512
    d_llr, _ = generate_frames(code, 36, 0, F, 1.45, 1.45, with_codewords=False)
    outs = outputs(ctx, F, spec.k, extra=1)
    lib = _lib.load()
    dec.decode_device(d_llr, F, 50, *outs)
    ctx.synchronize()
    ref = snapshot(outs)
    say("profiling level", level)
    _lib.check(ctx.handle, lib.gf2l_profile_enable(ctx.handle, level))
    try:
        dec.decode_device(d_llr, F, 50, *outs)
        ctx.synchronize()
        got = snapshot(outs)
        counts = {}
        for k in (_lib.KERNEL_CHECK, _lib.KERNEL_VARIABLE, _lib.KERNEL_LAYERED):
            ms, la, fi = ctypes.c_double(), ctypes.c_uint64(), ctypes.c_uint64()
            _lib.check(ctx.handle, lib.gf2l_profile_read_kernel(ctx.handle, k, ctypes.byref(ms), ctypes.byref(la),
                                                                ctypes.byref(fi)))
            counts[k] = (la.value, fi.value, ms.value)
    finally:
        _lib.check(ctx.handle, lib.gf2l_profile_enable(ctx.handle, 0))
    assert all(np.array_equal(x, y) for x, y in zip(ref, got))
    fb = outs[4].download(np.uint8, F)
    assert fb.any()  # the continuation ran
    layered, check = counts[_lib.KERNEL_LAYERED], counts[_lib.KERNEL_CHECK]
    if level == 0:
        assert layered[0] == 0 and check[0] == 0
    else:
        assert layered[0] == 1 and layered[2] > 0
        # iterations run by the layered kernel, summed over frames
        its = outs[1].download(np.uint32, F)
        assert layered[1] == int(np.minimum(its, dec.iterations).sum())
        assert (check[0] > 0) == (level == _lib.PROFILE_ALL)
