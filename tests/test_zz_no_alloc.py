"""gf2l_decode_batch_device_noalloc (the no-allocation variant of
gf2l_decode_batch_device, working buffers in a static device array) against
gf2l_decode_batch_device, bit for bit, through the C ABI: two calls in a row
per code, the second with a smaller batch, so the second call runs over what
the first left in the shared array.  Codes on each path it serves: the
N=64800 QC-IRA code (STATE path), the DVB-T2 normal rate-3/5 code (STATE path
with the permuted-LLR buffer) and the N=2304 code (LDS path, no scratch);
plus max_it = 0 and a batch too large for the array (a clear error, before
any work is queued)."""
import sys

import numpy as np
import pytest

from gf2_amd import _lib, codes
from gf2_amd.channel import dvb_random_frames
from gf2_amd.device import DeviceBuffer
from gf2_amd.ldpc import CodeRate, Context, LdpcCode
from tests import frames

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    return Context(0)


def decode(ctx, code, llr, max_it, noalloc):
    F, n = llr.shape
    kw = (code.k() + 63) // 64
    d_llr = DeviceBuffer.from_array(ctx, np.ascontiguousarray(llr, dtype=np.float32))
    outs = [DeviceBuffer(ctx, F * kw * 8), DeviceBuffer(ctx, F * 4), DeviceBuffer(ctx, F), DeviceBuffer(ctx, F)]
    lib = _lib.load()
    fn = lib.gf2l_decode_batch_device_noalloc if noalloc else lib.gf2l_decode_batch_device
    _lib.check(ctx.handle, fn(ctx.handle, code.handle, d_llr.ptr, F, max_it, *[o.ptr for o in outs]))
    ctx.synchronize()
    return (outs[0].download(np.uint64, (F, kw)), outs[1].download(np.uint32, F),
            outs[2].download(np.uint8, F), outs[3].download(np.uint8, F))


def same(a, b, what):
    for name, x, y in zip(["msg", "iterations", "converged", "syndrome_ok"], a, b):
        assert np.array_equal(x, y), f"{what}: {name} differs"


def two_calls(ctx, code, llr_big, llr_small, max_it, what):
    for tag, llr in (("first", llr_big), ("second, smaller", llr_small)):
        say(what, tag, llr.shape[0], "frames")
        ref = decode(ctx, code, llr, max_it, noalloc=False)
        got = decode(ctx, code, llr, max_it, noalloc=True)
        same(got, ref, f"{what} {tag}")
        assert ref[2].any(), "want some converged frames"


def test_n64800_state_path(ctx):
    spec = codes.n64800_r12()
    code = LdpcCode(spec.qc(), ctx)
    _, _, a = frames.random_frames(spec.base_matrix(), spec.z, 64, 1.55, seed=81)
    _, _, b = frames.random_frames(spec.base_matrix(), spec.z, 16, 1.6, seed=82)
    two_calls(ctx, code, a, b, 50, "n64800")


def test_dvb_normal_state_path(ctx):
    spec = codes.dvb_t2(1, 1)
    code = LdpcCode.dvb_t2_normal(CodeRate(1), ctx)
    _, _, a = dvb_random_frames(spec.table, spec.n, spec.k, spec.q, 32, 2.3, seed=83)
    _, _, b = dvb_random_frames(spec.table, spec.n, spec.k, spec.q, 8, 2.2, seed=84)
    two_calls(ctx, code, a, b, 50, "dvb normal 3/5")


def test_n2304_lds_path(ctx):
    spec = codes.n2304_r12()
    code = LdpcCode(spec.qc(), ctx)
    _, _, a = frames.random_frames(spec.base_matrix(), spec.z, 256, 2.0, seed=85)
    _, _, b = frames.random_frames(spec.base_matrix(), spec.z, 64, 1.8, seed=86)
    two_calls(ctx, code, a, b, 50, "n2304")


def test_zero_iterations(ctx):
    spec = codes.n64800_r12()
    code = LdpcCode(spec.qc(), ctx)
    _, _, a = frames.random_frames(spec.base_matrix(), spec.z, 4, 1.55, seed=87)
    say("max_it 0")
    same(decode(ctx, code, a, 0, True), decode(ctx, code, a, 0, False), "max_it 0")


def test_batch_above_scratch_is_rejected(ctx):
    spec = codes.n64800_r12()
    code = LdpcCode(spec.qc(), ctx)
    F = 1024  # about 0.7 MB of buffers per frame: well above 64 MiB
    kw = (spec.k + 63) // 64
    bufs = [DeviceBuffer(ctx, F * spec.n * 4), DeviceBuffer(ctx, F * kw * 8), DeviceBuffer(ctx, F * 4),
            DeviceBuffer(ctx, F), DeviceBuffer(ctx, F)]
    say("oversized batch")
    lib = _lib.load()
    rc = lib.gf2l_decode_batch_device_noalloc(ctx.handle, code.handle, bufs[0].ptr, F, 50,
                                              *[b.ptr for b in bufs[1:]])
    assert rc == _lib.GF2L_ERR_INVALID
    msg = lib.gf2l_last_error(ctx.handle).decode()
    assert "split the batch" in msg and str(64 << 20) in msg, msg
    # the largest batch named in the message decodes
    fit = int(msg.split("at most ")[1].split(" ")[0])
    assert 16 <= fit < F
    _, _, a = frames.random_frames(spec.base_matrix(), spec.z, fit, 1.6, seed=88)
    say("largest batch", fit)
    same(decode(ctx, code, a, 50, True), decode(ctx, code, a, 50, False), "largest batch")
