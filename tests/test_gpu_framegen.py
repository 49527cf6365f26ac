"""GPU frame generator against its oracle restatement."""
import numpy as np
import pytest

from gf2_amd import _lib, codes
from gf2_amd.device import generate_frames
from gf2_amd.ldpc import Context, LdpcCode, QuasiCyclicLdpc
from oracle import oracle as orc
from tests import frames

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    orc.build()
    return Context(0)


@pytest.mark.parametrize("spec", [codes.n2304_r12(), codes.n64800_r12()], ids=["n2304", "n64800"])
def test_framegen_matches_oracle(ctx, spec):
    base = spec.base_matrix()
    code = LdpcCode(QuasiCyclicLdpc(base, spec.z), ctx)
    F, first = 8, 1000
    print("[gpu-case] framegen", spec.name, flush=True)
    d_llr, d_cw = generate_frames(code, 12345, first, F, 1.0, 3.0)
    llr = d_llr.download(np.float32, (F, spec.n))
    cw = d_cw.download(np.uint64, (F, (spec.n + 63) // 64))
    oc = orc.Code.from_qc(base, spec.z)
    for f in range(F):
        rl, rcw = orc.framegen(base, spec.z, 12345, first + f, 1.0, 3.0)
        assert np.array_equal(cw[f], rcw), f                # bits are integer work: exact
        assert oc.syndrome_weight(frames.unpack_words(cw[f:f + 1], spec.n)[0]) == 0
        # LLRs go through double log/cos/sin/pow: equal up to the libm's last ulp
        np.testing.assert_allclose(llr[f], rl, rtol=2e-6, atol=1e-6)
        assert (llr[f] == rl).mean() > 0.99


def test_framegen_rejects_non_staircase(ctx):
    base = np.array([[0, 1, 0, -1], [2, 0, -1, 0]], dtype=np.int32)  # parity part not staircase
    code = LdpcCode(QuasiCyclicLdpc(base, 64), ctx)
    with pytest.raises(_lib.Gf2lError) as ei:
        generate_frames(code, 1, 0, 4, 2.0)
    assert ei.value.code == _lib.GF2L_ERR_UNSUPPORTED


def test_zero_codeword_frames_match_oracle(ctx):
    from gf2_amd.device import generate_zero_frames
    from gf2_amd.ldpc import CodeRate
    spec = codes.dvb_t2(0, 2)  # any code: a DVB-T2 code has no device encoder
    code = LdpcCode.dvb_t2_short(CodeRate.RATE_2_3, ctx)
    print("[gpu-case] zero-codeword frames", spec.name, flush=True)
    F, first = 6, 50
    llr = generate_zero_frames(code, 777, first, F, 1.0, 3.0).download(np.float32, (F, spec.n))
    for f in range(F):
        rl = orc.framegen_zero(spec.n, spec.k, 777, first + f, 1.0, 3.0)
        np.testing.assert_allclose(llr[f], rl, rtol=2e-6, atol=1e-6)
        assert (llr[f] == rl).mean() > 0.99


def test_fer_sweep_any_code(ctx):
    from gf2_amd import fer
    from gf2_amd.ldpc import CodeRate
    dvb = LdpcCode.dvb_t2_short(CodeRate.RATE_1_2, ctx)
    hi, lo = fer.sweep(dvb, [4.0, 0.0], 256, 512, 10**6, name="dvb short 1/2", zero_codeword=None)
    assert hi.frames == 512 and hi.frame_errors == 0 and hi.mean_iterations < 10
    assert lo.frame_errors == lo.frames  # far below threshold every frame fails
    # the QC-IRA code keeps encoded random messages; forcing all-zero gives the same statistics
    qc = LdpcCode(codes.n2304_r12().qc(), ctx)
    a = fer.sweep(qc, [2.0], 1024, 4096, 10**6, zero_codeword=False)[0]
    b = fer.sweep(qc, [2.0], 1024, 4096, 10**6, zero_codeword=True)[0]
    assert abs(a.fer - b.fer) < 5 * np.sqrt(max(a.fer, 1e-3) / 4096) + 0.01
