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
