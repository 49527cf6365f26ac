"""General sparse codes (LdpcCode::from_edges) on the CSR path vs the oracle."""
import numpy as np
import pytest

from gf2_amd import _lib, codes
from gf2_amd.ldpc import Context, LdpcCode, LdpcDecoder, QuasiCyclicLdpc, decode_batch_arrays
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


def test_reference_kats_through_from_edges(ctx):
    # core.rs:1150-1165, 1272-1309; backend_integration.rs:44-75: SPC[3,2]
    code = LdpcCode.from_edges(1, 3, [(0, 0), (0, 1), (0, 2)], ctx)
    res = LdpcDecoder.decode_batch(code, [[10.0] * 3, [-10.0, -10.0, 10.0], [10.0] * 3], 10)
    assert [r.converged for r in res] == [True] * 3
    assert [r.decoded_bits.count_ones() for r in res] == [0, 2, 0]
    assert res[0].iterations <= 2
    assert LdpcDecoder.decode_batch(code, [], 10) == []
    r = LdpcDecoder.decode_batch(code, [[-3.0, 1.0, 2.0]], 0)[0]
    assert r.iterations == 0 and not r.converged and r.syndrome_check_passed
    # duplicates cancel (sparse.rs doc example): (0,1) twice leaves only (1,2)
    c2 = LdpcCode.from_edges(2, 3, [(0, 1), (0, 1), (1, 2)], ctx)
    assert c2.nnz() == 1


def test_random_sparse_codes_match_oracle(ctx):
    rng = np.random.default_rng(41)
    for trial in range(4):
        m, n = int(rng.integers(20, 200)), int(rng.integers(250, 600))
        deg = rng.integers(1, 3 + 9 * trial, size=m)   # includes degree-1 checks, up to 30
        edges = [(r, int(c)) for r in range(m) for c in rng.choice(n, min(int(deg[r]), 32), replace=False)]
        edges += edges[:5]  # a few duplicates that cancel
        llr = rng.normal(0.7, 1.8, size=(48, n)).astype(np.float32)
        llr[rng.random(llr.shape) < 0.01] = -0.0
        print("[gpu-case] csr random", trial, m, n, flush=True)
        code = LdpcCode.from_edges(m, n, edges, ctx)
        oc = orc.Code(m, n, np.array(edges))
        same(decode_batch_arrays(code, llr, 25), oc.decode_batch(llr, 25, threads=8), f"csr{trial}")


def test_qc_code_through_edges_equals_qc_path(ctx):
    spec = codes.n2304_r12()
    base = spec.base_matrix()
    _, _, llr = frames.random_frames(base, spec.z, 256, 2.0, seed=42)
    qc = QuasiCyclicLdpc(base, spec.z)
    a = decode_batch_arrays(LdpcCode(qc, ctx), llr, 50)
    b = decode_batch_arrays(LdpcCode.from_edges(spec.mb * spec.z, spec.n, qc.to_edges(), ctx), llr, 50)
    same(b, a, "qc-vs-csr")
    c = LdpcCode.from_edges(spec.mb * spec.z, spec.n, qc.to_edges(), ctx)
    with pytest.raises(_lib.Gf2lError):
        from gf2_amd.nms import NmsDecoder
        NmsDecoder(c, 10)
