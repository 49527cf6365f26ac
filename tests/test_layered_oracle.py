"""Pins the layered learned decoder's CPU oracle (oracle/layered_oracle.c).

The reference has no layered or learned decoder; its anchors are the
alpha * min-sum rule (crates/gf2-coding/src/llr.rs:352-355), the min-sum
itself (crates/gf2-kernels-simd/src/llr.rs:117-180, pinned in
test_oracle.py) and the flooding decoder the plain continuation runs
(crates/gf2-coding/src/ldpc/core.rs:790-938, pinned in test_oracle.py).
What pins the layered oracle:
  * hand-computed beliefs on a two-layer code (z = 1): layer 1 must see
    layer 0's updates (a flooding schedule gives a different b[3]), the
    previous iteration's message is subtracted with the previous iteration's
    weight, and the early stop fires exactly when the hard decisions satisfy
    every check;
  * continuation = the reference flooding decoder: on a one-layer code a
    layered iteration with unit weights is exactly a flooding iteration, so
    one layered iteration plus K continuation iterations must give the
    flooding oracle's beliefs after every iteration, bit for bit; with zero
    weights the layered iterations leave the LLRs untouched and the
    continuation must be a fresh flooding decode;
  * the double-precision gradient against central finite differences of
    its own loss.
"""
import numpy as np
import pytest

from gf2_amd import codes
from oracle import oracle as orc
from tests import frames

TWO_LAYERS = np.array([[0, 0, 0, -1], [-1, -1, 0, 0]], dtype=np.int32)  # c0: v0 v1 v2, c1: v2 v3


def trace(llr, w, fb_it=0, base=TWO_LAYERS, z=1):
    return orc.layered_trace(base, z, np.array(llr, dtype=np.float32), np.array(w, dtype=np.float32), fb_it)


def test_layer_sees_the_previous_layer():
    tb, msg, (it, conv, synd, fb) = trace([1.0, -2.0, 3.0, 0.5], [[1.0] * 5])
    # layer 0: messages -2, 1, -1 -> b = -1, -1, 2; layer 1 from b[2] = 2:
    # messages 0.5 and 2 -> b[2] = b[3] = 2.5 (flooding would give b[3] = 3.5)
    assert tb[0].tolist() == [-1.0, -1.0, 2.5, 2.5]
    assert (it, conv, synd, fb) == (1, True, True, False)
    assert int(msg[0]) == 0b11


def test_second_iteration_subtracts_the_previous_message():
    tb, _, (it, conv, synd, fb) = trace([0.5, 0.5, 0.5, -2.0], [[1.0] * 5, [1.0] * 5])
    assert tb[0].tolist() == [1.0, 1.0, -1.0, -1.0]
    assert tb[1].tolist() == [0.0, 0.0, -1.0, -1.0]
    assert (it, conv, synd, fb) == (2, False, False, False)
    # weight 0.5 in iteration 2 scales only iteration 2's messages
    tb, _, _ = trace([0.5, 0.5, 0.5, -2.0], [[1.0] * 5, [0.5] * 5])
    assert tb[1].tolist() == [0.25, 0.25, -0.25, -1.625]
    # weight 0.5 in iteration 1: iteration 2 subtracts 0.5 * m_1 (w_{t-1}, not w_t)
    tb, _, _ = trace([0.5, 0.5, 0.5, -2.0], [[0.5] * 5, [1.0] * 5])
    assert tb[0].tolist() == [0.75, 0.75, -0.25, -1.625]
    assert tb[1].tolist() == [0.0, 0.0, -1.0, -1.0]


def test_continuation_subtracts_the_weighted_last_message():
    # T = 1 with unit weight: layered beliefs 1, 1, -1, -1 fail c0; the
    # flooding continuation subtracts m_1 (0.5 on c0, -2 and 1 on c1), sends
    # -0.5, -0.5, 0.5 and -2, 1, and rebuilds beliefs from the LLRs in check order
    tb, _, (it, conv, synd, fb) = trace([0.5, 0.5, 0.5, -2.0], [[1.0] * 5], 1)
    assert tb[0].tolist() == [1.0, 1.0, -1.0, -1.0]
    assert tb[1].tolist() == [0.0, 0.0, -1.0, -1.0]
    assert (it, conv, synd, fb) == (2, False, False, True)
    # weight 0.5: the continuation subtracts 0.5 * m_1 (0.25; -1 and 0.375)
    tb, _, (it, _, _, fb) = trace([0.5, 0.5, 0.5, -2.0], [[0.5] * 5], 1)
    assert tb[0].tolist() == [0.75, 0.75, -0.25, -1.625]
    assert tb[1].tolist() == [0.0, 0.0, -1.0, -1.25]
    assert (it, fb) == (2, True)


def test_zero_weights_then_continuation_is_a_fresh_flooding_decode():
    spec = codes.n2304_r12()
    base = spec.base_matrix()
    oc = orc.Code.from_qc(base, spec.z)
    _, _, llr = frames.random_frames(base, spec.z, 24, 1.6, seed=8)
    T, K = 3, 30
    w = np.zeros((T, int((base >= 0).sum())), dtype=np.float32)
    msg, it, conv, synd, fb = orc.layered_decode_batch(base, spec.z, llr, w, K)
    rmsg, rit, rconv, rsynd = oc.decode_batch(llr, K, threads=4)
    assert fb.all()  # sign(LLR) is not a codeword for any of these frames
    assert np.array_equal(msg, rmsg) and np.array_equal(it, T + rit)
    assert np.array_equal(conv, rconv) and np.array_equal(synd, rsynd)


def test_layered_converges_faster_than_flooding():
    """The schedule's purpose, on the N=2304 code: at equal (normalized) weights
    the layered decoder needs fewer iterations than the flooding one."""
    spec = codes.n2304_r12()
    base = spec.base_matrix()
    nbe = int((base >= 0).sum())
    _, _, llr = frames.random_frames(base, spec.z, 64, 2.0, seed=9)
    _, it_l, conv_l, _, _ = orc.layered_decode_batch(base, spec.z, llr, np.full((50, nbe), 0.8, np.float32), 0)
    oc = orc.Code.from_qc(base, spec.z)
    _, it_f, conv_f, _ = oc.decode_batch(llr, 50, edge_w=frames.edge_weights(
        base, spec.z, np.full((50, nbe), 0.8, np.float32), oc), threads=4)
    assert conv_l.sum() >= conv_f.sum()
    both = conv_l & conv_f
    assert it_l[both].mean() < 0.75 * it_f[both].mean()


@pytest.mark.parametrize("T", [1, 3])
def test_gradient_matches_finite_differences(T):
    spec = codes.n2304_r12()
    base = spec.base_matrix()
    nbe = int((base >= 0).sum())
    info, cw, llr = frames.random_frames(base, spec.z, 1, 1.5, seed=12)
    bits = frames.unpack_words(cw, spec.n)[0].astype(np.uint8) if cw.dtype == np.uint64 else cw[0].astype(np.uint8)
    rng = np.random.default_rng(T)
    w = rng.uniform(0.6, 1.0, size=(T, nbe))
    loss, g = orc.layered_loss_grad(base, spec.z, llr[0], bits, w)
    assert np.isfinite(loss) and loss > 0
    h = 1e-5
    for _ in range(24):
        t, e = int(rng.integers(T)), int(rng.integers(nbe))
        wp, wm = w.copy(), w.copy()
        wp[t, e] += h
        wm[t, e] -= h
        fd = (orc.layered_loss_grad(base, spec.z, llr[0], bits, wp)[0] -
              orc.layered_loss_grad(base, spec.z, llr[0], bits, wm)[0]) / (2 * h)
        assert abs(fd - g[t, e]) <= 1e-6 + 1e-4 * abs(fd), (t, e, fd, g[t, e])
