"""Pins the oracle's sum-product check rule and decoder (or_boxplus_n,
or_decode_spa: llr.rs:210-215, core.rs:762-787) before any GPU comparison.

- The reference's own boxplus_n tests (llr.rs:581-621, 810-822, 878-890).
- The box-plus identity: for two inputs 2 atanh(tanh(a/2) tanh(b/2)) equals
  log((1 + e^(a+b)) / (e^a + e^b)); for n inputs the n-fold box-plus is the
  LLR of the XOR of independent bits, computed here by brute force over all
  2^n bit patterns in double precision.
- The decoder: on a cycle-free Tanner graph, sum-product BP is exact after
  as many iterations as the graph's diameter, so its beliefs are the exact
  a-posteriori LLRs, computed here by brute force over all codewords.  A
  wrong message order, a dropped channel term or a transposed edge breaks it.
"""
import itertools

import numpy as np
import pytest

from oracle import oracle as orc


@pytest.fixture(scope="module", autouse=True)
def built():
    orc.build()


def bp(x):
    return float(orc.boxplus_n(np.asarray(x, dtype=np.float32)))


def test_reference_boxplus_n_cases():
    r = bp([3.0, 2.0, 4.0])
    assert 0.0 < r < 2.0                                   # llr.rs:581-588
    assert bp([3.0, 2.0, -4.0]) < 0.0                      # llr.rs:590-596
    assert bp([3.0, -2.0, -4.0, 5.0]) > 0.0                # llr.rs:598-604
    assert abs(bp([3.5]) - 3.5) < 0.1                      # llr.rs:606-612
    # two elements: the binary boxplus (llr.rs:144-148, 614-620)
    a, b = np.float32(3.0), np.float32(2.0)
    binary = 2.0 * np.arctanh(np.float32(np.tanh(a / 2)) * np.float32(np.tanh(b / 2)))
    assert abs(bp([3.0, 2.0]) - binary) < 1e-5
    rng = np.random.default_rng(5)
    for _ in range(200):                                   # llr.rs:810-822
        v = rng.uniform(-10, 10, size=rng.integers(2, 6)).astype(np.float32)
        assert abs(bp(v) - bp(v[::-1])) < 1e-2


def test_boxplus_two_inputs_closed_form():
    rng = np.random.default_rng(6)
    for _ in range(500):
        a, b = rng.uniform(-7, 7, size=2)
        a, b = float(np.float32(a)), float(np.float32(b))
        exact = np.log1p(np.exp(a + b)) - np.log(np.exp(a) + np.exp(b))
        got = bp([a, b])
        assert abs(got - exact) <= 1e-4 * max(1.0, abs(exact)), (a, b, got, exact)


def xor_llr_bruteforce(x):
    """LLR of b_1 xor ... xor b_n for independent bits with LLRs x (P(b=0)/P(b=1) = e^x)."""
    p0 = [1.0 / (1.0 + np.exp(-v)) for v in x]
    even = 0.0
    for bits in itertools.product([0, 1], repeat=len(x)):
        pr = 1.0
        for bit, q in zip(bits, p0):
            pr *= (1.0 - q) if bit else q
        if sum(bits) % 2 == 0:
            even += pr
    return np.log(even / (1.0 - even))


def test_boxplus_n_is_the_xor_llr():
    rng = np.random.default_rng(7)
    for n in range(1, 7):
        for _ in range(60):
            x = rng.uniform(-5, 5, size=n).astype(np.float32)
            exact = xor_llr_bruteforce([float(v) for v in x])
            assert abs(bp(x) - exact) <= 2e-4 * max(1.0, abs(exact)), (x, bp(x), exact)


def tree_code():
    """A cycle-free Tanner graph: 4 checks on 9 variables, a tree."""
    checks = [[0, 1, 2], [2, 3, 4], [4, 5, 6], [1, 7, 8]]
    edges = [(r, v) for r, vs in enumerate(checks) for v in vs]
    return 4, 9, checks, np.array(edges, dtype=np.uint32)


def exact_app(checks, n, llr):
    """A-posteriori LLRs of every bit over all codewords (uniform prior)."""
    lp1 = np.zeros(n)
    lp0 = np.zeros(n)
    logw = []
    words = []
    for bits in itertools.product([0, 1], repeat=n):
        if all(sum(bits[v] for v in c) % 2 == 0 for c in checks):
            words.append(bits)
            # p(y | x) up to a constant: exp(sum over bits of -x_v * llr_v / 2)
            logw.append(sum((0.5 if b == 0 else -0.5) * float(l) for b, l in zip(bits, llr)))
    logw = np.array(logw)
    w = np.exp(logw - logw.max())
    words = np.array(words)
    p1 = (w[:, None] * words).sum(0)
    p0 = (w[:, None] * (1 - words)).sum(0)
    return np.log(p0 / p1)


def test_spa_decoder_is_exact_on_a_tree():
    m, n, checks, edges = tree_code()
    code = orc.Code(m, n, edges)
    rng = np.random.default_rng(8)
    checked = 0
    for _ in range(40):
        llr = rng.normal(0.6, 1.2, size=n).astype(np.float32)
        tc, tb, msg, res = code.decode_trace_spa(llr, 8)
        # run all 8 iterations unless the syndrome stopped it: the last
        # recorded beliefs are those after res[0] iterations
        it = res[0]
        if it < 4:  # stopped before the messages crossed the tree's diameter
            continue
        app = exact_app(checks, n, llr)
        got = tb[it - 1].astype(np.float64)
        assert np.allclose(got, app, rtol=1e-4, atol=1e-4), (llr, got, app)
        checked += 1
    assert checked >= 10, checked


def test_spa_decoder_against_minsum_structure():
    """With every check of degree 2 the sum-product and min-sum messages are
    both the other input (box-plus of one value is that value up to
    atanh(tanh) rounding), so the two decoders agree on a repetition code."""
    n = 6
    edges = np.array([(r, v) for r in range(n - 1) for v in (r, r + 1)], dtype=np.uint32)
    code = orc.Code(n - 1, n, edges)
    rng = np.random.default_rng(9)
    llr = rng.normal(1.0, 1.0, size=(32, n)).astype(np.float32)
    a = code.decode_batch_spa(llr, 10)
    b = code.decode_batch(llr, 10)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
