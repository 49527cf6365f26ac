"""CPU-side checks of the C-ABI boundary and the host mirror (no GPU calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

from gf2_amd import _lib, codes
from gf2_amd.ldpc import BitVec, QuasiCyclicLdpc, qc_ira_base_matrix

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    src = open(os.path.join(ROOT, "include", "gf2ldpc.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(gf2l_\w+|gf2_ldpc_\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    syms = header_symbols()
    assert len(syms) >= 12
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(_lib.SIGNATURES), "binding table out of sync with the header"


def test_library_is_gfx950_code_object():
    blob = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in blob


def test_context_creation_fails_loudly_without_device():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    from gf2_amd.ldpc import Context
    with pytest.raises(_lib.Gf2lError):
        Context(0)


def test_qc_validation_messages():  # core.rs:519-540, 589-596
    with pytest.raises(ValueError, match="at least one row"):
        QuasiCyclicLdpc([], 4)
    with pytest.raises(ValueError, match="positive"):
        QuasiCyclicLdpc([[0]], 0)
    with pytest.raises(ValueError, match="same length"):
        QuasiCyclicLdpc([[0, 1], [0]], 3)
    with pytest.raises(ValueError, match=r"Shift value 5 at position \(0,1\)"):
        QuasiCyclicLdpc([[0, 5]], 5).to_edges()


def test_qc_mirror_to_edges_matches_reference_kat():  # qc_ldpc_tests.rs:109-118 analogue
    e = QuasiCyclicLdpc([[-1, -1, -1], [-1, -1, 1]], 3).to_edges()
    assert set(map(tuple, e.tolist())) == {(3, 7), (4, 8), (5, 6)}


def test_bitvec_layout():  # bitvec.rs:40-63
    bits = np.zeros(130, dtype=bool)
    bits[[0, 63, 64, 129]] = True
    bv = BitVec.from_bools(bits)
    w = bv.words()
    assert len(bv) == 130 and w.size == 3
    assert int(w[0]) == (1 | (1 << 63)) and int(w[1]) == 1 and int(w[2]) == 2
    assert bv.get(129) and not bv.get(128) and bv.count_ones() == 4
    assert BitVec.from_words(w, 130) == bv
    assert not (BitVec.from_words(w, 131) == bv)
    assert np.array_equal(bv.to_bools(), bits)


@pytest.mark.parametrize("spec", [codes.n2304_r12(), codes.n64800_r12()] + codes.mixed_twelve())
def test_qc_ira_structure(spec):
    base = spec.base_matrix()
    mb, nb, kb = spec.mb, spec.nb, spec.nb - spec.mb
    assert base.shape == (mb, nb)
    assert ((base == -1) | ((base >= 0) & (base < spec.z))).all()
    # info column weights follow the profile
    assert list((base[:, :kb] >= 0).sum(axis=0)) == spec.weights()
    # staircase parity part with identity blocks
    par = base[:, kb:]
    for j in range(mb):
        want = {j, j + 1} if j + 1 < mb else {j}
        assert set(np.nonzero(par[:, j] >= 0)[0]) == want
        assert (par[list(want), j] == 0).all()
    # balanced check degrees
    deg = (base >= 0).sum(axis=1)
    assert deg.max() - deg.min() <= 2
    assert deg.max() <= 32
    # deterministic for a seed
    assert np.array_equal(base, spec.base_matrix())


def test_n64800_edge_count_and_girth():
    spec = codes.n64800_r12()
    base = spec.base_matrix()
    assert int((base >= 0).sum()) == 629  # E = 226,440
    z = spec.z
    # no length-4 cycles: s_ac - s_bc + s_bd - s_ad != 0 mod z
    nz = [np.nonzero(base[r] >= 0)[0] for r in range(spec.mb)]
    for a in range(spec.mb):
        for b in range(a + 1, spec.mb):
            common = np.intersect1d(nz[a], nz[b])
            for i in range(len(common)):
                for j in range(i + 1, len(common)):
                    c, d = common[i], common[j]
                    t = (int(base[a, c]) - int(base[b, c]) + int(base[b, d]) - int(base[a, d])) % z
                    assert t != 0


def test_n2304_edge_count():
    base = codes.n2304_r12().base_matrix()
    assert int((base >= 0).sum()) == 71  # E = 6,816


def test_qc_ira_rejects_bad_arguments():
    with pytest.raises(ValueError):
        qc_ira_base_matrix(1, 4, 4, 8, [])
    with pytest.raises(ValueError):
        qc_ira_base_matrix(1, 2, 4, 8, [3, 1])  # weight above mb


def test_from_edges_bounds_messages():  # sparse.rs:170-171 panics
    from gf2_amd.ldpc import LdpcCode
    with pytest.raises(ValueError, match=r"row index 3 out of bounds \(rows=2\)"):
        LdpcCode.from_edges(2, 4, [(0, 0), (3, 1)])
    with pytest.raises(ValueError, match=r"col index 9 out of bounds \(cols=4\)"):
        LdpcCode.from_edges(2, 4, [(0, 9)])
