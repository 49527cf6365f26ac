"""A C11 program linked against libgf2ldpc.so (tests/native/c_abi_decode.c)
decodes a batch through gf2l_decode_batch; its outputs must equal the
oracle's decode_batch (crates/gf2-coding/src/ldpc/core.rs:729-756).  This is
the boundary a foreign-language binding sees, without Python in between.
The program is built by __graft_entry__.build() (make -C tests/native c_abi).
"""
import os
import subprocess

import numpy as np
import pytest

from gf2_amd import codes
from oracle import oracle as orc
from tests import frames

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native", "_bin", "c_abi_decode")


@pytest.mark.parametrize("name", ["n2304", "n64800"])
def test_c_program_decodes_like_oracle(tmp_path, name):
    assert os.path.exists(BIN), "tests/native/_bin/c_abi_decode is not built (make -C tests/native c_abi)"
    spec = codes.n2304_r12() if name == "n2304" else codes.n64800_r12()
    base = np.ascontiguousarray(spec.base_matrix(), dtype=np.int32)
    F = 64 if name == "n2304" else 8
    _, _, llr = frames.random_frames(base, spec.z, F, 2.0 if name == "n2304" else 1.55, seed=3)
    (tmp_path / "shifts.i32").write_bytes(base.tobytes())
    (tmp_path / "llr.f32").write_bytes(np.ascontiguousarray(llr, dtype=np.float32).tobytes())
    out = tmp_path / "out.bin"
    print(f"[gpu-case] c_abi_decode {name} {F} frames", flush=True)
    mb, nb = base.shape
    p = subprocess.run([BIN, str(mb), str(nb), str(spec.z), str(tmp_path / "shifts.i32"),
                        str(tmp_path / "llr.f32"), str(F), "50", str(out)],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    raw = out.read_bytes()
    kw = (spec.k + 63) // 64
    msg = np.frombuffer(raw, dtype=np.uint64, count=F * kw).reshape(F, kw)
    it = np.frombuffer(raw, dtype=np.uint32, count=F, offset=F * kw * 8)
    conv = np.frombuffer(raw, dtype=np.uint8, count=F, offset=F * kw * 8 + 4 * F).astype(bool)
    synd = np.frombuffer(raw, dtype=np.uint8, count=F, offset=F * kw * 8 + 5 * F).astype(bool)
    ref = orc.Code.from_qc(base, spec.z).decode_batch(llr, 50, threads=8)
    for what, a, b in zip(["msg", "iterations", "converged", "syndrome_ok"], (msg, it, conv, synd), ref):
        assert np.array_equal(a, b), f"{name}: {what} differs"
