"""The FER-sweep harness (gf2_amd.fer) against the oracle.

fer.sweep generates each batch on the device, decodes it and counts frame
errors (not converged, syndrome failed or message wrong: the criterion of
crates/gf2-coding/examples/ldpc_awgn.rs:157 plus undetected errors) and bit
errors.  Here the same batches are regenerated, downloaded and decoded by the
oracle; frame errors, bit errors and mean iterations must be identical.  The
CSV written from the points has results_to_csv's columns
(crates/gf2-coding/src/simulation.rs:209-229).
"""
import sys

import numpy as np
import pytest

from gf2_amd import codes, fer
from gf2_amd.device import generate_frames
from gf2_amd.ldpc import Context, LdpcCode
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    orc.build()
    return Context(0)


def test_sweep_matches_oracle(ctx, tmp_path):
    spec = codes.n2304_r12()
    code = LdpcCode(spec.qc(), ctx)
    F, total, seed = 1024, 2048, 5
    points = fer.sweep(code, [1.5, 2.0], F, total, 10**9, seed=seed, name="n2304", zero_codeword=False)
    oc = orc.Code.from_qc(spec.base_matrix(), spec.z)
    nw, kw = (spec.n + 63) // 64, (spec.k + 63) // 64
    for p in points:
        print("[gpu-case] fer sweep point", p.ebn0_db, flush=True)
        sys.stdout.flush()
        fe = be = iters = 0
        for b in range(total // F):
            fseed = seed * 1000003 + int(round(p.ebn0_db * 1000))
            d_llr, d_cw = generate_frames(code, fseed, b * F, F, p.ebn0_db, p.ebn0_db)
            llr = d_llr.download(np.float32, (F, spec.n))
            cw = d_cw.download(np.uint64, (F, nw))
            msg, it, conv, synd = oc.decode_batch(llr, 50, threads=8)
            e, w = fer.count_errors(msg.reshape(F, kw), cw, conv, synd, spec.k)
            fe += e
            be += w
            iters += int(it.sum())
        assert p.frames == total
        assert (p.frame_errors, p.bit_errors) == (fe, be), (p, fe, be)
        assert p.mean_iterations == iters / total
    assert points[0].frame_errors >= points[1].frame_errors
    path = tmp_path / "fer.csv"
    fer.write_csv(points, str(path), k=spec.k)
    lines = path.read_text().splitlines()
    assert lines[0] == "eb_n0_db,ber,num_bits,num_errors,fer,num_frames,num_frame_errors"
    first = lines[1].split(",")
    assert first[0] == "1.5" and int(first[2]) == total * spec.k and int(first[3]) == points[0].bit_errors
    assert int(first[5]) == total and int(first[6]) == points[0].frame_errors
