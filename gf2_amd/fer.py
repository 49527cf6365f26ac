"""Frame-error-rate sweep harness around the GPU decoder.

Mirrors the reference's FER loop (crates/gf2-coding/examples/ldpc_awgn.rs:121-175):
a frame is in error when the decoder did not converge or its syndrome check
failed (ldpc_awgn.rs:157); here a converged frame whose message differs from
the sent information bits also counts (undetected error).  Results are
reported per Eb/N0 like SimulationResult (crates/gf2-coding/src/simulation.rs:13-116):
frames, frame errors, bit errors, FER, BER, mean iterations, plus info
Gbit/s of the decode calls.

    python -m gf2_amd.fer --code n64800 --ebn0 1.5:2.5:0.25 --frames 4096 \
        --max-frames 65536 --min-errors 50 --json out.json
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from dataclasses import asdict, dataclass
from typing import List

import numpy as np

from . import channel, codes
from .ldpc import Context, LdpcCode, decode_batch_arrays


@dataclass
class FerPoint:
    ebn0_db: float
    frames: int
    frame_errors: int
    bit_errors: int
    fer: float
    ber: float
    mean_iterations: float
    decode_seconds: float
    info_gbps: float


def sweep(spec: codes.QcIraSpec, ebn0_list: List[float], frames_per_batch: int,
          max_frames: int, min_errors: int, max_it: int = 50, seed: int = 1,
          ctx: Context = None, log=sys.stderr) -> List[FerPoint]:
    ctx = ctx or Context(0)
    base = spec.base_matrix()
    code = LdpcCode(spec.qc(), ctx)
    out = []
    for ebn0 in ebn0_list:
        nf = fe = be = 0
        iters = 0
        t_dec = 0.0
        b = 0
        while nf < max_frames and fe < min_errors:
            info, _, llr = channel.random_frames(base, spec.z, frames_per_batch, ebn0,
                                                 seed=seed * 1000003 + b * 7919 + int(ebn0 * 1000))
            t0 = time.perf_counter()
            msg, it, conv, synd = decode_batch_arrays(code, llr, max_it)
            t_dec += time.perf_counter() - t0
            bits = channel.unpack_words(msg, spec.k)
            wrong = (bits != info).sum(axis=1)
            err = (~conv) | (~synd) | (wrong > 0)
            nf += frames_per_batch
            fe += int(err.sum())
            be += int(wrong.sum())
            iters += int(it.sum())
            b += 1
        p = FerPoint(ebn0, nf, fe, be, fe / nf, be / (nf * spec.k), iters / nf, t_dec,
                     nf * spec.k / t_dec / 1e9)
        print(f"[fer] {spec.name} Eb/N0={ebn0:.2f} frames={nf} errors={fe} FER={p.fer:.3e} "
              f"BER={p.ber:.3e} it={p.mean_iterations:.2f} {p.info_gbps:.3f} Gbit/s",
              file=log, flush=True)
        out.append(p)
    return out


def _parse_range(s: str) -> List[float]:
    if ":" in s:
        a, b, st = (float(x) for x in s.split(":"))
        return [round(a + i * st, 6) for i in range(int(round((b - a) / st)) + 1)]
    return [float(x) for x in s.split(",")]


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--code", default="n64800", choices=["n64800", "n2304"])
    ap.add_argument("--ebn0", default="1.5:2.5:0.25")
    ap.add_argument("--frames", type=int, default=4096)
    ap.add_argument("--max-frames", type=int, default=16384)
    ap.add_argument("--min-errors", type=int, default=50)
    ap.add_argument("--max-it", type=int, default=50)
    ap.add_argument("--json", default=None)
    a = ap.parse_args(argv)
    spec = codes.n64800_r12() if a.code == "n64800" else codes.n2304_r12()
    pts = sweep(spec, _parse_range(a.ebn0), a.frames, a.max_frames, a.min_errors, a.max_it)
    if a.json:
        with open(a.json, "w") as f:
            json.dump({"code": spec.name, "points": [asdict(p) for p in pts]}, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
