"""Learned / normalized min-sum decoders with plain-BP fallback against plain BP
on the N=64800 rate-1/2 QC-IRA code.

Every decoder runs on the same device-generated batch (seeded encoder +
BPSK/AWGN, LLRs resident in HBM) per Eb/N0; a frame is in error when it did
not converge, failed the syndrome or decoded a wrong message
(examples/ldpc_awgn.rs:157).  Decoders (comma-separated --decoders):
  plain             flooding min-sum, 50 iterations (bit-exact with the reference)
  <path>.npz        a learned decoder checkpoint (T weighted iterations) + fallback
  alpha<a>x<T>      T iterations of normalized min-sum with every weight = a
                    (the reference's alpha * min-sum rule, llr.rs:352-355) + fallback
The fallback is plain min-sum restarted from the LLRs, 50 iterations.  One
JSON line per decoder and Eb/N0: info Gbit/s, FER, mean iterations and the
fraction of frames that fell back.

    python -m gf2_amd.learned_bench [--ebn0 1.55,1.62] [--frames 4096]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import _lib, codes
from .device import DeviceBuffer, generate_frames
from .fer import count_errors
from .ldpc import Context, LdpcCode
from .nms import NmsDecoder

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT = ",".join(["plain", os.path.join(HERE, "weights", "nms_n64800_t10.npz"),
                    os.path.join(HERE, "weights", "nms_n64800_t25.npz"),
                    os.path.join(HERE, "weights", "nms_n64800_t50.npz"), "alpha0.9x50"])


def make_decoder(code: LdpcCode, name: str):
    if name == "plain":
        return None
    if name.startswith("alpha"):
        a, t = name[5:].split("x")
        return NmsDecoder(code, int(t), init=float(a))
    return NmsDecoder.load(code, name)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--ebn0", default="1.55,1.62")
    ap.add_argument("--frames", type=int, default=4096)
    ap.add_argument("--decoders", default=DEFAULT)
    ap.add_argument("--fallback", type=int, default=50)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--seed", type=int, default=7)
    a = ap.parse_args(argv)
    ctx = Context(0)
    lib = _lib.load()
    spec = codes.n64800_r12()
    code = LdpcCode(spec.qc(), ctx)
    F = a.frames
    kw, nw = (spec.k + 63) // 64, (spec.n + 63) // 64
    outs = [DeviceBuffer(ctx, F * kw * 8), DeviceBuffer(ctx, F * 4), DeviceBuffer(ctx, F), DeviceBuffer(ctx, F),
            DeviceBuffer(ctx, F)]
    names = a.decoders.split(",")
    decoders = [(n, make_decoder(code, n)) for n in names]
    for ebn0 in (float(x) for x in a.ebn0.split(",")):
        d_llr, d_cw = generate_frames(code, a.seed, 0, F, ebn0, ebn0)
        cw = d_cw.download(np.uint64, (F, nw))
        for name, dec in decoders:
            if dec is None:
                def run():
                    _lib.check(ctx.handle, lib.gf2l_decode_batch_device(
                        ctx.handle, code.handle, d_llr.ptr, F, a.fallback, *[o.ptr for o in outs[:4]]))
            else:
                def run(dec=dec):
                    dec.decode_device(d_llr, F, a.fallback, *outs)
            run()
            ctx.synchronize()
            t0 = time.perf_counter()
            for _ in range(a.reps):
                run()
            ctx.synchronize()
            dt = (time.perf_counter() - t0) / a.reps
            msg = outs[0].download(np.uint64, (F, kw))
            it = outs[1].download(np.uint32, F)
            conv = outs[2].download(np.uint8, F).astype(bool)
            synd = outs[3].download(np.uint8, F).astype(bool)
            fe, _ = count_errors(msg, cw, conv, synd, spec.k)
            line = {"ebn0_db": ebn0, "decoder": os.path.basename(name), "info_gbps": round(F * spec.k / dt / 1e9, 4),
                    "fer": fe / F, "mean_iterations": round(float(it.mean()), 3)}
            if dec is not None:
                line["weighted_iterations"] = dec.iterations
                line["fallback_fraction"] = float(outs[4].download(np.uint8, F).mean())
            print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
