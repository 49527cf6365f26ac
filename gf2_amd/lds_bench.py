"""Plain min-sum BP throughput on the N=2304 rate-1/2 QC-IRA code (Z=96).

The small-code workload: 1024 frames per batch (and larger batches that fill
the GPU) of BPSK/AWGN LLRs at Eb/N0 2.0 dB, 50 iterations with syndrome early
stop, bit-exact with the reference decoder (decode_batch,
crates/gf2-coding/src/ldpc/core.rs:729-756).  Frames are generated on the GPU
before timing; the LLRs are resident in HBM.  Each decoder path that can run
the code is timed on the same batch:
  lds     one workgroup per frame, messages / beliefs / LLRs in LDS, every
          iteration in one launch (the automatic choice for this code)
  state   compressed check states in HBM, three launches per iteration
  stream  per-edge messages in HBM, three launches per iteration
Per line: info Gbit/s, mean iterations, frames not converged, and for the
LDS path the LDS bytes it moves against the LDS peak (algorithmic bytes per
frame-iteration: check update reads a belief and a message and writes a
message per edge, 12E, and the same belief reads give the syndrome of the
previous iteration; variable update reads a message per edge, the LLR and
writes the belief, 4E + 8N).

    python -m gf2_amd.lds_bench [--frames 1024,16384] [--paths lds,state,stream] [--groups 0,4,5]
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import _lib, codes
from .channel import unpack_words
from .device import DeviceBuffer, generate_frames
from .ldpc import Context, LdpcCode

# LDS: 256 CUs x 128 bytes per clock x 2.4 GHz
LDS_PEAK_GBPS = 256 * 128 * 2.4
PATHS = {"lds": _lib.PATH_LDS, "state": _lib.PATH_STATE, "stream": _lib.PATH_STREAM}


def main(argv=None) -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--frames", default="1024,16384")
    ap.add_argument("--paths", default="lds,state,stream")
    ap.add_argument("--ebn0", type=float, default=2.0)
    ap.add_argument("--max-it", type=int, default=50)
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--groups", default="0", help="LDS path thread groups per workgroup to time (0: automatic)")
    a = ap.parse_args(argv)
    spec = codes.n2304_r12()
    ctx = Context(0)
    code = LdpcCode(spec.qc(), ctx)
    lib = _lib.load()
    base = spec.base_matrix()
    E = int((base >= 0).sum()) * spec.z
    lds_per_fi = 16 * E + 8 * spec.n
    kw = (spec.k + 63) // 64
    for F in [int(x) for x in a.frames.split(",")]:
        d_llr, d_cw = generate_frames(code, a.seed, 0, F, a.ebn0, a.ebn0)
        info = unpack_words(d_cw.download(np.uint64, (F, (spec.n + 63) // 64)), spec.n)[:, :spec.k]
        outs = [DeviceBuffer(ctx, F * kw * 8), DeviceBuffer(ctx, F * 4), DeviceBuffer(ctx, F),
                DeviceBuffer(ctx, F)]
        runs = [(name, gr) for name in a.paths.split(",")
                for gr in ([int(x) for x in a.groups.split(",")] if name == "lds" else [0])]
        for name, gr in runs:
            ctx.set_path(PATHS[name])
            ctx.set_tuning(_lib.TUNE_LDS_GROUPS, gr)

            def run():
                _lib.check(ctx.handle, lib.gf2l_decode_batch_device(
                    ctx.handle, code.handle, d_llr.ptr, F, a.max_it, *[o.ptr for o in outs]))
                ctx.synchronize()

            run()
            t0 = time.perf_counter()
            for _ in range(a.reps):
                run()
            dt = (time.perf_counter() - t0) / a.reps
            it = outs[1].download(np.uint32, F)
            conv = outs[2].download(np.uint8, F).astype(bool)
            bits = unpack_words(outs[0].download(np.uint64, (F, kw)), spec.k)
            fe = int(((~conv) | (bits != info).any(axis=1)).sum())
            line = {"code": "QC-IRA N=2304 K=1152 Z=96", "path": name, "frames": F, "ebn0_db": a.ebn0,
                    "gbps": round(F * spec.k / dt / 1e9, 4), "us_per_batch": round(dt * 1e6, 1),
                    "mean_iterations": round(float(it.mean()), 3), "frame_errors": fe}
            if name == "lds":
                line["groups"] = gr or "auto"
                # frame-iterations of the batch: each frame's iteration count
                lds = float(it.sum()) * lds_per_fi / dt / 1e9
                line["lds"] = {"achieved": round(lds, 1), "peak": round(LDS_PEAK_GBPS, 1), "unit": "GB/s",
                               "frac": round(lds / LDS_PEAK_GBPS, 4), "bytes_per_frame_iteration": lds_per_fi}
            print(json.dumps(line), flush=True)
    ctx.set_path(_lib.PATH_AUTO)
    ctx.set_tuning(_lib.TUNE_LDS_GROUPS, 0)


if __name__ == "__main__":
    main()
