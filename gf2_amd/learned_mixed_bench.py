"""Learned decoding plus plain-BP fallback on a mixed batch of the twelve Z=360 codes.

For each Eb/N0 of the sweep (default 0.8 ... 2.0 dB in 0.1 dB steps) a batch
of 16384 frames spread uniformly over the twelve codes is generated on the
GPU and decoded twice in one call each:
  learned  gf2l_decode_mixed_nms_device: one learned decoder per code
           (gf2_amd/weights/nms_n<N>-k<K>_t10.npz, trained by
           python -m gf2_amd.train_nms --all-mixed), 10 weighted iterations,
           then plain min-sum up to 50 iterations in all for frames that fail,
           restarted from the LLRs or (--mode continue) continued from the
           weighted state
  plain    gf2l_decode_mixed_device: plain 50-iteration min-sum
A frame is in error when it did not converge, its syndrome failed, or its
message differs from the sent information bits (ldpc_awgn.rs:157 plus
undetected errors).  One JSON line per Eb/N0: per-code FER of both decoders,
fallback fraction, and info Gbit/s of both calls.

    python -m gf2_amd.learned_mixed_bench [--frames 16384] [--ebn0 0.8:2.0:0.1] [--mode restart|continue]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import _lib, codes
from .device import DeviceBuffer
from .fer import _parse_range, count_errors
from .ldpc import Context, LdpcCode
from .mixed import CodeSet
from .nms import NmsDecoder
from .train_nms import default_path, mixed_name


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--frames", type=int, default=16384)
    ap.add_argument("--ebn0", default="0.8:2.0:0.1")
    ap.add_argument("--fallback", type=int, default=50)
    ap.add_argument("--seed", type=int, default=17)
    ap.add_argument("--mode", choices=["restart", "continue"], default="restart",
                    help="the learned decoders' fallback: from the LLRs or from the weighted state")
    ap.add_argument("--json", default=None, help="also write all points to this file")
    a = ap.parse_args(argv)
    specs = codes.mixed_twelve()
    ctx = Context(0)
    lib = _lib.load()
    gcodes = [LdpcCode(s.qc(), ctx) for s in specs]
    decs = [NmsDecoder.load(g, default_path(mixed_name(s))) for g, s in zip(gcodes, specs)]
    for d in decs:
        d.fallback = a.mode
    cs = CodeSet(gcodes)
    F = a.frames
    rng = np.random.default_rng(a.seed)
    ids = np.sort(rng.integers(0, len(specs), size=F)).astype(np.uint16)
    n = np.array([s.n for s in specs])
    k = np.array([s.k for s in specs])
    kw = (k + 63) // 64
    nw = (n + 63) // 64
    llr_off = np.concatenate([[0], np.cumsum(n[ids])])
    msg_off = np.concatenate([[0], np.cumsum(kw[ids])])
    cw_off = np.concatenate([[0], np.cumsum(nw[ids])])
    d_llr = DeviceBuffer(ctx, int(llr_off[-1]) * 4)
    d_cw = DeviceBuffer(ctx, int(cw_off[-1]) * 8)
    outs = [DeviceBuffer(ctx, int(msg_off[-1]) * 8), DeviceBuffer(ctx, F * 4), DeviceBuffer(ctx, F),
            DeviceBuffer(ctx, F), DeviceBuffer(ctx, F)]
    info_bits = int(k[ids].sum())
    first = [int(np.nonzero(ids == c)[0][0]) if (ids == c).any() else -1 for c in range(len(specs))]
    count = np.bincount(ids, minlength=len(specs))

    def run(learned: bool):
        ptrs = [o.ptr for o in outs]
        t0 = time.perf_counter()
        if learned:
            cs.decode_nms_device(decs, ids, d_llr.ptr, a.fallback, *ptrs)
        else:
            cs.decode_device(ids, d_llr.ptr, a.fallback, *ptrs[:4])
        ctx.synchronize()
        dt = time.perf_counter() - t0
        msg = outs[0].download(np.uint64, int(msg_off[-1]))
        it = outs[1].download(np.uint32, F)
        conv = outs[2].download(np.uint8, F).astype(bool)
        synd = outs[3].download(np.uint8, F).astype(bool)
        fb = outs[4].download(np.uint8, F).astype(bool) if learned else np.zeros(F, dtype=bool)
        cw = d_cw.download(np.uint64, int(cw_off[-1]))
        fer = {}
        for c, s in enumerate(specs):
            if first[c] < 0:
                continue
            f0, fc = first[c], int(count[c])
            m = msg[msg_off[f0]:msg_off[f0] + fc * kw[c]].reshape(fc, kw[c])
            w = cw[cw_off[f0]:cw_off[f0] + fc * nw[c]].reshape(fc, nw[c])
            fe, _ = count_errors(m, w, conv[f0:f0 + fc], synd[f0:f0 + fc], s.k)
            fer[mixed_name(s)] = fe / fc
        return dt, it, fb, fer

    points = []
    for e in _parse_range(a.ebn0):
        for c, g in enumerate(gcodes):
            if first[c] < 0:
                continue
            f0 = first[c]
            _lib.check(ctx.handle, lib.gf2l_frames_generate(
                ctx.handle, g.handle, a.seed * 1000 + c, int(round(e * 1000)) * F, int(count[c]), e, e,
                d_llr.ptr + int(llr_off[f0]) * 4, d_cw.ptr + int(cw_off[f0]) * 8))
        ctx.synchronize()
        run(False)  # warm-up of the set's workspace and graph
        dt_p, it_p, _, fer_p = run(False)
        dt_l, it_l, fb_l, fer_l = run(True)
        pt = {"ebn0_db": e, "frames": F, "fallback_mode": a.mode,
              "learned_gbps": round(info_bits / dt_l / 1e9, 4), "plain_gbps": round(info_bits / dt_p / 1e9, 4),
              "learned_over_plain": round(dt_p / dt_l, 4),
              "learned_mean_iterations": round(float(it_l.mean()), 3),
              "plain_mean_iterations": round(float(it_p.mean()), 3),
              "fallback_fraction": round(float(fb_l.mean()), 4),
              "fer_learned": {kk: round(v, 5) for kk, v in fer_l.items()},
              "fer_plain": {kk: round(v, 5) for kk, v in fer_p.items()}}
        print(json.dumps(pt), flush=True)
        points.append(pt)
    if a.json:
        with open(a.json, "w") as f:
            json.dump({"codes": [mixed_name(s) for s in specs], "points": points}, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
