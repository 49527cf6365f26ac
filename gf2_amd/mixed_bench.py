"""Mixed-code batch stream: the twelve Z=360 QC-IRA codes decoded together.

Batches of Poisson-distributed size (clipped to [256, 8192]) arrive one
after another; each holds frames of all twelve codes (N in {16200, 64800},
six rates each), every code at the Eb/N0 where plain min-sum BP has FER
about 1e-2 (codes.FER_1E2_EBN0, measured with gf2_amd.fer --target-fer 1e-2).  Frames
are generated on the GPU before the timed region.  Each batch is one
gf2l_decode_mixed_device call (one code set, per-frame early stop, graph-
replayed iteration loop); the same batches are also decoded one code at a
time with gf2l_decode_batch_device, the reference API's shape, and both
results are compared frame by frame.

    python -m gf2_amd.mixed_bench [--batches 12] [--mean-batch 4096] [--tune]
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from typing import List

import numpy as np

from . import _lib, codes
from .device import DeviceBuffer
from .ldpc import Context, LdpcCode
from .mixed import CodeSet

def poisson_batches(count: int, mean: int, seed: int, lo: int = 256, hi: int = 8192) -> List[int]:
    rng = np.random.default_rng(seed)
    return [int(x) for x in np.clip(rng.poisson(mean, size=count), lo, hi)]


class Stream:
    """Pre-generated batches resident in HBM."""

    def __init__(self, ctx: Context, gcodes: List[LdpcCode], sizes: List[int], ebn0: List[float], seed: int):
        lib = _lib.load()
        self.batches = []
        rng = np.random.default_rng(seed)
        n = np.array([c.n() for c in gcodes])
        for b, F in enumerate(sizes):
            ids = np.sort(rng.integers(0, len(gcodes), size=F)).astype(np.uint16)
            off = np.concatenate([[0], np.cumsum(n[ids])])
            d_llr = DeviceBuffer(ctx, int(off[-1]) * 4)
            for c in range(len(gcodes)):
                sel = np.nonzero(ids == c)[0]
                if sel.size == 0:
                    continue
                _lib.check(ctx.handle, lib.gf2l_frames_generate(
                    ctx.handle, gcodes[c].handle, seed * 7919 + b * 131 + c, 0, sel.size, ebn0[c], ebn0[c],
                    d_llr.ptr + int(off[sel[0]]) * 4, None))
            self.batches.append((ids, d_llr, off))
        ctx.synchronize()


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--batches", type=int, default=12)
    ap.add_argument("--mean-batch", type=int, default=4096)
    ap.add_argument("--max-it", type=int, default=50)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--tune", action="store_true", help="autotune the code set on the first batch")
    a = ap.parse_args(argv)
    specs = codes.mixed_twelve()
    ctx = Context(0)
    gcodes = [LdpcCode(s.qc(), ctx) for s in specs]
    cs = CodeSet(gcodes)
    lib = _lib.load()
    sizes = poisson_batches(a.batches, a.mean_batch, a.seed)
    stream = Stream(ctx, gcodes, sizes, [codes.FER_1E2_EBN0[(s.n, s.k)] for s in specs], a.seed)
    k = np.array([s.k for s in specs])
    kw = (k + 63) // 64
    Fmax = max(sizes)
    outs_m = [DeviceBuffer(ctx, Fmax * int(kw.max()) * 8), DeviceBuffer(ctx, Fmax * 4), DeviceBuffer(ctx, Fmax),
              DeviceBuffer(ctx, Fmax)]
    outs_p = [DeviceBuffer(ctx, Fmax * int(kw.max()) * 8), DeviceBuffer(ctx, Fmax * 4), DeviceBuffer(ctx, Fmax),
              DeviceBuffer(ctx, Fmax)]
    tuned = None
    if a.tune:
        from .autotune import tune_codeset
        ids0, d0, _ = stream.batches[0]
        tuned = tune_codeset(cs, ids0, d0, a.max_it)["config"]

    def mixed(b):
        ids, d_llr, _ = stream.batches[b]
        cs.decode_device(ids, d_llr.ptr, a.max_it, *[o.ptr for o in outs_m])

    def percode(b):
        ids, d_llr, off = stream.batches[b]
        moff = np.concatenate([[0], np.cumsum(kw[ids])])
        for c in range(len(specs)):
            sel = np.nonzero(ids == c)[0]
            if sel.size == 0:
                continue
            f0 = int(sel[0])
            _lib.check(ctx.handle, lib.gf2l_decode_batch_device(
                ctx.handle, gcodes[c].handle, d_llr.ptr + int(off[f0]) * 4, sel.size, a.max_it,
                outs_p[0].ptr + int(moff[f0]) * 8, outs_p[1].ptr + f0 * 4, outs_p[2].ptr + f0, outs_p[3].ptr + f0))

    info_bits = sum(int(k[ids].sum()) for ids, _, _ in stream.batches)
    mixed(0)
    percode(0)
    ctx.synchronize()
    res = {}
    same = True
    fer = np.zeros(len(specs))
    cnt = np.zeros(len(specs))
    its = np.zeros(len(specs))
    for name, fn in [("mixed", mixed), ("percode", percode)]:
        t0 = time.perf_counter()
        for b in range(len(sizes)):
            fn(b)
        ctx.synchronize()
        res[name] = time.perf_counter() - t0
    # results of every batch, both ways
    for b, (ids, _, _) in enumerate(stream.batches):
        mixed(b)
        percode(b)
        ctx.synchronize()
        F = ids.size
        nmsg = int(kw[ids].sum())
        got = [o.download(dt, s) for o, (dt, s) in zip(outs_m, [(np.uint64, nmsg), (np.uint32, F), (np.uint8, F), (np.uint8, F)])]
        ref = [o.download(dt, s) for o, (dt, s) in zip(outs_p, [(np.uint64, nmsg), (np.uint32, F), (np.uint8, F), (np.uint8, F)])]
        same &= all(np.array_equal(x, y) for x, y in zip(got, ref))
        np.add.at(fer, ids, 1 - got[2])
        np.add.at(cnt, ids, 1)
        np.add.at(its, ids, got[1])
    out = {
        "workload": "12 QC-IRA codes (N 16200/64800 x 6 rates, Z=360), Poisson batches, per-code Eb/N0 at "
                    "plain-BP FER ~1e-2",
        "batches": sizes, "frames": int(sum(sizes)), "max_iterations": a.max_it,
        "mixed_gbps": round(info_bits / res["mixed"] / 1e9, 4),
        "percode_gbps": round(info_bits / res["percode"] / 1e9, 4),
        "mixed_over_percode": round(res["percode"] / res["mixed"], 4),
        "identical_results": bool(same),
        "tuning": tuned,
        "per_code": {f"n{s.n}-k{s.k}": {"frames": int(cnt[c]), "not_converged_rate": round(fer[c] / max(cnt[c], 1), 4),
                                          "mean_iterations": round(its[c] / max(cnt[c], 1), 2)}
                     for c, s in enumerate(specs)},
    }
    print(json.dumps(out), flush=True)
    return 0 if same else 1


if __name__ == "__main__":
    sys.exit(main())
