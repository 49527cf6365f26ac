"""Trains the learned normalized min-sum decoder on the GPU and saves a checkpoint.

    python -m gf2_amd.train_nms --code n2304  [--out gf2_amd/weights/nms_n2304_t10.npz]
    python -m gf2_amd.train_nms --code n64800 [--out gf2_amd/weights/nms_n64800_t10.npz]

Recipes (defaults per code):
  n2304   N=2304 K=1152 Z=96 QC-IRA: 1024 steps x 4096 frames = 2^22 frames,
          Adam lr 1e-3, Eb/N0 uniform in [1, 3] dB per frame, 10 iterations,
          BCE on the soft output of every iteration, weights start at 1.0.
  n64800  N=64800 K=32400 Z=360 QC-IRA: 400 steps x 256 frames, Adam lr 1e-2,
          Eb/N0 uniform in [1.0, 2.0] dB, 10 iterations, weights start at 1.0.
Frames come from the on-device generator (seeded Philox, staircase encoder,
BPSK/AWGN); every step draws fresh frames.
"""
from __future__ import annotations

import argparse
import os
import sys
import time

from . import codes
from .ldpc import Context, LdpcCode
from .nms import NmsDecoder, train

RECIPES = {
    "n2304": dict(spec=codes.n2304_r12, steps=1024, batch=4096, lr=1e-3, lo=1.0, hi=3.0),
    "n64800": dict(spec=codes.n64800_r12, steps=400, batch=256, lr=1e-2, lo=1.0, hi=2.0),
}
HERE = os.path.dirname(os.path.abspath(__file__))


def default_path(code: str, iterations: int = 10) -> str:
    return os.path.join(HERE, "weights", f"nms_{code}_t{iterations}.npz")


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--code", choices=sorted(RECIPES), default="n2304")
    ap.add_argument("--iterations", type=int, default=10)
    ap.add_argument("--steps", type=int)
    ap.add_argument("--batch", type=int)
    ap.add_argument("--lr", type=float)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--out")
    a = ap.parse_args(argv)
    r = dict(RECIPES[a.code])
    for k in ("steps", "batch", "lr"):
        if getattr(a, k) is not None:
            r[k] = getattr(a, k)
    spec = r["spec"]()
    code = LdpcCode(spec.qc(), Context(0))
    dec = NmsDecoder(code, a.iterations, init=1.0)
    t0 = time.perf_counter()
    losses = train(dec, r["steps"], r["batch"], r["lr"], r["lo"], r["hi"], seed=a.seed,
                   log_every=max(1, r["steps"] // 8))
    dt = time.perf_counter() - t0
    out = a.out or default_path(a.code, a.iterations)
    os.makedirs(os.path.dirname(out), exist_ok=True)
    dec.save(out, meta={"code_name": spec.name, "recipe": {k: v for k, v in r.items() if k != "spec"},
                        "seed": a.seed, "final_loss": float(losses[-10:].mean()),
                        "train_seconds": dt})
    print(f"[nms-train] {spec.name}: {r['steps']} steps x {r['batch']} frames in {dt:.1f}s, "
          f"loss {losses[:10].mean():.4f} -> {losses[-10:].mean():.4f}; saved {out}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
