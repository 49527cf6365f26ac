"""Trains the learned normalized min-sum decoder on the GPU and saves a checkpoint.

    python -m gf2_amd.train_nms --code n2304   [--out gf2_amd/weights/nms_n2304_t10.npz]
    python -m gf2_amd.train_nms --code n64800  [--out gf2_amd/weights/nms_n64800_t10.npz]
    python -m gf2_amd.train_nms --code n16200-k9720   (any of the twelve Z=360 codes)
    python -m gf2_amd.train_nms --all-mixed   (one checkpoint per Z=360 code)
    python -m gf2_amd.train_nms --code n64800 --iterations 50 --steps 200 --init 0.9 \
        --lr 3e-3 --lo 1.4 --hi 1.8
    python -m gf2_amd.train_nms --code n64800 --schedule layered
        [--out gf2_amd/weights/nms_n64800_layered_t10.npz]
    python -m gf2_amd.train_nms --code n2304 --resume <checkpoint> --steps 512 [--out <new file>]
        (continues a checkpoint: its weights, Adam moments and step, recipe
        and seed; the frames continue the stream where it stopped)

Recipes (defaults per code):
  n2304   N=2304 K=1152 Z=96 QC-IRA: 1024 steps x 4096 frames = 2^22 frames,
          Adam lr 1e-3, Eb/N0 uniform in [1, 3] dB per frame, 10 iterations,
          BCE on the soft output of every iteration, weights start at 1.0.
  n64800  N=64800 K=32400 Z=360 QC-IRA: 400 steps x 256 frames, Adam lr 1e-2,
          Eb/N0 uniform in [1.0, 2.0] dB, 10 iterations, weights start at 1.0.
  Z=360 codes of the mixed set: 300 steps x 256 frames (N=64800) or x 1024
          frames (N=16200), Adam lr 1e-2, Eb/N0 uniform in [e - 1.0, e + 0.5]
          dB around the code's plain-BP FER-1e-2 point e (codes.FER_1E2_EBN0).
  layered schedule (--schedule layered): 1024 steps x 4096 frames = 2^22
          frames, Adam lr 1e-3, weights start at 0.85; Eb/N0 uniform in
          [1, 3] dB (n2304) or [1.4, 1.8] dB (n64800).
Frames come from the on-device generator (seeded Philox, staircase encoder,
BPSK/AWGN); every step draws fresh frames.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import codes
from .ldpc import Context, LdpcCode
from .nms import NmsDecoder, train

RECIPES = {
    "n2304": dict(spec=codes.n2304_r12, steps=1024, batch=4096, lr=1e-3, lo=1.0, hi=3.0),
    "n64800": dict(spec=codes.n64800_r12, steps=400, batch=256, lr=1e-2, lo=1.0, hi=2.0),
}
# Layered schedule: 2^22 frames in batches of 4096 with Adam lr 1e-3 (the
# N=2304 recipe's scale), Eb/N0 around the code's FER-1e-3 region, weights
# starting at 0.85 (the best constant for this schedule).
LAYERED_RECIPES = {
    "n2304": dict(spec=codes.n2304_r12, steps=1024, batch=4096, lr=1e-3, lo=1.0, hi=3.0, init=0.85),
    "n64800": dict(spec=codes.n64800_r12, steps=1024, batch=4096, lr=1e-3, lo=1.4, hi=1.8, init=0.85),
}
HERE = os.path.dirname(os.path.abspath(__file__))


def mixed_name(spec: codes.QcIraSpec) -> str:
    return f"n{spec.n}-k{spec.k}"


def recipe(name: str) -> dict:
    if name in RECIPES:
        return dict(RECIPES[name])
    for s in codes.mixed_twelve():
        if mixed_name(s) == name:
            e = codes.FER_1E2_EBN0[(s.n, s.k)]
            return dict(spec=lambda s=s: s, steps=300, batch=256 if s.n == 64800 else 1024, lr=1e-2,
                        lo=round(e - 1.0, 3), hi=round(e + 0.5, 3))
    raise ValueError(f"unknown code {name!r}")


def default_path(code: str, iterations: int = 10, schedule: str = "flooding") -> str:
    tag = "_layered" if schedule == "layered" else ""
    return os.path.join(HERE, "weights", f"nms_{code}{tag}_t{iterations}.npz")


def train_one(name: str, iterations: int, ctx: Context, seed: int, out: str = None, init: float = None,
              schedule: str = "flooding", **override) -> str:
    r = dict(LAYERED_RECIPES[name]) if schedule == "layered" and name in LAYERED_RECIPES else recipe(name)
    r.update({k: v for k, v in override.items() if v is not None})
    init = init if init is not None else r.pop("init", 1.0)
    r.pop("init", None)
    spec = r["spec"]()
    code = LdpcCode(spec.qc(), ctx)
    dec = NmsDecoder(code, iterations, init=init, schedule=schedule)
    t0 = time.perf_counter()
    losses = train(dec, r["steps"], r["batch"], r["lr"], r["lo"], r["hi"], seed=seed,
                   log_every=max(1, r["steps"] // 4))
    dt = time.perf_counter() - t0
    out = out or default_path(name, iterations, schedule)
    os.makedirs(os.path.dirname(out), exist_ok=True)
    dec.save(out, meta={"code_name": spec.name, "init": init,
                        "recipe": {k: v for k, v in r.items() if k != "spec"},
                        "seed": seed, "final_loss": float(losses[-10:].mean()), "train_seconds": dt})
    print(f"[nms-train] {spec.name}: {r['steps']} steps x {r['batch']} frames in {dt:.1f}s, "
          f"loss {losses[:10].mean():.4f} -> {losses[-10:].mean():.4f}; saved {out}", flush=True)
    return out


def resume_one(name: str, path: str, steps: int, ctx: Context, out: str = None) -> str:
    """Continues a checkpoint's training for `steps` more Adam steps with its
    own recipe and seed: the weights, Adam moments and step come from the
    file, and the frames continue the seeded stream where the checkpoint's
    steps left it (frame step * batch onwards), so a resumed run equals an
    uninterrupted one bit for bit."""
    spec = recipe(name)["spec"]()
    code = LdpcCode(spec.qc(), ctx)
    dec = NmsDecoder.load(code, path)
    with np.load(path, allow_pickle=False) as z:
        info = json.loads(str(z["meta"]))
    r, seed = info["recipe"], int(info["seed"])
    done = dec.adam_state()[2]
    t0 = time.perf_counter()
    losses = train(dec, steps, r["batch"], r["lr"], r["lo"], r["hi"], seed=seed, first_frame=done * r["batch"],
                   log_every=max(1, steps // 4))
    dt = time.perf_counter() - t0
    out = out or path
    info.update({"final_loss": float(losses[-10:].mean()), "train_seconds": float(info.get("train_seconds", 0.0)) + dt,
                 "recipe": dict(r, steps=int(r.get("steps", done)) + steps)})
    for k in ("iterations", "base_edges", "n", "k", "shifts", "z", "schedule", "fallback"):
        info.pop(k, None)
    dec.save(out, meta=info)
    print(f"[nms-train] {spec.name}: resumed at step {done}, {steps} more steps in {dt:.1f}s, "
          f"loss {losses[-10:].mean():.4f}; saved {out}", flush=True)
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--code", default="n2304")
    ap.add_argument("--all-mixed", action="store_true", help="train one checkpoint per Z=360 code")
    ap.add_argument("--iterations", type=int, default=10)
    ap.add_argument("--steps", type=int)
    ap.add_argument("--batch", type=int)
    ap.add_argument("--lr", type=float)
    ap.add_argument("--lo", type=float, help="lowest training Eb/N0 (dB)")
    ap.add_argument("--hi", type=float, help="highest training Eb/N0 (dB)")
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--init", type=float, default=None,
                    help="initial value of every weight (default 1.0; the layered recipes' own value)")
    ap.add_argument("--schedule", choices=["flooding", "layered"], default="flooding")
    ap.add_argument("--resume", help="checkpoint to continue for --steps more steps (its recipe and seed)")
    ap.add_argument("--out")
    ap.add_argument("--out-dir", help="directory for the checkpoints (default gf2_amd/weights)")
    a = ap.parse_args(argv)
    ctx = Context(0)
    if a.resume:
        if not a.steps:
            ap.error("--resume needs --steps")
        resume_one(a.code, a.resume, a.steps, ctx, a.out)
        return 0
    names = [mixed_name(s) for s in codes.mixed_twelve()] if a.all_mixed else [a.code]
    for name in names:
        out = a.out if not a.all_mixed else None
        if a.out_dir:
            out = os.path.join(a.out_dir, os.path.basename(default_path(name, a.iterations, a.schedule)))
        train_one(name, a.iterations, ctx, a.seed, out, init=a.init, schedule=a.schedule, steps=a.steps,
                  batch=a.batch, lr=a.lr, lo=a.lo, hi=a.hi)
    return 0


if __name__ == "__main__":
    sys.exit(main())
