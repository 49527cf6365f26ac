"""Learned normalized min-sum decoder (one weight per base edge and iteration).

Mirrors the reference's alpha * min-sum rule (Llr::boxplus_normalized_minsum_n,
crates/gf2-coding/src/llr.rs:352-355) with a learned alpha per base-graph edge
and iteration, on top of the flooding schedule of decode_iterative
(crates/gf2-coding/src/ldpc/core.rs:889-938).  Training runs on the GPU with
the library's forward/backward kernels and Adam; frames come from the on-device
generator (staircase encoder + BPSK/AWGN).  Checkpoints are .npz files holding
weights, Adam moments and step (no pickled objects).
"""
from __future__ import annotations

import ctypes
import json
import sys
import time
from typing import Optional

import numpy as np

from . import _lib
from .device import DeviceBuffer, generate_frames
from .ldpc import LdpcCode


SCHEDULES = ("flooding", "layered")
FALLBACKS = ("restart", "continue")


class NmsDecoder:
    """schedule 'flooding' (gf2l_nms_create: all checks, then all variables,
    per iteration) or 'layered' (gf2l_nms_create_layered: block-rows in turn,
    beliefs updated in place).  fallback 'restart' re-decodes failing frames
    from their LLRs (flooding only), 'continue' runs plain min-sum on from the
    learned decoder's state (the layered decoder's only mode)."""

    def __init__(self, code: LdpcCode, iterations: int = 10, weights: Optional[np.ndarray] = None,
                 init: float = 1.0, schedule: str = "flooding", fallback: Optional[str] = None):
        if schedule not in SCHEDULES:
            raise ValueError(f"schedule must be one of {SCHEDULES}")
        self.code = code
        self.ctx = code.context
        self.schedule = schedule
        lib = _lib.load()
        h = ctypes.c_void_p()
        w = None
        if weights is not None:
            w = np.ascontiguousarray(weights, dtype=np.float32)
        create = lib.gf2l_nms_create_layered if schedule == "layered" else lib.gf2l_nms_create
        _lib.check(self.ctx.handle, create(
            self.ctx.handle, code.handle, iterations, w.ctypes.data if w is not None else None,
            init, ctypes.byref(h)))
        self._h = h
        self._fallback = "continue" if schedule == "layered" else "restart"
        if fallback is not None:
            self.fallback = fallback
        t, nbe = ctypes.c_uint32(), ctypes.c_uint32()
        lib.gf2l_nms_info(h, ctypes.byref(t), ctypes.byref(nbe))
        self.iterations, self.nbe = t.value, nbe.value
        if w is not None and w.size != self.iterations * self.nbe:
            raise ValueError("weights must have iterations * base-edges entries")

    @property
    def handle(self):
        return self._h

    @property
    def layered_columns(self) -> int:
        """Block-columns of beliefs a layered decode keeps in LDS under the
        context's current tuning (the information columns when the staircase
        kernel runs); 0 for a flooding decoder."""
        x = ctypes.c_uint32()
        _lib.check(self.ctx.handle, _lib.load().gf2l_nms_layered_columns(self.ctx.handle, self._h, ctypes.byref(x)))
        return x.value

    @property
    def fallback(self) -> str:
        return self._fallback

    @fallback.setter
    def fallback(self, mode: str) -> None:
        if mode not in FALLBACKS:
            raise ValueError(f"fallback must be one of {FALLBACKS}")
        _lib.check(self.ctx.handle, _lib.load().gf2l_nms_set_fallback(self.ctx.handle, self._h,
                                                                       FALLBACKS.index(mode)))
        self._fallback = mode

    # -- parameters -------------------------------------------------------
    @property
    def weights(self) -> np.ndarray:
        w = np.zeros((self.iterations, self.nbe), dtype=np.float32)
        _lib.check(self.ctx.handle, _lib.load().gf2l_nms_get_weights(self.ctx.handle, self._h,
                                                                      w.ctypes.data))
        return w

    @weights.setter
    def weights(self, w: np.ndarray) -> None:
        w = np.ascontiguousarray(w, dtype=np.float32).reshape(self.iterations, self.nbe)
        _lib.check(self.ctx.handle, _lib.load().gf2l_nms_set_weights(self.ctx.handle, self._h,
                                                                      w.ctypes.data))

    def adam_state(self):
        m = np.zeros((self.iterations, self.nbe), dtype=np.float32)
        v = np.zeros_like(m)
        step = ctypes.c_uint64()
        _lib.check(self.ctx.handle, _lib.load().gf2l_nms_get_adam(
            self.ctx.handle, self._h, m.ctypes.data, v.ctypes.data, ctypes.byref(step)))
        return m, v, step.value

    def set_adam_state(self, m, v, step: int) -> None:
        m = np.ascontiguousarray(m, dtype=np.float32)
        v = np.ascontiguousarray(v, dtype=np.float32)
        _lib.check(self.ctx.handle, _lib.load().gf2l_nms_set_adam(
            self.ctx.handle, self._h, m.ctypes.data, v.ctypes.data, int(step)))

    def save(self, path: str, meta: Optional[dict] = None) -> None:
        m, v, step = self.adam_state()
        info = {"iterations": self.iterations, "base_edges": self.nbe, "n": self.code.n(),
                "k": self.code.k(), "shifts": self.code.qc.base_matrix.tolist(),
                "z": self.code.qc.expansion_factor(), "schedule": self.schedule, "fallback": self.fallback}
        info.update(meta or {})
        np.savez_compressed(path, weights=self.weights, adam_m=m, adam_v=v, step=np.array(step, dtype=np.uint64),
                 meta=np.array(json.dumps(info)))

    @classmethod
    def load(cls, code: LdpcCode, path: str) -> "NmsDecoder":
        with np.load(path, allow_pickle=False) as z:
            info = json.loads(str(z["meta"]))
            if info["n"] != code.n() or info["base_edges"] != len(code.qc.to_edges()) // code.qc.expansion_factor():
                raise ValueError("checkpoint was trained for a different code")
            if not np.array_equal(np.array(info["shifts"]), code.qc.base_matrix):
                raise ValueError("checkpoint was trained for a different base matrix")
            dec = cls(code, int(info["iterations"]), z["weights"], schedule=info.get("schedule", "flooding"),
                      fallback=info.get("fallback"))
            dec.set_adam_state(z["adam_m"], z["adam_v"], int(z["step"]))
        return dec

    # -- training -----------------------------------------------------------
    def gradient(self, d_llr: DeviceBuffer, d_cw: DeviceBuffer, frames: int):
        g = np.zeros((self.iterations, self.nbe), dtype=np.float32)
        loss = ctypes.c_float()
        _lib.check(self.ctx.handle, _lib.load().gf2l_nms_gradient(
            self.ctx.handle, self._h, d_llr.ptr, d_cw.ptr, frames, g.ctypes.data, ctypes.byref(loss)))
        return g, loss.value

    def train_step(self, d_llr: DeviceBuffer, d_cw: DeviceBuffer, frames: int, lr: float) -> float:
        loss = ctypes.c_float()
        _lib.check(self.ctx.handle, _lib.load().gf2l_nms_train_step(
            self.ctx.handle, self._h, d_llr.ptr, d_cw.ptr, frames, lr, ctypes.byref(loss)))
        return loss.value

    # -- decoding -----------------------------------------------------------
    def decode_device(self, d_llr: DeviceBuffer, frames: int, fallback_max_it: int,
                      d_msg, d_it, d_conv, d_synd, d_fb=None) -> None:
        _lib.check(self.ctx.handle, _lib.load().gf2l_nms_decode_device(
            self.ctx.handle, self._h, d_llr.ptr, frames, fallback_max_it, d_msg.ptr, d_it.ptr,
            d_conv.ptr, d_synd.ptr, d_fb.ptr if d_fb is not None else None))

    def decode(self, llr: np.ndarray, fallback_max_it: int = 50):
        """Host arrays in and out: (msg words, iterations, converged, syndrome_ok, fallback)."""
        llr = np.ascontiguousarray(llr, dtype=np.float32).reshape(-1, self.code.n())
        F = llr.shape[0]
        kw = (self.code.k() + 63) // 64
        bufs = [DeviceBuffer.from_array(self.ctx, llr), DeviceBuffer(self.ctx, F * kw * 8),
                DeviceBuffer(self.ctx, F * 4), DeviceBuffer(self.ctx, F), DeviceBuffer(self.ctx, F),
                DeviceBuffer(self.ctx, F)]
        self.decode_device(bufs[0], F, fallback_max_it, *bufs[1:])
        self.ctx.synchronize()
        return (bufs[1].download(np.uint64, (F, kw)), bufs[2].download(np.uint32, F),
                bufs[3].download(np.uint8, F).astype(bool), bufs[4].download(np.uint8, F).astype(bool),
                bufs[5].download(np.uint8, F).astype(bool))

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                _lib.load().gf2l_nms_destroy(self._h)
                self._h = None
        except Exception:
            pass


def train(dec: NmsDecoder, steps: int, batch: int, lr: float = 1e-3, ebn0_lo: float = 1.0,
          ebn0_hi: float = 3.0, seed: int = 1, first_frame: int = 0, log_every: int = 0,
          log=sys.stderr):
    """Adam on freshly generated frames: step s uses frames
    first_frame + s*batch ... + batch-1 of the seeded generator."""
    d_llr, d_cw = None, None
    losses = []
    t0 = time.perf_counter()
    for s in range(steps):
        if d_llr is not None:
            d_llr.free()
            d_cw.free()
        d_llr, d_cw = generate_frames(dec.code, seed, first_frame + s * batch, batch, ebn0_lo, ebn0_hi)
        losses.append(dec.train_step(d_llr, d_cw, batch, lr))
        if log_every and (s + 1) % log_every == 0:
            print(f"[nms-train] step {s + 1}/{steps} loss {np.mean(losses[-log_every:]):.5f} "
                  f"{time.perf_counter() - t0:.1f}s", file=log, flush=True)
    return np.array(losses)
