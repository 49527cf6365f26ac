"""Launch-shape autotuner for the HBM decoder paths.

The decoder's results never depend on its launch shape (every setting is
bit-exact with the reference decoder); only its speed does.  The knobs are
the context tunables of gf2l_set_tuning / gf2l_set_path:

  wg_per_cu    persistent workgroups per compute unit: how many (frame,
               block-row) and (frame, block-column) items each CU keeps in
               flight; more hides latency, fewer keeps each XCD's L2 on fewer
               frames
  xcd_map      frames pinned to XCDs (their items stay in one L2) or spread
               frame-major over the whole grid
  graph_iters  iterations per graph replay of the mixed-code decoder (host
               round trips vs. empty launches after the last frame retires)
  chunk        consecutive items per shared-counter fetch (0: fixed-stride
               assignment): counters keep the items in flight close together
               (less L2 traffic) at the cost of an atomic and a barrier per
               chunk
  path         automatic (whole frame in LDS when it fits), compressed check
               state, or one stored message per edge (single-code decoder);
               a path that cannot run the code is skipped
  lds_groups   thread groups per workgroup of the LDS path (one frame per
               workgroup; group g takes block-rows and block-columns g,
               g + groups, ...): the block-row tiling of small codes, which
               trades threads per frame (latency hiding) against passes per
               iteration and lanes left idle by z; 0 is the library's choice
By default a code whose automatic path is the LDS path tunes lds_groups, any
other code path, wg_per_cu and chunk.

A tuning run decodes a batch of frames generated on the device at a given
Eb/N0 (the iteration profile of the operating point matters) once per
candidate, checks that every candidate returns the same results, and keeps
the fastest.  Results are cached per code-shape key in a JSON file
($GF2_AMD_TUNING, default ~/.cache/gf2_amd/tuning.json) so a later process
applies them without measuring.
"""
from __future__ import annotations

import itertools
import json
import os
import time
from typing import Dict, Iterable, List, Optional, Sequence

import numpy as np

from . import _lib
from .device import DeviceBuffer, generate_frames, generate_zero_frames
from .ldpc import Context, LdpcCode

CANDIDATES: Dict[str, List[int]] = {
    "wg_per_cu": [2, 3, 4, 5, 6],
    "xcd_map": [1, 0],
    "graph_iters": [2, 4, 8],
    "chunk": [0, 4, 8, 16],
    "path": [_lib.PATH_AUTO, _lib.PATH_STATE, _lib.PATH_STREAM],
    "lds_groups": [0, 2, 3, 4, 5, 8],
}
_TUNE_KEYS = {"wg_per_cu": _lib.TUNE_WG_PER_CU, "xcd_map": _lib.TUNE_XCD_MAP,
              "graph_iters": _lib.TUNE_GRAPH_ITERS, "chunk": _lib.TUNE_CHUNK,
              "lds_groups": _lib.TUNE_LDS_GROUPS}


def default_cache_path() -> str:
    return os.environ.get("GF2_AMD_TUNING",
                          os.path.join(os.path.expanduser("~"), ".cache", "gf2_amd", "tuning.json"))


def code_key(code: LdpcCode) -> str:
    """Shape signature of a code: sizes, lift and degree bounds."""
    qc = code.qc
    if qc is None:  # DVB-T2 or general sparse code
        return f"sparse-n{code.n()}-m{code.m()}-e{code.nnz()}"
    base = qc.base_matrix
    deg_c = int((base >= 0).sum(axis=1).max())
    deg_v = int((base >= 0).sum(axis=0).max())
    return (f"n{code.n()}-m{code.m()}-z{qc.expansion_factor()}-e{int((base >= 0).sum())}"
            f"-dc{deg_c}-dv{deg_v}")


def batch_bucket(frames: int) -> int:
    """Batch sizes share a tuning within a power of two."""
    return 1 << max(0, int(frames - 1).bit_length())


def shape_key(codes: Sequence[LdpcCode], frames: int, kind: str) -> str:
    return f"{kind}:" + "|".join(code_key(c) for c in codes) + f":F{batch_bucket(frames)}"


class TuningCache:
    """JSON map shape key -> {"config": {...}, "ms": float, "candidates": n}."""

    def __init__(self, path: Optional[str] = None):
        self.path = path or default_cache_path()
        self.data: Dict[str, dict] = {}
        if os.path.exists(self.path):
            with open(self.path) as f:
                self.data = json.load(f)

    def get(self, key: str) -> Optional[dict]:
        e = self.data.get(key)
        return dict(e["config"]) if e else None

    def put(self, key: str, config: dict, ms: float, candidates: int) -> None:
        self.data[key] = {"config": dict(config), "ms": round(float(ms), 4), "candidates": int(candidates)}

    def save(self) -> None:
        d = os.path.dirname(self.path)
        if d:
            os.makedirs(d, exist_ok=True)
        tmp = self.path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(self.data, f, indent=1, sort_keys=True)
        os.replace(tmp, self.path)


def apply(ctx: Context, config: dict) -> None:
    """Sets a tuning (any subset of the knobs) on a context."""
    for k, v in config.items():
        if k == "path":
            ctx.set_path(int(v))
        elif k in _TUNE_KEYS:
            ctx.set_tuning(_TUNE_KEYS[k], int(v))
        else:
            raise ValueError(f"unknown tuning knob {k!r}")


def candidates(knobs: Iterable[str]) -> List[dict]:
    knobs = list(knobs)
    return [dict(zip(knobs, vals)) for vals in itertools.product(*(CANDIDATES[k] for k in knobs))]


def _time(fn, repeats: int) -> float:
    fn()
    t0 = time.perf_counter()
    for _ in range(repeats):
        fn()
    return (time.perf_counter() - t0) / repeats * 1e3


def default_knobs(code: LdpcCode) -> tuple:
    """Knobs worth tuning for a code: the LDS path's thread groups when the
    automatic path keeps the frame in LDS, else the HBM paths' knobs.  Sets
    the code's context to the automatic path to find out."""
    ctx = code.context
    ctx.set_path(_lib.PATH_AUTO)
    if ctx.code_path(code) == _lib.PATH_LDS:
        return ("lds_groups",)
    return ("path", "wg_per_cu", "chunk")


def tune_code(code: LdpcCode, frames: int, ebn0_db: float, max_it: int = 50, repeats: int = 2,
              knobs: Optional[Sequence[str]] = None, cache: Optional[TuningCache] = None,
              seed: int = 11) -> dict:
    """Tunes the single-code decoder for batches of about `frames` frames
    (knobs: default_knobs(code) unless given).  Returns {"config", "ms",
    "results": [(config, ms), ...]} and applies the best config to the
    code's context."""
    ctx = code.context
    lib = _lib.load()
    if knobs is None:
        knobs = default_knobs(code)
    F = frames
    kw = (code.k() + 63) // 64
    try:
        d_llr, _ = generate_frames(code, seed, 0, F, ebn0_db, ebn0_db, with_codewords=False)
    except _lib.Gf2lError as e:  # no device encoder for this code: all-zero codewords
        if e.code != _lib.GF2L_ERR_UNSUPPORTED:
            raise
        d_llr = generate_zero_frames(code, seed, 0, F, ebn0_db, ebn0_db)
    outs = [DeviceBuffer(ctx, F * kw * 8), DeviceBuffer(ctx, F * 4), DeviceBuffer(ctx, F), DeviceBuffer(ctx, F)]

    def run():
        _lib.check(ctx.handle, lib.gf2l_decode_batch_device(ctx.handle, code.handle, d_llr.ptr, F, max_it,
                                                            *[o.ptr for o in outs]))
        ctx.synchronize()

    return _search(ctx, run, outs, [(np.uint64, (F, kw)), (np.uint32, F), (np.uint8, F), (np.uint8, F)],
                   candidates(knobs), repeats, cache, shape_key([code], F, "single"))


def tune_codeset(cs, code_ids: np.ndarray, d_llr: DeviceBuffer, max_it: int = 50, repeats: int = 2,
                 knobs: Sequence[str] = ("wg_per_cu", "chunk", "graph_iters"),
                 cache: Optional[TuningCache] = None) -> dict:
    """Tunes the mixed-code decoder of a code set on a resident batch (d_llr
    packed as gf2l_decode_mixed_device expects)."""
    ctx = cs.ctx
    ids = np.ascontiguousarray(code_ids, dtype=np.uint16)
    F = ids.size
    _, msg_off = cs.offsets(ids)
    nmsg = int(msg_off[-1])
    outs = [DeviceBuffer(ctx, max(nmsg, 1) * 8), DeviceBuffer(ctx, F * 4), DeviceBuffer(ctx, F),
            DeviceBuffer(ctx, F)]

    def run():
        cs.decode_device(ids, d_llr.ptr, max_it, *[o.ptr for o in outs])
        ctx.synchronize()

    return _search(ctx, run, outs, [(np.uint64, nmsg), (np.uint32, F), (np.uint8, F), (np.uint8, F)],
                   candidates(knobs), repeats, cache, shape_key(cs.codes, F, "mixed"))


def _search(ctx, run, outs, shapes, cands: List[dict], repeats: int, cache: Optional[TuningCache],
            key: str) -> dict:
    results = []
    ref = None
    for cfg in cands:
        apply(ctx, cfg)
        try:
            ms = _time(run, repeats)
        except _lib.Gf2lError as e:
            if e.code == _lib.GF2L_ERR_UNSUPPORTED:
                continue
            raise
        got = [o.download(dt, shp) for o, (dt, shp) in zip(outs, shapes)]
        if ref is None:
            ref = got
        elif not all(np.array_equal(a, b) for a, b in zip(ref, got)):
            raise RuntimeError(f"launch shape {cfg} changed the decoder's results")
        results.append((cfg, ms))
    if not results:
        raise RuntimeError("no launch shape could run this decode")
    best_cfg, best_ms = min(results, key=lambda r: r[1])
    apply(ctx, best_cfg)
    if cache is not None:
        cache.put(key, best_cfg, best_ms, len(results))
        cache.save()
    return {"key": key, "config": best_cfg, "ms": best_ms, "results": results}


def apply_cached(ctx: Context, codes: Sequence[LdpcCode], frames: int, kind: str = "single",
                 cache: Optional[TuningCache] = None) -> Optional[dict]:
    """Applies a cached tuning for this shape if one exists; returns it."""
    cache = cache or TuningCache()
    cfg = cache.get(shape_key(codes, frames, kind))
    if cfg:
        apply(ctx, cfg)
    return cfg
