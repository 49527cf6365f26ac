"""Frame-error-rate harness around the GPU decoders.

Mirrors the reference's FER loop (crates/gf2-coding/examples/ldpc_awgn.rs:121-175):
a frame is in error when the decoder did not converge or its syndrome check
failed (ldpc_awgn.rs:157); here a converged frame whose message differs from
the sent information bits also counts (undetected error).  Results are
reported per Eb/N0 like SimulationResult (crates/gf2-coding/src/simulation.rs:13-116):
frames, frame errors, bit errors, FER, BER, mean iterations, plus info
Gbit/s of the decode calls.  Frames come from the on-device generator
(gf2l_frames_generate: seeded Philox info bits, staircase encoder, BPSK/AWGN
LLRs), so a sweep never leaves the GPU except for the per-batch error count.

    python -m gf2_amd.fer --code n64800 --ebn0 1.5:2.5:0.25 --frames 4096 \\
        --max-frames 65536 --min-errors 50 --json out.json
    python -m gf2_amd.fer --code n16200-k9720 --target-fer 1e-2
    python -m gf2_amd.fer --code dvb-normal-3/5 --ebn0 2.0:2.6:0.1     (all-zero codewords)

A decoder is any callable (d_llr, frames, outs) -> None that fills the
device outputs [msg words, iterations, converged, syndrome_ok]; plain BP and
the learned decoder with fallback are provided.
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from dataclasses import asdict, dataclass
from typing import Callable, List, Optional

import numpy as np

from . import _lib, codes
from .device import DeviceBuffer, generate_frames, generate_zero_frames
from .ldpc import Context, LdpcCode


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


def plain_decoder(code: LdpcCode, max_it: int = 50) -> Callable:
    lib = _lib.load()
    ctx = code.context

    def run(d_llr, frames, outs):
        _lib.check(ctx.handle, lib.gf2l_decode_batch_device(
            ctx.handle, code.handle, d_llr.ptr, frames, max_it, outs[0].ptr, outs[1].ptr,
            outs[2].ptr, outs[3].ptr))
    return run


def learned_decoder(nms, fallback_max_it: int = 50) -> Callable:
    def run(d_llr, frames, outs):
        nms.decode_device(d_llr, frames, fallback_max_it, outs[0], outs[1], outs[2], outs[3], None)
    return run


def count_errors(msg: np.ndarray, cw: np.ndarray, conv: np.ndarray, synd: np.ndarray, k: int):
    """Frame and bit errors of decoded message words against the codewords'
    first k bits (systematic codes)."""
    kw = (k + 63) // 64
    diff = (msg[:, :kw] ^ cw[:, :kw]).copy()
    if k % 64:
        diff[:, -1] &= np.uint64((1 << (k % 64)) - 1)
    bad = diff.any(axis=1)
    wrong = np.zeros(len(msg), dtype=np.int64)
    if bad.any():  # bit counts only where a word differs
        wrong[bad] = np.unpackbits(diff[bad].view(np.uint8), axis=1).sum(axis=1)
    err = (~conv) | (~synd) | (wrong > 0)
    return int(err.sum()), int(wrong.sum())


def sweep(code: LdpcCode, ebn0_list: List[float], frames_per_batch: int, max_frames: int,
          min_errors: int, decoder: Optional[Callable] = None, seed: int = 1, name: str = "",
          log=sys.stderr, zero_codeword: Optional[bool] = None) -> List[FerPoint]:
    """FER/BER per Eb/N0: batches until max_frames or min_errors.
    zero_codeword: send all-zero codewords (the reference example's method,
    examples/ldpc_awgn.rs:121-175; valid for any code) instead of encoded
    random messages; None = encoded where the device encoder supports the
    code (staircase QC-IRA), all-zero otherwise."""
    ctx = code.context
    if zero_codeword is None:
        try:
            generate_frames(code, 0, 0, 1, 1.0, 1.0, with_codewords=False)
            zero_codeword = False
        except _lib.Gf2lError as e:
            if e.code != _lib.GF2L_ERR_UNSUPPORTED:
                raise
            zero_codeword = True
    decoder = decoder or plain_decoder(code)
    n, k = code.n(), code.k()
    F = frames_per_batch
    kw, nw = (k + 63) // 64, (n + 63) // 64
    outs = [DeviceBuffer(ctx, F * kw * 8), DeviceBuffer(ctx, F * 4), DeviceBuffer(ctx, F),
            DeviceBuffer(ctx, F)]
    out = []
    for ebn0 in ebn0_list:
        nf = fe = be = iters = 0
        t_dec = 0.0
        b = 0
        while nf < max_frames and fe < min_errors:
            fseed = seed * 1000003 + int(round(ebn0 * 1000))
            if zero_codeword:
                d_llr, d_cw = generate_zero_frames(code, fseed, b * F, F, ebn0, ebn0), None
            else:
                d_llr, d_cw = generate_frames(code, fseed, b * F, F, ebn0, ebn0)
            t0 = time.perf_counter()
            decoder(d_llr, F, outs)
            ctx.synchronize()
            t_dec += time.perf_counter() - t0
            msg = outs[0].download(np.uint64, (F, kw))
            it = outs[1].download(np.uint32, F)
            conv = outs[2].download(np.uint8, F).astype(bool)
            synd = outs[3].download(np.uint8, F).astype(bool)
            cw = d_cw.download(np.uint64, (F, nw)) if d_cw is not None else np.zeros((F, nw), dtype=np.uint64)
            e, w = count_errors(msg, cw, conv, synd, k)
            nf += F
            fe += e
            be += w
            iters += int(it.sum())
            b += 1
        p = FerPoint(ebn0, nf, fe, be, fe / nf, be / (nf * k), iters / nf, t_dec, nf * k / t_dec / 1e9)
        print(f"[fer] {name} Eb/N0={ebn0:.2f} frames={nf} errors={fe} FER={p.fer:.3e} "
              f"BER={p.ber:.3e} it={p.mean_iterations:.2f} {p.info_gbps:.3f} Gbit/s", file=log, flush=True)
        out.append(p)
    return out


def _rust_float(x: float) -> str:
    """Rust's `{}` Display of an f64: the shortest digits that round-trip,
    positional, no exponent, no trailing '.0' (3.0 -> "3", 1e-5 -> "0.00001")."""
    return np.format_float_positional(float(x), unique=True, trim="-")


@dataclass
class SimulationResult:
    """crates/gf2-coding/src/simulation.rs:41-96: one Eb/N0 point."""
    eb_n0_db: float
    ber: float
    num_bits: int
    num_errors: int
    fer: Optional[float] = None
    num_frames: Optional[int] = None
    num_frame_errors: Optional[int] = None

    @classmethod
    def from_point(cls, p: "FerPoint", k: int) -> "SimulationResult":
        return cls(p.ebn0_db, p.ber, p.frames * k, p.bit_errors, p.fer, p.frames, p.frame_errors)

    def is_complete(self, min_errors: int) -> bool:
        """simulation.rs:79-81: enough bit errors collected."""
        return self.num_errors >= min_errors

    def to_csv_row(self) -> str:
        """simulation.rs:83-96."""
        head = [_rust_float(self.eb_n0_db), _rust_float(self.ber), str(self.num_bits), str(self.num_errors)]
        if self.fer is not None and self.num_frames is not None and self.num_frame_errors is not None:
            head += [_rust_float(self.fer), str(self.num_frames), str(self.num_frame_errors)]
        return ",".join(head)


def results_to_csv(results: List[SimulationResult], include_header: bool = True) -> str:
    """SimulationRunner::results_to_csv (simulation.rs:209-229)."""
    out = ""
    if include_header:
        if any(r.fer is not None for r in results):
            out += "eb_n0_db,ber,num_bits,num_errors,fer,num_frames,num_frame_errors\n"
        else:
            out += "eb_n0_db,ber,num_bits,num_errors\n"
    for r in results:
        out += r.to_csv_row() + "\n"
    return out


def write_csv(points: List[FerPoint], path: str, k: int) -> None:
    """Writes sweep points of a code with k information bits per frame as
    results_to_csv with a header."""
    with open(path, "w") as f:
        f.write(results_to_csv([SimulationResult.from_point(p, k) for p in points], True))


def calibrate(code: LdpcCode, target_fer: float, lo: float, hi: float, frames: int = 2048,
              steps: int = 7, decoder: Optional[Callable] = None, seed: int = 5, log=sys.stderr) -> float:
    """Eb/N0 (dB) where the FER crosses target_fer, by bisection on [lo, hi]
    with `frames` frames per probe (FER is decreasing in Eb/N0)."""
    for _ in range(steps):
        mid = 0.5 * (lo + hi)
        p = sweep(code, [mid], frames, frames, frames + 1, decoder, seed, log=log)[0]
        if p.fer > target_fer:
            lo = mid
        else:
            hi = mid
    return round(0.5 * (lo + hi), 3)


def _parse_range(s: str) -> List[float]:
    if ":" in s:
        a, b, st = (float(x) for x in s.split(":"))
        return [round(a + i * st, 6) for i in range(int(round((b - a) / st)) + 1)]
    return [float(x) for x in s.split(",")]


def spec_by_name(name: str) -> codes.QcIraSpec:
    """n2304, n64800, or n<N>-k<K> for one of the twelve Z=360 codes."""
    if name == "n2304":
        return codes.n2304_r12()
    if name == "n64800":
        return codes.n64800_r12()
    for s in codes.mixed_twelve():
        if name == f"n{s.n}-k{s.k}":
            return s
    raise ValueError(f"unknown code {name!r}")


def code_by_name(name: str, ctx: Context):
    """(label, LdpcCode): a QC-IRA name as spec_by_name, dvb-normal-<rate> /
    dvb-short-<rate> (rate 1/2, 3/5, 2/3, 3/4, 4/5, 5/6) for a DVB-T2 code,
    or the path of a GF2DATA parity-check matrix file."""
    if name.startswith("dvb-"):
        _, frame, rate = name.split("-", 2)
        r = codes.DVB_RATE_NAMES.index(rate)
        code = (LdpcCode.dvb_t2_normal if frame == "normal" else LdpcCode.dvb_t2_short)(r, ctx)
        return codes.dvb_t2(1 if frame == "normal" else 0, r).name, code
    if name.endswith(".gf2") or "/" in name:
        return name, LdpcCode.from_gf2data(name, ctx)
    spec = spec_by_name(name)
    return spec.name, LdpcCode(spec.qc(), ctx)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--code", default="n64800")
    ap.add_argument("--ebn0", default="1.5:2.5:0.25")
    ap.add_argument("--frames", type=int, default=4096)
    ap.add_argument("--max-frames", type=int, default=16384)
    ap.add_argument("--min-errors", type=int, default=50)
    ap.add_argument("--max-it", type=int, default=50)
    ap.add_argument("--weights", default=None, help="learned decoder weights (.npz); plain BP if absent")
    ap.add_argument("--target-fer", type=float, default=None,
                    help="find the Eb/N0 of this FER by bisection over the --ebn0 range ends")
    ap.add_argument("--json", default=None)
    ap.add_argument("--csv", default=None, help="also write the points with results_to_csv's columns")
    ap.add_argument("--zero", action="store_true",
                    help="all-zero codewords (automatic for codes the device encoder does not cover)")
    a = ap.parse_args(argv)
    ctx = Context(0)
    label, code = code_by_name(a.code, ctx)
    if a.weights:
        from .nms import NmsDecoder
        decoder = learned_decoder(NmsDecoder.load(code, a.weights), a.max_it)
    else:
        decoder = plain_decoder(code, a.max_it)
    if a.target_fer is not None:
        r = _parse_range(a.ebn0)
        e = calibrate(code, a.target_fer, min(r), max(r), a.frames, decoder=decoder)
        print(json.dumps({"code": label, "target_fer": a.target_fer, "ebn0_db": e}))
        return 0
    pts = sweep(code, _parse_range(a.ebn0), a.frames, a.max_frames, a.min_errors, decoder, name=label,
                zero_codeword=True if a.zero else None)
    if a.json:
        with open(a.json, "w") as f:
            json.dump({"code": label, "points": [asdict(p) for p in pts]}, f, indent=1)
    if a.csv:
        write_csv(pts, a.csv, code.k())
    return 0


if __name__ == "__main__":
    sys.exit(main())
