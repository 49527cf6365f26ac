"""DVB-T2 decode throughput on the GPU (bit-exact min-sum, quasi-cyclic kernels).

Two workloads per code:
  clean   the reference's own benchmark input (crates/gf2-coding/benches/
          ldpc_throughput.rs:97-124, tests/dvb_t2_ldpc_verification_suite.rs:
          150-228): codewords sent as error-free +-10 LLRs, so every frame
          stops after one iteration; batches of 202 frames (the reference's
          TP06->TP05 block count) and of --frames frames
  awgn    BPSK/AWGN frames at --ebn0 dB (default: near each code's
          waterfall), 50 iterations
Frames are encoded on the host (ETSI accumulator encoder) and uploaded
before timing.  Reported per line: info Mbit/s of the device entry point
(LLRs resident in HBM), of the host entry point for the 202-frame batch
(PCIe copies included), mean iterations and frames not converged.

    python -m gf2_amd.dvb_bench [--codes normal-3/5,normal-1/2] [--frames 4096]
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import _lib, codes
from .channel import awgn_llr, encode_dvb
from .device import DeviceBuffer
from .ldpc import Context, LdpcCode

# Eb/N0 (dB) near each code's min-sum waterfall (50 iterations)
WATERFALL_DB = {(1, 0): 1.6, (1, 1): 2.3, (1, 2): 2.1, (1, 3): 2.5, (1, 4): 2.9, (1, 5): 3.2,
                (0, 0): 1.4, (0, 1): 2.0, (0, 2): 2.2, (0, 3): 2.7, (0, 4): 3.0, (0, 5): 3.4}


def parse_codes(text: str):
    out = []
    for item in text.split(","):
        frame, rate = item.strip().split("-")
        out.append((1 if frame == "normal" else 0, codes.DVB_RATE_NAMES.index(rate)))
    return out


def time_device(ctx, code, d_llr, frames, max_it, reps):
    lib = _lib.load()
    kw = (code.k() + 63) // 64
    outs = [DeviceBuffer(ctx, frames * kw * 8), DeviceBuffer(ctx, frames * 4), DeviceBuffer(ctx, frames),
            DeviceBuffer(ctx, frames)]

    def run():
        _lib.check(ctx.handle, lib.gf2l_decode_batch_device(ctx.handle, code.handle, d_llr.ptr, frames, max_it,
                                                            *[o.ptr for o in outs]))
        ctx.synchronize()

    run()
    t0 = time.perf_counter()
    for _ in range(reps):
        run()
    dt = (time.perf_counter() - t0) / reps
    it = outs[1].download(np.uint32, frames)
    conv = outs[2].download(np.uint8, frames)
    return dt, float(it.mean()), int(frames - conv.sum())


def time_host(ctx, code, llr, max_it, reps):
    lib = _lib.load()
    F = llr.shape[0]
    kw = (code.k() + 63) // 64
    msg = np.zeros((F, kw), dtype=np.uint64)
    it = np.zeros(F, dtype=np.uint32)
    conv = np.zeros(F, dtype=np.uint8)
    synd = np.zeros(F, dtype=np.uint8)

    def run():
        _lib.check(ctx.handle, lib.gf2l_decode_batch(ctx.handle, code.handle, llr.ctypes.data, F, max_it,
                                                     msg.ctypes.data, it.ctypes.data, conv.ctypes.data,
                                                     synd.ctypes.data))

    run()
    t0 = time.perf_counter()
    for _ in range(reps):
        run()
    return (time.perf_counter() - t0) / reps


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--codes", default="normal-3/5,normal-1/2,short-1/2")
    ap.add_argument("--frames", type=int, default=4096)
    ap.add_argument("--ebn0", type=float, default=None)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--csr", action="store_true", help="also time the CSR path on the awgn frames")
    a = ap.parse_args(argv)
    ctx = Context(0)
    for frame, rate in parse_codes(a.codes):
        spec = codes.dvb_t2(frame, rate)
        code = LdpcCode(None, ctx, _build=lambda c, h: _lib.load().gf2l_code_dvb_t2(c.handle, frame, rate, h))
        rng = np.random.default_rng(a.seed)
        F = a.frames
        info = rng.integers(0, 2, size=(F, spec.k), dtype=np.uint8)
        cw = encode_dvb(spec.table, spec.n, spec.k, spec.q, info)
        clean = np.where(cw == 0, 10.0, -10.0).astype(np.float32)
        ebn0 = a.ebn0 if a.ebn0 is not None else WATERFALL_DB[(frame, rate)]
        noisy = awgn_llr(cw, ebn0, spec.k / spec.n, rng)
        del cw
        line = {"code": spec.name, "n": spec.n, "k": spec.k, "nnz": code.nnz(), "frames": F}
        if a.csr:  # the same code on the general sparse path, for comparison
            d = DeviceBuffer.from_array(ctx, noisy)
            csr = LdpcCode.from_edges(spec.m, spec.n, spec.edges(), ctx)
            dt, mit, bad = time_device(ctx, csr, d, F, 50, a.reps)
            line["awgn_csr_path_mbps"] = round(spec.k * F / dt / 1e6, 1)
            d.free()
            del csr
        for name, llr, max_it in (("clean", clean, 50), ("awgn", noisy, 50)):
            d = DeviceBuffer.from_array(ctx, llr)
            dt, mit, bad = time_device(ctx, code, d, F, max_it, a.reps)
            line[f"{name}_mbps"] = round(spec.k * F / dt / 1e6, 1)
            line[f"{name}_mean_iterations"] = round(mit, 3)
            line[f"{name}_not_converged"] = bad
            if name == "clean":
                d202 = DeviceBuffer.from_array(ctx, np.ascontiguousarray(llr[:202]))
                dt2, _, _ = time_device(ctx, code, d202, 202, max_it, a.reps)
                line["clean_202_mbps"] = round(spec.k * 202 / dt2 / 1e6, 1)
                dth = time_host(ctx, code, np.ascontiguousarray(llr[:202]), max_it, a.reps)
                line["clean_202_host_entry_mbps"] = round(spec.k * 202 / dth / 1e6, 1)
                d202.free()
            else:
                line["awgn_ebn0_db"] = ebn0
            d.free()
        print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
