"""Benchmark: information Gbit/s of batched LDPC decoding on one MI355X.

Workload: the N=64800, K=32400 (rate 1/2) QC-IRA code with Z=360 and the
DVB-T2 normal-1/2 degree profile (90x180 base graph, E=226,440 edges), 4096
frames per batch of BPSK/AWGN channel LLRs (FP32) at Eb/N0 = 1.62 dB, where
plain min-sum BP reaches FER about 1e-3.  Frames are generated on the GPU
(seeded Philox, staircase encoder) before the timed region; a step decodes the
whole resident batch and leaves its outputs in HBM.

Two decoders are timed on the same batch:
  learned  the learned normalized min-sum decoder, 10 weighted flooding
           iterations (weights trained on the GPU, gf2_amd/weights/), frames
           failing the syndrome re-decoded by plain 50-iteration BP  -> value
  plain    flooding min-sum BP, 50 iterations, bit-exact with the reference
           decoder                                                  -> config
and, when its checkpoint exists, a 50-iteration weighted network with the
same fallback (config.deeper_learned), which converges most frames itself.
The JSON line carries
  roofline      achieved HBM bandwidth of the dominant kernel of the learned
                run, from HIP events recorded by the library around every
                launch in the timed region, against the 8 TB/s HBM3E peak
  cpu_baseline  the structural C++ restatement of the reference's rayon
                decode_batch (oracle/ref_port.cpp) on a bounded sample of the
                same frames, on this host's cores; beside it the same
                arithmetic on flat arrays (oracle/gf2_oracle.c)

    python bench.py [--steps K] [--warmup W] [--frames F] [--ebn0 X]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
# HBM bytes per frame-iteration of the check-state kernels on this workload,
# measured with rocprofv3 FETCH_SIZE / WRITE_SIZE (weighted-iteration
# dispatches, all 1024 frames running) and scaled by the counts of a plain
# copy of known size (bytes read = 2 x FETCH_SIZE, written = 1 x WRITE_SIZE):
# profiles/n64800_state_kernels_bench.md.
MEASURED_TRAFFIC_PER_FI = {"check": 827_374 + 389_200, "variable": 1_173_078 + 271_543}
# Plain min-sum FER on this code is about 1e-3 here: the gf2_amd.fer sweep
# measured 2.1e-3 at 1.60 dB (16384 frames) and 0 errors in 16384 at 1.65 dB.
DEFAULT_EBN0 = 1.62


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--frames", type=int, default=4096)
    ap.add_argument("--ebn0", type=float, default=DEFAULT_EBN0)
    ap.add_argument("--max-it", type=int, default=50)
    ap.add_argument("--weights", default=os.path.join(ROOT, "gf2_amd", "weights", "nms_n64800_t10.npz"))
    ap.add_argument("--deep-weights", default=os.path.join(ROOT, "gf2_amd", "weights", "nms_n64800_t50.npz"),
                    help="a deeper learned decoder timed beside the headline one ('' to skip)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--wg-per-cu", type=int, default=0, help="override the launch tuning")
    ap.add_argument("--xcd-map", type=int, default=-1, help="override the launch tuning")
    ap.add_argument("--path", default="auto", choices=["auto", "state", "stream"])
    return ap.parse_args()


def cpu_baseline(spec, base, llr, max_it, seconds):
    """Time the structural restatement of the rayon path on a bounded sample."""
    from oracle import oracle as orc
    threads = max(1, min(16, os.cpu_count() or 1))
    oc = orc.Code.from_qc(base, spec.z)
    done = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds and done + threads <= llr.shape[0]:
        oc.port_decode_batch(llr[done:done + threads], max_it, threads)
        done += threads
    dt = time.perf_counter() - t0
    out = {"value": done * spec.k / dt / 1e9, "unit": "Gbit/s", "cores": threads, "kind": "port",
           "sample": f"first {done} frames of the same batch, {dt:.1f} s, plain min-sum "
                     f"{max_it} iterations, oracle/ref_port.cpp with {threads} worker threads"}
    # the same arithmetic on flat CSR/CSC arrays (the oracle's decoder), so the
    # GPU/CPU ratio does not rest on the reference's per-frame allocations
    n2 = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds / 2 and n2 + 4 * threads <= llr.shape[0]:
        oc.decode_batch(llr[n2:n2 + 4 * threads], max_it, threads=threads)
        n2 += 4 * threads
    dt2 = time.perf_counter() - t0
    if n2:
        out["flat_array_c"] = {"value": n2 * spec.k / dt2 / 1e9, "cores": threads,
                               "sample": f"first {n2} frames, {dt2:.1f} s, oracle/gf2_oracle.c decode_batch"}
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(frames would shard across GPUs with no exchange between them)"}))
        return 2
    from gf2_amd import _lib, channel, codes
    from gf2_amd.device import DeviceBuffer, generate_frames
    from gf2_amd.ldpc import Context, LdpcCode
    from gf2_amd.nms import NmsDecoder

    spec = codes.n64800_r12()
    base = spec.base_matrix()
    ctx = Context(0)
    code = LdpcCode(spec.qc(), ctx)
    lib = _lib.load()
    if a.wg_per_cu:
        _lib.check(ctx.handle, lib.gf2l_set_tuning(ctx.handle, _lib.TUNE_WG_PER_CU, a.wg_per_cu))
    if a.xcd_map >= 0:
        _lib.check(ctx.handle, lib.gf2l_set_tuning(ctx.handle, _lib.TUNE_XCD_MAP, a.xcd_map))
    ctx.set_path({"auto": _lib.PATH_AUTO, "state": _lib.PATH_STATE,
                  "stream": _lib.PATH_STREAM}[a.path])
    F, kw = a.frames, (spec.k + 63) // 64
    d_llr, d_cw = generate_frames(code, a.seed, 0, F, a.ebn0, a.ebn0)
    info = channel.unpack_words(d_cw.download(np.uint64, (F, (spec.n + 63) // 64)), spec.n)[:, :spec.k]
    outs = [DeviceBuffer(ctx, F * kw * 8), DeviceBuffer(ctx, F * 4), DeviceBuffer(ctx, F),
            DeviceBuffer(ctx, F), DeviceBuffer(ctx, F)]
    learned = NmsDecoder.load(code, a.weights)

    def plain_step():
        _lib.check(ctx.handle, lib.gf2l_decode_batch_device(
            ctx.handle, code.handle, d_llr.ptr, F, a.max_it, outs[0].ptr, outs[1].ptr, outs[2].ptr,
            outs[3].ptr))
        ctx.synchronize()

    def learned_step():
        learned.decode_device(d_llr, F, a.max_it, *outs)
        ctx.synchronize()

    def run(step, profile):
        for _ in range(a.warmup):
            step()
        ctx.synchronize()
        if profile:
            _lib.check(ctx.handle, lib.gf2l_profile_enable(ctx.handle, 1))
        t0 = time.perf_counter()
        for _ in range(a.steps):
            step()
        ctx.synchronize()
        dt = time.perf_counter() - t0
        prof = None
        if profile:
            ms = (ctypes.c_double * 3)()
            la = (ctypes.c_uint64 * 3)()
            fi = ctypes.c_uint64()
            lib.gf2l_profile_read(ctx.handle, ms, la, ctypes.byref(fi))
            lib.gf2l_profile_enable(ctx.handle, 0)
            prof = (list(ms), list(la), fi.value)
        msg = outs[0].download(np.uint64, (F, kw))
        it = outs[1].download(np.uint32, F)
        conv = outs[2].download(np.uint8, F).astype(bool)
        fb = outs[4].download(np.uint8, F).astype(bool)
        bits = channel.unpack_words(msg, spec.k)
        fe = int(((~conv) | (bits != info).any(axis=1)).sum())
        return dt, it, fe, fb, prof

    dt_p, it_p, fe_p, _, _ = run(plain_step, False)
    deep = None
    if a.deep_weights and os.path.exists(a.deep_weights):
        # a deeper weighted network with the same fallback, reported beside
        # the specified 10-iteration one (not the headline value)
        deep_dec = NmsDecoder.load(code, a.deep_weights)

        def deep_step():
            deep_dec.decode_device(d_llr, F, a.max_it, *outs)
            ctx.synchronize()

        dt_d, it_d, fe_d, fb_d, _ = run(deep_step, False)
        deep = {"weights": os.path.relpath(a.deep_weights, ROOT), "weighted_iterations": deep_dec.iterations,
                "gbps": round(a.steps * F * spec.k / dt_d / 1e9, 4), "fer": fe_d / F,
                "mean_iterations": round(float(it_d.mean()), 3),
                "fallback_fraction": round(float(fb_d.mean()), 4),
                "over_plain": round(dt_p / dt_d, 4)}
    dt_l, it_l, fe_l, fb_l, (ms, launches, frame_iters) = run(learned_step, True)
    gbps_p = a.steps * F * spec.k / dt_p / 1e9
    gbps_l = a.steps * F * spec.k / dt_l / 1e9

    N, M = spec.n, spec.n - spec.k
    # Algorithmic bytes per frame-iteration of each kernel of the compressed
    # check-state path (the unique data each kernel must move):
    #   check     read the frame's beliefs once (4N), read and write the
    #             12-byte state of every check (24M)
    #   variable  read every check state once (12M), read L (4N), write
    #             beliefs (4N) and hard-decision bits (N/8)
    #   syndrome  read the hard-decision bits (N/8)
    per_fi = {"check": 4 * N + 24 * M, "variable": 12 * M + 8 * N + N // 8, "syndrome": N // 8}
    names = ["check", "variable", "syndrome"]
    kernels = ["cn_state_kernel", "vn_state_kernel", "syn_words_kernel"]
    dom = max(range(3), key=lambda k: ms[k])
    achieved = per_fi[names[dom]] * frame_iters / (ms[dom] * 1e-3) / 1e9
    out = {
        "metric": "info Gbit/s on 1 MI355X at FER 1e-3, learned vs plain BP, N=64800 r=1/2",
        "value": round(gbps_l, 4),
        "unit": "Gbit/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(dt_l / a.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: on-GPU seeded Philox info bits, staircase encoder, BPSK/AWGN LLRs",
        "config": {
            "workload": "QC-IRA N=64800 K=32400 Z=360 (E=226440), 4096 frames/batch, Eb/N0 "
                        f"{a.ebn0} dB",
            "decoder": "learned NMS, 10 weighted flooding iterations + plain BP fallback (50)",
            "weights": os.path.relpath(a.weights, ROOT),
            "learned_gbps": round(gbps_l, 4), "learned_fer": fe_l / F,
            "learned_mean_iterations": round(float(it_l.mean()), 3),
            "learned_fallback_fraction": round(float(fb_l.mean()), 4),
            "plain_bp_gbps": round(gbps_p, 4), "plain_bp_fer": fe_p / F,
            "plain_bp_mean_iterations": round(float(it_p.mean()), 3),
            "plain_bp_ms_per_step": round(dt_p / a.steps * 1e3, 3),
            "learned_over_plain": round(gbps_l / gbps_p, 4),
            "frames": F, "ebn0_db": a.ebn0, "max_iterations": a.max_it,
            "deeper_learned": deep,
        },
        "roofline": {
            "bound": "hbm", "kernel": kernels[dom],
            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4),
            "traffic": (round(MEASURED_TRAFFIC_PER_FI[names[dom]] * frame_iters / launches[dom])
                        if a.path in ("auto", "state") and names[dom] in MEASURED_TRAFFIC_PER_FI else None),
            "traffic_unit": "bytes per launch (measured bytes per frame-iteration x mean frame-iterations per launch)",
            "avg_launch_us": round(ms[dom] / launches[dom] * 1e3, 2),
            "launches": int(launches[dom]),
            "bytes_per_frame_iteration": per_fi[names[dom]],
            "kernel_ms": {names[k]: round(ms[k], 3) for k in range(3)},
        },
        "cpu_baseline": None,
    }
    if not a.no_cpu:
        n_sample = min(F, 1024)
        llr = d_llr.download(np.float32, (n_sample, spec.n))
        out["cpu_baseline"] = cpu_baseline(spec, base, llr, a.max_it, a.cpu_seconds)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
