"""Benchmark: information Gbit/s of batched LDPC decoding on one MI355X.

Workload: the N=64800, K=32400 (rate 1/2) QC-IRA code with Z=360 and the
DVB-T2 normal-1/2 degree profile (90x180 base graph, E=226,440 edges), 4096
frames per batch of BPSK/AWGN channel LLRs (FP32) at Eb/N0 = 1.62 dB, where
plain min-sum BP reaches FER about 1e-3 (profiles/n64800_fer_sweep.md).
Frames are generated on the GPU (seeded Philox, staircase encoder) before
the timed region; a step decodes the whole resident batch and leaves its
outputs in HBM.

Decoders timed on the same batch:
  learned   the layered learned normalized min-sum decoder: 10 weighted
            row-sequential iterations with early stop (weights trained on the
            GPU, gf2_amd/weights/nms_n64800_layered_t10.npz); frames still
            failing continue with plain flooding min-sum from that state for
            up to 50 iterations                                      -> value
  plain     flooding min-sum BP, 50 iterations, bit-exact with the reference
            decoder                                                  -> config
  flooding  the 10-iteration flooding network (gf2_amd/weights/
            nms_n64800_t10.npz) with the restarting and with the continuing
            fallback                                                 -> config
The JSON line carries
  roofline        the dominant kernel of the learned run, from HIP events the
                  library records around every launch in the timed region
                  (per-kernel units from gf2l_profile_read_kernel), against
                  the 8 TB/s HBM3E peak; its LDS figure beside it
  roofline_plain  the same for the plain decoder's dominant kernel
  cpu_baseline    the structural C++ restatement of the reference's rayon
                  decode_batch (oracle/ref_port.cpp) on a bounded sample of
                  the same frames, on this host's CPU share and on one thread;
                  beside it the same arithmetic on flat arrays, and the port
                  on the reference benchmark's own workload
  reference_workload  that workload (DVB-T2 normal r3/5, 202 error-free
                  frames, benches/ldpc_throughput.rs:97-124) on the GPU,
                  beside the reference's published CPU figure

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
# LDS: 256 CUs x 128 bytes per clock x 2.4 GHz
LDS_PEAK_GBPS = 256 * 128 * 2.4
# HBM bytes per frame-iteration measured with rocprofv3 FETCH_SIZE and
# WRITE_SIZE on this workload (python bench.py --steps 1 --warmup 0 --no-cpu,
# profiles/n64800_layered_bench.md), scaled by what the counters read for a
# plain copy of known size (FETCH_SIZE counts half the bytes read, WRITE_SIZE
# the bytes written: the calibration in profiles/n64800_state_kernels_bench.md).
MEASURED_TRAFFIC_PER_FI = {"layered_stair": 1036154, "layered_general": 1832398, "check": 1201488,
                           "variable": 1477744}
# Vector-ALU issue: 256 CUs x 4 SIMDs, one 64-lane FP32/INT32 instruction per
# SIMD every 4 clocks (16 lanes per clock) at 2.4 GHz; wave-level VALU
# instructions per frame-iteration measured with rocprofv3 SQ_INSTS_VALU on
# this workload (profiles/n64800_layered_bench.md).
VALU_PEAK_GINST = 256 * 4 * 2.4 / 4
MEASURED_VALU_PER_FI = {"layered_stair": 91482}
# CPU threads of one GPU's share of the box
CPU_SHARE = 16
DEFAULT_EBN0 = 1.62
# The reference's own decode benchmark (crates/gf2-coding/benches/ldpc_throughput.rs:97-124):
# DVB-T2 normal rate-3/5 codewords sent as error-free +-10 LLRs, 202 blocks, published at
# 8.29 Mbit/s on a 24-core CPU (crates/gf2-coding/docs/PARALLELIZATION.md:46-47).
REF_FRAMES, REF_PUBLISHED_MBPS, REF_PUBLISHED_CORES = 202, 8.29, 24


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--frames", type=int, default=4096)
    ap.add_argument("--ebn0", type=float, default=DEFAULT_EBN0)
    ap.add_argument("--max-it", type=int, default=50)
    ap.add_argument("--weights", default=os.path.join(ROOT, "gf2_amd", "weights", "nms_n64800_layered_t10.npz"))
    ap.add_argument("--flooding-weights", default=os.path.join(ROOT, "gf2_amd", "weights", "nms_n64800_t10.npz"),
                    help="the 10-iteration flooding network timed beside the headline one ('' to skip)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--layered-columns", type=int, default=0, help="override GF2L_TUNE_LAYERED_COLUMNS")
    return ap.parse_args()


def reference_frames():
    """The reference benchmark's input: DVB-T2 normal r3/5 codewords of random
    messages as error-free LLRs (+10 for a 0 bit, -10 for a 1 bit)."""
    from gf2_amd import codes
    from gf2_amd.channel import encode_dvb
    s = codes.dvb_t2(1, 1)
    info = np.random.default_rng(97).integers(0, 2, size=(REF_FRAMES, s.k), dtype=np.uint8)
    cw = encode_dvb(s.table, s.n, s.k, s.q, info)
    return s, info, np.where(cw == 1, -10.0, 10.0).astype(np.float32)


def reference_workload_gpu(ctx, s, info, llr, max_it):
    """The reference benchmark's workload decoded on the GPU (device entry
    point, LLRs resident in HBM)."""
    from gf2_amd import _lib, channel
    from gf2_amd.device import DeviceBuffer
    from gf2_amd.ldpc import CodeRate, LdpcCode
    lib = _lib.load()
    code = LdpcCode.dvb_t2_normal(CodeRate.RATE_3_5, ctx)
    F, kw = llr.shape[0], (s.k + 63) // 64
    d_llr = DeviceBuffer.from_array(ctx, llr)
    outs = [DeviceBuffer(ctx, F * kw * 8), DeviceBuffer(ctx, F * 4), DeviceBuffer(ctx, F), DeviceBuffer(ctx, F)]

    def run():
        _lib.check(ctx.handle, lib.gf2l_decode_batch_device(ctx.handle, code.handle, d_llr.ptr, F, max_it,
                                                            *[o.ptr for o in outs]))
        ctx.synchronize()

    run()
    reps = 20
    t0 = time.perf_counter()
    for _ in range(reps):
        run()
    dt = (time.perf_counter() - t0) / reps
    bits = channel.unpack_words(outs[0].download(np.uint64, (F, kw)), s.k)
    return {"workload": f"DVB-T2 normal r3/5 (n={s.n}, k={s.k}), {F} codewords as error-free +-10 LLRs, "
                        f"max {max_it} iterations (crates/gf2-coding/benches/ldpc_throughput.rs:97-124)",
            "gpu_mbps": round(F * s.k / dt / 1e6, 1), "us_per_batch": round(dt * 1e6, 1),
            "mean_iterations": round(float(outs[1].download(np.uint32, F).mean()), 3),
            "decoded_correctly": int((bits == info).all(axis=1).sum()),
            "published_cpu_mbps": REF_PUBLISHED_MBPS, "published_cpu_cores": REF_PUBLISHED_CORES}


def cpu_baseline(spec, base, llr, max_it, seconds, ref=None):
    """Time the structural restatement of the rayon path on a bounded sample
    (and on the reference benchmark's own workload when ref is given)."""
    from oracle import oracle as orc
    nproc = os.cpu_count() or 1
    threads = max(1, min(CPU_SHARE, nproc))
    oc = orc.Code.from_qc(base, spec.z)

    def port(th, budget):
        done = 0
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < budget and done + th <= llr.shape[0]:
            oc.port_decode_batch(llr[done:done + th], max_it, th)
            done += th
        return done, time.perf_counter() - t0

    done, dt = port(threads, seconds)
    out = {"value": done * spec.k / dt / 1e9, "unit": "Gbit/s", "cores": threads, "kind": "port",
           "nproc": nproc,
           "sample": f"first {done} frames of the same batch, {dt:.1f} s, plain min-sum {max_it} iterations, "
                     f"oracle/ref_port.cpp with {threads} worker threads (the box's CPU share for one GPU; "
                     f"nproc = {nproc})"}
    d1, t1 = port(1, seconds / 3)
    if d1:
        out["single_thread"] = {"value": d1 * spec.k / t1 / 1e9, "cores": 1,
                                "sample": f"first {d1} frames, {t1:.1f} s, oracle/ref_port.cpp on one thread"}
    # the same arithmetic on flat CSR/CSC arrays (the oracle's decoder), so the
    # GPU/CPU ratio does not rest on the reference's per-frame allocations
    n2 = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds / 3 and n2 + 4 * threads <= llr.shape[0]:
        oc.decode_batch(llr[n2:n2 + 4 * threads], max_it, threads=threads)
        n2 += 4 * threads
    dt2 = time.perf_counter() - t0
    if n2:
        out["flat_array_c"] = {"value": n2 * spec.k / dt2 / 1e9, "cores": threads,
                               "sample": f"first {n2} frames, {dt2:.1f} s, oracle/gf2_oracle.c decode_batch"}
    if ref is not None:  # the reference benchmark's workload on the port, all of it and one thread
        s, rllr = ref
        rc = orc.Code(s.m, s.n, orc.dvb_edges(s.table, s.n, s.k, s.q))
        t0 = time.perf_counter()
        rc.port_decode_batch(rllr, max_it, threads)
        tr = time.perf_counter() - t0
        n1 = 8
        t0 = time.perf_counter()
        rc.port_decode_batch(rllr[:n1], max_it, 1)
        t1r = time.perf_counter() - t0
        out["reference_workload"] = {
            "value_mbps": round(rllr.shape[0] * s.k / tr / 1e6, 2), "cores": threads,
            "single_thread_mbps": round(n1 * s.k / t1r / 1e6, 2),
            "published_mbps": REF_PUBLISHED_MBPS, "published_cores": REF_PUBLISHED_CORES,
            "sample": f"all {rllr.shape[0]} frames of the reference benchmark's workload on oracle/ref_port.cpp "
                      f"with {threads} threads ({tr:.2f} s); one thread: first {n1} frames ({t1r:.2f} s)"}
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
    if a.layered_columns:
        ctx.set_tuning(_lib.TUNE_LAYERED_COLUMNS, a.layered_columns)
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

    def nms_step(dec):
        def step():
            dec.decode_device(d_llr, F, a.max_it, *outs)
            ctx.synchronize()
        return step

    def run(step, profile):
        for _ in range(a.warmup):
            step()
        ctx.synchronize()
        if profile:
            _lib.check(ctx.handle, lib.gf2l_profile_enable(ctx.handle, profile))
        t0 = time.perf_counter()
        for _ in range(a.steps):
            step()
        ctx.synchronize()
        dt = time.perf_counter() - t0
        prof = None
        if profile:
            prof = {}
            for k, name in enumerate(["check", "variable", "syndrome", "layered"]):
                ms, la, fi = ctypes.c_double(), ctypes.c_uint64(), ctypes.c_uint64()
                lib.gf2l_profile_read_kernel(ctx.handle, k, ctypes.byref(ms), ctypes.byref(la), ctypes.byref(fi))
                prof[name] = (ms.value, la.value, fi.value)
            lib.gf2l_profile_enable(ctx.handle, 0)
        it = outs[1].download(np.uint32, F)
        conv = outs[2].download(np.uint8, F).astype(bool)
        fb = outs[4].download(np.uint8, F).astype(bool)
        bits = channel.unpack_words(outs[0].download(np.uint64, (F, kw)), spec.k)
        fe = int(((~conv) | (bits != info).any(axis=1)).sum())
        return dt, it, fe, fb, prof

    gbps = lambda dt: a.steps * F * spec.k / dt / 1e9
    # plain BP timed unprofiled (its Gbit/s), then once more with events around
    # every launch (its roofline; profiling reads the running count back after
    # every iteration)
    dt_p, it_p, fe_p, _, _ = run(plain_step, 0)
    _, _, _, _, prof_p = run(plain_step, _lib.PROFILE_ALL)
    flooding = None
    if a.flooding_weights and os.path.exists(a.flooding_weights):
        fdec = NmsDecoder.load(code, a.flooding_weights)
        flooding = {"weights": os.path.relpath(a.flooding_weights, ROOT), "weighted_iterations": fdec.iterations}
        for mode in ("restart", "continue"):
            fdec.fallback = mode
            dt_f, it_f, fe_f, fb_f, _ = run(nms_step(fdec), False)
            flooding[mode] = {"gbps": round(gbps(dt_f), 4), "fer": fe_f / F,
                              "mean_iterations": round(float(it_f.mean()), 3),
                              "fallback_fraction": round(float(fb_f.mean()), 4),
                              "over_plain": round(dt_p / dt_f, 4)}
    # events around the layered kernel only: the plain continuation after it
    # runs as in an unprofiled decode (per-iteration events and readbacks
    # cost about 0.2 ms per step)
    dt_l, it_l, fe_l, fb_l, prof_l = run(nms_step(learned), _lib.PROFILE_LAYERED)

    cv_n, M, z = spec.n, spec.n - spec.k, spec.z
    nb = spec.n // z
    xc = ctypes.c_uint32()
    _lib.check(ctx.handle, lib.gf2l_nms_layered_columns(ctx.handle, learned.handle, ctypes.byref(xc)))
    X = xc.value
    kb = nb - M // z
    # the staircase kernel runs when every information column is in LDS and no
    # column count was forced (GF2L_TUNE_LAYERED_KERNEL automatic)
    stair = X == kb and not a.layered_columns
    edges_lds = int((base[:, :X] >= 0).sum()) * z
    E = int((base >= 0).sum()) * z
    # Algorithmic bytes per frame-iteration of each kernel (unique data it must move):
    #   check     read the frame's beliefs once (4N), read and write the 12-byte state of every check (24M)
    #   variable  read every check state once (12M), read L (4N), write beliefs (4N) and hard bits (N/8)
    #   syndrome  read the hard bits (N/8)
    #   layered, staircase kernel: HBM: read and write the 12-byte state of every check (24M), read
    #             and write each parity belief once (8 z (nb - kb)); LDS: read and write the
    #             information beliefs once per edge (8 E_info) and read them once for the hard
    #             decisions (4 z kb)
    #   layered, general kernel: HBM: 24M, read and write the beliefs of block-columns outside LDS
    #             once per edge (8 E_glob) and read them once for the hard decisions (4 z (nb - X));
    #             LDS: the same for the X columns in LDS
    per_fi = {"check": 4 * cv_n + 24 * M, "variable": 12 * M + 8 * cv_n + cv_n // 8, "syndrome": cv_n // 8,
              "layered": (24 * M + 8 * z * (nb - kb)) if stair else
                         (24 * M + 8 * (E - edges_lds) + 4 * z * (nb - X))}
    lds_per_fi = 8 * edges_lds + 4 * z * X
    kernels = {"check": "cn_state_kernel", "variable": "vn_state_kernel", "syndrome": "syn_words_kernel",
               "layered": "lnms_stair_kernel" if stair else "lnms_fwd_kernel"}
    traffic_key = {"layered": "layered_stair" if stair else "layered_general"}

    def roofline(prof):
        dom = max(prof, key=lambda k: prof[k][0])
        ms, launches, units = prof[dom]
        achieved = per_fi[dom] * units / (ms * 1e-3) / 1e9
        tk = traffic_key.get(dom, dom)
        r = {"bound": "hbm", "kernel": kernels[dom], "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS,
             "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4),
             "traffic": (round(MEASURED_TRAFFIC_PER_FI[tk] * units / launches)
                         if tk in MEASURED_TRAFFIC_PER_FI and launches else None),
             "traffic_unit": "HBM bytes per launch (rocprofv3 bytes per frame-iteration on this workload x "
                             "frame-iterations per launch)",
             "avg_launch_us": round(ms / launches * 1e3, 2), "launches": int(launches),
             "frame_iterations": int(units), "bytes_per_frame_iteration": per_fi[dom],
             "kernel_ms": {k: round(v[0], 3) for k, v in prof.items() if v[1]}}
        if dom == "layered":
            lds = lds_per_fi * units / (ms * 1e-3) / 1e9
            r["lds"] = {"achieved": round(lds, 1), "peak": round(LDS_PEAK_GBPS, 1), "unit": "GB/s",
                        "frac": round(lds / LDS_PEAK_GBPS, 4), "bytes_per_frame_iteration": lds_per_fi,
                        "lds_columns": X}
            if tk in MEASURED_VALU_PER_FI:
                # the issue rate that bounds the staircase kernel: measured
                # instructions per frame-iteration x frame-iterations / time
                v = MEASURED_VALU_PER_FI[tk] * units / (ms * 1e-3) / 1e9
                r["valu"] = {"achieved": round(v, 1), "peak": round(VALU_PEAK_GINST, 1), "unit": "G wave-inst/s",
                             "frac": round(v / VALU_PEAK_GINST, 4),
                             "inst_per_frame_iteration": MEASURED_VALU_PER_FI[tk]}
        if dom == "variable":
            # the whole iteration against the uncompressed message model 16E + 4N
            fi = prof["variable"][2]
            t_it = sum(prof[k][0] for k in ("check", "variable", "syndrome"))
            r["message_model_effective_gbps"] = round((16 * E + 4 * cv_n) * fi / (t_it * 1e-3) / 1e9, 1)
        return r

    out = {
        "metric": "info Gbit/s on 1 MI355X at FER 1e-3, learned vs plain BP, N=64800 r=1/2",
        "value": round(gbps(dt_l), 4),
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
            "decoder": f"learned NMS, {learned.iterations} weighted layered iterations + plain BP "
                       f"continuing from that state ({a.max_it})",
            "weights": os.path.relpath(a.weights, ROOT),
            "learned_gbps": round(gbps(dt_l), 4), "learned_fer": fe_l / F,
            "learned_mean_iterations": round(float(it_l.mean()), 3),
            "learned_fallback_fraction": round(float(fb_l.mean()), 4),
            "plain_bp_gbps": round(gbps(dt_p), 4), "plain_bp_fer": fe_p / F,
            "plain_bp_mean_iterations": round(float(it_p.mean()), 3),
            "plain_bp_ms_per_step": round(dt_p / a.steps * 1e3, 3),
            "learned_over_plain": round(dt_p / dt_l, 4),
            "flooding_learned": flooding,
            "frames": F, "ebn0_db": a.ebn0, "max_iterations": a.max_it,
        },
        "roofline": roofline({k: v for k, v in prof_l.items() if v[1]}),
        "roofline_plain": roofline({k: v for k, v in prof_p.items() if v[1]}),
        "cpu_baseline": None,
    }
    rs, rinfo, rllr = reference_frames()
    out["reference_workload"] = reference_workload_gpu(ctx, rs, rinfo, rllr, a.max_it)
    if not a.no_cpu:
        n_sample = min(F, 1024)
        llr = d_llr.download(np.float32, (n_sample, spec.n))
        out["cpu_baseline"] = cpu_baseline(spec, base, llr, a.max_it, a.cpu_seconds, ref=(rs, rllr))
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
