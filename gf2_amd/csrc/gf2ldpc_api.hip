// C ABI of libgf2ldpc: contexts, quasi-cyclic codes and the batch decoder.
// Declarations and the reference interfaces each call replaces are in
// include/gf2ldpc.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "gf2ldpc.h"
#include "internal.h"

using gf2l::fail;

namespace {

void free_workspace(gf2l::Workspace &ws) {
    hipFree(ws.c2v);
    hipFree(ws.state);
    hipFree(ws.belief);
    hipFree(ws.hb);
    hipFree(ws.lists);
    hipFree(ws.counts);
    hipFree(ws.llr_perm);
    if (ws.h_count) hipHostFree(ws.h_count);
    ws = gf2l::Workspace{};
}

}  // namespace

namespace gf2l {

// Grows the context's workspace so every buffer holds at least the given
// element counts (never shrinks; one reallocation of all buffers).
int ensure_buffers(gf2l_ctx *ctx, size_t frames, size_t c2v, size_t state, size_t bel, size_t hb, size_t llr) {
    gf2l::Workspace &ws = ctx->ws;
    if (ws.lists && ws.cap_frames >= frames && ws.cap_c2v >= c2v && ws.cap_state >= state && ws.cap_bel >= bel &&
        ws.cap_hb >= hb && ws.cap_llr >= llr)
        return GF2L_OK;
    const size_t f = std::max(frames, ws.cap_frames);
    c2v = std::max(c2v, ws.cap_c2v);
    state = std::max(state, ws.cap_state);
    bel = std::max(bel, ws.cap_bel);
    hb = std::max(hb, ws.cap_hb);
    llr = std::max(llr, ws.cap_llr);
    free_workspace(ws);
    if (llr) GF2L_HIP(ctx, hipMalloc(&ws.llr_perm, llr * sizeof(float)));
    if (c2v) GF2L_HIP(ctx, hipMalloc(&ws.c2v, c2v * sizeof(float)));
    if (state) GF2L_HIP(ctx, hipMalloc(&ws.state, state * sizeof(uint32_t)));
    GF2L_HIP(ctx, hipMalloc(&ws.belief, std::max<size_t>(bel, 1) * sizeof(float)));
    GF2L_HIP(ctx, hipMalloc(&ws.hb, std::max<size_t>(hb, 1) * sizeof(uint64_t)));
    GF2L_HIP(ctx, hipMalloc(&ws.lists, 2 * std::max<size_t>(f, 1) * sizeof(uint32_t)));
    GF2L_HIP(ctx, hipMalloc(&ws.counts, gf2l::kCountsWords * sizeof(uint32_t)));  // counts + item counters
    GF2L_HIP(ctx, hipHostMalloc((void **)&ws.h_count, gf2l::kProfileSlots * sizeof(uint32_t), hipHostMallocDefault));
    ws.cap_frames = f;
    ws.cap_c2v = c2v;
    ws.cap_state = state;
    ws.cap_bel = bel;
    ws.cap_hb = hb;
    ws.cap_llr = llr;
    return GF2L_OK;
}

}  // namespace gf2l

namespace {

// Grows the HBM-path workspace to hold `frames` frames of `code` on `path`.
int ensure_workspace(gf2l_ctx *ctx, const gf2l::CodeView &cv, int path, size_t frames) {
    const size_t c2v = path == gf2l::kStream ? (size_t)cv.nbe * cv.z * frames : 0;
    const size_t state = path == gf2l::kState ? 3ull * cv.mb * cv.zp * frames : 0;
    const size_t bel = (path == gf2l::kState ? (size_t)cv.nb * cv.zp : (size_t)cv.n) * frames;
    const size_t hb = (size_t)cv.nb * cv.zw * frames;
    const size_t llr = cv.dvb_q ? (size_t)cv.n * frames : 0;
    return gf2l::ensure_buffers(ctx, frames, c2v, state, bel, hb, llr);
}

}  // namespace

namespace gf2l {

// Resolves the decoder path for a code (GF2L_PATH_*), or -1 if the forced
// path cannot run it.
int resolve_path(const gf2l_ctx *ctx, const gf2l::CodeView &cv) {
    if (cv.dvb_q)  // the DVB lane rules live in the state kernels only
        return (ctx->path == gf2l::kAuto || ctx->path == gf2l::kState) && cv.max_dc <= gf2l::kMaxStateDegree
                   ? gf2l::kState
                   : -1;
    const size_t bytes = gf2l::lds_path_bytes(cv);
    switch (ctx->path) {
        case gf2l::kLds: return bytes <= ctx->lds_limit ? gf2l::kLds : -1;
        case gf2l::kStream: return gf2l::kStream;
        case gf2l::kState: return cv.max_dc <= gf2l::kMaxStateDegree ? gf2l::kState : -1;
        default:
            // keep at least two frames resident per CU on the LDS path
            if (bytes <= std::min<size_t>(ctx->lds_limit, 80 * 1024)) return gf2l::kLds;
            return cv.max_dc <= gf2l::kMaxStateDegree ? gf2l::kState : gf2l::kStream;
    }
}

// Runs the decoder on device buffers.  Without weights: plain min-sum for up
// to max_it iterations.  With weights d_w ([max_it][nbe]): max_it weighted
// iterations, then frames still failing the syndrome get up to fb_it plain
// iterations, from their LLRs or (fb_continue) from the weighted state
// (d_fb marks them).  trace_* (host) and trace_iters are only used by
// gf2l_decode_trace.
int decode_device(gf2l_ctx *ctx, const gf2l_code *code, const float *d_llr, size_t frames,
                  uint32_t max_it, const float *d_w, uint32_t fb_it, bool fb_continue, uint64_t *d_msg,
                  uint32_t *d_iters, uint8_t *d_conv, uint8_t *d_synd, uint8_t *d_fb,
                  float *h_trace_c2v, float *h_trace_bel, uint32_t trace_iters) {
    const gf2l::CodeView &cv = code->view;
    if (frames == 0) return GF2L_OK;
    if (frames > 0xffffffffull) return fail(ctx, GF2L_ERR_INVALID, "too many frames in one batch");
    const uint32_t F = (uint32_t)frames;
    const size_t kw = ((size_t)cv.k + 63) / 64;
    const size_t E = (size_t)cv.nbe * cv.z;
    hipStream_t s = ctx->stream;
    if (!d_w) fb_it = 0;
    if (max_it == 0) {  // fresh decoder: zero beliefs, no iterations (core.rs:906-937)
        GF2L_HIP(ctx, hipMemsetAsync(d_msg, 0, F * kw * sizeof(uint64_t), s));
        GF2L_HIP(ctx, hipMemsetAsync(d_iters, 0, F * sizeof(uint32_t), s));
        GF2L_HIP(ctx, hipMemsetAsync(d_conv, 0, F, s));
        GF2L_HIP(ctx, hipMemsetAsync(d_synd, 1, F, s));
        if (d_fb) GF2L_HIP(ctx, hipMemsetAsync(d_fb, 0, F, s));
        return GF2L_OK;
    }
    if (code->is_csr) {
        if (d_w || h_trace_c2v)
            return fail(ctx, GF2L_ERR_UNSUPPORTED, "weights and traces need a quasi-cyclic code");
        if (ctx->rule != GF2L_RULE_MINSUM)
            return fail(ctx, GF2L_ERR_UNSUPPORTED, "the sum-product rule needs a quasi-cyclic code");
        if (d_fb) GF2L_HIP(ctx, hipMemsetAsync(d_fb, 0, F, s));
        return decode_csr(ctx, code, d_llr, F, max_it, d_msg, d_iters, d_conv, d_synd);
    }
    if (cv.dvb_q && (d_w || h_trace_c2v))
        return fail(ctx, GF2L_ERR_UNSUPPORTED, "weights and traces are not supported for DVB codes");
    int path = resolve_path(ctx, cv);
    if (ctx->rule == GF2L_RULE_SPA) {  // the sum-product rule runs on the STREAM path only
        if (d_w || cv.dvb_q)
            return fail(ctx, GF2L_ERR_UNSUPPORTED, "the sum-product rule takes no weights and no DVB code");
        if (ctx->path != gf2l::kAuto && ctx->path != gf2l::kStream)
            return fail(ctx, GF2L_ERR_UNSUPPORTED, "the sum-product rule needs the STREAM path");
        path = gf2l::kStream;
    }
    if (path < 0) return fail(ctx, GF2L_ERR_UNSUPPORTED, "the forced decoder path cannot run this code");
    if (path == gf2l::kLds) {
        float *d_tc = nullptr, *d_tb = nullptr;
        if (h_trace_c2v && trace_iters) {
            GF2L_HIP(ctx, hipMalloc(&d_tc, (size_t)trace_iters * F * E * sizeof(float)));
            GF2L_HIP(ctx, hipMalloc(&d_tb, (size_t)trace_iters * F * cv.n * sizeof(float)));
        }
        GF2L_HIP(ctx, gf2l::launch_lds_decode(cv, d_llr, F, max_it, d_w, fb_it, fb_continue, d_msg, d_iters,
                                              d_conv, d_synd, d_fb, d_tc, d_tb, trace_iters, ctx->tune.lds_groups, s));
        if (d_tc) {
            GF2L_HIP(ctx, hipMemcpyAsync(h_trace_c2v, d_tc, (size_t)trace_iters * F * E * sizeof(float),
                                         hipMemcpyDeviceToHost, s));
            GF2L_HIP(ctx, hipMemcpyAsync(h_trace_bel, d_tb,
                                         (size_t)trace_iters * F * cv.n * sizeof(float),
                                         hipMemcpyDeviceToHost, s));
            GF2L_HIP(ctx, hipStreamSynchronize(s));
            hipFree(d_tc);
            hipFree(d_tb);
        }
        return GF2L_OK;
    }
    int rc = ensure_workspace(ctx, cv, path, F);
    if (rc != GF2L_OK) return rc;
    gf2l::Workspace &ws = ctx->ws;
    float *d_tc = nullptr, *d_tb = nullptr;  // explicit-layout trace staging (state path)
    if (h_trace_c2v && trace_iters && path == gf2l::kState) {
        GF2L_HIP(ctx, hipMalloc(&d_tc, (size_t)F * E * sizeof(float)));
        GF2L_HIP(ctx, hipMalloc(&d_tb, (size_t)F * cv.n * sizeof(float)));
    }
    struct Free {
        float *a, *b;
        ~Free() { hipFree(a); hipFree(b); }
    } free_trace{d_tc, d_tb};
    return run_hbm(ctx, cv, path, ws, d_llr, F, max_it, d_w, fb_it, fb_continue, d_msg, d_iters, d_conv,
                   d_synd, d_fb, h_trace_c2v, h_trace_bel, trace_iters, d_tc, d_tb);
}

// Captures `unroll` (even) plain iterations of the final phase as a graph:
// list parity starting at `parity`, iteration numbers from the device word
// ws.counts[kItcWord].  Reuses the context's graph when it was captured for
// the same code, sizes, buffers and launch shape.
int hbm_graph(gf2l_ctx *ctx, const CodeView &cv, int path, Workspace &ws, const IterArgs &base, uint32_t unroll,
              uint32_t parity) {
    const Tuning &t = ctx->tune;
    // the code view by value (sizes, degrees and table pointers are all baked
    // into the captured kernel arguments), then everything else the graph uses
    std::vector<uintptr_t> key((sizeof(CodeView) + sizeof(uintptr_t) - 1) / sizeof(uintptr_t), 0);
    std::memcpy(key.data(), &cv, sizeof(CodeView));
    key.insert(key.end(), {
        (uintptr_t)path, base.frames, (uintptr_t)base.llr, (uintptr_t)base.msg,
        (uintptr_t)base.iters, (uintptr_t)base.conv, (uintptr_t)base.synd, (uintptr_t)base.fb,
        (uintptr_t)ws.lists, (uintptr_t)ws.counts, (uintptr_t)ws.state, (uintptr_t)ws.c2v, (uintptr_t)ws.belief,
        (uintptr_t)ws.hb, (uintptr_t)ws.cap_frames, base.itc_last, base.itc_fb_from, (uintptr_t)base.spa, unroll,
        parity, (uintptr_t)(intptr_t)t.wg_per_cu, (uintptr_t)(intptr_t)t.chunk, (uintptr_t)(intptr_t)t.xcd_map,
        (uintptr_t)ctx->num_cus});
    if (ctx->hbm_exec && ctx->hbm_key == key) return GF2L_OK;
    if (ctx->hbm_exec) hipGraphExecDestroy(ctx->hbm_exec);
    ctx->hbm_exec = nullptr;
    ctx->hbm_key.clear();
    hipStream_t s = ctx->stream;
    hipGraph_t graph = nullptr;
    GF2L_HIP(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    hipError_t launch = hipSuccess;
    for (uint32_t u = 0; u < unroll && launch == hipSuccess; ++u) {
        IterArgs a = base;
        a.cur = (parity + u) & 1u;
        launch = launch_hbm_iteration(cv, ws, path, a, ctx->num_cus, ctx->tune, s, nullptr);
    }
    const hipError_t end = hipStreamEndCapture(s, &graph);
    if (launch != hipSuccess || end != hipSuccess) {
        if (graph) hipGraphDestroy(graph);
        return fail(ctx, GF2L_ERR_HIP,
                    std::string("graph capture: ") + hipGetErrorString(launch != hipSuccess ? launch : end));
    }
    const hipError_t inst = hipGraphInstantiate(&ctx->hbm_exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (inst != hipSuccess) {
        ctx->hbm_exec = nullptr;
        return fail(ctx, GF2L_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(inst));
    }
    ctx->hbm_key = std::move(key);
    return GF2L_OK;
}

// The iteration loop of the HBM paths over a prepared workspace: LLR
// permutation of DVB codes, then phase 0 (max_it iterations, weighted by d_w
// when given) and, with fb_it > 0, phase 1 (fb_it plain iterations for
// frames still failing, restarted from the LLRs or continued).  Iterations
// are enqueued in batches of GF2L_TUNE_GRAPH_ITERS (rounded down to an even
// count, at least 2); the host reads the running count once per batch (every
// iteration when tracing) and stops once no frame runs.  Iterations enqueued
// after the last frame retired find an empty list and do nothing.  The plain
// iterations of the final phase after its first are replayed from a
// captured graph (graphs: not when tracing or profiling); the rest are
// launched one by one.  Allocates nothing when graphs is false: ws.h_count
// may then point at pageable memory.
int run_hbm(gf2l_ctx *ctx, const CodeView &cv, int path, Workspace &ws, const float *d_llr, uint32_t F,
            uint32_t max_it, const float *d_w, uint32_t fb_it, bool fb_continue, uint64_t *d_msg,
            uint32_t *d_iters, uint8_t *d_conv, uint8_t *d_synd, uint8_t *d_fb, float *h_trace_c2v,
            float *h_trace_bel, uint32_t trace_iters, float *d_tc, float *d_tb, bool graphs) {
    hipStream_t s = ctx->stream;
    const size_t E = (size_t)cv.nbe * cv.z;
    gf2l::Profile &pf = ctx->prof;
    GF2L_HIP(ctx, gf2l::launch_hbm_init(cv, ws, path, d_llr, F, s));
    if (cv.dvb_q) {  // the kernels read LLRs in the code's quasi-cyclic order
        GF2L_HIP(ctx, gf2l::launch_dvb_permute(cv, d_llr, ws.llr_perm, F, s));
        d_llr = ws.llr_perm;
    }
    const uint32_t batch = h_trace_c2v ? 1u : (uint32_t)std::max(1, std::min(ctx->tune.graph_iters, kProfileSlots));
    const uint32_t unroll = std::max(2u, (uint32_t)std::min(ctx->tune.graph_iters, kProfileSlots) & ~1u);
    const bool ph = pf.on && pf.all;  // time these launches
    const bool use_graph = graphs && ctx->tune.hbm_graph && !h_trace_c2v && !ph;
    uint32_t running = F;  // frames entering the first iteration of the current batch
    uint32_t queued = 0;   // iterations enqueued since the last readback
    // per-kernel times of the iterations of one batch (stream synchronized)
    auto collect = [&](uint32_t iters) -> int {
        for (uint32_t b = 0; b < iters; ++b)
            for (int k = 0; k < 3; ++k) {
                float ms = 0.0f;
                GF2L_HIP(ctx, hipEventElapsedTime(&ms, pf.ev[4 * b + k], pf.ev[4 * b + k + 1]));
                pf.ms[k] += ms;
                pf.launches[k] += 1;
            }
        return GF2L_OK;
    };
    const uint32_t n_phase = fb_it ? 2 : 1;
    const uint32_t last_it = max_it + (fb_it ? fb_it : 0);
    uint32_t g = 0;  // iterations launched so far: parity of the running list
    for (uint32_t phase = 0; phase < n_phase; ++phase) {
        const uint32_t n_it = phase == 0 ? max_it : fb_it;
        const float *w = phase == 0 ? d_w : nullptr;
        if (phase == 1 && path == gf2l::kStream && !fb_continue) {  // restart the survivors from their LLRs
            GF2L_HIP(ctx, hipMemcpyAsync(ws.belief, d_llr, (size_t)F * cv.n * sizeof(float),
                                         hipMemcpyDeviceToDevice, s));
            GF2L_HIP(ctx, hipMemsetAsync(ws.c2v, 0, (size_t)F * E * sizeof(float), s));
        }
        // a continued fallback never runs the first-iteration kernels, and its
        // first check update subtracts the last weighted messages
        const bool cont = phase == 1 && fb_continue;
        const bool final_phase = phase + 1 == n_phase;
        uint32_t it = 1;
        while (it <= n_it) {
            gf2l::IterArgs a{};
            a.llr = d_llr;
            a.frames = F;
            a.it = cont ? max_it + it : it;
            a.cur = g & 1u;
            a.it_report = (phase == 0 ? 0 : max_it) + it;
            a.last = it == n_it && final_phase;
            a.fb_phase = (int)phase;
            a.w_prev = (w && it >= 2) ? w + (size_t)(it - 2) * cv.nbe : nullptr;
            a.w_cur = w ? w + (size_t)(it - 1) * cv.nbe : nullptr;
            if (cont && it == 1) a.w_prev = d_w + (size_t)(max_it - 1) * cv.nbe;
            a.spa = ctx->rule == GF2L_RULE_SPA && path == gf2l::kStream && !w;
            a.msg = d_msg;
            a.iters = d_iters;
            a.conv = d_conv;
            a.synd = d_synd;
            a.fb = d_fb;
            a.itc = ws.counts + kItcWord;
            a.itc_last = last_it;
            a.itc_fb_from = fb_it ? max_it : last_it;
            if (use_graph && final_phase && it >= 2 && !w && queued == 0) {
                // unroll iterations replayed from the graph (those past the
                // phase's end find no frame running: the last one retired all)
                a.itc_graph = 1;
                int rc = hbm_graph(ctx, cv, path, ws, a, unroll, a.cur);
                if (rc != GF2L_OK) return rc;
                GF2L_HIP(ctx, hipGraphLaunch(ctx->hbm_exec, s));
                g += unroll;
                it += unroll;
                GF2L_HIP(ctx, hipMemcpyAsync(ws.h_count, ws.counts + (g & 1u), sizeof(uint32_t),
                                             hipMemcpyDeviceToHost, s));
                GF2L_HIP(ctx, hipStreamSynchronize(s));
                running = ws.h_count[0];
                if (running == 0) return GF2L_OK;
                continue;
            }
            GF2L_HIP(ctx, gf2l::launch_hbm_iteration(cv, ws, path, a, ctx->num_cus, ctx->tune, s,
                                                     ph ? pf.ev + 4 * queued : nullptr));
            if (phase == 0 && h_trace_c2v && it <= trace_iters) {
                const float *src_c = ws.c2v, *src_b = ws.belief;
                if (path == gf2l::kState) {
                    GF2L_HIP(ctx, gf2l::launch_state_expand(cv, ws, F, a.w_cur, d_tc, d_tb, s));
                    src_c = d_tc;
                    src_b = d_tb;
                }
                GF2L_HIP(ctx, hipMemcpyAsync(h_trace_c2v + (size_t)(it - 1) * F * E, src_c,
                                             (size_t)F * E * sizeof(float), hipMemcpyDeviceToHost, s));
                GF2L_HIP(ctx, hipMemcpyAsync(h_trace_bel + (size_t)(it - 1) * F * cv.n, src_b,
                                             (size_t)F * cv.n * sizeof(float), hipMemcpyDeviceToHost, s));
            }
            // count of frames still running after this iteration: every
            // iteration's when profiling (frames processed per launch), else
            // the batch's last one, read back to stop launching early; the
            // first iteration of a phase that continues in a graph is read
            // back too (the graph starts a batch)
            const bool to_graph = use_graph && final_phase && !w && it + 1 <= n_it;
            const bool readback = queued + 1 >= batch || (it == n_it && final_phase) || to_graph;
            if (ph || readback)
                GF2L_HIP(ctx, hipMemcpyAsync(ws.h_count + (ph ? queued : 0), ws.counts + ((g + 1) & 1u),
                                             sizeof(uint32_t), hipMemcpyDeviceToHost, s));
            ++queued;
            ++g;
            ++it;
            if (!readback) continue;
            GF2L_HIP(ctx, hipStreamSynchronize(s));
            if (ph) {
                const int rc = collect(queued);
                if (rc != GF2L_OK) return rc;
                // frames entering each iteration of the batch
                pf.frame_iters += running;
                for (uint32_t b = 0; b + 1 < queued; ++b) pf.frame_iters += ws.h_count[b];
            }
            running = ws.h_count[ph ? queued - 1 : 0];
            queued = 0;
            if (running == 0) {
                const uint32_t done = it - 1;  // the iteration that retired the last frame
                if (phase == 0 && h_trace_c2v)
                    for (uint32_t t = done + 1; t <= trace_iters; ++t) {
                        std::memcpy(h_trace_c2v + (size_t)(t - 1) * F * E,
                                    h_trace_c2v + (size_t)(done - 1) * F * E, (size_t)F * E * sizeof(float));
                        std::memcpy(h_trace_bel + (size_t)(t - 1) * F * cv.n,
                                    h_trace_bel + (size_t)(done - 1) * F * cv.n,
                                    (size_t)F * cv.n * sizeof(float));
                    }
                return GF2L_OK;
            }
        }
    }
    return GF2L_OK;
}

}  // namespace gf2l

namespace {

using gf2l::decode_device;

int ensure_io(gf2l_ctx *ctx, size_t bytes) {
    if (ctx->d_io_bytes >= bytes) return GF2L_OK;
    hipFree(ctx->d_io);
    ctx->d_io = nullptr;
    ctx->d_io_bytes = 0;
    GF2L_HIP(ctx, hipMalloc(&ctx->d_io, bytes));
    ctx->d_io_bytes = bytes;
    return GF2L_OK;
}

int decode_host(gf2l_ctx *ctx, const gf2l_code *code, const float *llr, size_t frames,
                uint32_t max_it, uint64_t *msg, uint32_t *iters, uint8_t *conv, uint8_t *synd,
                float *trace_c2v, float *trace_bel, uint32_t trace_iters) {
    if (!ctx) return GF2L_ERR_INVALID;
    if (!code) return fail(ctx, GF2L_ERR_INVALID, "null code");
    if (frames == 0) return GF2L_OK;
    if (!llr || !msg || !iters || !conv || !synd)
        return fail(ctx, GF2L_ERR_INVALID, "null buffer");
    GF2L_HIP(ctx, hipSetDevice(ctx->device));
    const gf2l::CodeView &cv = code->view;
    const size_t kw = ((size_t)cv.k + 63) / 64;
    auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t b_llr = align(frames * cv.n * sizeof(float)), b_msg = align(frames * kw * 8),
                 b_it = align(frames * 4), b_c = align(frames);
    int rc = ensure_io(ctx, b_llr + b_msg + b_it + 2 * b_c);
    if (rc != GF2L_OK) return rc;
    char *base = static_cast<char *>(ctx->d_io);
    float *d_llr = reinterpret_cast<float *>(base);
    uint64_t *d_msg = reinterpret_cast<uint64_t *>(base + b_llr);
    uint32_t *d_it = reinterpret_cast<uint32_t *>(base + b_llr + b_msg);
    uint8_t *d_conv = reinterpret_cast<uint8_t *>(base + b_llr + b_msg + b_it);
    uint8_t *d_synd = d_conv + b_c;
    hipStream_t s = ctx->stream;
    GF2L_HIP(ctx, hipMemcpyAsync(d_llr, llr, frames * cv.n * sizeof(float), hipMemcpyHostToDevice, s));
    rc = decode_device(ctx, code, d_llr, frames, max_it, nullptr, 0, false, d_msg, d_it, d_conv, d_synd,
                       nullptr, trace_c2v, trace_bel, trace_iters);
    if (rc != GF2L_OK) return rc;
    GF2L_HIP(ctx, hipMemcpyAsync(msg, d_msg, frames * kw * 8, hipMemcpyDeviceToHost, s));
    GF2L_HIP(ctx, hipMemcpyAsync(iters, d_it, frames * 4, hipMemcpyDeviceToHost, s));
    GF2L_HIP(ctx, hipMemcpyAsync(conv, d_conv, frames, hipMemcpyDeviceToHost, s));
    GF2L_HIP(ctx, hipMemcpyAsync(synd, d_synd, frames, hipMemcpyDeviceToHost, s));
    GF2L_HIP(ctx, hipStreamSynchronize(s));
    return GF2L_OK;
}

}  // namespace

namespace gf2l {

int build_qc_view(gf2l_ctx *ctx, gf2l_code *c, uint32_t mb, uint32_t nb, uint32_t z, uint32_t dvb_q) {
    uint32_t max_dc = 0, max_dv = 0;
    for (uint32_t br = 0; br < mb; ++br) max_dc = std::max(max_dc, c->h_row_ptr[br + 1] - c->h_row_ptr[br]);
    // column order: block-rows ascending (the CSC order of every lifted
    // variable); DVB: standard check index br + q*((z - shift) mod z) of the
    // column's lane 0 ascending
    c->h_col_ptr.assign(nb + 1, 0);
    std::vector<std::vector<std::pair<uint64_t, uint4>>> cols(nb);
    for (uint32_t br = 0; br < mb; ++br)
        for (uint32_t e = c->h_row_ptr[br]; e < c->h_row_ptr[br + 1]; ++e) {
            const uint2 re = c->h_row_edge[e];
            const uint64_t key = dvb_q ? br + (uint64_t)dvb_q * ((z - re.y) % z) : ((uint64_t)br << 32) | e;
            cols[re.x].push_back({key, make_uint4(br, e - c->h_row_ptr[br], re.y, e)});
        }
    for (uint32_t bc = 0; bc < nb; ++bc) {
        std::stable_sort(cols[bc].begin(), cols[bc].end(),
                         [](const std::pair<uint64_t, uint4> &a, const std::pair<uint64_t, uint4> &b) {
                             return a.first < b.first;
                         });
        for (const auto &kv : cols[bc]) {
            c->h_col_edge.push_back(make_uint2(kv.second.w, kv.second.z));
            c->h_col_info.push_back(kv.second);
        }
        c->h_col_ptr[bc + 1] = (uint32_t)c->h_col_edge.size();
        max_dv = std::max(max_dv, c->h_col_ptr[bc + 1] - c->h_col_ptr[bc]);
    }
    if (max_dc > gf2l::kMaxCheckDegree) {
        delete c;
        return fail(ctx, GF2L_ERR_UNSUPPORTED, "check degree above 32");
    }
    const uint32_t nbe = (uint32_t)c->h_row_edge.size();
    const uint32_t zp = ((z + 31) / 32) * 32;
    std::vector<uint2> col_pack(nbe);  // {first state record of the block-row, shift << 5 | position}
    for (uint32_t q = 0; q < nbe; ++q) {
        const uint4 ci = c->h_col_info[q];
        col_pack[q] = make_uint2(ci.x * zp, (ci.z << 5) | ci.y);
    }
    const size_t b_rp = (mb + 1) * 4, b_cp = (nb + 1) * 4, b_e = (size_t)std::max(nbe, 1u) * 8,
                 b_ci = (size_t)std::max(nbe, 1u) * 16;
    auto al = [](size_t x) { return (x + 15) & ~(size_t)15; };
    const size_t o_re = al(b_rp), o_cp = al(o_re + b_e), o_ce = al(o_cp + b_cp), o_ci = al(o_ce + b_e),
                 o_pk = al(o_ci + b_ci);
    const size_t total = o_pk + b_e;
    std::vector<char> host(total, 0);
    std::memcpy(host.data(), c->h_row_ptr.data(), b_rp);
    if (nbe) std::memcpy(host.data() + o_re, c->h_row_edge.data(), nbe * 8);
    std::memcpy(host.data() + o_cp, c->h_col_ptr.data(), b_cp);
    if (nbe) std::memcpy(host.data() + o_ce, c->h_col_edge.data(), nbe * 8);
    if (nbe) std::memcpy(host.data() + o_ci, c->h_col_info.data(), nbe * 16);
    if (nbe) std::memcpy(host.data() + o_pk, col_pack.data(), nbe * 8);
    hipError_t e = hipMalloc(&c->d_tables, total);
    if (e == hipSuccess) e = hipMemcpy(c->d_tables, host.data(), total, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        hipFree(c->d_tables);
        delete c;
        return fail(ctx, GF2L_ERR_HIP, std::string("code upload: ") + hipGetErrorString(e));
    }
    char *d = static_cast<char *>(c->d_tables);
    gf2l::CodeView &v = c->view;
    v.mb = mb;
    v.nb = nb;
    v.z = z;
    v.n = nb * z;
    v.m = mb * z;
    v.k = v.n > v.m ? v.n - v.m : 0;
    v.nbe = nbe;
    v.max_dc = max_dc;
    v.max_dv = max_dv;
    v.zw = (z + 63) / 64;
    v.row_ptr = reinterpret_cast<const uint32_t *>(d);
    v.row_edge = reinterpret_cast<const uint2 *>(d + o_re);
    v.col_ptr = reinterpret_cast<const uint32_t *>(d + o_cp);
    v.col_edge = reinterpret_cast<const uint2 *>(d + o_ce);
    v.col_info = reinterpret_cast<const uint4 *>(d + o_ci);
    v.col_pack = reinterpret_cast<const uint2 *>(d + o_pk);
    v.zp = zp;
    if (dvb_q) {
        v.dvb_q = dvb_q;
        v.dvb_kb = nb - mb;
        v.wrap_bc = nb - 1;
    }
    return GF2L_OK;
}

}  // namespace gf2l

extern "C" {

int gf2l_create(int device, gf2l_ctx **out) {
    if (!out) return GF2L_ERR_INVALID;
    *out = nullptr;
    return gf2l::guard(nullptr, [&]() -> int {
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
            return GF2L_ERR_INVALID;
        if (hipSetDevice(device) != hipSuccess) return GF2L_ERR_HIP;
        gf2l_ctx *ctx = new gf2l_ctx;
        ctx->device = device;
        if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
            delete ctx;
            return GF2L_ERR_HIP;
        }
        int cus = 0, lds = 0;
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
        hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device);
        ctx->num_cus = cus > 0 ? cus : 256;
        ctx->lds_limit = lds > 0 ? (size_t)lds : 65536;
        *out = ctx;
        return GF2L_OK;
    });
}

void gf2l_destroy(gf2l_ctx *ctx) {
    if (!ctx) return;
    hipSetDevice(ctx->device);
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    if (ctx->hbm_exec) hipGraphExecDestroy(ctx->hbm_exec);
    free_workspace(ctx->ws);
    for (auto &e : ctx->prof.ev)
        if (e) hipEventDestroy(e);
    for (auto &e : ctx->prof.lev)
        if (e) hipEventDestroy(e);
    hipFree(ctx->d_io);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *gf2l_last_error(const gf2l_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int gf2l_synchronize(gf2l_ctx *ctx) {
    if (!ctx) return GF2L_ERR_INVALID;
    return gf2l::guard(ctx, [&]() -> int {
        GF2L_HIP(ctx, hipSetDevice(ctx->device));
        GF2L_HIP(ctx, hipStreamSynchronize(ctx->stream));
        return GF2L_OK;
    });
}

void *gf2l_stream(gf2l_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int gf2l_profile_enable(gf2l_ctx *ctx, int on) {
    if (!ctx) return GF2L_ERR_INVALID;
    return gf2l::guard(ctx, [&]() -> int {
        GF2L_HIP(ctx, hipSetDevice(ctx->device));
        gf2l::Profile &pf = ctx->prof;
        if (on && !pf.ev[0]) {
            for (auto &e : pf.ev) GF2L_HIP(ctx, hipEventCreate(&e));
            for (auto &e : pf.lev) GF2L_HIP(ctx, hipEventCreate(&e));
        }
        for (int k = 0; k < 4; ++k) {
            pf.ms[k] = 0;
            pf.launches[k] = 0;
        }
        pf.frame_iters = 0;
        pf.layered_frame_iters = 0;
        pf.on = on != 0;
        pf.all = on != GF2L_PROFILE_LAYERED;
        return GF2L_OK;
    });
}

int gf2l_profile_read(gf2l_ctx *ctx, double *ms, uint64_t *launches, uint64_t *frame_iterations) {
    if (!ctx) return GF2L_ERR_INVALID;
    for (int k = 0; k < 3; ++k) {
        if (ms) ms[k] = ctx->prof.ms[k];
        if (launches) launches[k] = ctx->prof.launches[k];
    }
    if (frame_iterations) *frame_iterations = ctx->prof.frame_iters;
    return GF2L_OK;
}

int gf2l_profile_read_kernel(gf2l_ctx *ctx, int kernel, double *ms, uint64_t *launches,
                             uint64_t *frame_iterations) {
    if (!ctx || kernel < 0 || kernel > GF2L_KERNEL_LAYERED) return GF2L_ERR_INVALID;
    const gf2l::Profile &pf = ctx->prof;
    if (ms) *ms = pf.ms[kernel];
    if (launches) *launches = pf.launches[kernel];
    if (frame_iterations) *frame_iterations = kernel == GF2L_KERNEL_LAYERED ? pf.layered_frame_iters : pf.frame_iters;
    return GF2L_OK;
}

int gf2l_set_path(gf2l_ctx *ctx, int path) {
    if (!ctx) return GF2L_ERR_INVALID;
    return gf2l::guard(ctx, [&]() -> int {
        if (path < GF2L_PATH_AUTO || path > GF2L_PATH_STATE)
            return fail(ctx, GF2L_ERR_INVALID, "unknown decoder path");
        ctx->path = path;
        return GF2L_OK;
    });
}

int gf2l_code_path(gf2l_ctx *ctx, const gf2l_code *code, int *path) {
    if (!ctx) return GF2L_ERR_INVALID;
    return gf2l::guard(ctx, [&]() -> int {
        if (!code || !path) return fail(ctx, GF2L_ERR_INVALID, "null argument");
        if (code->is_csr) {
            if (ctx->rule != GF2L_RULE_MINSUM)
                return fail(ctx, GF2L_ERR_UNSUPPORTED, "the sum-product rule needs a quasi-cyclic code");
            *path = GF2L_PATH_CSR;
            return GF2L_OK;
        }
        int p = gf2l::resolve_path(ctx, code->view);
        if (ctx->rule == GF2L_RULE_SPA) {  // as decode_device: the STREAM path only
            if (code->view.dvb_q || (ctx->path != gf2l::kAuto && ctx->path != gf2l::kStream))
                return fail(ctx, GF2L_ERR_UNSUPPORTED, "the sum-product rule needs the STREAM path");
            p = gf2l::kStream;
        }
        if (p < 0) return fail(ctx, GF2L_ERR_UNSUPPORTED, "the forced decoder path cannot run this code");
        *path = p;
        return GF2L_OK;
    });
}

int gf2l_set_check_rule(gf2l_ctx *ctx, int rule) {
    if (!ctx) return GF2L_ERR_INVALID;
    return gf2l::guard(ctx, [&]() -> int {
        if (rule != GF2L_RULE_MINSUM && rule != GF2L_RULE_SPA) return fail(ctx, GF2L_ERR_INVALID, "unknown check rule");
        ctx->rule = rule;
        return GF2L_OK;
    });
}

int gf2l_device_alloc(gf2l_ctx *ctx, size_t bytes, void **out) {
    if (!ctx || !out) return GF2L_ERR_INVALID;
    *out = nullptr;
    return gf2l::guard(ctx, [&]() -> int {
        GF2L_HIP(ctx, hipSetDevice(ctx->device));
        GF2L_HIP(ctx, hipMalloc(out, bytes ? bytes : 1));
        return GF2L_OK;
    });
}

int gf2l_device_free(gf2l_ctx *ctx, void *ptr) {
    if (!ctx) return GF2L_ERR_INVALID;
    return gf2l::guard(ctx, [&]() -> int {
        GF2L_HIP(ctx, hipSetDevice(ctx->device));
        GF2L_HIP(ctx, hipStreamSynchronize(ctx->stream));
        GF2L_HIP(ctx, hipFree(ptr));
        return GF2L_OK;
    });
}

int gf2l_copy_to_device(gf2l_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx) return GF2L_ERR_INVALID;
    return gf2l::guard(ctx, [&]() -> int {
        GF2L_HIP(ctx, hipSetDevice(ctx->device));
        GF2L_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        GF2L_HIP(ctx, hipStreamSynchronize(ctx->stream));
        return GF2L_OK;
    });
}

int gf2l_copy_to_host(gf2l_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx) return GF2L_ERR_INVALID;
    return gf2l::guard(ctx, [&]() -> int {
        GF2L_HIP(ctx, hipSetDevice(ctx->device));
        GF2L_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        GF2L_HIP(ctx, hipStreamSynchronize(ctx->stream));
        return GF2L_OK;
    });
}

// True when the code is a staircase QC-IRA code: block-column nb-mb+j holds
// identity blocks exactly at block-rows j and j+1.
static bool is_staircase_ira(const gf2l_code *c) {
    const uint32_t mb = c->view.mb, nb = c->view.nb;
    if (nb <= mb || c->shifts.size() != (size_t)mb * nb) return false;
    const uint32_t kb = nb - mb;
    for (uint32_t br = 0; br < mb; ++br)
        for (uint32_t j = 0; j < mb; ++j) {
            const int32_t s = c->shifts[(size_t)br * nb + kb + j];
            const bool want = br == j || br == j + 1;
            if (want ? s != 0 : s != -1) return false;
        }
    return true;
}

int gf2l_frames_generate(gf2l_ctx *ctx, const gf2l_code *code, uint64_t seed, uint64_t first_frame,
                         size_t frames, float ebn0_lo_db, float ebn0_hi_db, float *d_llr,
                         uint64_t *d_codeword_words) {
    if (!ctx) return GF2L_ERR_INVALID;
    return gf2l::guard(ctx, [&]() -> int {
        if (!code || (!d_llr && frames)) return fail(ctx, GF2L_ERR_INVALID, "null argument");
        if (frames > 0xffffffffull) return fail(ctx, GF2L_ERR_INVALID, "too many frames");
        if (!is_staircase_ira(code) || code->view.z < 64)
            return fail(ctx, GF2L_ERR_UNSUPPORTED,
                        "frame generation needs a staircase QC-IRA code with lift >= 64");
        GF2L_HIP(ctx, hipSetDevice(ctx->device));
        GF2L_HIP(ctx, gf2l::launch_framegen(code->view, seed, first_frame, (uint32_t)frames, ebn0_lo_db,
                                            ebn0_hi_db, d_llr, d_codeword_words, ctx->num_cus,
                                            ctx->stream));
        return GF2L_OK;
    });
}

int gf2l_frames_generate_zero(gf2l_ctx *ctx, const gf2l_code *code, uint64_t seed, uint64_t first_frame,
                              size_t frames, float ebn0_lo_db, float ebn0_hi_db, float *d_llr) {
    if (!ctx) return GF2L_ERR_INVALID;
    return gf2l::guard(ctx, [&]() -> int {
        if (!code || (!d_llr && frames)) return fail(ctx, GF2L_ERR_INVALID, "null argument");
        if (frames > 0xffffffffull) return fail(ctx, GF2L_ERR_INVALID, "too many frames");
        const uint32_t n = code->view.n, k = code->view.k;
        if (n == 0 || k == 0) return fail(ctx, GF2L_ERR_INVALID, "code rate must be positive");
        GF2L_HIP(ctx, hipSetDevice(ctx->device));
        GF2L_HIP(ctx, gf2l::launch_framegen_zero(n, k, seed, first_frame, (uint32_t)frames, ebn0_lo_db, ebn0_hi_db,
                                                 d_llr, ctx->num_cus, ctx->stream));
        return GF2L_OK;
    });
}

int gf2l_set_tuning(gf2l_ctx *ctx, int key, int value) {
    if (!ctx) return GF2L_ERR_INVALID;
    return gf2l::guard(ctx, [&]() -> int {
        switch (key) {
            case GF2L_TUNE_WG_PER_CU:
                if (value < 0 || value > 32) return fail(ctx, GF2L_ERR_INVALID, "wg_per_cu out of [0,32]");
                ctx->tune.wg_per_cu = value;
                return GF2L_OK;
            case GF2L_TUNE_XCD_MAP:
                ctx->tune.xcd_map = value != 0;
                return GF2L_OK;
            case GF2L_TUNE_GRAPH_ITERS:
                if (value < 1 || value > 64) return fail(ctx, GF2L_ERR_INVALID, "graph_iters out of [1,64]");
                ctx->tune.graph_iters = value;
                return GF2L_OK;
            case GF2L_TUNE_CHUNK:
                if (value < -1 || value > 64) return fail(ctx, GF2L_ERR_INVALID, "chunk out of [-1,64]");
                ctx->tune.chunk = value;
                return GF2L_OK;
            case GF2L_TUNE_LAYERED_COLUMNS:
                if (value < 0) return fail(ctx, GF2L_ERR_INVALID, "layered columns must be >= 0");
                ctx->tune.layered_cols = value;
                return GF2L_OK;
            case GF2L_TUNE_LAYERED_KERNEL:
                if (value < 0 || value > 2) return fail(ctx, GF2L_ERR_INVALID, "layered kernel out of [0,2]");
                ctx->tune.layered_kernel = value;
                return GF2L_OK;
            case GF2L_TUNE_HBM_GRAPH:
                if (value < 0 || value > 1) return fail(ctx, GF2L_ERR_INVALID, "hbm graph out of [0,1]");
                ctx->tune.hbm_graph = value;
                return GF2L_OK;
            case GF2L_TUNE_LDS_GROUPS:
                if (value < 0 || value > 512) return fail(ctx, GF2L_ERR_INVALID, "lds groups out of [0,512]");
                ctx->tune.lds_groups = value;
                return GF2L_OK;
            default:
                return fail(ctx, GF2L_ERR_INVALID, "unknown tuning key");
        }
    });
}

int gf2l_code_qc(gf2l_ctx *ctx, uint32_t mb, uint32_t nb, uint32_t z, const int32_t *shifts,
                 gf2l_code **out) {
    if (!ctx || !out) return GF2L_ERR_INVALID;
    *out = nullptr;
    return gf2l::guard(ctx, [&]() -> int {
        // QuasiCyclicLdpc::new panics (core.rs:519-540)
        if (mb == 0) return fail(ctx, GF2L_ERR_INVALID, "Base matrix must have at least one row");
        if (z == 0) return fail(ctx, GF2L_ERR_INVALID, "Expansion factor must be positive");
        if (nb == 0 || !shifts) return fail(ctx, GF2L_ERR_INVALID, "empty base matrix");
        if ((uint64_t)mb * z > (1ull << 28) || (uint64_t)nb * z > (1ull << 28))
            return fail(ctx, GF2L_ERR_UNSUPPORTED, "code too large (more than 2^28 checks or variables)");
        if (z > (1u << 26)) return fail(ctx, GF2L_ERR_UNSUPPORTED, "lift above 2^26");
        for (uint32_t br = 0; br < mb; ++br)
            for (uint32_t bc = 0; bc < nb; ++bc) {
                const int32_t s = shifts[(size_t)br * nb + bc];
                if (s != -1 && (s < 0 || (uint32_t)s >= z))  // core.rs:589-596
                    return fail(ctx, GF2L_ERR_INVALID,
                                "Shift value " + std::to_string(s) + " at position (" + std::to_string(br) +
                                    "," + std::to_string(bc) + ") must be -1 or in range [0," +
                                    std::to_string(z) + ")");
            }
        GF2L_HIP(ctx, hipSetDevice(ctx->device));
        gf2l_code *c = new gf2l_code;
        c->device = ctx->device;
        c->shifts.assign(shifts, shifts + (size_t)mb * nb);
        c->h_row_ptr.assign(mb + 1, 0);
        for (uint32_t br = 0; br < mb; ++br) {
            for (uint32_t bc = 0; bc < nb; ++bc) {
                const int32_t s = shifts[(size_t)br * nb + bc];
                if (s >= 0) c->h_row_edge.push_back(make_uint2(bc, (uint32_t)s));
            }
            c->h_row_ptr[br + 1] = (uint32_t)c->h_row_edge.size();
        }
        const int rc = gf2l::build_qc_view(ctx, c, mb, nb, z, 0);
        if (rc != GF2L_OK) return rc;
        *out = c;
        return GF2L_OK;
    });
}

int gf2l_code_info(const gf2l_code *code, uint32_t *n, uint32_t *m, uint32_t *k, uint64_t *nnz) {
    if (!code) return GF2L_ERR_INVALID;
    if (n) *n = code->view.n;
    if (m) *m = code->view.m;
    if (k) *k = code->view.k;
    if (nnz)
        *nnz = code->is_csr ? (uint64_t)code->csr.nnz
                            : (uint64_t)code->view.nbe * code->view.z - (code->view.dvb_q ? 1u : 0u);
    return GF2L_OK;
}

void gf2l_code_destroy(gf2l_code *code) {
    if (!code) return;
    hipSetDevice(code->device);
    hipFree(code->d_tables);
    delete code;
}

int gf2l_decode_batch(gf2l_ctx *ctx, const gf2l_code *code, const float *llr, size_t frames,
                      uint32_t max_it, uint64_t *msg_words, uint32_t *iterations,
                      uint8_t *converged, uint8_t *syndrome_ok) {
    return gf2l::guard(ctx, [&]() -> int {
        return decode_host(ctx, code, llr, frames, max_it, msg_words, iterations, converged, syndrome_ok,
                           nullptr, nullptr, 0);
    });
}

int gf2l_decode_batch_device(gf2l_ctx *ctx, const gf2l_code *code, const float *d_llr,
                             size_t frames, uint32_t max_it, uint64_t *d_msg_words,
                             uint32_t *d_iterations, uint8_t *d_converged,
                             uint8_t *d_syndrome_ok) {
    if (!ctx) return GF2L_ERR_INVALID;
    return gf2l::guard(ctx, [&]() -> int {
        if (!code) return fail(ctx, GF2L_ERR_INVALID, "null code");
        if (frames == 0) return GF2L_OK;
        if (!d_llr || !d_msg_words || !d_iterations || !d_converged || !d_syndrome_ok)
            return fail(ctx, GF2L_ERR_INVALID, "null buffer");
        GF2L_HIP(ctx, hipSetDevice(ctx->device));
        return decode_device(ctx, code, d_llr, frames, max_it, nullptr, 0, false, d_msg_words, d_iterations,
                             d_converged, d_syndrome_ok, nullptr, nullptr, nullptr, 0);
    });
}

int gf2l_decode_trace(gf2l_ctx *ctx, const gf2l_code *code, const float *llr, size_t frames,
                      uint32_t max_it, uint32_t trace_iters, float *c2v_out, float *belief_out,
                      uint64_t *msg_words, uint32_t *iterations, uint8_t *converged,
                      uint8_t *syndrome_ok) {
    if (!ctx) return GF2L_ERR_INVALID;
    return gf2l::guard(ctx, [&]() -> int {
        if (trace_iters > max_it) return fail(ctx, GF2L_ERR_INVALID, "trace_iters > max_it");
        if (trace_iters && (!c2v_out || !belief_out)) return fail(ctx, GF2L_ERR_INVALID, "null trace");
        return decode_host(ctx, code, llr, frames, max_it, msg_words, iterations, converged, syndrome_ok,
                           trace_iters ? c2v_out : nullptr, trace_iters ? belief_out : nullptr,
                           trace_iters);
    });
}

}  // extern "C"
