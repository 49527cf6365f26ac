// Mixed-code batches: frames of up to 16 quasi-cyclic codes (different
// length, rate and lift) decoded together, every frame with the reference's
// plain min-sum flooding decoder and its own early stop (crates/gf2-coding/
// src/ldpc/core.rs:889-938), bit-exact with decoding each code on its own.
//
// Each iteration is three launches over all running frames of all codes:
// check update, variable update (STATE layout: compressed check state, padded
// beliefs) and syndrome/retirement.  Every value that changes during a decode
// lives on the device -- each frame's completed-iteration count, one running
// list per code (double-buffered), the list parity -- and every value that
// changes between calls is read from a device argument block, so U iterations
// are captured once per code set as a hipGraph and replayed until no frame is
// running; the host reads the running counts back once per replay.
//
// Per frame f (element offsets in FrameInfo):
//   llr     n floats of the caller's packed input (frame after frame)
//   state   [3][mb][zp] u32 (min1, min2, argmin | output signs << 5)
//   belief  [nb][zp] floats
//   hb      [nb][zw] hard-decision words
//   msg     ceil(k/64) words of the caller's packed output
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "device_common.h"
#include "gf2ldpc.h"

namespace gf2l {

constexpr uint32_t kMaxCodes = 16;

struct FrameInfo {
    uint32_t code, pad;
    uint64_t llr, st, bel, hb, msg;
};

// Per-call arguments, uploaded before each decode (the graph reads them).
struct CallArgs {
    const float *llr;
    uint64_t *msg;
    uint32_t *iters;
    uint8_t *conv, *synd;
    uint32_t max_it, frames;
};

// Device pointers of a code set's workspace; fixed until the workspace grows.
struct MixShared {
    const CodeView *codes;
    uint32_t ncodes;
    const CallArgs *args;
    const FrameInfo *fi;
    const uint32_t *list_off;  // [kMaxCodes] start of each code's list
    uint32_t *lists;           // [2][list_cap]
    uint32_t list_cap;
    uint32_t *counts;          // [2][kMaxCodes] running frames per code
    uint32_t *fit;             // completed iterations per frame
    uint32_t *misc;            // [0] list parity, [1] finished syndrome workgroups
    uint32_t *state;
    float *belief;
    uint64_t *hb;
};

}  // namespace gf2l

struct gf2l_codeset {
    std::vector<const gf2l_code *> codes;
    gf2l::CodeView *d_views = nullptr;
    uint32_t max_dc = 0, max_dv = 0, max_z = 0, max_words = 0;
    int device = 0;
    // workspace (one device allocation) and pinned upload / readback staging
    char *d_base = nullptr;
    size_t cap_frames = 0, cap_state = 0, cap_bel = 0, cap_hb = 0;
    char *h_stage = nullptr;
    gf2l::MixShared m{};
    // captured iteration graph and the launch shape it was captured with
    hipGraphExec_t exec = nullptr;
    uint32_t g_grid = 0, g_unroll = 0;
};

namespace gf2l {
namespace {

// Maps item j of a pass with `per` items per frame (mb for check items, nb
// for variable items) to (code, frame, block index); false past the end.
template <bool ROWS>
__device__ __forceinline__ bool locate(const MixShared &m, const uint32_t *cnt, const uint32_t *list,
                                       uint32_t j, uint32_t &c, uint32_t &f, uint32_t &r) {
    for (c = 0; c < m.ncodes; ++c) {
        const uint32_t per = ROWS ? m.codes[c].mb : m.codes[c].nb;
        const uint32_t n = cnt[c] * per;
        if (j < n) {
            f = list[m.list_off[c] + j / per];
            r = j % per;
            return true;
        }
        j -= n;
    }
    return false;
}

template <bool ROWS>
__device__ __forceinline__ uint32_t total_items(const MixShared &m, const uint32_t *cnt) {
    uint32_t t = 0;
    for (uint32_t c = 0; c < m.ncodes; ++c) t += cnt[c] * (ROWS ? m.codes[c].mb : m.codes[c].nb);
    return t;
}

template <int MAXD>
__global__ void __launch_bounds__(512) cn_mix_kernel(MixShared m) {
    const uint32_t cur = m.misc[0];
    const uint32_t *cnt = m.counts + cur * kMaxCodes;
    const uint32_t *list = m.lists + (size_t)cur * m.list_cap;
    const float *llr = m.args->llr;
    const uint32_t items = total_items<true>(m, cnt);
    for (uint32_t j = blockIdx.x; j < items; j += gridDim.x) {
        uint32_t c, f, br;
        if (!locate<true>(m, cnt, list, j, c, f, br)) break;
        const CodeView &cv = m.codes[c];
        const FrameInfo fi = m.fi[f];
        uint32_t *st = m.state + fi.st + (size_t)br * cv.zp;
        if (m.fit[f] == 0)
            cn_state_item<MAXD, true, false>(cv, br, st, st, llr + fi.llr, nullptr);
        else
            cn_state_item<MAXD, false, false>(cv, br, st, st, m.belief + fi.bel, nullptr);
    }
}

template <int MAXD>
__global__ void __launch_bounds__(512) vn_mix_kernel(MixShared m) {
    const uint32_t cur = m.misc[0];
    const uint32_t *cnt = m.counts + cur * kMaxCodes;
    const uint32_t *list = m.lists + (size_t)cur * m.list_cap;
    const float *llr = m.args->llr;
    // the syndrome pass of this iteration appends to the other list
    if (blockIdx.x == 0 && threadIdx.x < kMaxCodes) m.counts[(cur ^ 1u) * kMaxCodes + threadIdx.x] = 0;
    const uint32_t items = total_items<false>(m, cnt);
    for (uint32_t j = blockIdx.x; j < items; j += gridDim.x) {
        uint32_t c, f, bc;
        if (!locate<false>(m, cnt, list, j, c, f, bc)) break;
        const CodeView &cv = m.codes[c];
        const FrameInfo fi = m.fi[f];
        vn_state_item<MAXD, false>(cv, bc, m.state + fi.st, llr + fi.llr + (size_t)bc * cv.z,
                                   m.belief + fi.bel + (size_t)bc * cv.zp, m.hb + fi.hb + (size_t)bc * cv.zw,
                                   nullptr);
    }
}

// One 64-lane workgroup per running frame: syndrome, then retire the frame
// (converged, or at the iteration cap) or append it to its code's next list.
// The last workgroup to finish flips the list parity for the next iteration.
__global__ void __launch_bounds__(64) syn_mix_kernel(MixShared m) {
    extern __shared__ uint64_t sh[];
    const uint32_t cur = m.misc[0];
    const uint32_t *cnt = m.counts + cur * kMaxCodes;
    const uint32_t *list = m.lists + (size_t)cur * m.list_cap;
    uint32_t *next_cnt = m.counts + (cur ^ 1u) * kMaxCodes;
    uint32_t *next_list = m.lists + (size_t)(cur ^ 1u) * m.list_cap;
    const CallArgs a = *m.args;
    uint32_t total = 0;
    for (uint32_t c = 0; c < m.ncodes; ++c) total += cnt[c];
    for (uint32_t item = blockIdx.x; item < total; item += gridDim.x) {
        uint32_t c = 0, j = item;
        while (j >= cnt[c]) j -= cnt[c++];
        const uint32_t f = list[m.list_off[c] + j];
        const CodeView &cv = m.codes[c];
        const FrameInfo fi = m.fi[f];
        const int unsat = frame_unsat(cv, sh, m.hb + fi.hb);
        const uint32_t it = m.fit[f] + 1;
        if (!unsat || it >= a.max_it) {
            frame_message(cv, sh, a.msg + fi.msg);
            if (threadIdx.x == 0) {
                a.iters[f] = it;
                a.conv[f] = (uint8_t)!unsat;
                a.synd[f] = (uint8_t)!unsat;
            }
        } else if (threadIdx.x == 0) {
            m.fit[f] = it;
            next_list[m.list_off[c] + atomicAdd(next_cnt + c, 1u)] = f;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(m.misc + 1, 1u) == gridDim.x - 1) {
            m.misc[1] = 0;
            m.misc[0] = cur ^ 1u;
        }
    }
}

// Start of a call: every frame on its code's list, zero iterations done.
__global__ void mix_init_kernel(MixShared m, const uint32_t *__restrict__ init_list,
                                const uint32_t *__restrict__ init_counts) {
    const uint32_t frames = m.args->frames;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < frames; t += gridDim.x * blockDim.x) {
        m.lists[t] = init_list[t];
        m.fit[t] = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x < kMaxCodes) {
        m.counts[threadIdx.x] = init_counts[threadIdx.x];
        m.counts[kMaxCodes + threadIdx.x] = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        m.misc[0] = 0;
        m.misc[1] = 0;
    }
}

uint64_t align256(uint64_t x) { return (x + 255) & ~(uint64_t)255; }

// Staging layout, the same on the device and in the pinned host mirror:
//   CallArgs | list_off[16] | init_counts[16] | FrameInfo[cap] | init_list[cap]
// followed on the host only by the readback area (counts[2][16], parity).
struct Upload {
    uint64_t args, off, icnt, fi, ilist, bytes, rb, host_bytes;
    explicit Upload(uint64_t cap) {
        args = 0;
        off = align256(sizeof(CallArgs));
        icnt = off + kMaxCodes * 4;
        fi = align256(icnt + kMaxCodes * 4);
        ilist = align256(fi + cap * sizeof(FrameInfo));
        bytes = align256(ilist + cap * 4);
        rb = bytes;
        host_bytes = rb + 256;
    }
};

void free_work(gf2l_codeset *cs) {
    if (cs->exec) hipGraphExecDestroy(cs->exec);
    cs->exec = nullptr;
    if (cs->d_base) hipFree(cs->d_base);
    if (cs->h_stage) hipHostFree(cs->h_stage);
    cs->d_base = nullptr;
    cs->h_stage = nullptr;
    cs->cap_frames = cs->cap_state = cs->cap_bel = cs->cap_hb = 0;
}

// Grows the workspace to hold `frames` frames with the given element totals
// (a quarter of headroom, so batches of varying size rarely regrow it).
int ensure_work(gf2l_ctx *ctx, gf2l_codeset *cs, uint64_t frames, uint64_t st, uint64_t bel, uint64_t hb) {
    if (frames <= cs->cap_frames && st <= cs->cap_state && bel <= cs->cap_bel && hb <= cs->cap_hb)
        return GF2L_OK;
    GF2L_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const uint64_t cf = std::max<uint64_t>(frames + frames / 4, cs->cap_frames);
    const uint64_t cst = std::max<uint64_t>(st + st / 4, cs->cap_state);
    const uint64_t cbe = std::max<uint64_t>(bel + bel / 4, cs->cap_bel);
    const uint64_t chb = std::max<uint64_t>(hb + hb / 4, cs->cap_hb);
    free_work(cs);
    const Upload up(cf);
    const uint64_t b_lists = align256(2 * cf * 4), b_cnt = align256(2 * kMaxCodes * 4),
                   b_fit = align256(cf * 4), b_misc = 256, b_st = align256(cst * 4),
                   b_bel = align256(cbe * 4), b_hb = align256(chb * 8);
    const uint64_t total = up.bytes + b_lists + b_cnt + b_fit + b_misc + b_st + b_bel + b_hb;
    GF2L_HIP(ctx, hipMalloc(&cs->d_base, total));
    const hipError_t e = hipHostMalloc(&cs->h_stage, up.host_bytes, hipHostMallocDefault);
    if (e != hipSuccess) {
        cs->h_stage = nullptr;
        free_work(cs);
        return fail(ctx, GF2L_ERR_NOMEM, std::string("pinned staging: ") + hipGetErrorString(e));
    }
    char *p = cs->d_base;
    MixShared &m = cs->m;
    m.codes = cs->d_views;
    m.ncodes = (uint32_t)cs->codes.size();
    m.args = reinterpret_cast<const CallArgs *>(p + up.args);
    m.list_off = reinterpret_cast<const uint32_t *>(p + up.off);
    m.fi = reinterpret_cast<const FrameInfo *>(p + up.fi);
    p += up.bytes;
    m.lists = reinterpret_cast<uint32_t *>(p), p += b_lists;
    m.list_cap = (uint32_t)cf;
    m.counts = reinterpret_cast<uint32_t *>(p), p += b_cnt;
    m.fit = reinterpret_cast<uint32_t *>(p), p += b_fit;
    m.misc = reinterpret_cast<uint32_t *>(p), p += b_misc;
    m.state = reinterpret_cast<uint32_t *>(p), p += b_st;
    m.belief = reinterpret_cast<float *>(p), p += b_bel;
    m.hb = reinterpret_cast<uint64_t *>(p);
    cs->cap_frames = cf;
    cs->cap_state = cst;
    cs->cap_bel = cbe;
    cs->cap_hb = chb;
    return GF2L_OK;
}

// Captures `unroll` iterations (check, variable, syndrome launches) as a
// graph over the code set's current workspace.
int capture(gf2l_ctx *ctx, gf2l_codeset *cs, uint32_t grid, uint32_t unroll) {
    if (cs->exec && cs->g_grid == grid && cs->g_unroll == unroll) return GF2L_OK;
    if (cs->exec) hipGraphExecDestroy(cs->exec);
    cs->exec = nullptr;
    const uint32_t block = std::min<uint32_t>(512, ((cs->max_z + 63) / 64) * 64);
    const uint32_t g_syn = (uint32_t)ctx->num_cus * 8u;
    const size_t sh = (size_t)cs->max_words * sizeof(uint64_t);
    const MixShared m = cs->m;
    hipStream_t s = ctx->stream;
    hipGraph_t graph = nullptr;
    GF2L_HIP(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    for (uint32_t u = 0; u < unroll; ++u) {
        if (cs->max_dc > 16)
            hipLaunchKernelGGL(cn_mix_kernel<27>, dim3(grid), dim3(block), 0, s, m);
        else
            hipLaunchKernelGGL(cn_mix_kernel<16>, dim3(grid), dim3(block), 0, s, m);
        if (cs->max_dv > 16)
            hipLaunchKernelGGL(vn_mix_kernel<32>, dim3(grid), dim3(block), 0, s, m);
        else
            hipLaunchKernelGGL(vn_mix_kernel<16>, dim3(grid), dim3(block), 0, s, m);
        hipLaunchKernelGGL(syn_mix_kernel, dim3(g_syn), dim3(64), sh, s, m);
    }
    const hipError_t launch = hipGetLastError();
    const hipError_t end = hipStreamEndCapture(s, &graph);
    if (launch != hipSuccess || end != hipSuccess) {
        if (graph) hipGraphDestroy(graph);
        return fail(ctx, GF2L_ERR_HIP,
                    std::string("graph capture: ") + hipGetErrorString(launch != hipSuccess ? launch : end));
    }
    const hipError_t inst = hipGraphInstantiate(&cs->exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (inst != hipSuccess) {
        cs->exec = nullptr;
        return fail(ctx, GF2L_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(inst));
    }
    cs->g_grid = grid;
    cs->g_unroll = unroll;
    return GF2L_OK;
}

}  // namespace
}  // namespace gf2l

using gf2l::fail;

extern "C" {

int gf2l_codeset_create(gf2l_ctx *ctx, const gf2l_code *const *codes, uint32_t count, gf2l_codeset **out) {
    if (!ctx || !out) return GF2L_ERR_INVALID;
    *out = nullptr;
    if (!codes || count == 0 || count > gf2l::kMaxCodes)
        return fail(ctx, GF2L_ERR_INVALID, "a code set holds 1 to 16 codes");
    std::vector<gf2l::CodeView> views;
    uint32_t max_dc = 0, max_dv = 0, max_z = 0, max_words = 0;
    for (uint32_t c = 0; c < count; ++c) {
        const gf2l_code *code = codes[c];
        if (!code) return fail(ctx, GF2L_ERR_INVALID, "null code in a code set");
        if (code->is_csr) return fail(ctx, GF2L_ERR_UNSUPPORTED, "code sets hold quasi-cyclic codes");
        if (code->device != ctx->device) return fail(ctx, GF2L_ERR_INVALID, "code built on another device");
        if (code->view.max_dc > gf2l::kMaxStateDegree)
            return fail(ctx, GF2L_ERR_UNSUPPORTED, "check degree above 27 in a code set");
        views.push_back(code->view);
        max_dc = std::max(max_dc, code->view.max_dc);
        max_dv = std::max(max_dv, code->view.max_dv);
        max_z = std::max(max_z, code->view.z);
        max_words = std::max(max_words, code->view.nb * code->view.zw);
    }
    if ((size_t)max_words * 8 > 65536)
        return fail(ctx, GF2L_ERR_UNSUPPORTED, "hard decisions of a frame exceed 64 KiB of LDS");
    GF2L_HIP(ctx, hipSetDevice(ctx->device));
    gf2l_codeset *cs = new gf2l_codeset;
    cs->codes.assign(codes, codes + count);
    cs->max_dc = max_dc;
    cs->max_dv = max_dv;
    cs->max_z = max_z;
    cs->max_words = max_words;
    cs->device = ctx->device;
    hipError_t e = hipMalloc(&cs->d_views, count * sizeof(gf2l::CodeView));
    if (e == hipSuccess)
        e = hipMemcpy(cs->d_views, views.data(), count * sizeof(gf2l::CodeView), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (cs->d_views) hipFree(cs->d_views);
        delete cs;
        return fail(ctx, GF2L_ERR_HIP, std::string("code set upload: ") + hipGetErrorString(e));
    }
    *out = cs;
    return GF2L_OK;
}

void gf2l_codeset_destroy(gf2l_codeset *cs) {
    if (!cs) return;
    hipSetDevice(cs->device);
    hipDeviceSynchronize();
    gf2l::free_work(cs);
    hipFree(cs->d_views);
    delete cs;
}

int gf2l_decode_mixed_device(gf2l_ctx *ctx, gf2l_codeset *cs, const uint16_t *code_ids, size_t frames,
                             const float *d_llr, uint32_t max_it, uint64_t *d_msg_words,
                             uint32_t *d_iterations, uint8_t *d_converged, uint8_t *d_syndrome_ok) {
    using namespace gf2l;
    if (!ctx) return GF2L_ERR_INVALID;
    if (!cs) return fail(ctx, GF2L_ERR_INVALID, "null code set");
    if (cs->device != ctx->device) return fail(ctx, GF2L_ERR_INVALID, "code set built on another device");
    if (frames == 0) return GF2L_OK;
    if (!code_ids || !d_llr || !d_msg_words || !d_iterations || !d_converged || !d_syndrome_ok)
        return fail(ctx, GF2L_ERR_INVALID, "null buffer");
    if (frames > (1u << 24)) return fail(ctx, GF2L_ERR_INVALID, "more than 2^24 frames in one call");
    GF2L_HIP(ctx, hipSetDevice(ctx->device));
    const uint32_t F = (uint32_t)frames, K = (uint32_t)cs->codes.size();
    hipStream_t s = ctx->stream;
    // per-code frame counts and workspace totals
    uint32_t per_code[kMaxCodes] = {0};
    uint64_t t_st = 0, t_bel = 0, t_hb = 0, t_msg = 0, items_c = 0, items_v = 0;
    for (uint32_t f = 0; f < F; ++f) {
        const uint32_t c = code_ids[f];
        if (c >= K)
            return fail(ctx, GF2L_ERR_INVALID, "code id " + std::to_string(c) + " of frame " + std::to_string(f) +
                                                   " outside the code set");
        const CodeView &v = cs->codes[c]->view;
        t_st += 3ull * v.mb * v.zp;
        t_bel += (uint64_t)v.nb * v.zp;
        t_hb += (uint64_t)v.nb * v.zw;
        t_msg += (v.k + 63) / 64;
        items_c += v.mb;
        items_v += v.nb;
        per_code[c]++;
    }
    if (items_c > 0xffffffffull || items_v > 0xffffffffull)
        return fail(ctx, GF2L_ERR_UNSUPPORTED, "too many block items in one call");
    if (max_it == 0) {  // fresh decoders: zero word, which satisfies every check (core.rs:906-937)
        GF2L_HIP(ctx, hipMemsetAsync(d_msg_words, 0, t_msg * sizeof(uint64_t), s));
        GF2L_HIP(ctx, hipMemsetAsync(d_iterations, 0, (size_t)F * sizeof(uint32_t), s));
        GF2L_HIP(ctx, hipMemsetAsync(d_converged, 0, F, s));
        GF2L_HIP(ctx, hipMemsetAsync(d_syndrome_ok, 1, F, s));
        return GF2L_OK;
    }
    int rc = ensure_work(ctx, cs, F, t_st, t_bel, t_hb);
    if (rc != GF2L_OK) return rc;
    // fill the pinned staging block (the previous call has synchronized)
    const Upload up(cs->cap_frames);
    char *h = cs->h_stage;
    *reinterpret_cast<CallArgs *>(h + up.args) =
        CallArgs{d_llr, d_msg_words, d_iterations, d_converged, d_syndrome_ok, max_it, F};
    uint32_t *h_off = reinterpret_cast<uint32_t *>(h + up.off);
    uint32_t *h_icnt = reinterpret_cast<uint32_t *>(h + up.icnt);
    FrameInfo *h_fi = reinterpret_cast<FrameInfo *>(h + up.fi);
    uint32_t *h_il = reinterpret_cast<uint32_t *>(h + up.ilist);
    uint32_t fill[kMaxCodes];
    for (uint32_t c = 0; c < kMaxCodes; ++c) {
        h_off[c] = c == 0 ? 0 : h_off[c - 1] + per_code[c - 1];
        h_icnt[c] = per_code[c];
        fill[c] = 0;
    }
    uint64_t o_llr = 0, o_st = 0, o_bel = 0, o_hb = 0, o_msg = 0;
    for (uint32_t f = 0; f < F; ++f) {
        const uint32_t c = code_ids[f];
        const CodeView &v = cs->codes[c]->view;
        h_fi[f] = FrameInfo{c, 0, o_llr, o_st, o_bel, o_hb, o_msg};
        h_il[h_off[c] + fill[c]++] = f;
        o_llr += v.n;
        o_st += 3ull * v.mb * v.zp;
        o_bel += (uint64_t)v.nb * v.zp;
        o_hb += (uint64_t)v.nb * v.zw;
        o_msg += (v.k + 63) / 64;
    }
    char *d_up = cs->d_base;
    GF2L_HIP(ctx, hipMemcpyAsync(d_up, h, up.fi + (size_t)F * sizeof(FrameInfo), hipMemcpyHostToDevice, s));
    GF2L_HIP(ctx, hipMemcpyAsync(d_up + up.ilist, h + up.ilist, (size_t)F * 4, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(mix_init_kernel, dim3(std::min<uint32_t>((F + 255) / 256, 1024)), dim3(256), 0, s, cs->m,
                       reinterpret_cast<const uint32_t *>(d_up + up.ilist),
                       reinterpret_cast<const uint32_t *>(d_up + up.icnt));
    GF2L_HIP(ctx, hipGetLastError());
    const uint32_t grid = ((uint32_t)ctx->num_cus * (uint32_t)ctx->tune.wg_per_cu + 7u) & ~7u;
    rc = capture(ctx, cs, grid, (uint32_t)ctx->tune.graph_iters);
    if (rc != GF2L_OK) return rc;
    // replay until every frame has retired (at most ceil(max_it / U) replays)
    uint32_t *h_rb = reinterpret_cast<uint32_t *>(h + up.rb);
    for (uint32_t done_it = 0; done_it < max_it; done_it += cs->g_unroll) {
        GF2L_HIP(ctx, hipGraphLaunch(cs->exec, s));
        GF2L_HIP(ctx, hipMemcpyAsync(h_rb, cs->m.counts, 2 * kMaxCodes * 4, hipMemcpyDeviceToHost, s));
        GF2L_HIP(ctx, hipMemcpyAsync(h_rb + 2 * kMaxCodes, cs->m.misc, 4, hipMemcpyDeviceToHost, s));
        GF2L_HIP(ctx, hipStreamSynchronize(s));
        const uint32_t *cur = h_rb + (h_rb[2 * kMaxCodes] & 1u) * kMaxCodes;
        uint32_t running = 0;
        for (uint32_t c = 0; c < K; ++c) running += cur[c];
        if (running == 0) break;
    }
    return GF2L_OK;
}

}  // extern "C"
