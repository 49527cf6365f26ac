// Mixed-code batches: frames of up to 16 quasi-cyclic codes (different
// length, rate and lift) decoded together, every frame with the reference's
// plain min-sum flooding decoder and its own early stop (crates/gf2-coding/
// src/ldpc/core.rs:889-938), bit-exact with decoding each code on its own.
//
// Layout.  At the start of a call the frames of code c get the slots
// list_off[c] .. list_off[c] + frames_c - 1 (input order within the code), and
// their LLRs are gathered into a per-code workspace segment, so every per-frame
// array is segment base + rank * stride:
//   llr     [frames_c][n] floats
//   state   [frames_c][mb][zp][3] u32 (min1, min2, argmin | output signs << 5)
//   belief  [frames_c][nb][zp] floats
//   hb      [frames_c][nb][zw] hard-decision words
// Results go to the caller's packed arrays (frames back to back in input
// order) through the slot table.
//
// Iterations.  All frames start together, so one device counter holds the
// iteration number of every running frame.  With learned decoders (one weight
// set per code) the first T iterations are the weighted rule, and frames
// still failing the syndrome after T restart from their LLRs together at
// iteration T + 1 with plain min-sum, or continue from the weighted state
// (the single-code learned decoder's two fallbacks, gf2l_nms_decode_device).  An iteration is a check-update
// and a variable-update launch per degree class present in the set
// (registers sized per class) and one syndrome/retirement launch, each over
// the running frames of every code; the running lists (one per code,
// double-buffered) and their counts live on the device.  The first iteration
// (inputs = channel LLRs) is launched directly; later iterations replay a
// hipGraph of U iterations captured once per code set, and the host reads the
// running counts back once per replay.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "device_common.h"
#include "gf2ldpc.h"

namespace gf2l {

constexpr uint32_t kMaxCodes = 16;
// misc[kSchedBase + (8 * slot + xcd) * kCtrStride]: item counters of the
// check (slots 0-2) and variable (slots 3-5) launches of one iteration, one
// per degree class.
constexpr uint32_t kSchedBase = 64;
constexpr uint32_t kSchedCounters = 6 * 8;
constexpr uint32_t kMiscWords = kSchedBase + kSchedCounters * 64;

// Workspace segment of one code for one call (element offsets).
struct CodeSeg {
    uint64_t llr, st, bel, hb;
    uint32_t list_off, frames;
};

// Per-call arguments, uploaded before each decode (the graph reads them).
struct CallArgs {
    const float *llr;  // caller's packed LLRs
    uint64_t *msg;
    uint32_t *iters;
    uint8_t *conv, *synd;
    uint8_t *fb;                  // per frame: re-decoded by the fallback (may be null)
    uint32_t max_total;           // a frame retires at this iteration count
    uint32_t weighted;            // T: weighted iterations (0: plain min-sum only)
    uint32_t frames, pad;
    CodeSeg seg[kMaxCodes];
    const float *w[kMaxCodes];    // per code [T][nbe] weights (weighted decodes)
};

// Caller-side position of the frame in a slot.
struct SlotInfo {
    uint32_t frame, code;
    uint64_t llr, msg;  // element offsets in the caller's packed arrays
};

// Device pointers of a code set's workspace; fixed until the workspace grows.
struct MixShared {
    const CodeView *codes;
    uint32_t ncodes;
    const CallArgs *args;
    const SlotInfo *slots;
    uint32_t *lists;   // [2][list_cap] slots of running frames, code c from list_off[c]
    uint32_t list_cap;
    uint32_t *counts;  // [2][kMaxCodes] running frames per code
    uint32_t *misc;    // [0] list parity, [1] finished syndrome workgroups, [2] iterations done,
                       // item counters from kSchedBase (kMiscWords in all)
    float *llr;
    uint32_t *state;
    float *belief;
    uint64_t *hb;
};

}  // namespace gf2l

struct gf2l_codeset {
    std::vector<const gf2l_code *> codes;
    gf2l::CodeView *d_views = nullptr;
    uint32_t max_dc = 0, max_dv = 0, max_z = 0, max_words = 0;
    uint32_t cn_mask[3] = {0, 0, 0};  // codes per check-degree class (8, 16, 27)
    uint32_t vn_mask[3] = {0, 0, 0};  // codes per variable-degree class (8, 16, 32)
    int device = 0;
    // workspace (one device allocation) and pinned upload / readback staging
    char *d_base = nullptr;
    size_t cap_frames = 0, cap_llr = 0, cap_state = 0, cap_bel = 0, cap_hb = 0;
    char *h_stage = nullptr;
    gf2l::MixShared m{};
    // captured iteration graph and the launch shape it was captured with
    hipGraphExec_t exec = nullptr;
    uint32_t g_grid = 0, g_unroll = 0, g_chunk = 0;
    int g_xcd = -1;
};

namespace gf2l {
namespace {

// The code table, the argument block and the running list a launch reads are
// not written while it runs: they are read with scalar loads (ldc), so the
// per-frame base pointers built from them stay in scalar registers and the
// item bodies use scalar-base addressing as with a kernel-argument CodeView.
__device__ __forceinline__ CodeView uniform_view(const CodeView *g) {
    CodeView v;
    v.mb = ldc(&g->mb);
    v.nb = ldc(&g->nb);
    v.z = ldc(&g->z);
    v.n = ldc(&g->n);
    v.m = ldc(&g->m);
    v.k = ldc(&g->k);
    v.nbe = ldc(&g->nbe);
    v.max_dc = ldc(&g->max_dc);
    v.max_dv = ldc(&g->max_dv);
    v.zw = ldc(&g->zw);
    v.zp = ldc(&g->zp);
    v.row_ptr = ldc(&g->row_ptr);
    v.row_edge = ldc(&g->row_edge);
    v.col_ptr = ldc(&g->col_ptr);
    v.col_edge = ldc(&g->col_edge);
    v.col_info = ldc(&g->col_info);
    v.col_pack = ldc(&g->col_pack);
    return v;
}
__device__ __forceinline__ CodeSeg uniform_seg(const CodeSeg *g) {
    CodeSeg s;
    s.llr = ldc(&g->llr);
    s.st = ldc(&g->st);
    s.bel = ldc(&g->bel);
    s.hb = ldc(&g->hb);
    s.list_off = ldc(&g->list_off);
    s.frames = ldc(&g->frames);
    return s;
}

// Work split of one launch: with xcd != 0 the running frames of each code
// are dealt round-robin to the 8 XCDs (list index x, x+8, ... to XCD
// x = blockIdx % 8) and the workgroups of an XCD stride over their items in
// frame order, as in the single-code kernels (needs gridDim.x % 8 == 0);
// with xcd == 0 the whole grid strides over all items.
struct MixSplit {
    uint32_t x, q, Q, step;
    __device__ explicit MixSplit(int xcd) {
        if (xcd) {
            x = blockIdx.x & 7u;
            q = blockIdx.x >> 3;
            Q = gridDim.x >> 3;
            step = 8u;
        } else {
            x = 0;
            q = blockIdx.x;
            Q = gridDim.x;
            step = 1u;
        }
    }
    __device__ uint32_t share(uint32_t count) const {
        return step == 1u ? count : (count > x ? (count - x + 7u) >> 3 : 0u);
    }
};

enum ItemKind { kRows, kCols, kFrames };

// First item >= start that the workgroup at grid position q (stride Q) owns.
__device__ __forceinline__ uint32_t first_item(uint32_t start, uint32_t q, uint32_t Q) {
    return start <= q ? q : q + (start - q + Q - 1) / Q * Q;
}

// Item prefix over the codes of one launch in LDS: s_start[c] = first item of
// code c; returns the total.  Items per running frame: mb, nb or 1.
template <ItemKind KIND>
__device__ uint32_t build_prefix(const MixShared &m, uint32_t mask, const uint32_t *cnt, const MixSplit &sp,
                                 uint32_t *s_start) {
    if (threadIdx.x < kMaxCodes) {
        const uint32_t c = threadIdx.x;
        uint32_t n = 0;
        if (c < m.ncodes && ((mask >> c) & 1u)) {
            const uint32_t per = KIND == kRows ? m.codes[c].mb : KIND == kCols ? m.codes[c].nb : 1u;
            n = sp.share(cnt[c]) * per;
        }
        s_start[c + 1] = n;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        s_start[0] = 0;
        for (uint32_t c = 0; c < kMaxCodes; ++c) s_start[c + 1] += s_start[c];
    }
    __syncthreads();
    return uni(s_start[kMaxCodes]);
}

// Check update of the running frames of the codes in `mask` (one degree
// class: the kernel's registers are sized for MAXD).  FIRST: inputs are the
// channel LLRs, no previous messages.  NMS: the previous messages were
// weighted by iteration t-1's weights (t = this iteration, 1-based).
template <int MAXD, bool FIRST, bool NMS, bool DYN>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(kCheckWaves<MAXD>))) cn_mix_kernel(MixShared m, uint32_t mask, int xcd, uint32_t t, uint32_t slot,
                                                     uint32_t chunk) {
    __shared__ uint32_t s_start[kMaxCodes + 1];
    const uint32_t cur = ldc(m.misc);
    const uint32_t *cnt = m.counts + cur * kMaxCodes;
    const uint32_t *list = m.lists + (size_t)cur * m.list_cap;
    const MixSplit sp(xcd);
    const uint32_t items = build_prefix<kRows>(m, mask, cnt, sp, s_start);
    auto item = [&](const CodeView &cv, const CodeSeg &seg, const float *w_prev, uint32_t local) {
        const uint32_t br = local % cv.mb;
        const uint32_t slot_f = ldc(list + seg.list_off + sp.x + sp.step * (local / cv.mb));
        const uint64_t r = slot_f - seg.list_off;
        uint32_t *fs = m.state + seg.st + r * (3ull * cv.mb * cv.zp);
        if constexpr (FIRST)
            cn_state_item<MAXD, true, false>(cv, br, fs, fs, m.llr + seg.llr + r * cv.n, nullptr);
        else
            cn_state_item<MAXD, false, NMS>(cv, br, fs, fs, m.belief + seg.bel + r * ((uint64_t)cv.nb * cv.zp),
                                            w_prev);
    };
    if constexpr (DYN) {
        __shared__ uint32_t s_next[2];
        ItemStream it(sp.q, sp.Q, m.misc + kSchedBase + (slot * 8 + sp.x) * kCtrStride, chunk);
        uint32_t c = 0, cc = ~0u;
        CodeView cv{};
        CodeSeg seg{};
        const float *w_prev = nullptr;
        while (it.j < items) {
            while (it.j >= uni(s_start[c + 1])) ++c;
            if (c != cc) {
                cv = uniform_view(m.codes + c);
                seg = uniform_seg(m.args->seg + c);
                if constexpr (NMS) w_prev = ldc(m.args->w + c) + (size_t)(t - 2) * cv.nbe;
                cc = c;
            }
            item(cv, seg, w_prev, it.j - uni(s_start[c]));
            it.advance(s_next);
        }
    } else {
        for (uint32_t c = 0; c < m.ncodes; ++c) {
            const uint32_t start = uni(s_start[c]), end = uni(s_start[c + 1]);
            if (start == end) continue;
            const CodeView cv = uniform_view(m.codes + c);
            const CodeSeg seg = uniform_seg(m.args->seg + c);
            const float *w_prev = nullptr;
            if constexpr (NMS) w_prev = ldc(m.args->w + c) + (size_t)(t - 2) * cv.nbe;
            for (uint32_t j = first_item(start, sp.q, sp.Q); j < end; j += sp.Q) item(cv, seg, w_prev, j - start);
        }
    }
}

// Variable update; NMS: messages weighted by iteration t's weights.
template <int MAXD, bool NMS, bool DYN>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(kVarWaves<MAXD>))) vn_mix_kernel(MixShared m, uint32_t mask, int xcd, uint32_t t, uint32_t slot,
                                                     uint32_t chunk) {
    __shared__ uint32_t s_start[kMaxCodes + 1];
    const uint32_t cur = ldc(m.misc);
    const uint32_t *cnt = m.counts + cur * kMaxCodes;
    const uint32_t *list = m.lists + (size_t)cur * m.list_cap;
    const MixSplit sp(xcd);
    const uint32_t items = build_prefix<kCols>(m, mask, cnt, sp, s_start);
    auto item = [&](const CodeView &cv, const CodeSeg &seg, const float *w_cur, uint32_t local) {
        const uint32_t bc = local % cv.nb;
        const uint32_t slot_f = ldc(list + seg.list_off + sp.x + sp.step * (local / cv.nb));
        const uint64_t r = slot_f - seg.list_off;
        vn_state_item<MAXD, NMS>(cv, bc, m.state + seg.st + r * (3ull * cv.mb * cv.zp),
                                 m.llr + seg.llr + r * cv.n + (size_t)bc * cv.z,
                                 m.belief + seg.bel + r * ((uint64_t)cv.nb * cv.zp) + (size_t)bc * cv.zp,
                                 m.hb + seg.hb + r * ((uint64_t)cv.nb * cv.zw) + (size_t)bc * cv.zw, w_cur);
    };
    if constexpr (DYN) {
        __shared__ uint32_t s_next[2];
        ItemStream it(sp.q, sp.Q, m.misc + kSchedBase + (slot * 8 + sp.x) * kCtrStride, chunk);
        uint32_t c = 0, cc = ~0u;
        CodeView cv{};
        CodeSeg seg{};
        const float *w_cur = nullptr;
        while (it.j < items) {
            while (it.j >= uni(s_start[c + 1])) ++c;
            if (c != cc) {
                cv = uniform_view(m.codes + c);
                seg = uniform_seg(m.args->seg + c);
                if constexpr (NMS) w_cur = ldc(m.args->w + c) + (size_t)(t - 1) * cv.nbe;
                cc = c;
            }
            item(cv, seg, w_cur, it.j - uni(s_start[c]));
            it.advance(s_next);
        }
    } else {
        for (uint32_t c = 0; c < m.ncodes; ++c) {
            const uint32_t start = uni(s_start[c]), end = uni(s_start[c + 1]);
            if (start == end) continue;
            const CodeView cv = uniform_view(m.codes + c);
            const CodeSeg seg = uniform_seg(m.args->seg + c);
            const float *w_cur = nullptr;
            if constexpr (NMS) w_cur = ldc(m.args->w + c) + (size_t)(t - 1) * cv.nbe;
            for (uint32_t j = first_item(start, sp.q, sp.Q); j < end; j += sp.Q) item(cv, seg, w_cur, j - start);
        }
    }
    // The syndrome pass of this iteration writes the other list and
    // mix_compact_kernel stores its counts: every variable launch of the
    // iteration clears them first (the syndrome pass follows them all).  Stored last: a global store ahead of the loop would
    // keep the table loads above from being scalar loads.
    if (blockIdx.x == 0 && threadIdx.x < kMaxCodes) m.counts[(cur ^ 1u) * kMaxCodes + threadIdx.x] = 0;
}

// One 256-thread workgroup per running frame: syndrome, then retire the frame
// (converged, or at the iteration cap) or keep its slot in its code's next list.
// A frame failing at the end of the weighted iterations (it == T < cap)
// stays on the list: the next iteration restarts it from its LLRs or
// continues from its state.
// The last workgroup to finish advances the iteration counter and flips the
// list parity.
__global__ void __launch_bounds__(256) syn_mix_kernel(MixShared m) {
    extern __shared__ uint64_t sh[];
    __shared__ uint32_t s_start[kMaxCodes + 1];
    const uint32_t cur = ldc(m.misc);
    const uint32_t it = ldc(m.misc + 2) + 1;
    const uint32_t *cnt = m.counts + cur * kMaxCodes;
    const uint32_t *list = m.lists + (size_t)cur * m.list_cap;
    uint32_t *next_list = m.lists + (size_t)(cur ^ 1u) * m.list_cap;
    const uint32_t max_total = ldc(&m.args->max_total);
    const uint32_t weighted = ldc(&m.args->weighted);
    const MixSplit sp(0);
    const uint32_t total = build_prefix<kFrames>(m, 0xffffu, cnt, sp, s_start);
    (void)total;
    for (uint32_t c = 0; c < m.ncodes; ++c) {
        const uint32_t start = uni(s_start[c]), end = uni(s_start[c + 1]);
        if (start == end) continue;
        const CodeView cv = uniform_view(m.codes + c);
        const CodeSeg seg = uniform_seg(m.args->seg + c);
        for (uint32_t item = first_item(start, blockIdx.x, gridDim.x); item < end; item += gridDim.x) {
            const uint32_t slot = ldc(list + seg.list_off + item - start);
            const uint64_t r = slot - seg.list_off;
            const int unsat = frame_unsat(cv, sh, m.hb + seg.hb + r * ((uint64_t)cv.nb * cv.zw));
            if (!unsat || it >= max_total) {
                const SlotInfo si = m.slots[slot];
                const CallArgs *a = m.args;
                frame_message(cv, sh, a->msg + si.msg);
                if (threadIdx.x == 0) {
                    a->iters[si.frame] = it;
                    a->conv[si.frame] = (uint8_t)!unsat;
                    a->synd[si.frame] = (uint8_t)!unsat;
                    if (a->fb) a->fb[si.frame] = (uint8_t)(weighted && it > weighted);
                }
            }
            // same position in the next list; mix_compact_kernel closes the
            // holes of retired frames in list order
            if (threadIdx.x == 0)
                next_list[seg.list_off + item - start] = (!unsat || it >= max_total) ? kRetiredSlot : slot;
            __syncthreads();
        }
    }
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(m.misc + 1, 1u) == gridDim.x - 1) {
            m.misc[1] = 0;
            m.misc[2] = it;
            m.misc[0] = cur ^ 1u;
            for (uint32_t k = 0; k < kSchedCounters; ++k) m.misc[kSchedBase + k * kCtrStride] = 0;
        }
    }
}

// After the syndrome pass (which has flipped the list parity): one workgroup
// per code closes the holes of its segment of the list just written, in list
// order, and stores the code's running count.
__global__ void __launch_bounds__(1024) mix_compact_kernel(MixShared m) {
    __shared__ uint32_t wave_kept[16];
    const uint32_t c = blockIdx.x;
    const uint32_t cur = m.misc[0];
    const uint32_t n = m.counts[(cur ^ 1u) * kMaxCodes + c];
    uint32_t *list = m.lists + (size_t)cur * m.list_cap + m.args->seg[c].list_off;
    const uint32_t kept = compact_list(list, n, wave_kept);
    if (threadIdx.x == 0) m.counts[cur * kMaxCodes + c] = kept;
}

// Start of a call: every slot on its code's list, lists and counters reset.
__global__ void mix_init_kernel(MixShared m) {
    const uint32_t frames = m.args->frames;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < frames; t += gridDim.x * blockDim.x)
        m.lists[t] = t;
    if (blockIdx.x == 0 && threadIdx.x < kMaxCodes) {
        m.counts[threadIdx.x] = threadIdx.x < m.ncodes ? m.args->seg[threadIdx.x].frames : 0u;
        m.counts[kMaxCodes + threadIdx.x] = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x < 3) m.misc[threadIdx.x] = 0;
    if (blockIdx.x == 0 && threadIdx.x < kSchedCounters) m.misc[kSchedBase + threadIdx.x * kCtrStride] = 0;
}

// Copies each slot's LLRs from the caller's packed array into its code's
// workspace segment (one workgroup per slot, grid-stride).
__global__ void __launch_bounds__(256) mix_gather_kernel(MixShared m) {
    const uint32_t frames = m.args->frames;
    const float *src = m.args->llr;
    for (uint32_t slot = blockIdx.x; slot < frames; slot += gridDim.x) {
        const SlotInfo si = m.slots[slot];
        const uint32_t n = m.codes[si.code].n;
        const CodeSeg &seg = m.args->seg[si.code];
        const float *s = src + si.llr;
        float *d = m.llr + seg.llr + (uint64_t)(slot - seg.list_off) * n;
        for (uint32_t v = threadIdx.x; v < n; v += blockDim.x) d[v] = s[v];
    }
}

uint64_t align256(uint64_t x) { return (x + 255) & ~(uint64_t)255; }

// Staging layout, the same on the device and in the pinned host mirror:
//   CallArgs | SlotInfo[cap]
// followed on the host only by the readback area (counts[2][16], parity).
struct Upload {
    uint64_t args, slots, bytes, rb, host_bytes;
    explicit Upload(uint64_t cap) {
        args = 0;
        slots = align256(sizeof(CallArgs));
        bytes = align256(slots + cap * sizeof(SlotInfo));
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
    cs->cap_frames = cs->cap_llr = cs->cap_state = cs->cap_bel = cs->cap_hb = 0;
}

// Grows the workspace to hold `frames` frames with the given element totals
// (a quarter of headroom, so batches of varying size rarely regrow it).
int ensure_work(gf2l_ctx *ctx, gf2l_codeset *cs, uint64_t frames, uint64_t llr, uint64_t st, uint64_t bel,
                uint64_t hb) {
    if (frames <= cs->cap_frames && llr <= cs->cap_llr && st <= cs->cap_state && bel <= cs->cap_bel &&
        hb <= cs->cap_hb)
        return GF2L_OK;
    GF2L_HIP(ctx, hipStreamSynchronize(ctx->stream));
    auto grow = [](uint64_t need, uint64_t have) { return std::max<uint64_t>(need + need / 4, have); };
    const uint64_t cf = grow(frames, cs->cap_frames), cl = grow(llr, cs->cap_llr), cst = grow(st, cs->cap_state),
                   cbe = grow(bel, cs->cap_bel), chb = grow(hb, cs->cap_hb);
    free_work(cs);
    const Upload up(cf);
    const uint64_t b_lists = align256(2 * cf * 4), b_cnt = align256(2 * kMaxCodes * 4), b_misc = align256(kMiscWords * 4),
                   b_llr = align256(cl * 4), b_st = align256(cst * 4), b_bel = align256(cbe * 4),
                   b_hb = align256(chb * 8);
    const uint64_t total = up.bytes + b_lists + b_cnt + b_misc + b_llr + b_st + b_bel + b_hb;
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
    m.slots = reinterpret_cast<const SlotInfo *>(p + up.slots);
    p += up.bytes;
    m.lists = reinterpret_cast<uint32_t *>(p), p += b_lists;
    m.list_cap = (uint32_t)cf;
    m.counts = reinterpret_cast<uint32_t *>(p), p += b_cnt;
    m.misc = reinterpret_cast<uint32_t *>(p), p += b_misc;
    m.llr = reinterpret_cast<float *>(p), p += b_llr;
    m.state = reinterpret_cast<uint32_t *>(p), p += b_st;
    m.belief = reinterpret_cast<float *>(p), p += b_bel;
    m.hb = reinterpret_cast<uint64_t *>(p);
    cs->cap_frames = cf;
    cs->cap_llr = cl;
    cs->cap_state = cst;
    cs->cap_bel = cbe;
    cs->cap_hb = chb;
    return GF2L_OK;
}

struct Shape {
    uint32_t grid, block, g_syn, chunk;
    size_t sh;
    int xcd;
};

Shape launch_shape(const gf2l_ctx *ctx, const gf2l_codeset *cs) {
    Shape s;
    s.block = std::min<uint32_t>(512, ((cs->max_z + 63) / 64) * 64);
    s.g_syn = (uint32_t)ctx->num_cus * 8u;
    s.sh = (size_t)cs->max_words * sizeof(uint64_t);
    s.xcd = ctx->tune.xcd_map;
    // automatic: counter fetch (8 items) when every code's variable degrees fit
    // the smallest class, where the lower L2 traffic pays for it
    s.chunk = ctx->tune.chunk >= 0 ? (uint32_t)ctx->tune.chunk : (cs->max_dv <= 8 ? 8u : 0u);
    const uint32_t wg = ctx->tune.wg_per_cu ? (uint32_t)ctx->tune.wg_per_cu : (s.chunk ? 4u : 3u);
    s.grid = ((uint32_t)ctx->num_cus * wg + 7u) & ~7u;
    return s;
}

// One class launch: counter-fed items (sh.chunk > 0) or a fixed stride.
template <int MAXD, bool FIRST, bool NMS>
void launch_cn(const MixShared &m, uint32_t mask, const Shape &sh, hipStream_t s, uint32_t t, uint32_t slot) {
    if (!mask) return;
    if (sh.chunk)
        hipLaunchKernelGGL((cn_mix_kernel<MAXD, FIRST, NMS, true>), dim3(sh.grid), dim3(sh.block), 0, s, m, mask,
                           sh.xcd, t, slot, sh.chunk);
    else
        hipLaunchKernelGGL((cn_mix_kernel<MAXD, FIRST, NMS, false>), dim3(sh.grid), dim3(sh.block), 0, s, m, mask,
                           sh.xcd, t, slot, 1u);
}
template <int MAXD, bool NMS>
void launch_vn(const MixShared &m, uint32_t mask, const Shape &sh, hipStream_t s, uint32_t t, uint32_t slot) {
    if (!mask) return;
    if (sh.chunk)
        hipLaunchKernelGGL((vn_mix_kernel<MAXD, NMS, true>), dim3(sh.grid), dim3(sh.block), 0, s, m, mask, sh.xcd,
                           t, slot, sh.chunk);
    else
        hipLaunchKernelGGL((vn_mix_kernel<MAXD, NMS, false>), dim3(sh.grid), dim3(sh.block), 0, s, m, mask, sh.xcd,
                           t, slot, 1u);
}

// Enqueues iteration t: check and variable launches per degree class, then
// the syndrome launch.  FIRST: inputs are the LLRs; NMS: weighted rule.
template <bool FIRST, bool NMS>
void enqueue_iteration(const gf2l_codeset *cs, const Shape &sh, hipStream_t s, uint32_t t) {
    const MixShared m = cs->m;
    constexpr bool CN_NMS = NMS && !FIRST;
    launch_cn<8, FIRST, CN_NMS>(m, cs->cn_mask[0], sh, s, t, 0u);
    launch_cn<16, FIRST, CN_NMS>(m, cs->cn_mask[1], sh, s, t, 1u);
    launch_cn<27, FIRST, CN_NMS>(m, cs->cn_mask[2], sh, s, t, 2u);
    launch_vn<8, NMS>(m, cs->vn_mask[0], sh, s, t, 3u);
    launch_vn<16, NMS>(m, cs->vn_mask[1], sh, s, t, 4u);
    launch_vn<32, NMS>(m, cs->vn_mask[2], sh, s, t, 5u);
    hipLaunchKernelGGL(syn_mix_kernel, dim3(sh.g_syn), dim3(256), sh.sh, s, m);
    hipLaunchKernelGGL(mix_compact_kernel, dim3(m.ncodes), dim3(1024), 0, s, m);
}

// The continuing fallback's first iteration (t = T + 1): the check update
// subtracts iteration T's weighted messages from the weighted phase's
// beliefs; the variable update is plain.
void enqueue_continue(const gf2l_codeset *cs, const Shape &sh, hipStream_t s, uint32_t t) {
    const MixShared m = cs->m;
    launch_cn<8, false, true>(m, cs->cn_mask[0], sh, s, t, 0u);
    launch_cn<16, false, true>(m, cs->cn_mask[1], sh, s, t, 1u);
    launch_cn<27, false, true>(m, cs->cn_mask[2], sh, s, t, 2u);
    launch_vn<8, false>(m, cs->vn_mask[0], sh, s, t, 3u);
    launch_vn<16, false>(m, cs->vn_mask[1], sh, s, t, 4u);
    launch_vn<32, false>(m, cs->vn_mask[2], sh, s, t, 5u);
    hipLaunchKernelGGL(syn_mix_kernel, dim3(sh.g_syn), dim3(256), sh.sh, s, m);
    hipLaunchKernelGGL(mix_compact_kernel, dim3(m.ncodes), dim3(1024), 0, s, m);
}

// Captures `unroll` non-first iterations as a graph over the code set's
// current workspace.
int capture(gf2l_ctx *ctx, gf2l_codeset *cs, const Shape &sh, uint32_t unroll) {
    if (cs->exec && cs->g_grid == sh.grid && cs->g_unroll == unroll && cs->g_xcd == sh.xcd &&
        cs->g_chunk == sh.chunk)
        return GF2L_OK;
    if (cs->exec) hipGraphExecDestroy(cs->exec);
    cs->exec = nullptr;
    hipStream_t s = ctx->stream;
    hipGraph_t graph = nullptr;
    GF2L_HIP(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    for (uint32_t u = 0; u < unroll; ++u) enqueue_iteration<false, false>(cs, sh, s, 0);
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
    cs->g_grid = sh.grid;
    cs->g_unroll = unroll;
    cs->g_xcd = sh.xcd;
    cs->g_chunk = sh.chunk;
    return GF2L_OK;
}

// Frames still running after the work enqueued so far (synchronizes).
int running_frames(gf2l_ctx *ctx, gf2l_codeset *cs, uint32_t *out) {
    const Upload up(cs->cap_frames);
    uint32_t *h_rb = reinterpret_cast<uint32_t *>(cs->h_stage + up.rb);
    hipStream_t s = ctx->stream;
    GF2L_HIP(ctx, hipMemcpyAsync(h_rb, cs->m.counts, 2 * kMaxCodes * 4, hipMemcpyDeviceToHost, s));
    GF2L_HIP(ctx, hipMemcpyAsync(h_rb + 2 * kMaxCodes, cs->m.misc, 4, hipMemcpyDeviceToHost, s));
    GF2L_HIP(ctx, hipStreamSynchronize(s));
    const uint32_t *cur = h_rb + (h_rb[2 * kMaxCodes] & 1u) * kMaxCodes;
    uint32_t running = 0;
    for (uint32_t c = 0; c < kMaxCodes; ++c) running += cur[c];
    *out = running;
    return GF2L_OK;
}

}  // namespace
}  // namespace gf2l

namespace gf2l {

// Decodes a mixed batch.  Plain (decoders == NULL): min-sum for up to
// tail_it iterations.  Learned: decoders[c] (code c's learned decoder, all
// with the same T) for T weighted iterations, then plain min-sum from the
// LLRs for up to tail_it iterations for frames still failing.
int decode_mixed(gf2l_ctx *ctx, gf2l_codeset *cs, const gf2l_nms *const *decoders, const uint16_t *code_ids,
                 size_t frames, const float *d_llr, uint32_t tail_it, uint64_t *d_msg_words,
                 uint32_t *d_iterations, uint8_t *d_converged, uint8_t *d_syndrome_ok, uint8_t *d_fallback) {
    if (!ctx) return GF2L_ERR_INVALID;
    if (!cs) return fail(ctx, GF2L_ERR_INVALID, "null code set");
    if (cs->device != ctx->device) return fail(ctx, GF2L_ERR_INVALID, "code set built on another device");
    if (frames == 0) return GF2L_OK;
    if (!code_ids || !d_llr || !d_msg_words || !d_iterations || !d_converged || !d_syndrome_ok)
        return fail(ctx, GF2L_ERR_INVALID, "null buffer");
    if (frames > (1u << 24)) return fail(ctx, GF2L_ERR_INVALID, "more than 2^24 frames in one call");
    GF2L_HIP(ctx, hipSetDevice(ctx->device));
    const uint32_t F = (uint32_t)frames, K = (uint32_t)cs->codes.size();
    // learned decoders: one per code of the set, same weighted iteration count
    uint32_t T = 0;
    bool cont = false;  // the fallback continues from the weighted state (else restarts from the LLRs)
    const float *w[kMaxCodes] = {nullptr};
    if (decoders) {
        for (uint32_t c = 0; c < K; ++c) {
            if (!decoders[c]) return fail(ctx, GF2L_ERR_INVALID, "null learned decoder for code " + std::to_string(c));
            uint32_t Tc = 0;
            const gf2l_code *code = nullptr;
            bool layered = false, cc = false;
            w[c] = nms_weights(decoders[c], &Tc, &code, &layered, &cc);
            if (code != cs->codes[c])
                return fail(ctx, GF2L_ERR_INVALID, "learned decoder " + std::to_string(c) + " is for another code");
            if (layered) return fail(ctx, GF2L_ERR_UNSUPPORTED, "mixed batches take flooding learned decoders");
            if (c > 0 && Tc != T)
                return fail(ctx, GF2L_ERR_INVALID, "learned decoders differ in their iteration count");
            if (c > 0 && cc != cont)
                return fail(ctx, GF2L_ERR_INVALID, "learned decoders differ in their fallback mode");
            T = Tc;
            cont = cc;
        }
    }
    const uint32_t max_total = T + tail_it;
    hipStream_t s = ctx->stream;
    // per-code frame counts and workspace totals
    uint32_t per_code[kMaxCodes] = {0};
    uint64_t t_llr = 0, t_st = 0, t_bel = 0, t_hb = 0, t_msg = 0, items_c = 0, items_v = 0;
    for (uint32_t f = 0; f < F; ++f) {
        const uint32_t c = code_ids[f];
        if (c >= K)
            return fail(ctx, GF2L_ERR_INVALID, "code id " + std::to_string(c) + " of frame " + std::to_string(f) +
                                                   " outside the code set");
        const CodeView &v = cs->codes[c]->view;
        t_llr += v.n;
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
    if (max_total == 0) {  // fresh decoders: zero word, which satisfies every check (core.rs:906-937)
        GF2L_HIP(ctx, hipMemsetAsync(d_msg_words, 0, t_msg * sizeof(uint64_t), s));
        GF2L_HIP(ctx, hipMemsetAsync(d_iterations, 0, (size_t)F * sizeof(uint32_t), s));
        GF2L_HIP(ctx, hipMemsetAsync(d_converged, 0, F, s));
        GF2L_HIP(ctx, hipMemsetAsync(d_syndrome_ok, 1, F, s));
        if (d_fallback) GF2L_HIP(ctx, hipMemsetAsync(d_fallback, 0, F, s));
        return GF2L_OK;
    }
    int rc = ensure_work(ctx, cs, F, t_llr, t_st, t_bel, t_hb);
    if (rc != GF2L_OK) return rc;
    // fill the pinned staging block (the previous call has synchronized)
    const Upload up(cs->cap_frames);
    char *h = cs->h_stage;
    CallArgs *ha = reinterpret_cast<CallArgs *>(h + up.args);
    *ha = CallArgs{};
    ha->llr = d_llr;
    ha->msg = d_msg_words;
    ha->iters = d_iterations;
    ha->conv = d_converged;
    ha->synd = d_syndrome_ok;
    ha->fb = d_fallback;
    ha->max_total = max_total;
    ha->weighted = T;
    ha->frames = F;
    for (uint32_t c = 0; c < K; ++c) ha->w[c] = w[c];
    uint64_t o_llr = 0, o_st = 0, o_bel = 0, o_hb = 0;
    uint32_t o_slot = 0, fill[kMaxCodes];
    for (uint32_t c = 0; c < K; ++c) {
        const CodeView &v = cs->codes[c]->view;
        ha->seg[c] = CodeSeg{o_llr, o_st, o_bel, o_hb, o_slot, per_code[c]};
        o_llr += (uint64_t)per_code[c] * v.n;
        o_st += (uint64_t)per_code[c] * 3ull * v.mb * v.zp;
        o_bel += (uint64_t)per_code[c] * v.nb * v.zp;
        o_hb += (uint64_t)per_code[c] * v.nb * v.zw;
        o_slot += per_code[c];
        fill[c] = 0;
    }
    SlotInfo *h_slots = reinterpret_cast<SlotInfo *>(h + up.slots);
    uint64_t f_llr = 0, f_msg = 0;
    for (uint32_t f = 0; f < F; ++f) {
        const uint32_t c = code_ids[f];
        const CodeView &v = cs->codes[c]->view;
        h_slots[ha->seg[c].list_off + fill[c]++] = SlotInfo{f, c, f_llr, f_msg};
        f_llr += v.n;
        f_msg += (v.k + 63) / 64;
    }
    GF2L_HIP(ctx, hipMemcpyAsync(cs->d_base, h, up.slots + (size_t)F * sizeof(SlotInfo), hipMemcpyHostToDevice, s));
    const uint32_t g_init = std::min<uint32_t>((F + 255) / 256, 1024);
    hipLaunchKernelGGL(mix_init_kernel, dim3(g_init), dim3(256), 0, s, cs->m);
    hipLaunchKernelGGL(mix_gather_kernel, dim3(std::min<uint32_t>(F, 4096)), dim3(256), 0, s, cs->m);
    const Shape sh = launch_shape(ctx, cs);
    // weighted iterations 1..T (direct launches: each reads its own weights)
    for (uint32_t t = 1; t <= T; ++t) {
        if (t == 1)
            enqueue_iteration<true, true>(cs, sh, s, t);
        else
            enqueue_iteration<false, true>(cs, sh, s, t);
    }
    if (tail_it == 0) {
        GF2L_HIP(ctx, hipGetLastError());
        GF2L_HIP(ctx, hipStreamSynchronize(s));
        return GF2L_OK;
    }
    // plain iterations T+1 .. T+tail_it: the first from the LLRs (a fresh
    // decode, or the fallback's restart), later ones replayed from the graph
    uint32_t running = F;
    if (T) {
        rc = running_frames(ctx, cs, &running);
        if (rc != GF2L_OK) return rc;
    }
    if (running) {
        if (cont && T)
            enqueue_continue(cs, sh, s, T + 1);
        else
            enqueue_iteration<true, false>(cs, sh, s, T + 1);
    }
    GF2L_HIP(ctx, hipGetLastError());
    if (running && tail_it > 1) {
        rc = capture(ctx, cs, sh, (uint32_t)ctx->tune.graph_iters);
        if (rc != GF2L_OK) return rc;
        rc = running_frames(ctx, cs, &running);
        if (rc != GF2L_OK) return rc;
        // replay until every frame has retired (at most ceil((tail_it - 1) / U) replays)
        for (uint32_t done_it = 1; running && done_it < tail_it; done_it += cs->g_unroll) {
            GF2L_HIP(ctx, hipGraphLaunch(cs->exec, s));
            rc = running_frames(ctx, cs, &running);
            if (rc != GF2L_OK) return rc;
        }
    }
    GF2L_HIP(ctx, hipStreamSynchronize(s));
    return GF2L_OK;
}

}  // namespace gf2l

using gf2l::fail;

extern "C" {

int gf2l_codeset_create(gf2l_ctx *ctx, const gf2l_code *const *codes, uint32_t count, gf2l_codeset **out) {
    if (!ctx || !out) return GF2L_ERR_INVALID;
    *out = nullptr;
    return gf2l::guard(ctx, [&]() -> int {
        if (!codes || count == 0 || count > gf2l::kMaxCodes)
            return fail(ctx, GF2L_ERR_INVALID, "a code set holds 1 to 16 codes");
        std::vector<gf2l::CodeView> views;
        uint32_t max_dc = 0, max_dv = 0, max_z = 0, max_words = 0;
        for (uint32_t c = 0; c < count; ++c) {
            const gf2l_code *code = codes[c];
            if (!code) return fail(ctx, GF2L_ERR_INVALID, "null code in a code set");
            if (code->is_csr || code->view.dvb_q)
                return fail(ctx, GF2L_ERR_UNSUPPORTED, "code sets hold quasi-cyclic codes (not DVB)");
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
        for (uint32_t c = 0; c < count; ++c) {
            const uint32_t dc = codes[c]->view.max_dc, dv = codes[c]->view.max_dv;
            cs->cn_mask[dc <= 8 ? 0 : dc <= 16 ? 1 : 2] |= 1u << c;
            cs->vn_mask[dv <= 8 ? 0 : dv <= 16 ? 1 : 2] |= 1u << c;
        }
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
    });
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
    return gf2l::guard(ctx, [&]() -> int {
        if (ctx->rule != GF2L_RULE_MINSUM)
            return fail(ctx, GF2L_ERR_UNSUPPORTED, "the mixed-code decoder runs the min-sum rule only");
        return gf2l::decode_mixed(ctx, cs, nullptr, code_ids, frames, d_llr, max_it, d_msg_words, d_iterations,
                                  d_converged, d_syndrome_ok, nullptr);
    });
}

int gf2l_decode_mixed_nms_device(gf2l_ctx *ctx, gf2l_codeset *cs, const gf2l_nms *const *decoders,
                                 const uint16_t *code_ids, size_t frames, const float *d_llr,
                                 uint32_t fallback_max_it, uint64_t *d_msg_words, uint32_t *d_iterations,
                                 uint8_t *d_converged, uint8_t *d_syndrome_ok, uint8_t *d_fallback) {
    if (!ctx) return GF2L_ERR_INVALID;
    return gf2l::guard(ctx, [&]() -> int {
        if (!decoders) return fail(ctx, GF2L_ERR_INVALID, "null decoder list");
        if (ctx->rule != GF2L_RULE_MINSUM)
            return fail(ctx, GF2L_ERR_UNSUPPORTED, "the learned decoders run the min-sum rule only");
        return gf2l::decode_mixed(ctx, cs, decoders, code_ids, frames, d_llr, fallback_max_it, d_msg_words,
                                  d_iterations, d_converged, d_syndrome_ok, d_fallback);
    });
}

}  // extern "C"
