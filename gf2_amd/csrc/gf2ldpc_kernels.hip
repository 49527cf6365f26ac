// HIP kernels of the batched LDPC min-sum decoder for gfx950 (MI355X).
//
// The arithmetic contract (bit-exact with the reference decoder) is stated
// in device_common.h with the helpers that implement it.
//
// Three decoder paths:
//   LDS     one workgroup per frame, whole frame state in LDS, all
//           iterations in one launch (small codes).
//   STATE   per-iteration kernels over HBM; each check keeps a compressed
//           state (min1, min2, argmin | output sign bits: 12 bytes) from
//           which every check->variable message is rebuilt bit for bit.
//   STREAM  per-iteration kernels over HBM with one stored message per
//           edge (any check degree up to 32).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "device_common.h"

namespace gf2l {
namespace {

// XCD-aware split of a persistent grid: workgroups are dispatched round-robin
// over the 8 XCDs, so workgroup b runs on XCD b % 8.  Frames list[x],
// list[x+8], ... go to XCD x and the workgroups of that XCD stride over the
// items of those frames in frame order, so the few frames in flight on an
// XCD keep their beliefs / check states in that XCD's L2 while their
// block-rows (block-columns) re-read them.  Needs gridDim.x % 8 == 0.
// With xcd == 0 the whole grid strides over all items in frame order.
struct XcdSplit {
    uint32_t x, q, Q, items, step;
    __device__ XcdSplit(uint32_t count, uint32_t per_frame, int xcd) {
        if (xcd) {
            x = blockIdx.x & 7u;
            q = blockIdx.x >> 3;
            Q = gridDim.x >> 3;
            step = 8u;
            const uint32_t nfx = count > x ? (count - x + 7u) >> 3 : 0u;
            items = nfx * per_frame;
        } else {
            x = 0;
            q = blockIdx.x;
            Q = gridDim.x;
            step = 1u;
            items = count * per_frame;
        }
    }
    __device__ uint32_t list_index(uint32_t j, uint32_t per_frame) const {
        return x + step * (j / per_frame);
    }
};

// ---------------------------------------------------------------------------
// STATE path.

// Check update of one block-row of one frame.  Threads map to the z lifted
// checks; belief reads of a circulant are a rotated contiguous run, the
// check's own state is contiguous.  In the first iteration the inputs are
// the channel LLRs (belief = L, c2v = 0).
// DYN: items come in chunks of `chunk` from per-XCD counters at
// sched + (kind * 8 + xcd) * kCtrStride (kind 0: check, 1: variable kernel);
// otherwise a fixed stride.  Each kernel clears the other kind's counters
// after its items, for the kernel that follows it.
template <int MAXD, bool FIRST, bool NMS, bool DYN, bool DVB = false>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(kCheckWaves<MAXD>)))
cn_state_kernel(CodeView cv, const uint32_t *__restrict__ list, const uint32_t *__restrict__ count_ptr,
                const uint32_t *state_in, uint32_t *state_out, const float *__restrict__ belief,
                const float *__restrict__ llr, const float *__restrict__ w_prev, int xcd,
                uint32_t *sched, uint32_t chunk) {
    const XcdSplit xs(*count_ptr, cv.mb, xcd);
    const size_t fstride = 3ull * cv.mb * cv.zp;
    auto item = [&](uint32_t j) {
        const uint32_t f = list[xs.list_index(j, cv.mb)];
        const uint32_t br = j % cv.mb;
        const float *fb = FIRST ? llr + (size_t)f * cv.n : belief + (size_t)f * cv.nb * cv.zp;
        cn_state_item<MAXD, FIRST, NMS, DVB>(cv, br, state_in + f * fstride, state_out + f * fstride, fb, w_prev);
    };
    if constexpr (DYN) {
        __shared__ uint32_t s_next[2];
        ItemStream st(xs.q, xs.Q, sched + xs.x * kCtrStride, chunk);
        while (st.j < xs.items) {
            item(st.j);
            st.advance(s_next);
        }
        // the variable kernel that follows draws its items from these
        if (blockIdx.x == 0 && threadIdx.x < 8) sched[(8 + threadIdx.x) * kCtrStride] = 0;
    } else {
        for (uint32_t j = xs.q; j < xs.items; j += xs.Q) item(j);
    }
}

// Variable update of one block-column of one frame per item.
template <int MAXD, bool NMS, bool DYN, bool DVB = false>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(kVarWaves<MAXD>)))
vn_state_kernel(CodeView cv, const uint32_t *__restrict__ list, const uint32_t *__restrict__ count_ptr,
                const uint32_t *__restrict__ state, float *__restrict__ belief,
                const float *__restrict__ llr, uint64_t *__restrict__ hb,
                uint32_t *__restrict__ next_count, const float *__restrict__ w_cur, int xcd,
                uint32_t *sched, uint32_t chunk) {
    const XcdSplit xs(*count_ptr, cv.nb, xcd);
    const size_t fstride = 3ull * cv.mb * cv.zp;
    auto item = [&](uint32_t jj) {
        const uint32_t f = list[xs.list_index(jj, cv.nb)];
        const uint32_t bc = jj % cv.nb;
        vn_state_item<MAXD, NMS, DVB>(cv, bc, state + f * fstride, llr + (size_t)f * cv.n + (size_t)bc * cv.z,
                                 belief + ((size_t)f * cv.nb + bc) * cv.zp,
                                 hb ? hb + ((size_t)f * cv.nb + bc) * cv.zw : nullptr, w_cur);
    };
    if constexpr (DYN) {
        __shared__ uint32_t s_next[2];
        ItemStream st(xs.q, xs.Q, sched + (8 + xs.x) * kCtrStride, chunk);
        while (st.j < xs.items) {
            item(st.j);
            st.advance(s_next);
        }
        if (blockIdx.x == 0 && threadIdx.x < 8) sched[threadIdx.x * kCtrStride] = 0;
    } else {
        for (uint32_t jj = xs.q; jj < xs.items; jj += xs.Q) item(jj);
    }
    // stored last: the next running count starts at zero (the syndrome step
    // stores the survivor count after compacting the list)
    if (next_count && blockIdx.x == 0 && threadIdx.x == 0) *next_count = 0;
}

// ---------------------------------------------------------------------------
// STREAM path (one stored message per edge).

template <int MAXD, bool NMS>
__global__ void __launch_bounds__(512)
cn_stream_kernel(CodeView cv, const uint32_t *__restrict__ list, const uint32_t *__restrict__ count_ptr,
                 float *__restrict__ c2v, const float *__restrict__ belief,
                 const float *__restrict__ w_it, int xcd) {
    const XcdSplit xs(*count_ptr, cv.mb, xcd);
    const size_t E = (size_t)cv.nbe * cv.z;
    for (uint32_t item = xs.q; item < xs.items; item += xs.Q) {
        const uint32_t f = list[xs.list_index(item, cv.mb)];
        const uint32_t br = item % cv.mb;
        const uint32_t e0 = cv.row_ptr[br];
        const uint32_t d = cv.row_ptr[br + 1] - e0;
        float *fc = c2v + (size_t)f * E + (size_t)e0 * cv.z;
        const float *fb = belief + (size_t)f * cv.n;
        dispatch_degree<MAXD>(d, [&](auto DC) {
            constexpr int D = decltype(DC)::value, DA = D > 0 ? D : 1;
            uint2 re[DA];
            float wc[DA];
#pragma unroll
            for (int p = 0; p < D; ++p) {
                re[p] = cv.row_edge[e0 + p];
                if constexpr (NMS) wc[p] = w_it[e0 + p];
            }
            for (uint32_t i = threadIdx.x; i < cv.z; i += blockDim.x) {
                float bl[DA], co[DA];
#pragma unroll
                for (int p = 0; p < D; ++p) {
                    bl[p] = fb[re[p].x * cv.z + wrap_add(i, re[p].y, cv.z)];
                    co[p] = fc[(size_t)p * cv.z + i];
                }
                CheckAcc acc;
                acc.init();
#pragma unroll
                for (int p = 0; p < D; ++p) acc.add(p, bl[p] - co[p]);
#pragma unroll
                for (int p = 0; p < D; ++p) {
                    float o = acc.out(D, p);
                    if constexpr (NMS) o = wc[p] * o;
                    fc[(size_t)p * cv.z + i] = o;
                }
            }
        });
    }
}

// The sum-product rule of check_node_update_spa (core.rs:762-787) in place
// of min-sum: the message to position p is Llr::boxplus_n of the other
// inputs (llr.rs:210-215), 2 atanh(prod tanh(x_q / 2)) with the product
// folded from 1.0 in ascending position order; a lone input gets +0.0.
// Loops run to MAXD under a runtime-degree guard (one copy of the
// transcendental code per class rather than one per degree).
template <int MAXD>
__global__ void __launch_bounds__(512)
cn_stream_spa_kernel(CodeView cv, const uint32_t *__restrict__ list, const uint32_t *__restrict__ count_ptr,
                     float *__restrict__ c2v, const float *__restrict__ belief, int xcd) {
    const XcdSplit xs(*count_ptr, cv.mb, xcd);
    const size_t E = (size_t)cv.nbe * cv.z;
    for (uint32_t item = xs.q; item < xs.items; item += xs.Q) {
        const uint32_t f = list[xs.list_index(item, cv.mb)];
        const uint32_t br = item % cv.mb;
        const uint32_t e0 = cv.row_ptr[br];
        const uint32_t d = cv.row_ptr[br + 1] - e0;
        float *fc = c2v + (size_t)f * E + (size_t)e0 * cv.z;
        const float *fb = belief + (size_t)f * cv.n;
        for (uint32_t i = threadIdx.x; i < cv.z; i += blockDim.x) {
            float th[MAXD];
#pragma unroll
            for (int p = 0; p < MAXD; ++p) {
                th[p] = 1.0f;
                if (p < (int)d) {
                    const uint2 re = cv.row_edge[e0 + p];
                    const float x = fb[re.x * cv.z + wrap_add(i, re.y, cv.z)] - fc[(size_t)p * cv.z + i];
                    th[p] = tanhf(x / 2.0f);
                }
            }
#pragma unroll
            for (int p = 0; p < MAXD; ++p) {
                if (p < (int)d) {
                    float prod = 1.0f;
#pragma unroll
                    for (int q = 0; q < MAXD; ++q)
                        if (q != p && q < (int)d) prod = prod * th[q];
                    fc[(size_t)p * cv.z + i] = d > 1 ? 2.0f * atanhf(prod) : 0.0f;
                }
            }
        }
    }
}

template <int MAXD>
__global__ void __launch_bounds__(512)
vn_stream_kernel(CodeView cv, const uint32_t *__restrict__ list, const uint32_t *__restrict__ count_ptr,
                 const float *__restrict__ c2v, float *__restrict__ belief,
                 const float *__restrict__ llr, uint64_t *__restrict__ hb,
                 uint32_t *__restrict__ next_count, int xcd) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *next_count = 0;
    const XcdSplit xs(*count_ptr, cv.nb, xcd);
    const size_t E = (size_t)cv.nbe * cv.z;
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (uint32_t item = xs.q; item < xs.items; item += xs.Q) {
        const uint32_t f = list[xs.list_index(item, cv.nb)];
        const uint32_t bc = item % cv.nb;
        const uint32_t c0 = cv.col_ptr[bc];
        const uint32_t dv = cv.col_ptr[bc + 1] - c0;
        const float *fc = c2v + (size_t)f * E;
        const size_t vbase = (size_t)f * cv.n + (size_t)bc * cv.z;
        uint64_t *fh = hb + ((size_t)f * cv.nb + bc) * cv.zw;
        dispatch_degree<MAXD>(dv, [&](auto DC) {
            constexpr int D = decltype(DC)::value, DA = D > 0 ? D : 1;
            uint2 ce[DA];
#pragma unroll
            for (int q = 0; q < D; ++q) ce[q] = cv.col_edge[c0 + q];
            for (uint32_t j0 = 0; j0 < cv.z; j0 += blockDim.x) {
                const uint32_t j = j0 + threadIdx.x;
                const bool valid = j < cv.z;
                float bel = 0.0f;
                if (valid) {
                    float c[DA];
#pragma unroll
                    for (int q = 0; q < D; ++q)
                        c[q] = fc[(size_t)ce[q].x * cv.z + wrap_sub(j, ce[q].y, cv.z)];
                    bel = llr[vbase + j];
#pragma unroll
                    for (int q = 0; q < D; ++q) bel = bel + c[q];
                    belief[vbase + j] = bel;
                }
                const uint64_t mask = __ballot(valid && bel < 0.0f);
                const uint32_t word_bit = j0 + wave * 64;
                if (lane == 0 && word_bit < cv.z) fh[word_bit >> 6] = mask;
            }
        });
    }
}

// ---------------------------------------------------------------------------
// Syndrome + retirement, shared by both HBM paths: one workgroup of
// kSynThreads per frame (the check-row words are independent, so four waves
// share a frame's rows and the pass is not latency-bound on one wave).

constexpr uint32_t kSynThreads = 256;

__global__ void __launch_bounds__(kSynThreads)
syn_words_kernel(CodeView cv, const uint32_t *__restrict__ list, const uint32_t *__restrict__ count_ptr,
                 uint32_t *__restrict__ next_list, uint32_t *__restrict__ next_count,
                 const uint64_t *__restrict__ hb, uint32_t it_report, int last, int fb_phase,
                 const uint32_t *__restrict__ itc, uint32_t itc_last, uint32_t itc_fb_from,
                 uint64_t *__restrict__ msg, uint32_t *__restrict__ iters,
                 uint8_t *__restrict__ conv, uint8_t *__restrict__ synd, uint8_t *__restrict__ fb) {
    extern __shared__ uint64_t sh[];
    const uint32_t count = *count_ptr;
    if (itc) {  // replayed from a graph: the iteration number is on the device
        it_report = *itc;
        last = it_report >= itc_last;
        fb_phase = it_report > itc_fb_from;
    }
    const uint32_t words = cv.nb * cv.zw;
    const uint32_t kw = (cv.k + 63) / 64;
    for (uint32_t item = blockIdx.x; item < count; item += gridDim.x) {
        const uint32_t f = list[item];
        const int unsat = frame_unsat(cv, sh, hb + (size_t)f * words);
        if (!unsat || last) {
            frame_message(cv, sh, msg + (size_t)f * kw);
            if (threadIdx.x == 0) {
                iters[f] = it_report;
                conv[f] = (uint8_t)!unsat;
                synd[f] = (uint8_t)!unsat;
                if (fb) fb[f] = (uint8_t)fb_phase;
            }
        }
        // the slot is kept in place (retired frames leave a hole) so that
        // list_compact_kernel keeps the survivors in list order
        if (threadIdx.x == 0) next_list[item] = (!unsat || last) ? kRetiredSlot : f;
        __syncthreads();
    }
}

// Closes the holes the syndrome pass left in next_list (survivors stay in
// list order) and stores the survivor count.  One workgroup: the list holds
// at most one slot per frame.  itc (may be null): the device iteration
// word, set to itc_next (a direct launch) or, with itc_next 0, advanced by
// one (an iteration replayed from a graph).
__global__ void __launch_bounds__(1024)
list_compact_kernel(const uint32_t *__restrict__ count_ptr, uint32_t *__restrict__ next_list,
                    uint32_t *__restrict__ next_count, uint32_t *__restrict__ itc, uint32_t itc_next) {
    __shared__ uint32_t wave_kept[16];
    const uint32_t kept = compact_list(next_list, *count_ptr, wave_kept);
    if (threadIdx.x == 0) {
        *next_count = kept;
        if (itc) *itc = itc_next ? itc_next : *itc + 1u;
    }
}

// counts: [0..1] running-list counts; from kSchedOffset the check-kernel
// (8) and variable-kernel (8) item counters, one per XCD, kCtrStride apart.
__global__ void hbm_init_kernel(uint32_t *list, uint32_t *counts, uint32_t frames) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < frames) list[t] = t;
    if (t == 0) {
        counts[0] = frames;
        counts[1] = 0;
    }
    if (t < 16) counts[kSchedOffset + t * kCtrStride] = 0;
}

// Trace support for the STATE path: explicit messages and unpadded beliefs.
__global__ void state_expand_kernel(CodeView cv, const uint32_t *__restrict__ state,
                                    const float *__restrict__ belief, uint32_t frames,
                                    const float *__restrict__ w, float *__restrict__ c2v_out,
                                    float *__restrict__ bel_out) {
    const size_t E = (size_t)cv.nbe * cv.z, fstride = 3ull * cv.mb * cv.zp;
    const size_t total = (size_t)frames * E;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total;
         t += (size_t)gridDim.x * blockDim.x) {
        const uint32_t f = (uint32_t)(t / E);
        const uint32_t r = (uint32_t)(t - (size_t)f * E), e = r / cv.z, i = r - e * cv.z;
        uint32_t br = 0;
        while (cv.row_ptr[br + 1] <= e) ++br;
        const uint32_t *s = state + f * fstride + 3 * ((size_t)br * cv.zp + i);
        float v = recon(__uint_as_float(s[0]), __uint_as_float(s[1]), s[2], e - cv.row_ptr[br]);
        if (w) v = w[e] * v;
        c2v_out[t] = v;
    }
    const size_t nt = (size_t)frames * cv.n;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < nt;
         t += (size_t)gridDim.x * blockDim.x) {
        const uint32_t f = (uint32_t)(t / cv.n), v = (uint32_t)(t - (size_t)f * cv.n);
        const uint32_t bc = v / cv.z, j = v - bc * cv.z;
        bel_out[t] = belief[((size_t)f * cv.nb + bc) * cv.zp + j];
    }
}

// Dynamic LDS of the LDS path: messages, beliefs, LLRs, then the code
// tables (row and column edges, row and column pointers).
size_t lds_decode_bytes(const CodeView &cv) {
    return (size_t)lds_table_offset(cv) * sizeof(float) + (size_t)cv.nbe * 2 * sizeof(uint2) +
           (size_t)(cv.mb + 1 + cv.nb + 1) * sizeof(uint32_t);
}

// ---------------------------------------------------------------------------
// LDS path: one workgroup owns one frame for all its iterations.  Messages,
// beliefs and LLRs live in LDS, and the code tables are copied there once
// per workgroup (a table read is an LDS read, not a dependent global load).
// The threads form G groups of ZT lanes: group g updates block-rows g, g + G, ... and
// block-columns g, g + G, ..., lane i the lifted checks / variables i,
// i + ZT, ... of each, so no thread divides an index inside the loops.
// An iteration is a check pass and a variable pass.  The check pass also
// takes the parity of the hard decisions b < 0 of every check's inputs,
// which is the syndrome of the previous iteration's beliefs: a frame whose
// previous iteration satisfied every check stops there (the messages that
// pass computed are dropped; the beliefs, the output, are untouched).  The
// last iteration of a phase is checked by a pass of its own.
// Phase 0 runs max_it iterations (weighted by w[it-1][e]: NMS, w given);
// a frame that still fails the syndrome is, when fb_it > 0, decoded again
// from its channel LLRs by fb_it plain iterations (phase 1), or with
// fb_continue by fb_it plain iterations continued from the phase-0 beliefs
// and (weighted) messages.  The plain instantiation (NMS false) carries no
// weight code; for degrees up to 8 it is built for 80 VGPRs (6 waves per
// SIMD, no spills).
template <int DMAX, int DVMAX, bool NMS>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(!NMS && DMAX <= 8 && DVMAX <= 8 ? 6 : 1)))
lds_decode_kernel(CodeView cv, LdsShape shp, const float *__restrict__ llr, uint32_t frames, uint32_t max_it,
                  const float *__restrict__ w, uint32_t fb_it, int fb_continue, uint64_t *__restrict__ msg,
                  uint32_t *__restrict__ iters, uint8_t *__restrict__ conv,
                  uint8_t *__restrict__ synd, uint8_t *__restrict__ fb,
                  float *__restrict__ trace_c2v, float *__restrict__ trace_bel,
                  uint32_t trace_iters) {
    extern __shared__ float sm[];
    const uint32_t E = cv.nbe * cv.z, n = cv.n, z = cv.z, mb = cv.mb, nb = cv.nb;
    float *s_c2v = sm;
    float *s_bel = sm + E;
    float *s_llr = s_bel + n;
    uint2 *s_re = reinterpret_cast<uint2 *>(sm + lds_table_offset(cv));
    uint2 *s_ce = s_re + cv.nbe;
    uint32_t *s_rp = reinterpret_cast<uint32_t *>(s_ce + cv.nbe);
    uint32_t *s_cp = s_rp + mb + 1;
    const uint32_t G = shp.groups, ZT = shp.lanes;
    const uint32_t g = threadIdx.x / ZT, lane = threadIdx.x - g * ZT;
    const bool act = g < G;
    for (uint32_t t = threadIdx.x; t < cv.nbe; t += blockDim.x) {
        s_re[t] = cv.row_edge[t];
        s_ce[t] = cv.col_edge[t];
    }
    for (uint32_t t = threadIdx.x; t <= mb; t += blockDim.x) s_rp[t] = cv.row_ptr[t];
    for (uint32_t t = threadIdx.x; t <= nb; t += blockDim.x) s_cp[t] = cv.col_ptr[t];
    // (the barrier after the first frame's reset orders the tables)
    const uint32_t kw = (cv.k + 63) / 64;
    const uint32_t wave = threadIdx.x >> 6, wlane = threadIdx.x & 63u, nwaves = blockDim.x >> 6;

    // messages of one iteration from the current beliefs (update), and
    // whether one of this thread's checks is unsatisfied by their signs
    auto check_pass = [&](const float *w_it, bool update) -> int {
        int unsat = 0;
        if (!act) return 0;
        for (uint32_t i = lane; i < z; i += ZT) {
            for (uint32_t br = g; br < mb; br += G) {
                const uint32_t e0 = s_rp[br], d = s_rp[br + 1] - e0;
                float x[DMAX];
                uint32_t par = 0;
#pragma unroll
                for (int p = 0; p < DMAX; ++p) {
                    if (p < (int)d) {
                        const uint2 re = s_re[e0 + p];
                        const float b = s_bel[re.x * z + wrap_add(i, re.y, z)];
                        par ^= (uint32_t)(b < 0.0f);
                        x[p] = b - s_c2v[(e0 + p) * z + i];
                    }
                }
                unsat |= (int)par;
                if (!update) continue;
                CheckAcc acc;
                acc.init();
#pragma unroll
                for (int p = 0; p < DMAX; ++p)
                    if (p < (int)d) acc.add_sel(p, x[p]);
                // output signs of all positions at once (acc.out's rule)
                const uint32_t sg = acc.signs<DMAX>(d);
#pragma unroll
                for (int p = 0; p < DMAX; ++p) {
                    if (p < (int)d) {
                        const float mag = (uint32_t)p == acc.idx ? acc.min2 : acc.min1;
                        float o = d == 1 ? 0.0f : ((sg >> p) & 1u) ? -mag : mag;
                        if constexpr (NMS) {  // phase 1 (the plain fallback) has no weights
                            if (w_it) o = w_it[e0 + p] * o;
                        }
                        s_c2v[(e0 + p) * z + i] = o;
                    }
                }
            }
        }
        return unsat;
    };

    for (uint32_t f = blockIdx.x; f < frames; f += gridDim.x) {
        const float *fl = llr + (size_t)f * n;
        for (uint32_t v = threadIdx.x; v < n; v += blockDim.x) s_llr[v] = fl[v];
        // beliefs := LLRs, messages := 0 (each thread reads what it wrote above)
        auto reset = [&]() {
            for (uint32_t v = threadIdx.x; v < n; v += blockDim.x) s_bel[v] = s_llr[v];
            for (uint32_t e = threadIdx.x; e < E; e += blockDim.x) s_c2v[e] = 0.0f;
        };
        uint32_t it_done = 0;
        int converged = 0, fell_back = 0;
        for (int phase = 0; phase < 2 && !converged; ++phase) {
            const uint32_t n_it = phase == 0 ? max_it : fb_it;
            if (phase == 1) {
                if (fb_it == 0) break;
                fell_back = 1;
            }
            if (phase == 0 || !fb_continue) reset();
            __syncthreads();
            const uint32_t base = phase == 0 ? 0 : max_it;
            uint32_t it = 1;
            for (; it <= n_it; ++it) {
                const float *w_it = (phase == 0 && w) ? w + (size_t)(it - 1) * cv.nbe : nullptr;
                const int unsat = __syncthreads_or(check_pass(w_it, true));
                if (it > 1 && !unsat) {  // iteration it - 1 satisfied every check
                    converged = 1;
                    break;
                }
                if (act) {  // beliefs: L + the messages in ascending check order
                    for (uint32_t j = lane; j < z; j += ZT) {
                        for (uint32_t bc = g; bc < nb; bc += G) {
                            const uint32_t c0 = s_cp[bc], dv = s_cp[bc + 1] - c0;
                            float c[DVMAX];
#pragma unroll
                            for (int q = 0; q < DVMAX; ++q) {
                                if (q < (int)dv) {
                                    const uint2 ce = s_ce[c0 + q];
                                    c[q] = s_c2v[ce.x * z + wrap_sub(j, ce.y, z)];
                                }
                            }
                            float bel = s_llr[bc * z + j];
#pragma unroll
                            for (int q = 0; q < DVMAX; ++q)
                                if (q < (int)dv) bel = bel + c[q];
                            s_bel[bc * z + j] = bel;
                        }
                    }
                }
                __syncthreads();
                if (phase == 0 && trace_c2v && it <= trace_iters) {
                    for (uint32_t e = threadIdx.x; e < E; e += blockDim.x)
                        trace_c2v[((size_t)(it - 1) * frames + f) * E + e] = s_c2v[e];
                    for (uint32_t v = threadIdx.x; v < n; v += blockDim.x)
                        trace_bel[((size_t)(it - 1) * frames + f) * n + v] = s_bel[v];
                }
            }
            if (converged) {
                it_done = base + it - 1;
            } else if (n_it > 0) {  // the phase's last iteration
                it_done = base + n_it;
                if (!__syncthreads_or(check_pass(nullptr, false))) converged = 1;
            }
            // frames that stop keep their final state in the later trace slots
            // (copied from the slot this thread wrote: the messages in LDS may
            // already be those of the iteration after)
            if (phase == 0 && trace_c2v && converged && it_done >= 1 && it_done < trace_iters) {
                const size_t src = (size_t)(it_done - 1) * frames + f;
                for (uint32_t t = it_done + 1; t <= trace_iters; ++t) {
                    const size_t dst = (size_t)(t - 1) * frames + f;
                    for (uint32_t e = threadIdx.x; e < E; e += blockDim.x)
                        trace_c2v[dst * E + e] = trace_c2v[src * E + e];
                    for (uint32_t v = threadIdx.x; v < n; v += blockDim.x)
                        trace_bel[dst * n + v] = trace_bel[src * n + v];
                }
            }
        }
        // message words: the first k hard decisions, one ballot per wave and word
        // (max_it == 0: a fresh decoder's beliefs are all zero, core.rs:680-693)
        for (uint32_t wd = wave; wd < kw; wd += nwaves) {
            const uint32_t v = wd * 64 + wlane;
            const uint64_t word = __ballot(max_it > 0 && v < cv.k && s_bel[v] < 0.0f);
            if (wlane == 0) msg[(size_t)f * kw + wd] = word;
        }
        if (threadIdx.x == 0) {
            iters[f] = it_done;
            conv[f] = (uint8_t)converged;
            synd[f] = (uint8_t)(converged || max_it == 0);
            if (fb) fb[f] = (uint8_t)fell_back;
        }
        __syncthreads();
    }
}

template <typename K>
hipError_t set_lds(K kernel, size_t bytes) {
    if (bytes <= 65536) return hipSuccess;
    return hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

int pick_dmax(uint32_t d) { return d <= 8 ? 8 : d <= 16 ? 16 : 32; }
int pick_dvmax(uint32_t d) { return d <= 4 ? 4 : d <= 8 ? 8 : d <= 16 ? 16 : 32; }

// One thread per lifted check / variable of a block, in whole waves; larger
// lifts loop inside the block.
uint32_t round_block(uint32_t z) {
    uint32_t b = ((z + 63) / 64) * 64;
    return b > 512 ? 512 : b;
}

// Persistent grid of a multiple of 8 workgroups, no larger than the work.
uint32_t hbm_grid(uint32_t items, int num_cus, int wg_per_cu) {
    uint32_t g = (uint32_t)num_cus * (uint32_t)wg_per_cu;
    g = ((g + 7u) / 8u) * 8u;
    const uint32_t need = ((items + 7u) / 8u) * 8u;
    if (need < g) g = need;
    return g < 8u ? 8u : g;
}

// Degree classes of the STATE kernels: registers are allocated for the
// largest degree a kernel dispatches, so a code runs the smallest class that
// covers its check (variable) degrees.
#define GF2L_STATE_CLASS(DEG, LAUNCH, A, B) \
    do {                                     \
        if ((DEG) <= 8)                      \
            LAUNCH(8, A, B);                 \
        else if ((DEG) <= 16)                \
            LAUNCH(16, A, B);                \
        else                                 \
            LAUNCH(27, A, B);                \
    } while (0)
#define GF2L_VAR_CLASS(DEG, LAUNCH, A, B) \
    do {                                   \
        if ((DEG) <= 8)                    \
            LAUNCH(8, A, B);               \
        else if ((DEG) <= 16)              \
            LAUNCH(16, A, B);              \
        else                               \
            LAUNCH(32, A, B);              \
    } while (0)

template <int DMAX, int DVMAX, bool NMS>
hipError_t launch_lds_t(const CodeView &cv, LdsShape shp, const float *llr, uint32_t frames, uint32_t max_it,
                        const float *w, uint32_t fb_it, bool fb_continue, uint64_t *msg, uint32_t *iters,
                        uint8_t *conv, uint8_t *synd, uint8_t *fb, float *tc, float *tb, uint32_t ti,
                        hipStream_t s) {
    const size_t bytes = lds_decode_bytes(cv);
    auto k = lds_decode_kernel<DMAX, DVMAX, NMS>;
    hipError_t err = set_lds(k, bytes);
    if (err != hipSuccess) return err;
    const uint32_t block = std::max<uint32_t>(64u, (shp.groups * shp.lanes + 63u) & ~63u);
    hipLaunchKernelGGL(k, dim3(frames), dim3(block), bytes, s, cv, shp, llr, frames, max_it, w, fb_it,
                       (int)fb_continue, msg, iters, conv, synd, fb, tc, tb, ti);
    return hipGetLastError();
}

// LLRs in the standard's bit order -> quasi-cyclic order of a DVB code:
// parity bit k + a + q*b moves to k + a*z + b (a < q, b < z).
__global__ void dvb_permute_kernel(CodeView cv, const float *__restrict__ llr, float *__restrict__ out,
                                   uint32_t frames) {
    const size_t total = (size_t)frames * cv.n;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const size_t f = t / cv.n;
        const uint32_t u = (uint32_t)(t - f * cv.n);
        uint32_t v = u;
        if (u >= cv.k) {
            const uint32_t w = u - cv.k, a = w / cv.z, b = w - a * cv.z;
            v = cv.k + a + cv.dvb_q * b;
        }
        out[t] = llr[f * cv.n + v];
    }
}

}  // namespace

hipError_t launch_dvb_permute(const CodeView &cv, const float *llr, float *out, uint32_t frames, hipStream_t s) {
    if (frames == 0) return hipSuccess;
    const size_t total = (size_t)frames * cv.n;
    const uint32_t grid = (uint32_t)std::min<size_t>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(dvb_permute_kernel, dim3(grid), dim3(256), 0, s, cv, llr, out, frames);
    return hipGetLastError();
}

size_t lds_path_bytes(const CodeView &cv) { return lds_decode_bytes(cv); }

// Groups of the LDS path: lanes = min(z, 512); automatic groups: the count up
// to 512 / lanes (and at most max(mb, nb)) with the fewest check plus
// variable passes, the smaller count on a tie.
LdsShape lds_shape(const CodeView &cv, int groups) {
    LdsShape shp;
    shp.lanes = std::min<uint32_t>(cv.z, 512u);
    const uint32_t gmax = std::max<uint32_t>(1u, std::min<uint32_t>(512u / shp.lanes, std::max(cv.mb, cv.nb)));
    if (groups > 0) {
        shp.groups = std::min<uint32_t>((uint32_t)groups, gmax);
        return shp;
    }
    uint32_t best = 1, best_passes = ~0u;
    for (uint32_t g = 1; g <= gmax; ++g) {
        const uint32_t passes = (cv.mb + g - 1) / g + (cv.nb + g - 1) / g;
        if (passes < best_passes) {
            best = g;
            best_passes = passes;
        }
    }
    shp.groups = best;
    return shp;
}

hipError_t launch_lds_decode(const CodeView &cv, const float *llr, uint32_t frames, uint32_t max_it,
                             const float *w, uint32_t fb_it, bool fb_continue, uint64_t *msg, uint32_t *iters,
                             uint8_t *conv, uint8_t *synd, uint8_t *fb, float *trace_c2v,
                             float *trace_bel, uint32_t trace_iters, int groups, hipStream_t s) {
    if (frames == 0) return hipSuccess;
    const LdsShape shp = lds_shape(cv, groups);
    const int dm = pick_dmax(cv.max_dc), dvm = pick_dvmax(cv.max_dv);
#define GF2L_LDS(DM, DV)                                                                                       \
    if (dm == DM && dvm == DV)                                                                                 \
        return w ? launch_lds_t<DM, DV, true>(cv, shp, llr, frames, max_it, w, fb_it, fb_continue, msg, iters,  \
                                              conv, synd, fb, trace_c2v, trace_bel, trace_iters, s)             \
                 : launch_lds_t<DM, DV, false>(cv, shp, llr, frames, max_it, w, fb_it, fb_continue, msg, iters, \
                                               conv, synd, fb, trace_c2v, trace_bel, trace_iters, s);
    GF2L_LDS(8, 4) GF2L_LDS(8, 8) GF2L_LDS(8, 16) GF2L_LDS(8, 32)
    GF2L_LDS(16, 4) GF2L_LDS(16, 8) GF2L_LDS(16, 16) GF2L_LDS(16, 32)
    GF2L_LDS(32, 4) GF2L_LDS(32, 8) GF2L_LDS(32, 16) GF2L_LDS(32, 32)
#undef GF2L_LDS
    return hipErrorInvalidValue;
}

hipError_t launch_hbm_init(const CodeView &cv, Workspace &ws, int path, const float *llr,
                           uint32_t frames, hipStream_t s) {
    if (path == kStream) {
        hipError_t err = hipMemcpyAsync(ws.belief, llr, (size_t)frames * cv.n * sizeof(float),
                                        hipMemcpyDeviceToDevice, s);
        if (err != hipSuccess) return err;
        err = hipMemsetAsync(ws.c2v, 0, (size_t)frames * cv.nbe * cv.z * sizeof(float), s);
        if (err != hipSuccess) return err;
    }
    hipLaunchKernelGGL(hbm_init_kernel, dim3((frames + 255) / 256), dim3(256), 0, s, ws.lists,
                       ws.counts, frames);
    return hipGetLastError();
}

hipError_t launch_hbm_iteration(const CodeView &cv, Workspace &ws, int path, const IterArgs &a,
                                int num_cus, const Tuning &tune, hipStream_t s, hipEvent_t *ev) {
    const uint32_t cur = a.cur & 1u, nxt = cur ^ 1u;
    const uint32_t *list = ws.lists + (size_t)cur * ws.cap_frames;
    uint32_t *next_list = ws.lists + (size_t)nxt * ws.cap_frames;
    const uint32_t *count = ws.counts + cur;
    uint32_t *next_count = ws.counts + nxt;
    // items from shared counters (tune.chunk per fetch), or a fixed stride;
    // automatic (tune.chunk < 0): counters with 8 items per fetch when the
    // code's variable degrees fit the smallest class (state path only)
    const int ch = path != kState ? 0 : tune.chunk >= 0 ? tune.chunk : (cv.max_dv <= 8 ? 8 : 0);
    const bool dyn = ch > 0;
    uint32_t *sched = ws.counts + kSchedOffset;
    const uint32_t chunk = dyn ? (uint32_t)ch : 1u;
    // counter-fed items go with the smallest degree class, whose variable kernel
    // is built for 8 waves per SIMD: 5 workgroups of 6 waves fit a CU
    const int wg = tune.wg_per_cu ? tune.wg_per_cu : (dyn ? 5 : 3);
    const uint32_t block = round_block(cv.z);
    const uint32_t g_cn = hbm_grid(a.frames * cv.mb, num_cus, wg);
    const uint32_t g_vn = hbm_grid(a.frames * cv.nb, num_cus, wg);
    const int xcd = tune.xcd_map;
    const uint32_t g_syn = a.frames < (uint32_t)num_cus * 8u ? a.frames : (uint32_t)num_cus * 8u;
    const int first = a.it == 1;

    if (ev) hipEventRecord(ev[0], s);
    const bool nms_prev = a.w_prev != nullptr, nms_cur = a.w_cur != nullptr;
    const bool wide_c = cv.max_dc > 16, wide_v = cv.max_dv > 16;
    if (path == kState && cv.dvb_q) {  // DVB codes: plain min-sum, lane rules of the standard's order
#define GF2L_CN_DYN(MD, FI, UNUSED, DY)                                                            \
    hipLaunchKernelGGL((cn_state_kernel<MD, FI, false, DY, true>), dim3(g_cn), dim3(block), 0, s, cv, list, \
                       count, ws.state, ws.state, ws.belief, a.llr, nullptr, xcd, sched, chunk)
#define GF2L_CN(MD, FI, NM) \
    do { if (dyn) GF2L_CN_DYN(MD, FI, NM, true); else GF2L_CN_DYN(MD, FI, NM, false); } while (0)
        if (first)
            GF2L_STATE_CLASS(cv.max_dc, GF2L_CN, true, false);
        else
            GF2L_STATE_CLASS(cv.max_dc, GF2L_CN, false, false);
#undef GF2L_CN
#undef GF2L_CN_DYN
        if (ev) hipEventRecord(ev[1], s);
#define GF2L_VN_DYN(MD, UNUSED, DY)                                                                \
    hipLaunchKernelGGL((vn_state_kernel<MD, false, DY, true>), dim3(g_vn), dim3(block), 0, s, cv, list, count, \
                       ws.state, ws.belief, a.llr, ws.hb, next_count, nullptr, xcd, sched, chunk)
#define GF2L_VN(MD, NM, UNUSED) \
    do { if (dyn) GF2L_VN_DYN(MD, NM, true); else GF2L_VN_DYN(MD, NM, false); } while (0)
        GF2L_VAR_CLASS(cv.max_dv, GF2L_VN, false, 0);
#undef GF2L_VN
#undef GF2L_VN_DYN
    } else if (path == kState) {
#define GF2L_CN_DYN(MD, FI, NM, DY)                                                                \
    hipLaunchKernelGGL((cn_state_kernel<MD, FI, NM, DY>), dim3(g_cn), dim3(block), 0, s, cv, list, count, \
                       ws.state, ws.state, ws.belief, a.llr, a.w_prev, xcd, sched, chunk)
#define GF2L_CN(MD, FI, NM) \
    do { if (dyn) GF2L_CN_DYN(MD, FI, NM, true); else GF2L_CN_DYN(MD, FI, NM, false); } while (0)
        if (first) {
            GF2L_STATE_CLASS(cv.max_dc, GF2L_CN, true, false);
        } else if (nms_prev) {
            GF2L_STATE_CLASS(cv.max_dc, GF2L_CN, false, true);
        } else {
            GF2L_STATE_CLASS(cv.max_dc, GF2L_CN, false, false);
        }
#undef GF2L_CN
#undef GF2L_CN_DYN
        if (ev) hipEventRecord(ev[1], s);
#define GF2L_VN_DYN(MD, NM, DY)                                                                    \
    hipLaunchKernelGGL((vn_state_kernel<MD, NM, DY>), dim3(g_vn), dim3(block), 0, s, cv, list, count, \
                       ws.state, ws.belief, a.llr, ws.hb, next_count, a.w_cur, xcd, sched, chunk)
#define GF2L_VN(MD, NM, UNUSED) \
    do { if (dyn) GF2L_VN_DYN(MD, NM, true); else GF2L_VN_DYN(MD, NM, false); } while (0)
        if (nms_cur) {
            GF2L_VAR_CLASS(cv.max_dv, GF2L_VN, true, 0);
        } else {
            GF2L_VAR_CLASS(cv.max_dv, GF2L_VN, false, 0);
        }
#undef GF2L_VN
#undef GF2L_VN_DYN
    } else {
#define GF2L_CN(MD, NM)                                                                        \
    hipLaunchKernelGGL((cn_stream_kernel<MD, NM>), dim3(g_cn), dim3(block), 0, s, cv, list, count, \
                       ws.c2v, ws.belief, a.w_cur, xcd)
#define GF2L_CN_SPA(MD)                                                                                \
    hipLaunchKernelGGL((cn_stream_spa_kernel<MD>), dim3(g_cn), dim3(block), 0, s, cv, list, count, ws.c2v, \
                       ws.belief, xcd)
        if (a.spa) {
            if (wide_c)
                GF2L_CN_SPA(32);
            else if (cv.max_dc > 8)
                GF2L_CN_SPA(16);
            else
                GF2L_CN_SPA(8);
        } else if (nms_cur) {
            if (wide_c) GF2L_CN(32, true); else GF2L_CN(16, true);
        } else {
            if (wide_c) GF2L_CN(32, false); else GF2L_CN(16, false);
        }
#undef GF2L_CN_SPA
#undef GF2L_CN
        if (ev) hipEventRecord(ev[1], s);
        if (wide_v)
            hipLaunchKernelGGL((vn_stream_kernel<32>), dim3(g_vn), dim3(block), 0, s, cv, list, count,
                               ws.c2v, ws.belief, a.llr, ws.hb, next_count, xcd);
        else
            hipLaunchKernelGGL((vn_stream_kernel<16>), dim3(g_vn), dim3(block), 0, s, cv, list, count,
                               ws.c2v, ws.belief, a.llr, ws.hb, next_count, xcd);
    }
    if (ev) hipEventRecord(ev[2], s);
    const size_t sh = (size_t)cv.nb * cv.zw * sizeof(uint64_t);
    hipLaunchKernelGGL(syn_words_kernel, dim3(g_syn), dim3(kSynThreads), sh, s, cv, list, count, next_list,
                       next_count, ws.hb, a.it_report, a.last, a.fb_phase,
                       a.itc_graph ? (const uint32_t *)a.itc : nullptr, a.itc_last, a.itc_fb_from, a.msg, a.iters,
                       a.conv, a.synd, a.fb);
    hipLaunchKernelGGL(list_compact_kernel, dim3(1), dim3(1024), 0, s, count, next_list, next_count, a.itc,
                       a.itc_graph ? 0u : a.it_report + 1u);
    if (ev) hipEventRecord(ev[3], s);
    return hipGetLastError();
}

hipError_t launch_state_forward(const CodeView &cv, const uint32_t *list, const uint32_t *count,
                                uint32_t frames, bool first, const uint32_t *state_in,
                                uint32_t *state_out, const float *belief_in, float *belief_out,
                                const float *llr, const float *w_prev, const float *w_cur,
                                int num_cus, const Tuning &tune, hipStream_t s) {
    const uint32_t block = round_block(cv.z);
    const int wg = tune.wg_per_cu ? tune.wg_per_cu : 3;
    const uint32_t g_cn = hbm_grid(frames * cv.mb, num_cus, wg);
    const uint32_t g_vn = hbm_grid(frames * cv.nb, num_cus, wg);
    const int xcd = tune.xcd_map;
    const bool nms_prev = w_prev != nullptr, nms_cur = w_cur != nullptr;
#define GF2L_CN(MD, FI, NM)                                                                    \
    hipLaunchKernelGGL((cn_state_kernel<MD, FI, NM, false>), dim3(g_cn), dim3(block), 0, s, cv, list, count, \
                       state_in, state_out, belief_in, llr, w_prev, xcd, nullptr, 1u)
    if (first) {
        GF2L_STATE_CLASS(cv.max_dc, GF2L_CN, true, false);
    } else if (nms_prev) {
        GF2L_STATE_CLASS(cv.max_dc, GF2L_CN, false, true);
    } else {
        GF2L_STATE_CLASS(cv.max_dc, GF2L_CN, false, false);
    }
#undef GF2L_CN
#define GF2L_VN(MD, NM, UNUSED)                                                                \
    hipLaunchKernelGGL((vn_state_kernel<MD, NM, false>), dim3(g_vn), dim3(block), 0, s, cv, list, count, \
                       state_out, belief_out, llr, nullptr, nullptr, w_cur, xcd, nullptr, 1u)
    if (nms_cur) {
        GF2L_VAR_CLASS(cv.max_dv, GF2L_VN, true, 0);
    } else {
        GF2L_VAR_CLASS(cv.max_dv, GF2L_VN, false, 0);
    }
#undef GF2L_VN
    return hipGetLastError();
}

hipError_t launch_state_expand(const CodeView &cv, const Workspace &ws, uint32_t frames,
                               const float *w, float *c2v_out, float *bel_out, hipStream_t s) {
    hipLaunchKernelGGL(state_expand_kernel, dim3(1024), dim3(256), 0, s, cv, ws.state, ws.belief,
                       frames, w, c2v_out, bel_out);
    return hipGetLastError();
}

}  // namespace gf2l
