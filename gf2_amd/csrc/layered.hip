// Layered (row-sequential) learned normalized min-sum decoder: forward
// (decoding, and the training forward with a tape) and backward kernels.
//
// The reference has no learned decoder (ROADMAP.md:93 plans one); the check
// rule is its alpha * min-sum (Llr::boxplus_normalized_minsum_n,
// crates/gf2-coding/src/llr.rs:352-355) with one alpha per base edge and
// iteration, on the min-sum of crates/gf2-kernels-simd/src/llr.rs:117-180
// (CheckAcc).  Schedule: the block-rows of a quasi-cyclic code are the
// layers, processed in ascending order; the z checks of a layer touch
// disjoint variables, so each one updates its variables' beliefs in place:
//   x_p    = b[v_p] - w_{t-1}(e_p) * m_{t-1}(e_p, i)      (x_p = b[v_p] at t = 1)
//   m_t(p) = sign * min over the other inputs of the check
//   b[v_p] = x_p + w_t(e_p) * m_t(p)
// After each iteration the hard decisions b < 0 are checked against every
// parity check (word-parallel, as the syndrome pass of the flooding
// decoder); a frame that satisfies them stops.  A frame failing after T
// iterations is handed to the plain flooding decoder, which continues from
// its beliefs and compressed check states (the first check update subtracts
// w_T * m_T).
//
// One workgroup owns one frame at a time (frames from an atomic counter).
// The beliefs of block-columns 0..X-1 live in LDS, the others in the frame's
// global belief rows; the compressed check states (min1, min2, argmin |
// output signs, the flooding decoder's 12-byte format) are in global memory,
// read and written once per check and iteration.  A barrier separates the
// layers.  The training forward runs all T iterations and writes a tape
// (per check and iteration: min1, min2, meta, second argmin and the input
// signs at the two argmins; beliefs after each iteration); the backward
// kernel walks the layers in reverse and accumulates the weight gradient.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "device_common.h"

namespace gf2l {
namespace {

// Beliefs of one frame: block-columns below X in LDS ([c][z]), the others in
// the frame's global rows ([c][zp]).  The column is wave-uniform at every
// access, so the choice is a uniform branch.
struct Beliefs {
    float *lds;
    float *glob;
    uint32_t X, z, zp;
    __device__ float ld(uint32_t c, uint32_t j) const { return c < X ? lds[c * z + j] : glob[c * zp + j]; }
    __device__ void st(uint32_t c, uint32_t j, float v) const {
        if (c < X)
            lds[c * z + j] = v;
        else
            glob[c * zp + j] = v;
    }
};

// Hard decisions b < 0 of every block-column as nb*zw words in LDS (bits
// at z and above zero).  Needs blockDim.x % 64 == 0.
__device__ __forceinline__ void hard_words(const CodeView &cv, const Beliefs &B, uint64_t *sh) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (uint32_t s = wave; s < cv.nb * cv.zw; s += nw) {
        const uint32_t c = s / cv.zw, j = (s - c * cv.zw) * 64u + lane;
        const bool neg = j < cv.z && B.ld(c, j) < 0.0f;
        const uint64_t m = __ballot(neg);
        if (lane == 0) sh[s] = m;
    }
}

// One layer (block-row br) of iteration t.  st: the block-row's check
// states; tape (TAPE only): its tape records of this iteration.
template <int MAXD, bool TAPE>
__device__ __forceinline__ void layer_fwd(const LayeredArgs &a, const Beliefs &B, uint32_t br, uint32_t t,
                                          uint3 *st, uint4 *tape) {
    const CodeView &cv = a.cv;
    const uint32_t e0 = uni(ldc(cv.row_ptr + br));
    const uint32_t d = uni(ldc(cv.row_ptr + br + 1)) - e0;
    dispatch_degree<MAXD>(d, [&](auto DC) {
        constexpr int D = decltype(DC)::value, DA = D > 0 ? D : 1;
        uint2 re[DA];
        float wt[DA], wp[DA];
        const float *w_t = a.w + (size_t)(t - 1) * cv.nbe + e0;
#pragma unroll
        for (int p = 0; p < D; ++p) {
            const uint2 r = ldc(cv.row_edge + e0 + p);
            re[p] = make_uint2(uni(r.x), uni(r.y));
            wt[p] = __uint_as_float(uni(__float_as_uint(ldc(w_t + p))));
            wp[p] = t >= 2 ? __uint_as_float(uni(__float_as_uint(ldc(w_t - cv.nbe + p)))) : 0.0f;
        }
        for (uint32_t i = threadIdx.x; i < cv.z; i += blockDim.x) {
            float m1o = 0.0f, m2o = 0.0f;
            uint32_t mo = 0;
            if (t >= 2) {
                const uint3 o = st[i];
                m1o = __uint_as_float(o.x);
                m2o = __uint_as_float(o.y);
                mo = o.z;
            }
            uint32_t jv[DA];
            float x[DA];
#pragma unroll
            for (int p = 0; p < D; ++p) {
                jv[p] = wrap_add(i, re[p].y, cv.z);
                const float bv = B.ld(re[p].x, jv[p]);
                x[p] = t >= 2 ? bv - wp[p] * recon(m1o, m2o, mo, p) : bv;
            }
            CheckAcc acc;
            acc.init();
            uint32_t idx2 = 0;
#pragma unroll
            for (int p = 0; p < D; ++p) {
                if constexpr (TAPE) {  // second argmin (first on ties), before the update
                    const float ax = fabsf(x[p]);
                    idx2 = ax < acc.min1 ? acc.idx : (ax < acc.min2 ? (uint32_t)p : idx2);
                }
                acc.add(p, x[p]);
            }
            float m1 = acc.min1, m2 = acc.min2;
            uint32_t meta = acc.idx | (acc.signs<D>(D) << 5);
            if constexpr (D == 1) {  // the lone message is +0.0
                m1 = 0.0f;
                m2 = 0.0f;
                meta = 0;
            }
#pragma unroll
            for (int p = 0; p < D; ++p) B.st(re[p].x, jv[p], x[p] + wt[p] * recon(m1, m2, meta, p));
            st[i] = make_uint3(__float_as_uint(m1), __float_as_uint(m2), meta);
            if constexpr (TAPE) {
                uint32_t n1 = 0, n2 = 0;  // input signs (x < 0) at the two argmins
#pragma unroll
                for (int p = 0; p < D; ++p) {
                    n1 |= ((uint32_t)p == acc.idx && x[p] < 0.0f) ? 1u : 0u;
                    n2 |= ((uint32_t)p == idx2 && x[p] < 0.0f) ? 1u : 0u;
                }
                tape[i] = make_uint4(__float_as_uint(m1), __float_as_uint(m2), meta, idx2 | (n1 << 5) | (n2 << 6));
            }
        }
    });
}

template <int MAXD, bool TAPE>
__global__ void __launch_bounds__(512) lnms_fwd_kernel(LayeredArgs a) {
    extern __shared__ uint64_t smem[];
    __shared__ uint32_t s_f;
    const CodeView &cv = a.cv;
    uint64_t *s_hb = smem;
    float *s_bel = reinterpret_cast<float *>(smem + cv.nb * cv.zw);
    const uint32_t kw = (cv.k + 63) / 64;
    for (;;) {
        __syncthreads();  // the previous frame is done with s_f and the LDS
        if (threadIdx.x == 0) s_f = atomicAdd(a.ctr, 1u);
        __syncthreads();
        const uint32_t f = s_f;
        if (f >= a.F) break;
        float *gb = a.belief + (size_t)f * cv.nb * cv.zp;
        const Beliefs B{s_bel, gb, a.X, cv.z, cv.zp};
        const float *fl = a.llr + (size_t)f * cv.n;
        for (uint32_t c = 0; c < cv.nb; ++c)  // beliefs start at the channel LLRs
            for (uint32_t j = threadIdx.x; j < cv.z; j += blockDim.x) B.st(c, j, fl[c * cv.z + j]);
        __syncthreads();
        uint3 *fst = reinterpret_cast<uint3 *>(a.state) + (size_t)f * cv.mb * cv.zp;
        int conv = 0;
        uint32_t t = 1;
        for (; t <= a.T; ++t) {
            for (uint32_t br = 0; br < cv.mb; ++br) {
                uint4 *tp = nullptr;
                if constexpr (TAPE)
                    tp = reinterpret_cast<uint4 *>(a.tape_state) + (((size_t)(t - 1) * a.F + f) * cv.mb + br) * cv.zp;
                layer_fwd<MAXD, TAPE>(a, B, br, t, fst + br * cv.zp, tp);
                __syncthreads();
            }
            if constexpr (TAPE) {  // beliefs after iteration t, all block-columns
                float *tb = a.tape_bel + ((size_t)(t - 1) * a.F + f) * cv.nb * cv.zp;
                for (uint32_t c = 0; c < cv.nb; ++c)
                    for (uint32_t j = threadIdx.x; j < cv.z; j += blockDim.x) tb[c * cv.zp + j] = B.ld(c, j);
                continue;
            }
            hard_words(cv, B, s_hb);
            __syncthreads();
            if (!words_unsat(cv, s_hb)) {
                conv = 1;
                break;
            }
        }
        if constexpr (TAPE) continue;
        const uint32_t its = conv ? t : a.T;
        if (threadIdx.x == 0) atomicAdd(a.ctr + 1, its);
        if (conv || !a.fallback) {
            frame_message(cv, s_hb, a.msg + (size_t)f * kw);
            if (threadIdx.x == 0) {
                a.iters[f] = its;
                a.conv[f] = (uint8_t)conv;
                a.synd[f] = (uint8_t)conv;
                if (a.fb) a.fb[f] = 0;
            }
        } else {
            // hand the frame to the plain continuation: every belief in its
            // global rows, the frame on the running list
            for (uint32_t c = 0; c < a.X; ++c)
                for (uint32_t j = threadIdx.x; j < cv.z; j += blockDim.x) gb[c * cv.zp + j] = s_bel[c * cv.z + j];
            if (threadIdx.x == 0) a.list[atomicAdd(a.count, 1u)] = f;
        }
    }
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Reverse of one layer: G (dL/db, LDS below X and global above) holds the
// gradient w.r.t. the beliefs after the layer and leaves the gradient
// w.r.t. the beliefs before it; H[e][i] holds dL/d(weighted message of
// iteration t) from iteration t+1 and leaves -dL/dx (the message of
// iteration t-1 was subtracted from the beliefs).  Weight gradients of the
// block-row's edges, reduced over the lanes in a fixed order, go to
// gpart[(t-1)*nbe + e][f].
template <int MAXD>
__device__ __forceinline__ void layer_bwd(const LayeredBwdArgs &a, const Beliefs &G, uint32_t f, uint32_t br,
                                          uint32_t t, const uint4 *tape, float *H, float (*red)[16]) {
    const CodeView &cv = a.cv;
    const uint32_t e0 = uni(ldc(cv.row_ptr + br));
    const uint32_t d = uni(ldc(cv.row_ptr + br + 1)) - e0;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
    dispatch_degree<MAXD>(d, [&](auto DC) {
        constexpr int D = decltype(DC)::value, DA = D > 0 ? D : 1;
        uint2 re[DA];
        float wt[DA], gw[DA];
        const float *w_t = a.w + (size_t)(t - 1) * cv.nbe + e0;
#pragma unroll
        for (int p = 0; p < D; ++p) {
            const uint2 r = ldc(cv.row_edge + e0 + p);
            re[p] = make_uint2(uni(r.x), uni(r.y));
            wt[p] = __uint_as_float(uni(__float_as_uint(ldc(w_t + p))));
            gw[p] = 0.0f;
        }
        for (uint32_t i = threadIdx.x; i < cv.z; i += blockDim.x) {
            const uint4 tp = tape[i];
            const float m1 = __uint_as_float(tp.x), m2 = __uint_as_float(tp.y);
            const uint32_t meta = tp.z, idx1 = meta & 31u, idx2 = tp.w & 31u;
            const float s1 = (tp.w >> 5) & 1u ? -1.0f : 1.0f, s2 = (tp.w >> 6) & 1u ? -1.0f : 1.0f;
            uint32_t jv[DA];
            float gb[DA], gc[DA];
#pragma unroll
            for (int p = 0; p < D; ++p) {
                jv[p] = wrap_add(i, re[p].y, cv.z);
                gb[p] = G.ld(re[p].x, jv[p]);
                gc[p] = gb[p] + (t < a.T ? H[(size_t)(e0 + p) * cv.z + i] : 0.0f);
            }
            float g1 = 0.0f, g2 = 0.0f;  // to the argmin and the second argmin
#pragma unroll
            for (int p = 0; p < D; ++p) {
                if constexpr (D >= 2) {
                    const float mp = recon(m1, m2, meta, p);
                    gw[p] += gc[p] * mp;
                    const float v = gc[p] * wt[p] * (((meta >> (5 + p)) & 1u) ? -1.0f : 1.0f);
                    if ((uint32_t)p == idx1)
                        g2 += v * s2;
                    else
                        g1 += v * s1;
                }
            }
#pragma unroll
            for (int p = 0; p < D; ++p) {
                float gx = gb[p];
                if ((uint32_t)p == idx1) gx += g1;
                if ((uint32_t)p == idx2 && D >= 2) gx += g2;
                G.st(re[p].x, jv[p], gx);
                if (t >= 2) H[(size_t)(e0 + p) * cv.z + i] = -gx;
            }
        }
        // fixed-order reduction of the weight gradients over the lanes
#pragma unroll
        for (int p = 0; p < D; ++p) {
            const float s = wave_sum(gw[p]);
            if (lane == 0) red[p][wave] = s;
        }
        __syncthreads();
        if (threadIdx.x < (uint32_t)D) {
            float s = 0.0f;
            for (uint32_t w = 0; w < nwave; ++w) s += red[threadIdx.x][w];
            a.gpart[((size_t)(t - 1) * cv.nbe + e0 + threadIdx.x) * a.F + f] = s;
        }
    });
}

template <int MAXD>
__global__ void __launch_bounds__(512) lnms_bwd_kernel(LayeredBwdArgs a) {
    extern __shared__ float sg[];
    __shared__ float red[MAXD][16];
    __shared__ uint32_t s_f;
    const CodeView &cv = a.cv;
    const uint32_t nw = (cv.n + 63) / 64;
    float lsum = 0.0f;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_f = atomicAdd(a.ctr, 1u);
        __syncthreads();
        const uint32_t f = s_f;
        if (f >= a.F) break;
        const Beliefs G{sg, a.g_glob + (size_t)f * cv.nb * cv.zp, a.X, cv.z, cv.zp};
        float *H = a.H + (size_t)f * cv.nbe * cv.z;
        const uint64_t *cw = a.cw + (size_t)f * nw;
        for (uint32_t c = 0; c < cv.nb; ++c)
            for (uint32_t j = threadIdx.x; j < cv.z; j += blockDim.x) G.st(c, j, 0.0f);
        for (uint32_t t = a.T; t >= 1; --t) {
            // BCE of the beliefs after iteration t
            const float *tb = a.tape_bel + ((size_t)(t - 1) * a.F + f) * cv.nb * cv.zp;
            for (uint32_t c = 0; c < cv.nb; ++c)
                for (uint32_t j = threadIdx.x; j < cv.z; j += blockDim.x) {
                    const uint32_t v = c * cv.z + j;
                    const float y = ((cw[v >> 6] >> (v & 63u)) & 1ull) ? -1.0f : 1.0f;
                    const float ys = y * tb[c * cv.zp + j];
                    lsum += ys > 0.0f ? log1pf(expf(-ys)) : -ys + log1pf(expf(ys));
                    G.st(c, j, G.ld(c, j) - y * a.scale / (1.0f + expf(ys)));
                }
            __syncthreads();
            for (uint32_t br = cv.mb; br-- > 0;) {
                const uint4 *tp =
                    reinterpret_cast<const uint4 *>(a.tape_state) + (((size_t)(t - 1) * a.F + f) * cv.mb + br) * cv.zp;
                layer_bwd<MAXD>(a, G, f, br, t, tp, H, red);
                __syncthreads();
            }
        }
    }
    lsum = wave_sum(lsum);
    if ((threadIdx.x & 63u) == 0) red[0][threadIdx.x >> 6] = lsum;
    __syncthreads();
    if (threadIdx.x == 0) {
        float s = 0.0f;
        for (uint32_t w = 0; w < blockDim.x / 64; ++w) s += red[0][w];
        atomicAdd(a.loss, s);
    }
}

template <typename K>
hipError_t allow_lds(K kernel, size_t bytes) {
    if (bytes <= 65536) return hipSuccess;
    return hipFuncSetAttribute(reinterpret_cast<const void *>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)bytes);
}

uint32_t layered_block(uint32_t z) {
    const uint32_t b = ((z + 63) / 64) * 64;
    return b > 512 ? 512 : b;
}

}  // namespace

size_t layered_lds_bytes(const CodeView &cv, uint32_t X) {
    return (size_t)cv.nb * cv.zw * sizeof(uint64_t) + (size_t)X * cv.z * sizeof(float);
}

uint32_t layered_lds_columns(const CodeView &cv, size_t lds_limit) {
    const size_t fixed = (size_t)cv.nb * cv.zw * sizeof(uint64_t) + 32 * 16 * sizeof(float) + 64;
    if (lds_limit <= fixed) return 0;
    const size_t cols = (lds_limit - fixed) / ((size_t)cv.z * sizeof(float));
    return (uint32_t)std::min<size_t>(cols, cv.nb);
}

hipError_t launch_layered_fwd(const LayeredArgs &a, bool tape, int num_cus, hipStream_t s) {
    const size_t bytes = layered_lds_bytes(a.cv, a.X);
    const uint32_t block = layered_block(a.cv.z);
    const uint32_t per_cu = (uint32_t)std::max<size_t>(1, std::min<size_t>(8, (160 * 1024) / std::max<size_t>(bytes, 1)));
    const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(a.F, (uint32_t)num_cus * per_cu));
#define GF2L_LFWD(MD, TP)                                                               \
    do {                                                                                \
        auto k = lnms_fwd_kernel<MD, TP>;                                               \
        hipError_t e = allow_lds(k, bytes);                                             \
        if (e != hipSuccess) return e;                                                  \
        hipLaunchKernelGGL(k, dim3(grid), dim3(block), bytes, s, a);                    \
    } while (0)
    if (a.cv.max_dc <= 8) {
        if (tape) GF2L_LFWD(8, true); else GF2L_LFWD(8, false);
    } else if (a.cv.max_dc <= 16) {
        if (tape) GF2L_LFWD(16, true); else GF2L_LFWD(16, false);
    } else {
        if (tape) GF2L_LFWD(27, true); else GF2L_LFWD(27, false);
    }
#undef GF2L_LFWD
    return hipGetLastError();
}

hipError_t launch_layered_bwd(const LayeredBwdArgs &a, int num_cus, hipStream_t s) {
    const size_t bytes = (size_t)a.X * a.cv.z * sizeof(float);
    const uint32_t block = layered_block(a.cv.z);
    const uint32_t per_cu = (uint32_t)std::max<size_t>(1, std::min<size_t>(8, (160 * 1024) / std::max<size_t>(bytes + 2048, 1)));
    const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(a.F, (uint32_t)num_cus * per_cu));
#define GF2L_LBWD(MD)                                                                   \
    do {                                                                                \
        auto k = lnms_bwd_kernel<MD>;                                                   \
        hipError_t e = allow_lds(k, bytes);                                             \
        if (e != hipSuccess) return e;                                                  \
        hipLaunchKernelGGL(k, dim3(grid), dim3(block), bytes, s, a);                    \
    } while (0)
    if (a.cv.max_dc <= 8)
        GF2L_LBWD(8);
    else if (a.cv.max_dc <= 16)
        GF2L_LBWD(16);
    else
        GF2L_LBWD(27);
#undef GF2L_LBWD
    return hipGetLastError();
}

}  // namespace gf2l
