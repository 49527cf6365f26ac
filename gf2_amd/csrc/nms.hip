// Learned (neural) normalized min-sum decoder: per-iteration, per-base-edge
// weights, trained on the device with hand-written forward and backward
// kernels; decoding runs the weighted iterations and re-decodes frames that
// still fail the syndrome with plain min-sum.
//
// The check rule is the reference's alpha * min-sum (Llr::boxplus_normalized_
// minsum_n, crates/gf2-coding/src/llr.rs:352-355) with one alpha per base
// edge and iteration: c2v_t(e) = w_t(e) * sgn * min_{e' != e} |v2c_{t-1}(e')|.
// Everything else is the reference flooding schedule (core.rs:889-938).
//
// Loss: mean over frames, iterations t = 1..T and variables v of
// softplus(-y_v * belief_t[v]), y = +1 for bit 0 and -1 for bit 1: the BCE of
// P(bit = 1) = sigmoid(-belief) under the LLR sign of llr.rs:9-15.
// Min-sum is differentiated as a subgradient: only the extrinsic argmin of
// each message (first position on ties) receives gradient; signs are
// piecewise constant.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "device_common.h"
#include "gf2ldpc.h"

struct gf2l_nms {
    const gf2l_code *code = nullptr;
    uint32_t T = 0;           // weighted iterations
    uint32_t nbe = 0;         // weights per iteration (base-graph edges)
    bool layered = false;     // layered schedule (gf2l_nms_create_layered), else flooding
    bool fb_continue = false; // the plain fallback continues from the learned state (layered: always)
    uint32_t X = 0;           // layered: block-columns whose beliefs a workgroup keeps in LDS
    float *d_w = nullptr;     // [T][nbe]
    uint32_t *d_zero_state = nullptr;  // layered: [mb][zp][3] zeros, iteration 1's check states
    float *d_m = nullptr;     // Adam first moment
    float *d_v = nullptr;     // Adam second moment
    uint64_t step = 0;
    // training workspace, grown on demand
    size_t cap = 0;
    uint32_t *tape_state = nullptr;  // [T][F][mb][zp][3] (layered: [4])
    float *tape_bel = nullptr;       // [T][F][nb][zp]
    float *g_a = nullptr, *g_b = nullptr;  // [F][nbe][z] dL/dv2c (layered: g_a carries dL/dmessage)
    float *g_bel = nullptr;          // [F][nb][zp]
    float *gpart = nullptr;          // [T][nbe][F]
    float *grad = nullptr;           // [T][nbe]
    float *loss = nullptr;           // 1
    float *lpart = nullptr;          // loss partials, summed in a fixed order
    uint32_t *list = nullptr;        // [F] identity list
    uint32_t *count = nullptr;       // 4 words: running count; layered frame counters
    uint32_t *l_state = nullptr;     // layered training forward: [F][mb][zp][3] check states
    float *l_bel = nullptr;          // layered training forward: [F][nb][zp] beliefs outside LDS
};

namespace gf2l {
namespace {

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Variable pass of iteration t: dL/dbelief_t[v] = BCE'(belief_t[v]) * scale
// + sum of dL/dv2c_t over v's edges (absent at t = T).  Also writes the loss
// of the items this workgroup took (a fixed stride over the grid) to
// lpart[blockIdx.x], for a fixed-order sum.
template <int MAXD>
__global__ void __launch_bounds__(512)
vn_bwd_kernel(CodeView cv, uint32_t frames, const float *__restrict__ bel_t,
              const uint64_t *__restrict__ cw, const float *__restrict__ g_cur, int has_gcur,
              float *__restrict__ g_bel, float scale, float *__restrict__ lpart) {
    __shared__ float red[8];
    const uint32_t items = frames * cv.nb;
    const uint32_t nw = (cv.n + 63) / 64;
    const size_t E = (size_t)cv.nbe * cv.z;
    float lsum = 0.0f;
    for (uint32_t item = blockIdx.x; item < items; item += gridDim.x) {
        const uint32_t f = item / cv.nb, bc = item % cv.nb;
        const uint32_t c0 = cv.col_ptr[bc], dv = cv.col_ptr[bc + 1] - c0;
        const float *fg = g_cur + (size_t)f * E;
        dispatch_degree<MAXD>(dv, [&](auto DC) {
            constexpr int D = decltype(DC)::value, DA = D > 0 ? D : 1;
            uint2 ce[DA];
#pragma unroll
            for (int q = 0; q < D; ++q) ce[q] = cv.col_edge[c0 + q];
            for (uint32_t j = threadIdx.x; j < cv.z; j += blockDim.x) {
                const size_t bi = ((size_t)f * cv.nb + bc) * cv.zp + j;
                float gin[DA];
                if (has_gcur) {
#pragma unroll
                    for (int q = 0; q < D; ++q)
                        gin[q] = fg[(size_t)ce[q].x * cv.z + wrap_sub(j, ce[q].y, cv.z)];
                }
                const float s = bel_t[bi];
                const uint32_t v = bc * cv.z + j;
                const float y = ((cw[(size_t)f * nw + (v >> 6)] >> (v & 63u)) & 1ull) ? -1.0f : 1.0f;
                const float ys = y * s;
                lsum += ys > 0.0f ? log1pf(expf(-ys)) : -ys + log1pf(expf(ys));
                float g = -y * scale / (1.0f + expf(ys));  // d softplus(-y s) / ds
                if (has_gcur) {
#pragma unroll
                    for (int q = 0; q < D; ++q) g += gin[q];
                }
                g_bel[bi] = g;
            }
        });
    }
    lsum = wave_sum(lsum);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = lsum;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.0f;
        for (uint32_t w = 0; w < blockDim.x / 64; ++w) t += red[w];
        lpart[blockIdx.x] = t;
    }
}

// Check pass of iteration t (1-based).  Recomputes the check's inputs
// v2c_{t-1} from the tape exactly as the forward pass did, then
//   g_c2v(p) = dL/dbelief_t[v_p] - dL/dv2c_t(p)
//   dL/dw_t(e_p) += g_c2v(p) * m_t(p)            (m = unweighted min-sum value)
//   dL/dv2c_{t-1}(a) += g_c2v(p) * w_t(e_p) * sgn_p * sign(v2c_{t-1}(a)),
//       a = argmin over the other inputs (t >= 2 only; v2c_0 = LLR)
// and writes the per-frame weight gradients of the block-row's edges,
// reduced over the z lanes in a fixed order, to gpart[e][frame].
template <int MAXD>
__global__ void __launch_bounds__(512)
cn_bwd_kernel(CodeView cv, uint32_t frames, uint32_t t, const float *__restrict__ llr,
              const uint32_t *__restrict__ st_prev, const float *__restrict__ bel_prev,
              const float *__restrict__ w_prev, const float *__restrict__ w_t,
              const float *__restrict__ g_bel, const float *__restrict__ g_cur, int has_gcur,
              float *__restrict__ g_prev, float *__restrict__ gpart) {
    __shared__ float red[MAXD > 0 ? MAXD : 1][8];
    const uint32_t items = frames * cv.mb;
    const size_t fstride = 3ull * cv.mb * cv.zp;
    const size_t E = (size_t)cv.nbe * cv.z;
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nwave = blockDim.x >> 6;
    for (uint32_t item = blockIdx.x; item < items; item += gridDim.x) {
        const uint32_t f = item / cv.mb, br = item % cv.mb;
        const uint32_t e0 = cv.row_ptr[br], d = cv.row_ptr[br + 1] - e0;
        dispatch_degree<MAXD>(d, [&](auto DC) {
            constexpr int D = decltype(DC)::value, DA = D > 0 ? D : 1;
            uint2 re[DA];
            float wp[DA], wt[DA], gw[DA];
#pragma unroll
            for (int p = 0; p < D; ++p) {
                re[p] = cv.row_edge[e0 + p];
                wp[p] = (t >= 2) ? w_prev[e0 + p] : 0.0f;
                wt[p] = w_t[e0 + p];
                gw[p] = 0.0f;
            }
            const uint32_t *si = st_prev + f * fstride + 3ull * br * cv.zp;
            for (uint32_t i = threadIdx.x; i < cv.z; i += blockDim.x) {
                float x[DA], gb[DA], gc[DA];
                float m1o = 0.0f, m2o = 0.0f;
                uint32_t mo = 0;
                if (t >= 2) {
                    m1o = __uint_as_float(si[3 * i]);
                    m2o = __uint_as_float(si[3 * i + 1]);
                    mo = si[3 * i + 2];
                }
#pragma unroll
                for (int p = 0; p < D; ++p) {
                    const uint32_t col = wrap_add(i, re[p].y, cv.z);
                    const size_t bi = ((size_t)f * cv.nb + re[p].x) * cv.zp + col;
                    if (t >= 2) x[p] = bel_prev[bi] - wp[p] * recon(m1o, m2o, mo, p);
                    else x[p] = llr[(size_t)f * cv.n + (size_t)re[p].x * cv.z + col];
                    gb[p] = g_bel[bi];
                    gc[p] = has_gcur ? g_cur[(size_t)f * E + (size_t)(e0 + p) * cv.z + i] : 0.0f;
                }
                // argmin, second argmin and input signs
                float min1 = __builtin_inff(), min2 = __builtin_inff();
                uint32_t idx1 = 0, idx2 = 0, neg = 0;
#pragma unroll
                for (int p = 0; p < D; ++p) {
                    const float a = fabsf(x[p]);
                    if (a < min1) {
                        min2 = min1;
                        idx2 = idx1;
                        min1 = a;
                        idx1 = p;
                    } else if (a < min2) {
                        min2 = a;
                        idx2 = p;
                    }
                    neg |= (uint32_t)(x[p] < 0.0f) << p;
                }
                const uint32_t par = __popc(neg) & 1u;
                float gx[DA];
#pragma unroll
                for (int p = 0; p < D; ++p) gx[p] = 0.0f;
#pragma unroll
                for (int p = 0; p < D; ++p) {
                    const float sgn = (par ^ ((neg >> p) & 1u)) ? -1.0f : 1.0f;  // product of the others
                    const float g_c2v = gb[p] - gc[p];
                    if constexpr (D >= 2) {
                        const float m = sgn * ((p == (int)idx1) ? min2 : min1);
                        gw[p] += g_c2v * m;
                        if (t >= 2) {
                            const uint32_t a = (p == (int)idx1) ? idx2 : idx1;
                            const float sa = ((neg >> a) & 1u) ? -1.0f : 1.0f;
                            const float gv = g_c2v * wt[p] * sgn * sa;
#pragma unroll
                            for (int q = 0; q < D; ++q)
                                if ((uint32_t)q == a) gx[q] += gv;
                        }
                    }
                }
                if (t >= 2) {
#pragma unroll
                    for (int p = 0; p < D; ++p) g_prev[(size_t)f * E + (size_t)(e0 + p) * cv.z + i] = gx[p];
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
                gpart[(size_t)(e0 + threadIdx.x) * frames + f] = s;
            }
            __syncthreads();
        });
    }
}

// grad[te] = sum over frames of gpart[te][f], fixed-order tree.
__global__ void __launch_bounds__(256)
grad_reduce_kernel(const float *__restrict__ gpart, uint32_t frames, float *__restrict__ grad) {
    __shared__ float sh[256];
    const float *row = gpart + (size_t)blockIdx.x * frames;
    float s = 0.0f;
    for (uint32_t f = threadIdx.x; f < frames; f += blockDim.x) s += row[f];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t o = blockDim.x / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) grad[blockIdx.x] = sh[0];
}

// Adam (Kingma & Ba 2015), bias-corrected, beta1 0.9, beta2 0.999, eps 1e-8.
__global__ void adam_kernel(float *__restrict__ w, float *__restrict__ m, float *__restrict__ v,
                            const float *__restrict__ g, uint32_t n, float lr, float bc1, float bc2) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = 0.9f * m[i] + 0.1f * gi;
    const float vi = 0.999f * v[i] + 0.001f * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + 1e-8f);
}

__global__ void iota_kernel(uint32_t *list, uint32_t *count, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) list[i] = i;
    if (i == 0) *count = n;
}

uint32_t bwd_block(uint32_t z) {
    uint32_t b = ((z + 63) / 64) * 64;
    return b > 512 ? 512 : b;
}

void free_train(gf2l_nms *q) {
    hipFree(q->tape_state);
    hipFree(q->tape_bel);
    hipFree(q->g_a);
    hipFree(q->g_b);
    hipFree(q->g_bel);
    hipFree(q->gpart);
    hipFree(q->grad);
    hipFree(q->loss);
    hipFree(q->lpart);
    hipFree(q->list);
    hipFree(q->count);
    hipFree(q->l_state);
    hipFree(q->l_bel);
    q->tape_state = nullptr;
    q->tape_bel = nullptr;
    q->g_a = q->g_b = q->g_bel = q->gpart = q->grad = q->loss = q->lpart = nullptr;
    q->list = q->count = nullptr;
    q->l_state = nullptr;
    q->l_bel = nullptr;
    q->cap = 0;
}

// Loss partials: one per frame (layered backward) or one per workgroup and
// iteration (flooding variable pass, at most 4 workgroups per CU).
size_t loss_parts(const gf2l_ctx *ctx, size_t T, size_t F) {
    return std::max<size_t>(F, T * (size_t)ctx->num_cus * 4);
}

int ensure_train(gf2l_ctx *ctx, gf2l_nms *q, size_t F) {
    if (q->cap >= F) return GF2L_OK;
    free_train(q);
    const CodeView &cv = q->code->view;
    const size_t T = q->T, E = (size_t)cv.nbe * cv.z;
    const size_t rec = q->layered ? 4 : 3;  // words per tape record of a check
    GF2L_HIP(ctx, hipMalloc(&q->tape_state, T * F * rec * cv.mb * cv.zp * sizeof(uint32_t)));
    GF2L_HIP(ctx, hipMalloc(&q->tape_bel, T * F * cv.nb * cv.zp * sizeof(float)));
    GF2L_HIP(ctx, hipMalloc(&q->g_a, F * E * sizeof(float)));
    if (!q->layered) GF2L_HIP(ctx, hipMalloc(&q->g_b, F * E * sizeof(float)));
    GF2L_HIP(ctx, hipMalloc(&q->g_bel, F * cv.nb * cv.zp * sizeof(float)));
    GF2L_HIP(ctx, hipMalloc(&q->gpart, T * cv.nbe * F * sizeof(float)));
    GF2L_HIP(ctx, hipMalloc(&q->grad, T * cv.nbe * sizeof(float)));
    GF2L_HIP(ctx, hipMalloc(&q->loss, sizeof(float)));
    GF2L_HIP(ctx, hipMalloc(&q->lpart, loss_parts(ctx, T, F) * sizeof(float)));
    GF2L_HIP(ctx, hipMalloc(&q->list, F * sizeof(uint32_t)));
    GF2L_HIP(ctx, hipMalloc(&q->count, 4 * sizeof(uint32_t)));
    if (q->layered) {
        GF2L_HIP(ctx, hipMalloc(&q->l_state, F * 3 * cv.mb * cv.zp * sizeof(uint32_t)));
        GF2L_HIP(ctx, hipMalloc(&q->l_bel, F * cv.nb * cv.zp * sizeof(float)));
    }
    q->cap = F;
    return GF2L_OK;
}

// Block-columns of beliefs the layered kernels keep in LDS.  Default: as
// many as fit in a third of the LDS, so three frames share a compute unit
// and hide each other's memory latency (measured on the N=64800 code: 30
// columns and three workgroups per CU decode 1.5x faster than 107 columns
// and one); GF2L_TUNE_LAYERED_COLUMNS sets it (up to what fits).
uint32_t layered_x(const gf2l_ctx *ctx, const gf2l_nms *q) {
    const int t = ctx->tune.layered_cols;
    if (t > 0) return std::min<uint32_t>((uint32_t)t, q->X);
    return std::min<uint32_t>(q->X, layered_lds_columns(q->code->view, ctx->lds_limit / 3));
}

// Kernel of a layered decode: the staircase kernel's row degree, 0 for the
// general kernel, -1 when GF2L_TUNE_LAYERED_KERNEL asks for the staircase
// kernel and the code (or the LDS) does not allow it.  Automatic choice: the
// staircase kernel unless a column count was set for the general one.
int layered_stair(const gf2l_ctx *ctx, const gf2l_nms *q) {
    const int k = ctx->tune.layered_kernel;
    if (k == 1 || (k == 0 && ctx->tune.layered_cols > 0)) return 0;
    const uint32_t dh = layered_stair_degree(q->code);
    if (dh && layered_stair_bytes(q->code->view) <= ctx->lds_limit) return (int)dh;
    return k == 2 ? -1 : 0;
}

// Block-columns of beliefs a layered decode keeps in LDS.
uint32_t layered_decode_x(const gf2l_ctx *ctx, const gf2l_nms *q) {
    const CodeView &cv = q->code->view;
    return layered_stair(ctx, q) > 0 ? cv.nb - cv.mb : layered_x(ctx, q);
}

// Layered schedule: training forward (tape, all T iterations) and backward;
// leaves the summed gradient in q->grad and the summed loss in q->loss.
int layered_forward_backward(gf2l_ctx *ctx, gf2l_nms *q, const float *d_llr, const uint64_t *d_cw, size_t frames) {
    const CodeView &cv = q->code->view;
    if (frames == 0 || frames > 0xffffffffull) return fail(ctx, GF2L_ERR_INVALID, "bad frame count");
    const uint32_t F = (uint32_t)frames, T = q->T;
    int rc = ensure_train(ctx, q, F);
    if (rc != GF2L_OK) return rc;
    hipStream_t s = ctx->stream;
    GF2L_HIP(ctx, hipMemsetAsync(q->count, 0, 4 * sizeof(uint32_t), s));
    LayeredArgs a{};
    a.cv = cv;
    a.llr = d_llr;
    a.w = q->d_w;
    a.F = F;
    a.T = T;
    a.X = layered_x(ctx, q);
    a.fallback = 0;
    a.state = q->l_state;
    a.belief = q->l_bel;
    a.ctr = q->count;
    a.list = q->list;
    a.count = q->count + 3;
    a.tape_state = q->tape_state;
    a.tape_bel = q->tape_bel;
    GF2L_HIP(ctx, launch_layered_fwd(a, true, ctx->num_cus, s));
    LayeredBwdArgs b{};
    b.cv = cv;
    b.w = q->d_w;
    b.F = F;
    b.T = T;
    b.X = a.X;
    b.tape_state = q->tape_state;
    b.tape_bel = q->tape_bel;
    b.cw = d_cw;
    b.g_glob = q->g_bel;
    b.H = q->g_a;
    b.gpart = q->gpart;
    b.scale = 1.0f / ((float)T * (float)cv.n * (float)F);
    b.lpart = q->lpart;
    b.ctr = q->count + 2;
    GF2L_HIP(ctx, launch_layered_bwd(b, ctx->num_cus, s));
    hipLaunchKernelGGL(grad_reduce_kernel, dim3(T * cv.nbe), dim3(256), 0, s, q->gpart, F, q->grad);
    hipLaunchKernelGGL(grad_reduce_kernel, dim3(1), dim3(256), 0, s, q->lpart, F, q->loss);
    GF2L_HIP(ctx, hipGetLastError());
    return GF2L_OK;
}

__global__ void clear_words_kernel(uint32_t *p, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

// Layered decode: the layered kernel over all frames (early stop after any
// iteration), then plain flooding min-sum continued from the layered state
// for the frames it handed over, up to fb_it iterations (reported as T + u;
// d_fb marks them).  Device buffers; synchronizes the context's stream.
int layered_decode(gf2l_ctx *ctx, const gf2l_nms *q, const float *d_llr, uint32_t F, uint32_t fb_it,
                   uint64_t *d_msg, uint32_t *d_iters, uint8_t *d_conv, uint8_t *d_synd, uint8_t *d_fb) {
    const CodeView &cv = q->code->view;
    const int dh = layered_stair(ctx, q);
    if (dh < 0)
        return fail(ctx, GF2L_ERR_UNSUPPORTED, "the staircase layered kernel does not apply to this code");
    int rc = ensure_buffers(ctx, F, 0, 3ull * cv.mb * cv.zp * F, (size_t)cv.nb * cv.zp * F,
                            (size_t)cv.nb * cv.zw * F, 0);
    if (rc != GF2L_OK) return rc;
    Workspace &ws = ctx->ws;
    Profile &pf = ctx->prof;
    hipStream_t s = ctx->stream;
    // running counts, item counters and the layered frame counters start at zero
    hipLaunchKernelGGL(clear_words_kernel, dim3((kCountsWords + 255) / 256), dim3(256), 0, s, ws.counts, kCountsWords);
    GF2L_HIP(ctx, hipGetLastError());
    LayeredArgs a{};
    a.cv = cv;
    a.llr = d_llr;
    a.w = q->d_w;
    a.F = F;
    a.T = q->T;
    a.X = dh > 0 ? cv.nb - cv.mb : layered_x(ctx, q);
    if (dh > 0) layered_stair_degree(q->code, a.stair_deg, a.stair_sync);
    a.fallback = fb_it > 0;
    a.state = ws.state;
    a.zero_state = reinterpret_cast<const uint3 *>(q->d_zero_state);
    a.belief = ws.belief;
    a.ctr = ws.counts + 2;
    a.list = ws.lists;  // list 0: read by the first continuation iteration
    a.count = ws.counts;
    a.msg = d_msg;
    a.iters = d_iters;
    a.conv = d_conv;
    a.synd = d_synd;
    a.fb = d_fb;
    if (pf.on) GF2L_HIP(ctx, hipEventRecord(pf.lev[0], s));
    if (dh > 0)
        GF2L_HIP(ctx, launch_layered_stair(a, (uint32_t)dh, ctx->num_cus, s));
    else
        GF2L_HIP(ctx, launch_layered_fwd(a, false, ctx->num_cus, s));
    if (pf.on) GF2L_HIP(ctx, hipEventRecord(pf.lev[1], s));
    // frames handed over, and the layered frame-iterations run
    GF2L_HIP(ctx, hipMemcpyAsync(ws.h_count, ws.counts, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    GF2L_HIP(ctx, hipStreamSynchronize(s));
    uint32_t running = ws.h_count[0];
    if (pf.on) {
        float ms = 0.0f;
        GF2L_HIP(ctx, hipEventElapsedTime(&ms, pf.lev[0], pf.lev[1]));
        pf.ms[3] += ms;
        pf.launches[3] += 1;
        pf.layered_frame_iters += ws.h_count[3];
    }
    if (!running) return GF2L_OK;
    // plain flooding continuation over list 0 (never the first-iteration
    // kernels: beliefs and check states come from the layered decoder, whose
    // last messages were weighted by w_T)
    const uint32_t batch = (uint32_t)std::max(1, std::min(ctx->tune.graph_iters, kProfileSlots));
    const bool ph = pf.on && pf.all;  // time the continuation's launches
    uint32_t queued = 0;
    for (uint32_t u = 1; u <= fb_it; ++u) {
        IterArgs it{};
        it.llr = d_llr;
        it.frames = F;
        it.it = 2;
        it.cur = (u - 1) & 1u;
        it.it_report = q->T + u;
        it.last = u == fb_it;
        it.fb_phase = 1;
        it.w_prev = u == 1 ? q->d_w + (size_t)(q->T - 1) * cv.nbe : nullptr;
        it.w_cur = nullptr;
        it.msg = d_msg;
        it.iters = d_iters;
        it.conv = d_conv;
        it.synd = d_synd;
        it.fb = d_fb;
        GF2L_HIP(ctx, launch_hbm_iteration(cv, ws, kState, it, ctx->num_cus, ctx->tune, s,
                                           ph ? pf.ev + 4 * queued : nullptr));
        const bool readback = queued + 1 >= batch || u == fb_it;
        if (ph || readback)
            GF2L_HIP(ctx, hipMemcpyAsync(ws.h_count + (ph ? queued : 0), ws.counts + (u & 1u), sizeof(uint32_t),
                                         hipMemcpyDeviceToHost, s));
        ++queued;
        if (!readback) continue;
        GF2L_HIP(ctx, hipStreamSynchronize(s));
        if (ph) {
            for (uint32_t b = 0; b < queued; ++b)
                for (int k = 0; k < 3; ++k) {
                    float ms = 0.0f;
                    GF2L_HIP(ctx, hipEventElapsedTime(&ms, pf.ev[4 * b + k], pf.ev[4 * b + k + 1]));
                    pf.ms[k] += ms;
                    pf.launches[k] += 1;
                }
            pf.frame_iters += running;
            for (uint32_t b = 0; b + 1 < queued; ++b) pf.frame_iters += ws.h_count[b];
        }
        running = ws.h_count[ph ? queued - 1 : 0];
        queued = 0;
        if (!running) break;
    }
    return GF2L_OK;
}

// Forward pass with tape + backward pass; leaves the summed gradient in
// q->grad and the summed loss in q->loss.
int forward_backward(gf2l_ctx *ctx, gf2l_nms *q, const float *d_llr, const uint64_t *d_cw, size_t frames) {
    if (q->layered) return layered_forward_backward(ctx, q, d_llr, d_cw, frames);
    const CodeView &cv = q->code->view;
    if (cv.max_dc > kMaxStateDegree) return fail(ctx, GF2L_ERR_UNSUPPORTED, "check degree above 27");
    if (frames == 0 || frames > 0xffffffffull) return fail(ctx, GF2L_ERR_INVALID, "bad frame count");
    const uint32_t F = (uint32_t)frames, T = q->T;
    int rc = ensure_train(ctx, q, F);
    if (rc != GF2L_OK) return rc;
    hipStream_t s = ctx->stream;
    const size_t sstride = (size_t)F * 3 * cv.mb * cv.zp, bstride = (size_t)F * cv.nb * cv.zp;
    hipLaunchKernelGGL(iota_kernel, dim3((F + 255) / 256), dim3(256), 0, s, q->list, q->count, F);
    GF2L_HIP(ctx, hipGetLastError());
    for (uint32_t t = 1; t <= T; ++t) {
        const uint32_t *st_in = t >= 2 ? q->tape_state + (t - 2) * sstride : nullptr;
        const float *bel_in = t >= 2 ? q->tape_bel + (t - 2) * bstride : nullptr;
        GF2L_HIP(ctx, launch_state_forward(cv, q->list, q->count, F, t == 1, st_in,
                                           q->tape_state + (t - 1) * sstride, bel_in,
                                           q->tape_bel + (t - 1) * bstride, d_llr,
                                           t >= 2 ? q->d_w + (size_t)(t - 2) * cv.nbe : nullptr,
                                           q->d_w + (size_t)(t - 1) * cv.nbe, ctx->num_cus, ctx->tune, s));
    }
    const uint32_t block = bwd_block(cv.z);
    const uint32_t gv = std::min<uint32_t>(F * cv.nb, (uint32_t)ctx->num_cus * 4u);
    const uint32_t gc = std::min<uint32_t>(F * cv.mb, (uint32_t)ctx->num_cus * 4u);
    const float scale = 1.0f / ((float)T * (float)cv.n * (float)F);
    float *g_cur = q->g_a, *g_prev = q->g_b;
    for (uint32_t t = T; t >= 1; --t) {
        const int has_gcur = t < T;
        const float *bel_t = q->tape_bel + (t - 1) * bstride;
        if (cv.max_dv > 16)
            hipLaunchKernelGGL(vn_bwd_kernel<32>, dim3(gv), dim3(block), 0, s, cv, F, bel_t, d_cw, g_cur,
                               has_gcur, q->g_bel, scale, q->lpart + (size_t)(t - 1) * gv);
        else
            hipLaunchKernelGGL(vn_bwd_kernel<16>, dim3(gv), dim3(block), 0, s, cv, F, bel_t, d_cw, g_cur,
                               has_gcur, q->g_bel, scale, q->lpart + (size_t)(t - 1) * gv);
        GF2L_HIP(ctx, hipGetLastError());
        const uint32_t *st_prev = t >= 2 ? q->tape_state + (t - 2) * sstride : nullptr;
        const float *bel_prev = t >= 2 ? q->tape_bel + (t - 2) * bstride : nullptr;
        const float *w_prev = t >= 2 ? q->d_w + (size_t)(t - 2) * cv.nbe : nullptr;
        const float *w_t = q->d_w + (size_t)(t - 1) * cv.nbe;
        float *gpart_t = q->gpart + (size_t)(t - 1) * cv.nbe * F;
        if (cv.max_dc > 16)
            hipLaunchKernelGGL(cn_bwd_kernel<27>, dim3(gc), dim3(block), 0, s, cv, F, t, d_llr, st_prev,
                               bel_prev, w_prev, w_t, q->g_bel, g_cur, has_gcur, g_prev, gpart_t);
        else
            hipLaunchKernelGGL(cn_bwd_kernel<16>, dim3(gc), dim3(block), 0, s, cv, F, t, d_llr, st_prev,
                               bel_prev, w_prev, w_t, q->g_bel, g_cur, has_gcur, g_prev, gpart_t);
        GF2L_HIP(ctx, hipGetLastError());
        std::swap(g_cur, g_prev);
    }
    hipLaunchKernelGGL(grad_reduce_kernel, dim3(T * cv.nbe), dim3(256), 0, s, q->gpart, F, q->grad);
    hipLaunchKernelGGL(grad_reduce_kernel, dim3(1), dim3(256), 0, s, q->lpart, T * gv, q->loss);
    GF2L_HIP(ctx, hipGetLastError());
    return GF2L_OK;
}

}  // namespace
}  // namespace gf2l

namespace gf2l {
namespace {

// Shared constructor of the flooding and layered learned decoders.
int create_nms(gf2l_ctx *ctx, const gf2l_code *code, uint32_t iterations, const float *init_weights,
               float init_value, bool layered, gf2l_nms **out) {
    if (!code || iterations == 0) return fail(ctx, GF2L_ERR_INVALID, "need a code and iterations >= 1");
    if (code->is_csr || code->view.dvb_q)
        return fail(ctx, GF2L_ERR_UNSUPPORTED, "learned weights need a quasi-cyclic code (not DVB)");
    if (layered && code->view.max_dc > kMaxStateDegree)
        return fail(ctx, GF2L_ERR_UNSUPPORTED, "the layered decoder needs check degree <= 27");
    GF2L_HIP(ctx, hipSetDevice(ctx->device));
    const size_t nw = (size_t)iterations * code->view.nbe;
    std::vector<float> w(nw, init_value);
    if (init_weights) std::memcpy(w.data(), init_weights, nw * sizeof(float));
    gf2l_nms *q = new gf2l_nms;
    q->code = code;
    q->T = iterations;
    q->nbe = code->view.nbe;
    q->layered = layered;
    q->fb_continue = layered;
    q->X = layered ? layered_lds_columns(code->view, ctx->lds_limit) : 0;
    hipError_t e = hipMalloc(&q->d_w, nw * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&q->d_m, nw * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&q->d_v, nw * sizeof(float));
    if (e == hipSuccess) e = hipMemcpy(q->d_w, w.data(), nw * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(q->d_m, 0, nw * sizeof(float));
    if (e == hipSuccess) e = hipMemset(q->d_v, 0, nw * sizeof(float));
    const size_t zbytes = 3ull * code->view.mb * code->view.zp * sizeof(uint32_t);
    if (e == hipSuccess && layered) e = hipMalloc(&q->d_zero_state, zbytes);
    if (e == hipSuccess && layered) e = hipMemset(q->d_zero_state, 0, zbytes);
    if (e != hipSuccess) {
        hipFree(q->d_w);
        hipFree(q->d_m);
        hipFree(q->d_v);
        hipFree(q->d_zero_state);
        delete q;
        return fail(ctx, GF2L_ERR_HIP, std::string("nms create: ") + hipGetErrorString(e));
    }
    *out = q;
    return GF2L_OK;
}

}  // namespace
}  // namespace gf2l

using gf2l::fail;

extern "C" {

int gf2l_nms_create(gf2l_ctx *ctx, const gf2l_code *code, uint32_t iterations,
                    const float *init_weights, float init_value, gf2l_nms **out) {
    if (!ctx || !out) return GF2L_ERR_INVALID;
    *out = nullptr;
    return gf2l::guard(ctx, [&]() -> int { return gf2l::create_nms(ctx, code, iterations, init_weights, init_value, false, out); });
}

int gf2l_nms_create_layered(gf2l_ctx *ctx, const gf2l_code *code, uint32_t iterations, const float *init_weights,
                            float init_value, gf2l_nms **out) {
    if (!ctx || !out) return GF2L_ERR_INVALID;
    *out = nullptr;
    return gf2l::guard(ctx, [&]() -> int { return gf2l::create_nms(ctx, code, iterations, init_weights, init_value, true, out); });
}

int gf2l_nms_layered_columns(gf2l_ctx *ctx, const gf2l_nms *q, uint32_t *columns) {
    if (!ctx || !q || !columns) return GF2L_ERR_INVALID;
    *columns = q->layered ? gf2l::layered_decode_x(ctx, q) : 0;
    return GF2L_OK;
}

int gf2l_nms_set_fallback(gf2l_ctx *ctx, gf2l_nms *q, int mode) {
    if (!ctx || !q) return GF2L_ERR_INVALID;
    return gf2l::guard(ctx, [&]() -> int {
        if (mode != GF2L_FALLBACK_RESTART && mode != GF2L_FALLBACK_CONTINUE)
            return fail(ctx, GF2L_ERR_INVALID, "unknown fallback mode");
        if (q->layered && mode == GF2L_FALLBACK_RESTART)
            return fail(ctx, GF2L_ERR_UNSUPPORTED, "a layered decoder's fallback continues from its state");
        q->fb_continue = mode == GF2L_FALLBACK_CONTINUE;
        return GF2L_OK;
    });
}

void gf2l_nms_destroy(gf2l_nms *q) {
    if (!q) return;
    hipSetDevice(q->code->device);
    hipDeviceSynchronize();
    gf2l::free_train(q);
    hipFree(q->d_w);
    hipFree(q->d_m);
    hipFree(q->d_v);
    hipFree(q->d_zero_state);
    delete q;
}

int gf2l_nms_info(const gf2l_nms *q, uint32_t *iterations, uint32_t *weights_per_iteration) {
    if (!q) return GF2L_ERR_INVALID;
    if (iterations) *iterations = q->T;
    if (weights_per_iteration) *weights_per_iteration = q->nbe;
    return GF2L_OK;
}

int gf2l_nms_get_weights(gf2l_ctx *ctx, const gf2l_nms *q, float *w) {
    if (!ctx || !q || !w) return GF2L_ERR_INVALID;
    return gf2l::guard(ctx, [&]() -> int {
        GF2L_HIP(ctx, hipSetDevice(ctx->device));
        GF2L_HIP(ctx, hipStreamSynchronize(ctx->stream));
        GF2L_HIP(ctx, hipMemcpy(w, q->d_w, (size_t)q->T * q->nbe * sizeof(float), hipMemcpyDeviceToHost));
        return GF2L_OK;
    });
}

int gf2l_nms_set_weights(gf2l_ctx *ctx, gf2l_nms *q, const float *w) {
    if (!ctx || !q || !w) return GF2L_ERR_INVALID;
    return gf2l::guard(ctx, [&]() -> int {
        GF2L_HIP(ctx, hipSetDevice(ctx->device));
        GF2L_HIP(ctx, hipStreamSynchronize(ctx->stream));
        GF2L_HIP(ctx, hipMemcpy(q->d_w, w, (size_t)q->T * q->nbe * sizeof(float), hipMemcpyHostToDevice));
        return GF2L_OK;
    });
}

int gf2l_nms_get_adam(gf2l_ctx *ctx, const gf2l_nms *q, float *m, float *v, uint64_t *step) {
    if (!ctx || !q) return GF2L_ERR_INVALID;
    return gf2l::guard(ctx, [&]() -> int {
        GF2L_HIP(ctx, hipSetDevice(ctx->device));
        GF2L_HIP(ctx, hipStreamSynchronize(ctx->stream));
        const size_t b = (size_t)q->T * q->nbe * sizeof(float);
        if (m) GF2L_HIP(ctx, hipMemcpy(m, q->d_m, b, hipMemcpyDeviceToHost));
        if (v) GF2L_HIP(ctx, hipMemcpy(v, q->d_v, b, hipMemcpyDeviceToHost));
        if (step) *step = q->step;
        return GF2L_OK;
    });
}

int gf2l_nms_set_adam(gf2l_ctx *ctx, gf2l_nms *q, const float *m, const float *v, uint64_t step) {
    if (!ctx || !q || !m || !v) return GF2L_ERR_INVALID;
    return gf2l::guard(ctx, [&]() -> int {
        GF2L_HIP(ctx, hipSetDevice(ctx->device));
        GF2L_HIP(ctx, hipStreamSynchronize(ctx->stream));
        const size_t b = (size_t)q->T * q->nbe * sizeof(float);
        GF2L_HIP(ctx, hipMemcpy(q->d_m, m, b, hipMemcpyHostToDevice));
        GF2L_HIP(ctx, hipMemcpy(q->d_v, v, b, hipMemcpyHostToDevice));
        q->step = step;
        return GF2L_OK;
    });
}

int gf2l_nms_gradient(gf2l_ctx *ctx, gf2l_nms *q, const float *d_llr, const uint64_t *d_cw,
                      size_t frames, float *grad, float *loss) {
    if (!ctx || !q || !d_llr || !d_cw) return GF2L_ERR_INVALID;
    return gf2l::guard(ctx, [&]() -> int {
        GF2L_HIP(ctx, hipSetDevice(ctx->device));
        int rc = gf2l::forward_backward(ctx, q, d_llr, d_cw, frames);
        if (rc != GF2L_OK) return rc;
        GF2L_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (grad)
            GF2L_HIP(ctx, hipMemcpy(grad, q->grad, (size_t)q->T * q->nbe * sizeof(float), hipMemcpyDeviceToHost));
        if (loss) {
            float l = 0.0f;
            GF2L_HIP(ctx, hipMemcpy(&l, q->loss, sizeof(float), hipMemcpyDeviceToHost));
            *loss = l / ((float)q->T * (float)q->code->view.n * (float)frames);
        }
        return GF2L_OK;
    });
}

int gf2l_nms_train_step(gf2l_ctx *ctx, gf2l_nms *q, const float *d_llr, const uint64_t *d_cw,
                        size_t frames, float lr, float *loss) {
    if (!ctx || !q || !d_llr || !d_cw) return GF2L_ERR_INVALID;
    return gf2l::guard(ctx, [&]() -> int {
        GF2L_HIP(ctx, hipSetDevice(ctx->device));
        int rc = gf2l::forward_backward(ctx, q, d_llr, d_cw, frames);
        if (rc != GF2L_OK) return rc;
        q->step += 1;
        const uint32_t n = q->T * q->nbe;
        const float bc1 = 1.0f - std::pow(0.9f, (float)q->step);
        const float bc2 = 1.0f - std::pow(0.999f, (float)q->step);
        hipLaunchKernelGGL(gf2l::adam_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, q->d_w,
                           q->d_m, q->d_v, q->grad, n, lr, bc1, bc2);
        GF2L_HIP(ctx, hipGetLastError());
        if (loss) {
            float l = 0.0f;
            GF2L_HIP(ctx, hipMemcpyAsync(&l, q->loss, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
            GF2L_HIP(ctx, hipStreamSynchronize(ctx->stream));
            *loss = l / ((float)q->T * (float)q->code->view.n * (float)frames);
        }
        return GF2L_OK;
    });
}

int gf2l_nms_decode_device(gf2l_ctx *ctx, const gf2l_nms *q, const float *d_llr, size_t frames,
                           uint32_t fallback_max_it, uint64_t *d_msg_words, uint32_t *d_iterations,
                           uint8_t *d_converged, uint8_t *d_syndrome_ok, uint8_t *d_fallback) {
    if (!ctx || !q) return GF2L_ERR_INVALID;
    return gf2l::guard(ctx, [&]() -> int {
        if (frames == 0) return GF2L_OK;
        if (!d_llr || !d_msg_words || !d_iterations || !d_converged || !d_syndrome_ok)
            return fail(ctx, GF2L_ERR_INVALID, "null buffer");
        if (ctx->rule != GF2L_RULE_MINSUM)
            return fail(ctx, GF2L_ERR_UNSUPPORTED, "the learned decoders run the min-sum rule only");
        GF2L_HIP(ctx, hipSetDevice(ctx->device));
        if (q->layered) {
            if (frames > 0xffffffffull) return fail(ctx, GF2L_ERR_INVALID, "too many frames in one batch");
            return gf2l::layered_decode(ctx, q, d_llr, (uint32_t)frames, fallback_max_it, d_msg_words, d_iterations,
                                        d_converged, d_syndrome_ok, d_fallback);
        }
        return gf2l::decode_device(ctx, q->code, d_llr, frames, q->T, q->d_w, fallback_max_it, q->fb_continue, d_msg_words,
                                   d_iterations, d_converged, d_syndrome_ok, d_fallback, nullptr, nullptr, 0);
    });
}

}  // extern "C"

namespace gf2l {

const float *nms_weights(const gf2l_nms *q, uint32_t *iterations, const gf2l_code **code, bool *layered,
                         bool *fb_continue) {
    *iterations = q->T;
    *code = q->code;
    *layered = q->layered;
    *fb_continue = q->fb_continue;
    return q->d_w;
}

}  // namespace gf2l
