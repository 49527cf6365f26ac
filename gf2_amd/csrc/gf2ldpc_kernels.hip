// HIP kernels of the batched LDPC min-sum decoder for gfx950 (MI355X).
//
// Arithmetic contract (bit-exact with the reference decoder for finite
// LLRs, crates/gf2-coding/src/ldpc/core.rs:790-938):
//   * check message = sign_product * min |v2c| over the other d-1 edges,
//     exact (abs / compare / negate only); the sign of an input counts as
//     negative by its sign bit inside the AVX2 8-lane chunks and by
//     `value < 0` in the scalar remainder, as in
//     crates/gf2-kernels-simd/src/llr.rs:117-180 (matters only for -0.0);
//     a degree-1 check sends +0.0 (core.rs:804-806).
//   * belief = ((L + c2v_0) + c2v_1) + ... in ascending check order, then
//     v2c = belief - c2v (core.rs:819-838): sequential binary32 adds, built
//     with -ffp-contract=off so nothing fuses into an FMA.
//   * v2c is not stored: the check update recomputes belief - c2v from the
//     stored belief and its own previous message, the same two operands and
//     the same subtraction the reference performs.  With belief := L and
//     c2v := +0 before the first iteration this gives v2c = L exactly.
//   * hard decision belief < 0 (llr.rs:73-75); syndrome over GF(2) by XOR.
#include <hip/hip_runtime.h>

#include "internal.h"

namespace gf2l {
namespace {

// Running state of one check while its d inputs stream past.
struct CheckAcc {
    float min1, min2;
    uint32_t idx;
    uint64_t sb, lt;  // per-position sign-bit and (x < 0) masks
    __device__ void init() {
        min1 = __builtin_inff();
        min2 = __builtin_inff();
        idx = 0;
        sb = 0;
        lt = 0;
    }
    __device__ void add(uint32_t p, float x) {
        float a = fabsf(x);
        if (a < min1) {
            min2 = min1;
            min1 = a;
            idx = p;
        } else if (a < min2) {
            min2 = a;
        }
        sb |= (uint64_t)(__float_as_uint(x) >> 31) << p;
        lt |= (uint64_t)(x < 0.0f) << p;
    }
    // Message to position p of a degree-d check.
    __device__ float out(uint32_t d, uint32_t p) const {
        if (d == 1) return 0.0f;
        const float mag = (p == idx) ? min2 : min1;
        const uint32_t l8 = (d - 1) & ~7u;  // inputs covered by full 8-lane chunks
        const uint64_t pb = 1ull << p;
        const uint64_t all = (1ull << d) - 1;
        const uint64_t lane = (p < l8) ? (((1ull << (l8 + 1)) - 1) & ~pb) : ((1ull << l8) - 1);
        const uint64_t rem = all & ~lane & ~pb;
        const uint32_t neg = (__popcll(sb & lane) ^ __popcll(lt & rem)) & 1u;
        return neg ? -mag : mag;
    }
};

__device__ __forceinline__ uint32_t wrap_add(uint32_t a, uint32_t b, uint32_t z) {
    uint32_t s = a + b;
    return s >= z ? s - z : s;
}
__device__ __forceinline__ uint32_t wrap_sub(uint32_t a, uint32_t b, uint32_t z) {
    return a >= b ? a - b : a + z - b;
}

// ---------------------------------------------------------------------------
// Streaming path, one launch per phase per iteration.  Frames still running
// are listed in `list[0..*count)`; persistent workgroups stride over
// (frame, block-row) or (frame, block-column) items.

// Check update for all checks of one block-row of one frame.  Threads map to
// the z lifted checks; belief reads of a circulant are a rotated contiguous
// run, c2v reads/writes are contiguous.
template <int DMAX>
__global__ void __launch_bounds__(1024)
cn_stream_kernel(CodeView cv, const uint32_t *__restrict__ list, const uint32_t *__restrict__ count_ptr,
                 float *__restrict__ c2v, const float *__restrict__ belief,
                 const float *__restrict__ w_it) {
    const uint32_t items = *count_ptr * cv.mb;
    const size_t E = (size_t)cv.nbe * cv.z;
    for (uint32_t item = blockIdx.x; item < items; item += gridDim.x) {
        const uint32_t f = list[item / cv.mb];
        const uint32_t br = item % cv.mb;
        const uint32_t e0 = cv.row_ptr[br];
        const uint32_t d = cv.row_ptr[br + 1] - e0;
        float *fc = c2v + (size_t)f * E;
        const float *fb = belief + (size_t)f * cv.n;
        for (uint32_t i = threadIdx.x; i < cv.z; i += blockDim.x) {
            float x[DMAX];
#pragma unroll
            for (int p = 0; p < DMAX; ++p) {
                if (p < (int)d) {
                    const uint2 re = cv.row_edge[e0 + p];
                    x[p] = fb[re.x * cv.z + wrap_add(i, re.y, cv.z)] - fc[(size_t)(e0 + p) * cv.z + i];
                }
            }
            CheckAcc acc;
            acc.init();
#pragma unroll
            for (int p = 0; p < DMAX; ++p)
                if (p < (int)d) acc.add(p, x[p]);
#pragma unroll
            for (int p = 0; p < DMAX; ++p) {
                if (p < (int)d) {
                    float o = acc.out(d, p);
                    if (w_it) o = w_it[e0 + p] * o;
                    fc[(size_t)(e0 + p) * cv.z + i] = o;
                }
            }
        }
    }
}

// Variable update for the z variables of one block-column of one frame:
// belief = L + sum of its c2v in ascending block-row order; hard decisions
// packed by wave ballot into 64-bit words per block-column.
template <int DVMAX>
__global__ void __launch_bounds__(1024)
vn_stream_kernel(CodeView cv, const uint32_t *__restrict__ list, const uint32_t *__restrict__ count_ptr,
                 const float *__restrict__ c2v, float *__restrict__ belief,
                 const float *__restrict__ llr, uint64_t *__restrict__ hb,
                 uint32_t *__restrict__ next_count) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *next_count = 0;
    const uint32_t items = *count_ptr * cv.nb;
    const size_t E = (size_t)cv.nbe * cv.z;
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (uint32_t item = blockIdx.x; item < items; item += gridDim.x) {
        const uint32_t f = list[item / cv.nb];
        const uint32_t bc = item % cv.nb;
        const uint32_t c0 = cv.col_ptr[bc];
        const uint32_t dv = cv.col_ptr[bc + 1] - c0;
        const float *fc = c2v + (size_t)f * E;
        const size_t vbase = (size_t)f * cv.n + (size_t)bc * cv.z;
        for (uint32_t j0 = 0; j0 < cv.z; j0 += blockDim.x) {
            const uint32_t j = j0 + threadIdx.x;
            const bool valid = j < cv.z;
            float bel = 0.0f;
            if (valid) {
                float c[DVMAX];
#pragma unroll
                for (int q = 0; q < DVMAX; ++q) {
                    if (q < (int)dv) {
                        const uint2 ce = cv.col_edge[c0 + q];
                        c[q] = fc[(size_t)ce.x * cv.z + wrap_sub(j, ce.y, cv.z)];
                    }
                }
                bel = llr[vbase + j];
#pragma unroll
                for (int q = 0; q < DVMAX; ++q)
                    if (q < (int)dv) bel = bel + c[q];
                belief[vbase + j] = bel;
            }
            const uint64_t mask = __ballot(valid && bel < 0.0f);
            const uint32_t word_bit = j0 + wave * 64;
            if (lane == 0 && word_bit < cv.z)
                hb[((size_t)f * cv.nb + bc) * cv.zw + (word_bit >> 6)] = mask;
        }
    }
}

// Syndrome of each running frame's hard decisions; retires converged frames
// (and every frame at the iteration cap), writing DecoderResult fields and
// the first k hard bits; appends the rest to the next list.
__global__ void __launch_bounds__(256)
syn_stream_kernel(CodeView cv, const uint32_t *__restrict__ list, const uint32_t *__restrict__ count_ptr,
                  uint32_t *__restrict__ next_list, uint32_t *__restrict__ next_count,
                  const uint64_t *__restrict__ hb, uint32_t it, uint32_t max_it,
                  uint64_t *__restrict__ msg, uint32_t *__restrict__ iters,
                  uint8_t *__restrict__ conv, uint8_t *__restrict__ synd) {
    extern __shared__ uint64_t sh_hb[];
    const uint32_t count = *count_ptr;
    const uint32_t words = cv.nb * cv.zw;
    const uint32_t kw = (cv.k + 63) / 64;
    for (uint32_t item = blockIdx.x; item < count; item += gridDim.x) {
        const uint32_t f = list[item];
        for (uint32_t w = threadIdx.x; w < words; w += blockDim.x) sh_hb[w] = hb[(size_t)f * words + w];
        __syncthreads();
        int unsat = 0;
        uint32_t br = threadIdx.x / cv.z, i = threadIdx.x % cv.z;
        for (uint32_t c = threadIdx.x; c < cv.m; c += blockDim.x) {
            const uint32_t e0 = cv.row_ptr[br], e1 = cv.row_ptr[br + 1];
            uint32_t par = 0;
            for (uint32_t e = e0; e < e1; ++e) {
                const uint2 re = cv.row_edge[e];
                const uint32_t col = wrap_add(i, re.y, cv.z);
                par ^= (uint32_t)(sh_hb[re.x * cv.zw + (col >> 6)] >> (col & 63)) & 1u;
            }
            unsat |= (int)par;
            i += blockDim.x;
            while (i >= cv.z) { i -= cv.z; ++br; }
        }
        unsat = __syncthreads_or(unsat);
        if (!unsat || it >= max_it) {
            for (uint32_t w = threadIdx.x; w < kw; w += blockDim.x) {
                uint64_t word = 0;
                uint32_t v = w * 64, bc = v / cv.z, j = v - bc * cv.z;
                for (uint32_t b = 0; b < 64 && v < cv.k; ++b, ++v) {
                    word |= ((sh_hb[bc * cv.zw + (j >> 6)] >> (j & 63)) & 1ull) << b;
                    if (++j == cv.z) { j = 0; ++bc; }
                }
                msg[(size_t)f * kw + w] = word;
            }
            if (threadIdx.x == 0) {
                iters[f] = it;
                conv[f] = (uint8_t)!unsat;
                synd[f] = (uint8_t)!unsat;
            }
        } else if (threadIdx.x == 0) {
            next_list[atomicAdd(next_count, 1u)] = f;
        }
        __syncthreads();
    }
}

__global__ void stream_init_kernel(uint32_t *list, uint32_t *counts, uint32_t frames) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < frames) list[t] = t;
    if (t == 0) { counts[0] = frames; counts[1] = 0; }
}

// ---------------------------------------------------------------------------
// LDS path: one workgroup owns one frame for all its iterations; messages,
// beliefs and LLRs live in LDS.  Used when (E + 2n) floats fit.
template <int DMAX, int DVMAX>
__global__ void __launch_bounds__(512)
lds_decode_kernel(CodeView cv, const float *__restrict__ llr, uint32_t frames, uint32_t max_it,
                  const float *__restrict__ w, uint32_t w_iters, uint64_t *__restrict__ msg,
                  uint32_t *__restrict__ iters, uint8_t *__restrict__ conv,
                  uint8_t *__restrict__ synd, float *__restrict__ trace_c2v,
                  float *__restrict__ trace_bel, uint32_t trace_iters) {
    extern __shared__ float sm[];
    const uint32_t E = cv.nbe * cv.z, n = cv.n, z = cv.z;
    float *s_c2v = sm;
    float *s_bel = sm + E;
    float *s_llr = s_bel + n;
    const uint32_t kw = (cv.k + 63) / 64;
    for (uint32_t f = blockIdx.x; f < frames; f += gridDim.x) {
        for (uint32_t v = threadIdx.x; v < n; v += blockDim.x) {
            const float l = llr[(size_t)f * n + v];
            s_llr[v] = l;
            s_bel[v] = l;
        }
        for (uint32_t e = threadIdx.x; e < E; e += blockDim.x) s_c2v[e] = 0.0f;
        __syncthreads();
        uint32_t it_done = 0;
        int converged = 0;
        for (uint32_t it = 1; it <= max_it; ++it) {
            const float *w_it = (w && it <= w_iters) ? w + (size_t)(it - 1) * cv.nbe : nullptr;
            {   // check update
                uint32_t br = threadIdx.x / z, i = threadIdx.x % z;
                for (uint32_t c = threadIdx.x; c < cv.m; c += blockDim.x) {
                    const uint32_t e0 = cv.row_ptr[br], d = cv.row_ptr[br + 1] - e0;
                    float x[DMAX];
#pragma unroll
                    for (int p = 0; p < DMAX; ++p) {
                        if (p < (int)d) {
                            const uint2 re = cv.row_edge[e0 + p];
                            x[p] = s_bel[re.x * z + wrap_add(i, re.y, z)] - s_c2v[(e0 + p) * z + i];
                        }
                    }
                    CheckAcc acc;
                    acc.init();
#pragma unroll
                    for (int p = 0; p < DMAX; ++p)
                        if (p < (int)d) acc.add(p, x[p]);
#pragma unroll
                    for (int p = 0; p < DMAX; ++p) {
                        if (p < (int)d) {
                            float o = acc.out(d, p);
                            if (w_it) o = w_it[e0 + p] * o;
                            s_c2v[(e0 + p) * z + i] = o;
                        }
                    }
                    i += blockDim.x;
                    while (i >= z) { i -= z; ++br; }
                }
            }
            __syncthreads();
            {   // variable update
                uint32_t bc = threadIdx.x / z, j = threadIdx.x % z;
                for (uint32_t v = threadIdx.x; v < n; v += blockDim.x) {
                    const uint32_t c0 = cv.col_ptr[bc], dv = cv.col_ptr[bc + 1] - c0;
                    float c[DVMAX];
#pragma unroll
                    for (int q = 0; q < DVMAX; ++q) {
                        if (q < (int)dv) {
                            const uint2 ce = cv.col_edge[c0 + q];
                            c[q] = s_c2v[ce.x * z + wrap_sub(j, ce.y, z)];
                        }
                    }
                    float bel = s_llr[v];
#pragma unroll
                    for (int q = 0; q < DVMAX; ++q)
                        if (q < (int)dv) bel = bel + c[q];
                    s_bel[v] = bel;
                    j += blockDim.x;
                    while (j >= z) { j -= z; ++bc; }
                }
            }
            __syncthreads();
            if (trace_c2v && it <= trace_iters) {
                for (uint32_t e = threadIdx.x; e < E; e += blockDim.x)
                    trace_c2v[((size_t)(it - 1) * frames + f) * E + e] = s_c2v[e];
                for (uint32_t v = threadIdx.x; v < n; v += blockDim.x)
                    trace_bel[((size_t)(it - 1) * frames + f) * n + v] = s_bel[v];
            }
            int unsat = 0;
            {   // syndrome of the hard decisions
                uint32_t br = threadIdx.x / z, i = threadIdx.x % z;
                for (uint32_t c = threadIdx.x; c < cv.m; c += blockDim.x) {
                    const uint32_t e0 = cv.row_ptr[br], e1 = cv.row_ptr[br + 1];
                    uint32_t par = 0;
                    for (uint32_t e = e0; e < e1; ++e) {
                        const uint2 re = cv.row_edge[e];
                        par ^= (uint32_t)(s_bel[re.x * z + wrap_add(i, re.y, z)] < 0.0f);
                    }
                    unsat |= (int)par;
                    i += blockDim.x;
                    while (i >= z) { i -= z; ++br; }
                }
            }
            unsat = __syncthreads_or(unsat);
            it_done = it;
            if (!unsat) { converged = 1; break; }
        }
        if (trace_c2v) {  // frames that stopped early keep their final state
            for (uint32_t t = it_done + 1; t <= trace_iters && it_done > 0; ++t) {
                for (uint32_t e = threadIdx.x; e < E; e += blockDim.x)
                    trace_c2v[((size_t)(t - 1) * frames + f) * E + e] = s_c2v[e];
                for (uint32_t v = threadIdx.x; v < n; v += blockDim.x)
                    trace_bel[((size_t)(t - 1) * frames + f) * n + v] = s_bel[v];
            }
        }
        // max_it == 0: a fresh decoder's beliefs are all zero (core.rs:680-693)
        for (uint32_t wd = threadIdx.x; wd < kw; wd += blockDim.x) {
            uint64_t word = 0;
            if (max_it > 0)
                for (uint32_t b = 0; b < 64; ++b) {
                    const uint32_t v = wd * 64 + b;
                    if (v < cv.k && s_bel[v] < 0.0f) word |= 1ull << b;
                }
            msg[(size_t)f * kw + wd] = word;
        }
        if (threadIdx.x == 0) {
            iters[f] = it_done;
            conv[f] = (uint8_t)converged;
            synd[f] = (uint8_t)(converged || max_it == 0);
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

uint32_t round_block(uint32_t z) {
    uint32_t b = ((z + 63) / 64) * 64;
    return b > 1024 ? 1024 : b;
}

template <int DMAX, int DVMAX>
hipError_t launch_lds_t(const CodeView &cv, const float *llr, uint32_t frames, uint32_t max_it,
                        const float *w, uint32_t w_iters, uint64_t *msg, uint32_t *iters,
                        uint8_t *conv, uint8_t *synd, float *tc, float *tb, uint32_t ti,
                        hipStream_t s) {
    const size_t bytes = lds_path_bytes(cv);
    auto k = lds_decode_kernel<DMAX, DVMAX>;
    hipError_t err = set_lds(k, bytes);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(k, dim3(frames), dim3(512), bytes, s, cv, llr, frames, max_it, w, w_iters,
                       msg, iters, conv, synd, tc, tb, ti);
    return hipGetLastError();
}

}  // namespace

size_t lds_path_bytes(const CodeView &cv) {
    return ((size_t)cv.nbe * cv.z + 2 * (size_t)cv.n) * sizeof(float);
}

hipError_t launch_lds_decode(const CodeView &cv, const float *llr, uint32_t frames, uint32_t max_it,
                             const float *w, uint32_t w_iters, uint64_t *msg, uint32_t *iters,
                             uint8_t *conv, uint8_t *synd, float *trace_c2v, float *trace_bel,
                             uint32_t trace_iters, hipStream_t s) {
    if (frames == 0) return hipSuccess;
    const int dm = pick_dmax(cv.max_dc), dvm = pick_dvmax(cv.max_dv);
#define GF2L_LDS(DM, DV)                                                                     \
    if (dm == DM && dvm == DV)                                                               \
        return launch_lds_t<DM, DV>(cv, llr, frames, max_it, w, w_iters, msg, iters, conv, synd, \
                                    trace_c2v, trace_bel, trace_iters, s);
    GF2L_LDS(8, 4) GF2L_LDS(8, 8) GF2L_LDS(8, 16) GF2L_LDS(8, 32)
    GF2L_LDS(16, 4) GF2L_LDS(16, 8) GF2L_LDS(16, 16) GF2L_LDS(16, 32)
    GF2L_LDS(32, 4) GF2L_LDS(32, 8) GF2L_LDS(32, 16) GF2L_LDS(32, 32)
#undef GF2L_LDS
    return hipErrorInvalidValue;
}

hipError_t launch_stream_init(const CodeView &cv, Workspace &ws, const float *llr, uint32_t frames,
                              hipStream_t s) {
    const size_t E = (size_t)cv.nbe * cv.z;
    hipError_t err = hipMemcpyAsync(ws.belief, llr, (size_t)frames * cv.n * sizeof(float),
                                    hipMemcpyDeviceToDevice, s);
    if (err != hipSuccess) return err;
    err = hipMemsetAsync(ws.c2v, 0, (size_t)frames * E * sizeof(float), s);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(stream_init_kernel, dim3((frames + 255) / 256), dim3(256), 0, s, ws.lists,
                       ws.counts, frames);
    return hipGetLastError();
}

hipError_t launch_stream_iteration(const CodeView &cv, Workspace &ws, const float *llr,
                                   uint32_t frames, uint32_t it, uint32_t max_it, const float *w_it,
                                   uint64_t *msg, uint32_t *iters, uint8_t *conv, uint8_t *synd,
                                   int num_cus, hipStream_t s, hipEvent_t *ev) {
    const uint32_t cur = (it - 1) & 1u, nxt = it & 1u;
    const uint32_t *list = ws.lists + (size_t)cur * ws.cap_frames;
    uint32_t *next_list = ws.lists + (size_t)nxt * ws.cap_frames;
    const uint32_t *count = ws.counts + cur;
    uint32_t *next_count = ws.counts + nxt;
    const uint32_t block = round_block(cv.z);
    const uint32_t grid_cap = (uint32_t)num_cus * 8u;
    uint32_t g_cn = frames * cv.mb, g_vn = frames * cv.nb;
    if (g_cn > grid_cap) g_cn = grid_cap;
    if (g_vn > grid_cap) g_vn = grid_cap;
    uint32_t g_syn = frames < grid_cap ? frames : grid_cap;

    if (ev) hipEventRecord(ev[0], s);
    switch (pick_dmax(cv.max_dc)) {
        case 8:
            hipLaunchKernelGGL(cn_stream_kernel<8>, dim3(g_cn), dim3(block), 0, s, cv, list, count,
                               ws.c2v, ws.belief, w_it);
            break;
        case 16:
            hipLaunchKernelGGL(cn_stream_kernel<16>, dim3(g_cn), dim3(block), 0, s, cv, list, count,
                               ws.c2v, ws.belief, w_it);
            break;
        default:
            hipLaunchKernelGGL(cn_stream_kernel<32>, dim3(g_cn), dim3(block), 0, s, cv, list, count,
                               ws.c2v, ws.belief, w_it);
            break;
    }
    if (ev) hipEventRecord(ev[1], s);
    switch (pick_dvmax(cv.max_dv)) {
        case 4:
            hipLaunchKernelGGL(vn_stream_kernel<4>, dim3(g_vn), dim3(block), 0, s, cv, list, count,
                               ws.c2v, ws.belief, llr, ws.hb, next_count);
            break;
        case 8:
            hipLaunchKernelGGL(vn_stream_kernel<8>, dim3(g_vn), dim3(block), 0, s, cv, list, count,
                               ws.c2v, ws.belief, llr, ws.hb, next_count);
            break;
        case 16:
            hipLaunchKernelGGL(vn_stream_kernel<16>, dim3(g_vn), dim3(block), 0, s, cv, list, count,
                               ws.c2v, ws.belief, llr, ws.hb, next_count);
            break;
        default:
            hipLaunchKernelGGL(vn_stream_kernel<32>, dim3(g_vn), dim3(block), 0, s, cv, list, count,
                               ws.c2v, ws.belief, llr, ws.hb, next_count);
            break;
    }
    if (ev) hipEventRecord(ev[2], s);
    const size_t sh = (size_t)cv.nb * cv.zw * sizeof(uint64_t);
    hipLaunchKernelGGL(syn_stream_kernel, dim3(g_syn), dim3(256), sh, s, cv, list, count, next_list,
                       next_count, ws.hb, it, max_it, msg, iters, conv, synd);
    if (ev) hipEventRecord(ev[3], s);
    return hipGetLastError();
}

}  // namespace gf2l
