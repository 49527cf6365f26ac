// General sparse (CSR/CSC) path of the min-sum decoder, for codes built from
// an arbitrary edge list: LdpcCode::from_edges (crates/gf2-coding/src/ldpc/
// core.rs:96-104) -> SpBitMatrixDual::from_coo (crates/gf2-core/src/sparse.rs:
// 166-204, 297-331, 441-445).  Same arithmetic contract as the quasi-cyclic
// paths: check message = sign_product * min over the other inputs (AVX2
// lane / remainder sign rule), belief = L + c2v in ascending check order,
// v2c = belief - c2v recomputed from the stored belief and message.
//
// Per frame in HBM: c2v [nnz] in CSR order, beliefs [n], hard decisions
// [ceil(n/64)] in the linear BitVec layout.  One thread per check (check
// kernel) and per variable (variable kernel); degrees vary per thread, so
// the loops are runtime-bounded.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "device_common.h"
#include "gf2ldpc.h"
#include "host_codes.h"

namespace gf2l {
namespace {

__global__ void __launch_bounds__(256)
cn_csr_kernel(CsrView cv, const uint32_t *__restrict__ list, const uint32_t *__restrict__ count_ptr,
              float *__restrict__ c2v, const float *__restrict__ belief) {
    const uint32_t chunks = (cv.m + blockDim.x - 1) / blockDim.x;
    const uint32_t items = *count_ptr * chunks;
    for (uint32_t item = blockIdx.x; item < items; item += gridDim.x) {
        const uint32_t f = list[item / chunks];
        const uint32_t r = (item % chunks) * blockDim.x + threadIdx.x;
        if (r >= cv.m) continue;
        float *fc = c2v + (size_t)f * cv.nnz;
        const float *fb = belief + (size_t)f * cv.n;
        const uint32_t e0 = cv.row_ptr[r], d = cv.row_ptr[r + 1] - e0;
        CheckAcc acc;
        acc.init();
        for (uint32_t p = 0; p < d; ++p) acc.add(p, fb[cv.col_idx[e0 + p]] - fc[e0 + p]);
        for (uint32_t p = 0; p < d; ++p) fc[e0 + p] = acc.out(d, p);
    }
}

__global__ void __launch_bounds__(256)
vn_csr_kernel(CsrView cv, const uint32_t *__restrict__ list, const uint32_t *__restrict__ count_ptr,
              const float *__restrict__ c2v, float *__restrict__ belief, const float *__restrict__ llr,
              uint64_t *__restrict__ hb, uint32_t *__restrict__ next_count) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *next_count = 0;
    const uint32_t chunks = (cv.n + blockDim.x - 1) / blockDim.x;
    const uint32_t items = *count_ptr * chunks;
    const uint32_t nw = (cv.n + 63) / 64;
    for (uint32_t item = blockIdx.x; item < items; item += gridDim.x) {
        const uint32_t f = list[item / chunks];
        const uint32_t v = (item % chunks) * blockDim.x + threadIdx.x;
        float bel = 0.0f;
        if (v < cv.n) {
            const float *fc = c2v + (size_t)f * cv.nnz;
            bel = llr[(size_t)f * cv.n + v];
            for (uint32_t q = cv.col_ptr[v]; q < cv.col_ptr[v + 1]; ++q) bel = bel + fc[cv.csc_edge[q]];
            belief[(size_t)f * cv.n + v] = bel;
        }
        const uint64_t mask = __ballot(v < cv.n && bel < 0.0f);
        if ((threadIdx.x & 63) == 0 && v < cv.n) hb[(size_t)f * nw + (v >> 6)] = mask;
    }
}

// One wave per frame: syndrome from the linear hard-decision words, then
// retire (first k bits are the message) or keep the frame running.
__global__ void __launch_bounds__(64)
syn_csr_kernel(CsrView cv, const uint32_t *__restrict__ list, const uint32_t *__restrict__ count_ptr,
               uint32_t *__restrict__ next_list, uint32_t *__restrict__ next_count,
               const uint64_t *__restrict__ hb, uint32_t it_report, int last,
               uint64_t *__restrict__ msg, uint32_t *__restrict__ iters, uint8_t *__restrict__ conv,
               uint8_t *__restrict__ synd) {
    extern __shared__ uint64_t sh[];
    const uint32_t count = *count_ptr;
    const uint32_t nw = (cv.n + 63) / 64, kw = (cv.k + 63) / 64;
    for (uint32_t item = blockIdx.x; item < count; item += gridDim.x) {
        const uint32_t f = list[item];
        for (uint32_t w = threadIdx.x; w < nw; w += 64) sh[w] = hb[(size_t)f * nw + w];
        __syncthreads();
        uint32_t bad = 0;
        for (uint32_t r = threadIdx.x; r < cv.m; r += 64) {
            uint32_t par = 0;
            for (uint32_t e = cv.row_ptr[r]; e < cv.row_ptr[r + 1]; ++e) {
                const uint32_t v = cv.col_idx[e];
                par ^= (uint32_t)(sh[v >> 6] >> (v & 63u)) & 1u;
            }
            bad |= par;
        }
        const int unsat = __syncthreads_or(bad != 0);
        if (!unsat || last) {
            for (uint32_t w = threadIdx.x; w < kw; w += 64) {
                uint64_t word = sh[w];
                if (64u * w + 64u > cv.k) word &= (1ull << (cv.k - 64u * w)) - 1;
                msg[(size_t)f * kw + w] = word;
            }
            if (threadIdx.x == 0) {
                iters[f] = it_report;
                conv[f] = (uint8_t)!unsat;
                synd[f] = (uint8_t)!unsat;
            }
        } else if (threadIdx.x == 0) {
            next_list[atomicAdd(next_count, 1u)] = f;
        }
        __syncthreads();
    }
}

__global__ void csr_init_kernel(uint32_t *list, uint32_t *counts, uint32_t frames) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < frames) list[t] = t;
    if (t == 0) {
        counts[0] = frames;
        counts[1] = 0;
    }
}

}  // namespace

// Decoder driver of the CSR path (same semantics as decode_device's plain
// min-sum): device buffers, synchronizes the stream once per iteration.
int decode_csr(gf2l_ctx *ctx, const gf2l_code *code, const float *d_llr, uint32_t F, uint32_t max_it,
               uint64_t *d_msg, uint32_t *d_iters, uint8_t *d_conv, uint8_t *d_synd) {
    const CsrView &cv = code->csr;
    hipStream_t s = ctx->stream;
    const size_t nw = (cv.n + 63) / 64;
    // the context's reusable workspace: c2v [F][nnz], beliefs [F][n], hard bits [F][nw]
    const int rc = ensure_buffers(ctx, F, (size_t)F * std::max<uint64_t>(cv.nnz, 1), 0, (size_t)F * cv.n,
                                  (size_t)F * nw, 0);
    if (rc != GF2L_OK) return rc;
    Workspace &ws = ctx->ws;
    float *c2v = ws.c2v, *bel = ws.belief;
    uint64_t *hb = ws.hb;
    uint32_t *lists = ws.lists, *counts = ws.counts, *h_count = ws.h_count;
    GF2L_HIP(ctx, hipMemcpyAsync(bel, d_llr, (size_t)F * cv.n * sizeof(float), hipMemcpyDeviceToDevice, s));
    GF2L_HIP(ctx, hipMemsetAsync(c2v, 0, (size_t)F * cv.nnz * sizeof(float), s));
    hipLaunchKernelGGL(csr_init_kernel, dim3((F + 255) / 256), dim3(256), 0, s, lists, counts, F);
    GF2L_HIP(ctx, hipGetLastError());
    const uint32_t grid = (uint32_t)ctx->num_cus * 4u;
    for (uint32_t it = 1; it <= max_it; ++it) {
        const uint32_t cur = (it - 1) & 1u, nxt = it & 1u;
        const uint32_t *list = lists + (size_t)cur * F;
        hipLaunchKernelGGL(cn_csr_kernel, dim3(grid), dim3(256), 0, s, cv, list, counts + cur, c2v, bel);
        hipLaunchKernelGGL(vn_csr_kernel, dim3(grid), dim3(256), 0, s, cv, list, counts + cur, c2v, bel,
                           d_llr, hb, counts + nxt);
        const uint32_t g_syn = std::min<uint32_t>(F, (uint32_t)ctx->num_cus * 8u);
        hipLaunchKernelGGL(syn_csr_kernel, dim3(g_syn), dim3(64), nw * sizeof(uint64_t), s, cv, list,
                           counts + cur, lists + (size_t)nxt * F, counts + nxt, hb, it, (int)(it == max_it),
                           d_msg, d_iters, d_conv, d_synd);
        GF2L_HIP(ctx, hipGetLastError());
        GF2L_HIP(ctx, hipMemcpyAsync(h_count, counts + nxt, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        GF2L_HIP(ctx, hipStreamSynchronize(s));
        if (*h_count == 0) break;
    }
    GF2L_HIP(ctx, hipStreamSynchronize(s));
    return GF2L_OK;
}

}  // namespace gf2l

using gf2l::fail;

extern "C" int gf2l_code_edges(gf2l_ctx *ctx, uint32_t m, uint32_t n, const uint32_t *rc_pairs,
                               size_t nnz, gf2l_code **out) {
    if (!ctx || !out) return GF2L_ERR_INVALID;
    *out = nullptr;
    return gf2l::guard(ctx, [&]() -> int {
        // SpBitMatrix::from_coo (sparse.rs:166-204): bounds checks, per-row sort,
        // duplicates cancel in pairs; then the stable transpose
        gf2l::host::Csr h;
        std::string err;
        const int hrc = gf2l::host::build_csr(m, n, rc_pairs, nnz, h, err);
        if (hrc != GF2L_OK) return fail(ctx, hrc, err);
        if (h.max_dc > gf2l::kMaxCheckDegree) return fail(ctx, GF2L_ERR_UNSUPPORTED, "check degree above 32");
        const std::vector<uint32_t> &row_ptr = h.row_ptr, &col_idx = h.col_idx, &col_ptr = h.col_ptr,
                                    &csc_edge = h.csc_edge;
        const uint32_t max_dc = h.max_dc;
        const uint32_t E = (uint32_t)col_idx.size();
        GF2L_HIP(ctx, hipSetDevice(ctx->device));
        const size_t b_rp = ((size_t)m + 1) * 4, b_ci = (size_t)std::max(E, 1u) * 4, b_cp = ((size_t)n + 1) * 4;
        auto al = [](size_t x) { return (x + 15) & ~(size_t)15; };
        const size_t o_ci = al(b_rp), o_cp = al(o_ci + b_ci), o_ce = al(o_cp + b_cp), total = o_ce + b_ci;
        std::vector<char> host(total, 0);
        std::memcpy(host.data(), row_ptr.data(), b_rp);
        if (E) std::memcpy(host.data() + o_ci, col_idx.data(), (size_t)E * 4);
        std::memcpy(host.data() + o_cp, col_ptr.data(), b_cp);
        if (E) std::memcpy(host.data() + o_ce, csc_edge.data(), (size_t)E * 4);
        gf2l_code *c = new gf2l_code;
        c->device = ctx->device;
        c->is_csr = true;
        hipError_t he = hipMalloc(&c->d_tables, total);
        if (he == hipSuccess) he = hipMemcpy(c->d_tables, host.data(), total, hipMemcpyHostToDevice);
        if (he != hipSuccess) {
            hipFree(c->d_tables);
            delete c;
            return fail(ctx, GF2L_ERR_HIP, std::string("code upload: ") + hipGetErrorString(he));
        }
        char *d = static_cast<char *>(c->d_tables);
        gf2l::CsrView &v = c->csr;
        v.m = m;
        v.n = n;
        v.k = n > m ? n - m : 0;
        v.nnz = E;
        v.row_ptr = reinterpret_cast<const uint32_t *>(d);
        v.col_idx = reinterpret_cast<const uint32_t *>(d + o_ci);
        v.col_ptr = reinterpret_cast<const uint32_t *>(d + o_cp);
        v.csc_edge = reinterpret_cast<const uint32_t *>(d + o_ce);
        // the QC view stays empty except the sizes used by the generic helpers
        c->view = gf2l::CodeView{};
        c->view.n = n;
        c->view.m = m;
        c->view.k = v.k;
        c->view.max_dc = max_dc;
        *out = c;
        return GF2L_OK;
    });
}
