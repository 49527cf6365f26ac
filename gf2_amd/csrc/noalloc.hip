// gf2l_decode_batch_device_noalloc: the batch decoder of
// gf2l_decode_batch_device for callers that cannot allocate device memory
// (code under stream capture, tight serving loops, harnesses that forbid
// hipMalloc).  The decoder's working buffers -- compressed check states,
// beliefs, hard decisions, the running-frame lists and their counters, and
// the permuted LLRs of a DVB code -- are carved from one static device array
// of kScratchBytes instead of the context's workspace, so the call makes no
// allocation.  Each GPU the library runs on holds its own copy of the array
// from the first call on that device until the process ends.  Every call on
// a GPU shares that GPU's array: calls that overlap on one GPU (two contexts,
// two threads) race and must be serialized by the caller.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>

#include "gf2ldpc.h"
#include "internal.h"

namespace gf2l {
namespace {

constexpr size_t kScratchBytes = (size_t)GF2L_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` per device, looked up once per device with that
// device current (the array's name is only a host-side handle).
std::mutex g_mu;
void *g_scratch[kMaxDevices] = {};

int scratch_ptr(gf2l_ctx *ctx, unsigned char **out) {
    if (ctx->device < 0 || ctx->device >= kMaxDevices) return fail(ctx, GF2L_ERR_UNSUPPORTED, "device index above 63");
    std::lock_guard<std::mutex> lock(g_mu);
    void *&p = g_scratch[ctx->device];
    if (!p) {
        void *q = nullptr;
        GF2L_HIP(ctx, hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)));
        if (!q) return fail(ctx, GF2L_ERR_HIP, "hipGetSymbolAddress returned a null scratch address");
        p = q;
    }
    *out = static_cast<unsigned char *>(p);
    return GF2L_OK;
}

// Outputs of a fresh decoder that runs no iteration (core.rs:906-937):
// message 0, 0 iterations, not converged, syndrome passed.  A kernel rather
// than memsets, so callers may record the call into a graph.
__global__ void zero_iteration_kernel(uint64_t *msg, size_t msg_words, uint32_t *iters, uint8_t *conv,
                                      uint8_t *synd, uint32_t frames) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < msg_words; t += stride) msg[t] = 0;
    for (size_t f = (size_t)blockIdx.x * blockDim.x + threadIdx.x; f < frames; f += stride) {
        iters[f] = 0;
        conv[f] = 0;
        synd[f] = 1;
    }
}

constexpr size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Byte offsets of the STATE-path buffers of F frames inside the scratch
// array; returns the total.
struct Carve {
    size_t state, belief, hb, lists, counts, llr, total;
};
Carve carve(const CodeView &cv, size_t F) {
    Carve c{};
    c.state = 0;
    c.belief = c.state + align256(3ull * cv.mb * cv.zp * F * sizeof(uint32_t));
    c.hb = c.belief + align256((size_t)cv.nb * cv.zp * F * sizeof(float));
    c.lists = c.hb + align256((size_t)cv.nb * cv.zw * F * sizeof(uint64_t));
    c.counts = c.lists + align256(2 * F * sizeof(uint32_t));
    c.llr = c.counts + align256(kCountsWords * sizeof(uint32_t));
    c.total = c.llr + (cv.dvb_q ? align256((size_t)cv.n * F * sizeof(float)) : 0);
    return c;
}

}  // namespace
}  // namespace gf2l

using gf2l::fail;

extern "C" int gf2l_decode_batch_device_noalloc(gf2l_ctx *ctx, const gf2l_code *code, const float *d_llr,
                                                size_t frames, uint32_t max_it, uint64_t *d_msg_words,
                                                uint32_t *d_iterations, uint8_t *d_converged,
                                                uint8_t *d_syndrome_ok) {
    if (!ctx) return GF2L_ERR_INVALID;
    return gf2l::guard(ctx, [&]() -> int {
        if (!code) return fail(ctx, GF2L_ERR_INVALID, "null code");
        if (frames == 0) return GF2L_OK;
        if (!d_llr || !d_msg_words || !d_iterations || !d_converged || !d_syndrome_ok)
            return fail(ctx, GF2L_ERR_INVALID, "null buffer");
        if (frames > 0xffffffffull) return fail(ctx, GF2L_ERR_INVALID, "too many frames in one batch");
        if (ctx->rule != GF2L_RULE_MINSUM)
            return fail(ctx, GF2L_ERR_UNSUPPORTED, "the no-allocation decoder runs the min-sum rule only");
        const gf2l::CodeView &cv = code->view;
        if (code->is_csr || cv.max_dc > gf2l::kMaxStateDegree)
            return fail(ctx, GF2L_ERR_UNSUPPORTED,
                        "gf2l_decode_batch_device_noalloc needs a quasi-cyclic code with check degree <= 27");
        GF2L_HIP(ctx, hipSetDevice(ctx->device));
        const uint32_t F = (uint32_t)frames;
        hipStream_t s = ctx->stream;
        if (max_it == 0) {
            const size_t words = (size_t)F * ((cv.k + 63) / 64);
            const uint32_t grid = (uint32_t)std::min<size_t>((std::max<size_t>(words, F) + 255) / 256, 4096);
            hipLaunchKernelGGL(gf2l::zero_iteration_kernel, dim3(grid), dim3(256), 0, s, d_msg_words, words,
                               d_iterations, d_converged, d_syndrome_ok, F);
            GF2L_HIP(ctx, hipGetLastError());
            return GF2L_OK;
        }
        const int path = gf2l::resolve_path(ctx, cv);
        if (path == gf2l::kLds) {  // the whole frame state lives in LDS: no scratch at all
            GF2L_HIP(ctx, gf2l::launch_lds_decode(cv, d_llr, F, max_it, nullptr, 0, false, d_msg_words, d_iterations,
                                                  d_converged, d_syndrome_ok, nullptr, nullptr, nullptr, 0,
                                                  ctx->tune.lds_groups, s));
            return GF2L_OK;
        }
        if (path != gf2l::kState)
            return fail(ctx, GF2L_ERR_UNSUPPORTED,
                        "gf2l_decode_batch_device_noalloc runs the LDS and STATE paths only");
        const gf2l::Carve c = gf2l::carve(cv, F);
        if (c.total > gf2l::kScratchBytes) {
            size_t fit = 0;  // largest batch of this code that fits
            for (size_t lo = 1, hi = F; lo <= hi;) {
                const size_t mid = lo + (hi - lo) / 2;
                if (gf2l::carve(cv, mid).total <= gf2l::kScratchBytes) {
                    fit = mid;
                    lo = mid + 1;
                } else {
                    hi = mid - 1;
                }
            }
            return fail(ctx, GF2L_ERR_INVALID,
                        "gf2l_decode_batch_device_noalloc: " + std::to_string(F) + " frames need " +
                            std::to_string(c.total) + " bytes of scratch, more than its " +
                            std::to_string(gf2l::kScratchBytes) + "; split the batch (at most " +
                            std::to_string(fit) + " frames of this code per call)");
        }
        unsigned char *base = nullptr;
        int rc = gf2l::scratch_ptr(ctx, &base);
        if (rc != GF2L_OK) return rc;
        uint32_t h_counts[gf2l::kProfileSlots] = {};  // pageable: read back after a stream synchronize
        gf2l::Workspace ws{};
        ws.cap_frames = F;
        ws.state = reinterpret_cast<uint32_t *>(base + c.state);
        ws.belief = reinterpret_cast<float *>(base + c.belief);
        ws.hb = reinterpret_cast<uint64_t *>(base + c.hb);
        ws.lists = reinterpret_cast<uint32_t *>(base + c.lists);
        ws.counts = reinterpret_cast<uint32_t *>(base + c.counts);
        ws.llr_perm = cv.dvb_q ? reinterpret_cast<float *>(base + c.llr) : nullptr;
        ws.h_count = h_counts;
        return gf2l::run_hbm(ctx, cv, path, ws, d_llr, F, max_it, nullptr, 0, false, d_msg_words, d_iterations,
                             d_converged, d_syndrome_ok, nullptr, nullptr, nullptr, 0, nullptr, nullptr,
                             /*graphs=*/false);  // instantiating a graph allocates
    });
}
