// On-device frame generation for a staircase QC-IRA code: random information
// bits, systematic encoding, BPSK over AWGN and channel LLRs.
//
// Per frame g (global frame id) and seed s, with Philox4x32-10 keyed by
// (s_lo, s_hi) and counters {k, g_lo, g_hi, stream}:
//   information bits  block bc < kb, 64-bit word w of its z-bit vector =
//                     Philox({bc*zw + w, g, 0}) words 0,1 (LE), bits >= z
//                     cleared
//   Eb/N0 (dB)        lo + (hi - lo) * u,  u = (word0 + 0.5) / 2^32 of
//                     Philox({0, g, 2})
//   noise             variable v: pair q = v / 2 of Box-Muller on
//                     u1 = (word0 + 0.5)/2^32, u2 = (word1 + 0.5)/2^32 of
//                     Philox({q, g, 1}):  r = sqrt(-2 ln u1),
//                     n = r cos(2 pi u2) for even v, r sin(2 pi u2) for odd v
//   channel           bit 0 -> +1, bit 1 -> -1; y = x + sigma n with
//                     sigma^2 = 1 / (2 R 10^(EbN0/10)); LLR = (float)(2 y / sigma^2)
//                     in double (crates/gf2-coding/src/channel.rs:51-57,
//                     100-102, 153-158)
//   encoding          parity block-column kb+j carries p_j = s_j xor p_{j-1}
//                     (p_{-1} = 0), s_j = XOR over the info blocks (j, bc, sh)
//                     of the info block rotated by sh: the running XOR that
//                     solves the dual-diagonal staircase.
#include <hip/hip_runtime.h>

#include "internal.h"

namespace gf2l {
namespace {

__device__ __forceinline__ void philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                       uint32_t k1, uint32_t out[4]) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

__device__ __forceinline__ uint64_t cyc64g(const uint64_t *v, uint32_t z, uint32_t zw, uint32_t pos) {
    const uint32_t wi = pos >> 6, sh = pos & 63u;
    uint64_t r = v[wi] >> sh;
    if (sh && wi + 1 < zw) r |= v[wi + 1] << (64u - sh);
    const uint32_t avail = z - pos;
    if (avail < 64u) r |= v[0] << avail;
    return r;
}

__device__ __forceinline__ uint32_t getbit(const uint64_t *blk, uint32_t zw, uint32_t z, uint32_t v) {
    const uint32_t bc = v / z, j = v - bc * z;
    return (uint32_t)(blk[bc * zw + (j >> 6)] >> (j & 63u)) & 1u;
}

// One workgroup per frame; the frame's codeword lives in LDS as nb blocks of
// zw 64-bit words.  Requires z >= 64 (the word-parallel encoder) -- the host
// checks it.
__global__ void __launch_bounds__(256)
framegen_kernel(CodeView cv, uint32_t seed_lo, uint32_t seed_hi, uint64_t first_frame, uint32_t frames,
                float ebn0_lo, float ebn0_hi, float *__restrict__ llr, uint64_t *__restrict__ cw_words) {
    extern __shared__ uint64_t blk[];  // nb * zw
    const uint32_t z = cv.z, zw = cv.zw, kb = cv.nb - cv.mb;
    const double rate = (double)cv.k / (double)cv.n;
    const uint32_t nw = (cv.n + 63) / 64;
    for (uint32_t f = blockIdx.x; f < frames; f += gridDim.x) {
        const uint64_t g = first_frame + f;
        const uint32_t g_lo = (uint32_t)g, g_hi = (uint32_t)(g >> 32);
        // information blocks
        for (uint32_t t = threadIdx.x; t < kb * zw; t += blockDim.x) {
            uint32_t o[4];
            philox(t, g_lo, g_hi, 0u, seed_lo, seed_hi, o);
            uint64_t w = (uint64_t)o[0] | ((uint64_t)o[1] << 32);
            const uint32_t wi = t % zw;
            if (wi == zw - 1 && (z & 63u)) w &= (1ull << (z & 63u)) - 1;
            blk[t] = w;
        }
        __syncthreads();
        // block-row syndromes of the information part into the parity blocks
        for (uint32_t t = threadIdx.x; t < cv.mb * zw; t += blockDim.x) {
            const uint32_t br = t / zw, w = t - br * zw;
            uint64_t acc = 0;
            for (uint32_t e = cv.row_ptr[br]; e < cv.row_ptr[br + 1]; ++e) {
                const uint2 re = cv.row_edge[e];
                if (re.x >= kb) continue;
                uint32_t pos = 64u * w + re.y;
                if (pos >= z) pos -= z;
                acc ^= cyc64g(blk + re.x * zw, z, zw, pos);
            }
            if (w == zw - 1 && (z & 63u)) acc &= (1ull << (z & 63u)) - 1;
            blk[(kb + br) * zw + w] = acc;
        }
        __syncthreads();
        // staircase: p_j = s_j ^ p_{j-1}, word-parallel, block-rows in order
        if (threadIdx.x < zw)
            for (uint32_t br = 1; br < cv.mb; ++br)
                blk[(kb + br) * zw + threadIdx.x] ^= blk[(kb + br - 1) * zw + threadIdx.x];
        __syncthreads();
        // Eb/N0 of this frame
        uint32_t o[4];
        philox(0u, g_lo, g_hi, 2u, seed_lo, seed_hi, o);
        const double u = ((double)o[0] + 0.5) * (1.0 / 4294967296.0);
        const double ebn0 = (double)ebn0_lo + ((double)ebn0_hi - (double)ebn0_lo) * u;
        const double s2 = 1.0 / (2.0 * rate * pow(10.0, ebn0 / 10.0));
        const double sigma = sqrt(s2);
        // channel LLRs, two variables per Box-Muller pair
        for (uint32_t q = threadIdx.x; 2 * q < cv.n; q += blockDim.x) {
            philox(q, g_lo, g_hi, 1u, seed_lo, seed_hi, o);
            const double u1 = ((double)o[0] + 0.5) * (1.0 / 4294967296.0);
            const double u2 = ((double)o[1] + 0.5) * (1.0 / 4294967296.0);
            const double r = sqrt(-2.0 * log(u1));
            const double th = 6.283185307179586476925286766559 * u2;
            const uint32_t v0 = 2 * q;
            const double x0 = getbit(blk, zw, z, v0) ? -1.0 : 1.0;
            llr[(size_t)f * cv.n + v0] = (float)(2.0 * (x0 + sigma * (r * cos(th))) / s2);
            if (v0 + 1 < cv.n) {
                const double x1 = getbit(blk, zw, z, v0 + 1) ? -1.0 : 1.0;
                llr[(size_t)f * cv.n + v0 + 1] = (float)(2.0 * (x1 + sigma * (r * sin(th))) / s2);
            }
        }
        // codeword bits in the linear LE layout
        if (cw_words) {
            for (uint32_t w = threadIdx.x; w < nw; w += blockDim.x) {
                const uint32_t v0 = 64u * w, bc = v0 / z, j = v0 - bc * z;
                const uint64_t *b = blk + bc * zw;
                const uint32_t wi = j >> 6, s = j & 63u;
                uint64_t r = b[wi] >> s;
                if (s && wi + 1 < zw) r |= b[wi + 1] << (64u - s);
                const uint32_t avail = z - j;
                if (avail < 64u && bc + 1 < cv.nb) r |= b[zw] << avail;
                if (v0 + 64u > cv.n) r &= (1ull << (cv.n - v0)) - 1;
                cw_words[(size_t)f * nw + w] = r;
            }
        }
        __syncthreads();
    }
}

// All-zero codeword frames of any code (the reference's FER example sends
// all-zero codewords, crates/gf2-coding/examples/ldpc_awgn.rs): the same
// Eb/N0 and noise streams as framegen_kernel with every bit 0, so x = +1.
__global__ void __launch_bounds__(256)
zero_framegen_kernel(uint32_t n, double rate, uint32_t seed_lo, uint32_t seed_hi, uint64_t first_frame,
                     uint32_t frames, float ebn0_lo, float ebn0_hi, float *__restrict__ llr) {
    for (uint32_t f = blockIdx.x; f < frames; f += gridDim.x) {
        const uint64_t g = first_frame + f;
        const uint32_t g_lo = (uint32_t)g, g_hi = (uint32_t)(g >> 32);
        uint32_t o[4];
        philox(0u, g_lo, g_hi, 2u, seed_lo, seed_hi, o);
        const double u = ((double)o[0] + 0.5) * (1.0 / 4294967296.0);
        const double ebn0 = (double)ebn0_lo + ((double)ebn0_hi - (double)ebn0_lo) * u;
        const double s2 = 1.0 / (2.0 * rate * pow(10.0, ebn0 / 10.0));
        const double sigma = sqrt(s2);
        for (uint32_t q = threadIdx.x; 2 * q < n; q += blockDim.x) {
            philox(q, g_lo, g_hi, 1u, seed_lo, seed_hi, o);
            const double u1 = ((double)o[0] + 0.5) * (1.0 / 4294967296.0);
            const double u2 = ((double)o[1] + 0.5) * (1.0 / 4294967296.0);
            const double r = sqrt(-2.0 * log(u1));
            const double th = 6.283185307179586476925286766559 * u2;
            const uint32_t v0 = 2 * q;
            llr[(size_t)f * n + v0] = (float)(2.0 * (1.0 + sigma * (r * cos(th))) / s2);
            if (v0 + 1 < n) llr[(size_t)f * n + v0 + 1] = (float)(2.0 * (1.0 + sigma * (r * sin(th))) / s2);
        }
    }
}

}  // namespace

hipError_t launch_framegen(const CodeView &cv, uint64_t seed, uint64_t first_frame, uint32_t frames,
                           float ebn0_lo, float ebn0_hi, float *llr, uint64_t *cw_words, int num_cus,
                           hipStream_t s) {
    if (frames == 0) return hipSuccess;
    const uint32_t grid = frames < (uint32_t)num_cus * 8u ? frames : (uint32_t)num_cus * 8u;
    const size_t sh = (size_t)cv.nb * cv.zw * sizeof(uint64_t);
    hipLaunchKernelGGL(framegen_kernel, dim3(grid), dim3(256), sh, s, cv, (uint32_t)seed,
                       (uint32_t)(seed >> 32), first_frame, frames, ebn0_lo, ebn0_hi, llr, cw_words);
    return hipGetLastError();
}

hipError_t launch_framegen_zero(uint32_t n, uint32_t k, uint64_t seed, uint64_t first_frame, uint32_t frames,
                                float ebn0_lo, float ebn0_hi, float *llr, int num_cus, hipStream_t s) {
    if (frames == 0) return hipSuccess;
    const uint32_t grid = frames < (uint32_t)num_cus * 8u ? frames : (uint32_t)num_cus * 8u;
    hipLaunchKernelGGL(zero_framegen_kernel, dim3(grid), dim3(256), 0, s, n, (double)k / (double)n,
                       (uint32_t)seed, (uint32_t)(seed >> 32), first_frame, frames, ebn0_lo, ebn0_hi, llr);
    return hipGetLastError();
}

}  // namespace gf2l
