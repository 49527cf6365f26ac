// Device helpers shared by the decoder kernels: the min-sum check rule, the
// compressed check state, circulant index arithmetic, degree dispatch, and the
// per-item bodies of the check update, variable update and syndrome that the
// single-code and mixed-code kernels both run.
//
// Arithmetic contract (bit-exact with the reference decoder for finite LLRs,
// crates/gf2-coding/src/ldpc/core.rs:790-938):
//   * check message = sign_product * min |v2c| over the other d-1 edges, exact
//     (abs / compare / negate only); the sign of an input counts as negative
//     by its sign bit inside the AVX2 8-lane chunks and by `value < 0` in the
//     scalar remainder, as in crates/gf2-kernels-simd/src/llr.rs:117-180
//     (matters only for -0.0); a degree-1 check sends +0.0 (core.rs:804-806).
//   * belief = ((L + c2v_0) + c2v_1) + ... in ascending check order, then
//     v2c = belief - c2v (core.rs:819-838): sequential binary32 adds, built
//     with -ffp-contract=off so nothing fuses into an FMA.
//   * hard decision belief < 0 (llr.rs:73-75); syndrome over GF(2) by XOR.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

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
    // a < min1 moves min1 to min2; otherwise a < min2 replaces min2 (strict
    // comparisons, so the first of equal magnitudes keeps idx).
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
    // The same update as selects instead of branches: cheaper to issue for
    // small degrees, but it keeps more values live for large ones.
    __device__ void add_sel(uint32_t p, float x) {
        const float a = fabsf(x);
        const bool lt1 = a < min1, lt2 = a < min2;
        min2 = lt1 ? min1 : (lt2 ? a : min2);
        min1 = lt1 ? a : min1;
        idx = lt1 ? p : idx;
        sb |= (uint64_t)(__float_as_uint(x) >> 31) << p;
        lt |= (uint64_t)(x < 0.0f) << p;
    }
    // Sign of the message to position p: product of the other inputs' signs,
    // sign bit for inputs inside full 8-lane chunks of the extrinsic list,
    // `< 0` for the remainder.
    __device__ uint32_t neg(uint32_t d, uint32_t p) const {
        const uint32_t l8 = (d - 1) & ~7u;
        const uint64_t pb = 1ull << p;
        const uint64_t all = (1ull << d) - 1;
        const uint64_t lane = (p < l8) ? (((1ull << (l8 + 1)) - 1) & ~pb) : ((1ull << l8) - 1);
        const uint64_t rem = all & ~lane & ~pb;
        return (__popcll(sb & lane) ^ __popcll(lt & rem)) & 1u;
    }
    // Output sign bits of all d positions.  When sign bit and `< 0` agree
    // for every input (no -0.0) the rule reduces to parity(all) ^ own sign.
    template <int DMAX>
    __device__ uint32_t signs(uint32_t d) const {
        const uint64_t all = (1ull << d) - 1;
        if (sb == lt) return (uint32_t)(((__popcll(sb) & 1u) ? all : 0ull) ^ sb);
        uint32_t s = 0;
#pragma unroll
        for (int p = 0; p < DMAX; ++p)
            if (p < (int)d) s |= neg(d, p) << p;
        return s;
    }
    __device__ float out(uint32_t d, uint32_t p) const {
        if (d == 1) return 0.0f;
        const float mag = (p == idx) ? min2 : min1;
        return neg(d, p) ? -mag : mag;
    }
};

// Message of position p rebuilt from a compressed check state.
__device__ __forceinline__ float recon(float m1, float m2, uint32_t meta, uint32_t p) {
    const float mag = (p == (meta & 31u)) ? m2 : m1;
    return ((meta >> (5 + p)) & 1u) ? -mag : mag;
}
// The same value with the sign bit or-ed in: the magnitudes carry no sign
// bit (they are |x| minima), so output sign bit 5 + p of meta moves to bit
// 31 (p < 27).  One instruction fewer; used where it costs no registers.
__device__ __forceinline__ float recon_or(float m1, float m2, uint32_t meta, uint32_t p) {
    const float mag = (p == (meta & 31u)) ? m2 : m1;
    return __uint_as_float(__float_as_uint(mag) | ((meta << (26u - p)) & 0x80000000u));
}

__device__ __forceinline__ uint32_t wrap_add(uint32_t a, uint32_t b, uint32_t z) {
    uint32_t s = a + b;
    return s >= z ? s - z : s;
}
__device__ __forceinline__ uint32_t wrap_sub(uint32_t a, uint32_t b, uint32_t z) {
    return a >= b ? a - b : a + z - b;
}

// Element idx of a wave-uniform base pointer, addressed as base + 32-bit byte
// offset so the load / store takes the scalar-base form (one offset VGPR per
// access instead of a 64-bit address pair).  Per-frame arrays stay below
// 4 GiB (codes are limited to 2^28 checks and variables).
// Wave-uniform value moved to a scalar register (graph tables are read by
// every lane of a workgroup at the same address).
__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// Load through the constant address space: for a wave-uniform address this is
// a scalar load (s_load) even when the pointer itself came from memory.  Only
// for data no workgroup of the running kernel writes (code tables, weights,
// the running list being read).
template <typename T>
__device__ __forceinline__ T ldc(const T *p) {
    return *(const __attribute__((address_space(4))) T *)p;
}
// Vector types are read component-wise (one s_load_dwordx2 after merging).
__device__ __forceinline__ uint2 ldc(const uint2 *p) {
    const auto *q = (const __attribute__((address_space(4))) uint32_t *)p;
    return make_uint2(q[0], q[1]);
}

template <typename T>
__device__ __forceinline__ T ldu(const T *base, uint32_t idx) {
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + idx * (uint32_t)sizeof(T));
}
template <typename T>
__device__ __forceinline__ void stu(T *base, uint32_t idx, T v) {
    *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + idx * (uint32_t)sizeof(T)) = v;
}
// Same with the byte offset given (scaling folded into wave-uniform terms).
template <typename T>
__device__ __forceinline__ T ldb(const T *base, uint32_t off) {
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + off);
}

// Waves per SIMD the state kernels of a degree class are compiled for
// (amdgpu_waves_per_eu; 1 = no constraint).  The compiler's default
// allocation gave the class-8 variable kernel 106 scalar registers, which
// caps it at 7 waves; compiled for 8 it fits in 78 scalar and 56 vector
// registers without spills.  Class 16 also fits 8 (check) and 7 (variable)
// waves, but measured slower on the mixed-code stream, so it keeps the
// compiler's choice.
template <int MAXD> constexpr int kCheckWaves = MAXD <= 8 ? 8 : 1;
template <int MAXD> constexpr int kVarWaves = MAXD <= 8 ? 8 : 1;

// Calls f(std::integral_constant<int, d>) for a wave-uniform degree d in
// [0, MAXD], so per-edge loops are fully unrolled with no guards and every
// load of an item is issued before the first wait.  Degree 0 (a variable or
// check outside every edge) is a valid case.
#define GF2L_DEG_CASE(K)                                              \
    case K:                                                           \
        if constexpr (K <= MAXD) f(std::integral_constant<int, K>{}); \
        break;
template <int MAXD, typename F>
__device__ __forceinline__ void dispatch_degree(uint32_t d, F &&f) {
    switch (d) {
        GF2L_DEG_CASE(0) GF2L_DEG_CASE(1) GF2L_DEG_CASE(2) GF2L_DEG_CASE(3) GF2L_DEG_CASE(4)
        GF2L_DEG_CASE(5) GF2L_DEG_CASE(6) GF2L_DEG_CASE(7) GF2L_DEG_CASE(8)
        GF2L_DEG_CASE(9) GF2L_DEG_CASE(10) GF2L_DEG_CASE(11) GF2L_DEG_CASE(12)
        GF2L_DEG_CASE(13) GF2L_DEG_CASE(14) GF2L_DEG_CASE(15) GF2L_DEG_CASE(16)
        GF2L_DEG_CASE(17) GF2L_DEG_CASE(18) GF2L_DEG_CASE(19) GF2L_DEG_CASE(20)
        GF2L_DEG_CASE(21) GF2L_DEG_CASE(22) GF2L_DEG_CASE(23) GF2L_DEG_CASE(24)
        GF2L_DEG_CASE(25) GF2L_DEG_CASE(26) GF2L_DEG_CASE(27) GF2L_DEG_CASE(28)
        GF2L_DEG_CASE(29) GF2L_DEG_CASE(30) GF2L_DEG_CASE(31) GF2L_DEG_CASE(32)
        default: break;
    }
}
#undef GF2L_DEG_CASE

// Item stream of a persistent workgroup.  Items differ in cost (a variable's
// degree runs from 1 to 13), and with a fixed stride fast and slow
// workgroups drift apart until the items in flight span many frames and
// evict each other's check states from L2.  With a counter the workgroup
// takes chunks of G consecutive items: its first chunk is q*G, each later
// one the next value of the counter plus Q*G, fetched when the current chunk
// starts so the atomic's latency hides behind G items.  That keeps the items
// in flight within about G grid-widths.  ctr == nullptr: fixed stride Q.
// Counters of different XCDs sit kCtrStride words apart (separate cache
// lines and channels).  The caller's loop must be workgroup-uniform.
constexpr uint32_t kCtrStride = 64;

struct ItemStream {
    uint32_t j, cend, G, QG, Q, nxt, par;
    uint32_t *ctr;
    __device__ ItemStream(uint32_t q, uint32_t stride, uint32_t *counter, uint32_t chunk)
        : j(counter ? q * chunk : q), cend(counter ? q * chunk + chunk : 0), G(chunk), QG(stride * chunk),
          Q(stride), nxt(0), par(0), ctr(counter) {
        if (ctr && threadIdx.x == 0) nxt = atomicAdd(ctr, G) + QG;
    }
    // s_next: two LDS words (double-buffered: one barrier per chunk).
    __device__ void advance(uint32_t *s_next) {
        if (!ctr) {
            j += Q;
            return;
        }
        if (++j < cend) return;
        if (threadIdx.x == 0) s_next[par] = nxt;
        __syncthreads();
        j = uni(s_next[par]);
        par ^= 1u;
        cend = j + G;
        if (threadIdx.x == 0) nxt = atomicAdd(ctr, G) + QG;
    }
};

// ---------------------------------------------------------------------------
// STATE-path item bodies.  A frame's check state is [mb][zp] records of three
// u32 (min1, min2, argmin | output sign bits << 5), read and written as one
// 12-byte access; its beliefs are [nb][zp] floats.

// Output sign bits (indexed by table slot) of a DVB check with a -0.0
// input, where the AVX2 lane / remainder rule needs each input's place in
// the check's variable order at lane i.  The slots of `row` (d_tab of them)
// are in the standard's variable order except inside a block: slots of one
// block-column (circulants of one block) are adjacent and sorted by shift,
// and at lane i their order is that list rotated (slots with i + shift >= z
// wrapped and come first).  `skip` is a slot without an edge at this lane
// (or d_tab).  Out of line: only reached when an input is -0.0.
[[maybe_unused]] __device__ __noinline__ uint32_t dvb_signs_slow(const uint2 *row, uint32_t d_tab, uint32_t z, uint32_t i,
                                                uint32_t skip, uint64_t sb, uint64_t lt) {
    const uint32_t d = d_tab - (skip < d_tab ? 1u : 0u);
    auto pos = [&](uint32_t p) -> uint32_t {
        const uint32_t bc = row[p].x;
        uint32_t g0 = p, g1 = p + 1;
        while (g0 > 0 && row[g0 - 1].x == bc) --g0;
        while (g1 < d_tab && row[g1].x == bc) ++g1;
        const uint32_t g = g1 - g0, l = p - g0;
        uint32_t r = 0;  // slots of the group that did not wrap
        for (uint32_t q = g0; q < g1; ++q) r += row[q].y + i < z ? 1u : 0u;
        uint32_t x = l + g - r;
        if (x >= g) x -= g;
        return g0 + x - (skip < p ? 1u : 0u);
    };
    CheckAcc t;
    t.init();
    for (uint32_t p = 0; p < d_tab; ++p) {
        if (p == skip) continue;
        const uint32_t ps = pos(p);
        t.sb |= ((sb >> p) & 1ull) << ps;
        t.lt |= ((lt >> p) & 1ull) << ps;
    }
    uint32_t out = 0;
    for (uint32_t p = 0; p < d_tab; ++p)
        if (p != skip) out |= t.neg(d, pos(p)) << p;
    return out;
}

// Output sign bits (indexed by slot) of a DVB check: without a -0.0 input
// the sign of each output is the parity of the other inputs' signs,
// whatever their order.
template <int D>
__device__ __forceinline__ uint32_t dvb_signs(const CodeView &cv, const CheckAcc &acc, const uint2 *row,
                                              uint32_t i, uint32_t skip) {
    constexpr uint64_t all = D >= 64 ? ~0ull : (1ull << D) - 1;
    if (acc.sb == acc.lt) return (uint32_t)(((__popcll(acc.sb) & 1u) ? all : 0ull) ^ acc.sb);
    return dvb_signs_slow(row, D, cv.z, i, skip, acc.sb, acc.lt);
}

template <int MAXD, bool FIRST, bool NMS, bool DVB = false>
__device__ __forceinline__ void cn_state_item(const CodeView &cv, uint32_t br, const uint32_t *fsi,
                                              uint32_t *fso, const float *__restrict__ fb,
                                              const float *__restrict__ w_prev) {
    const uint3 *si = reinterpret_cast<const uint3 *>(fsi) + br * cv.zp;
    uint3 *st = reinterpret_cast<uint3 *>(fso) + br * cv.zp;
    const uint32_t bstride = FIRST ? cv.z : cv.zp;
    const uint32_t e0 = uni(ldc(cv.row_ptr + br));
    const uint32_t d = uni(ldc(cv.row_ptr + br + 1)) - e0;
    dispatch_degree<MAXD>(d, [&](auto DC) {
        constexpr int D = decltype(DC)::value, DA = D > 0 ? D : 1;
        uint2 re[DA];
        float wp[DA];
#pragma unroll
        for (int p = 0; p < D; ++p) {
            // uniform: scalar registers even where the compiler picks a
            // vector load (the state stores of earlier items may alias)
            const uint2 r = ldc(cv.row_edge + e0 + p);
            re[p] = make_uint2(uni(r.x), uni(r.y));
            if constexpr (NMS && !FIRST) wp[p] = __uint_as_float(uni(__float_as_uint(ldc(w_prev + e0 + p))));
        }
        // small degrees: byte offset of belief (i + s) mod z of block-column c
        // precomputed as 4*(c*stride + s) + 4i, less 4z once i >= z - s (two
        // more scalar registers per edge, which larger degrees cannot spare)
        constexpr bool kPre = MAXD <= 8;
        uint32_t boff[kPre ? DA : 1], bthr[kPre ? DA : 1];
        if constexpr (kPre) {
#pragma unroll
            for (int p = 0; p < D; ++p) {
                boff[p] = 4u * (re[p].x * bstride + re[p].y);
                bthr[p] = cv.z - re[p].y;
            }
        }
        const uint32_t z4 = 4u * cv.z;
        for (uint32_t i = threadIdx.x; i < cv.z; i += blockDim.x) {
            const uint32_t i4 = 4u * i;
            float m1o = 0.0f, m2o = 0.0f;
            uint32_t mo = 0;
            if constexpr (!FIRST) {  // first: the old state is needed by every edge
                const uint3 o = ldu(si, i);
                m1o = __uint_as_float(o.x);
                m2o = __uint_as_float(o.y);
                mo = o.z;
            }
            float xv[DA];
#pragma unroll
            for (int p = 0; p < D; ++p) {  // belief (i + shift) mod z of the slot's block-column
                if constexpr (kPre)
                    xv[p] = ldb(fb, i4 + boff[p] - (i >= bthr[p] ? z4 : 0u));
                else
                    xv[p] = ldu(fb, re[p].x * bstride + wrap_add(i, re[p].y, cv.z));
            }
            // small degrees: every load of the item issued before the first use
            // (larger ones keep the compiler's interleaving, which needs fewer registers)
            if constexpr (MAXD <= 8) __builtin_amdgcn_sched_barrier(0);
            // DVB: block-row 0 has no staircase edge at lane 0 (check 0 of
            // the standard code); `skip` is that slot, D when every slot is live
            uint32_t skip = D;
            if constexpr (DVB) {
                if (br == 0 && i == 0) {
#pragma unroll
                    for (int p = 0; p < D; ++p)
                        if (re[p].x == cv.wrap_bc) skip = p;
                }
            }
            CheckAcc acc;
            acc.init();
#pragma unroll
            for (int p = 0; p < D; ++p) {
                if (DVB && (uint32_t)p == skip) continue;
                float v = xv[p];
                if constexpr (!FIRST) {
                    float c = recon_or(m1o, m2o, mo, p);
                    if constexpr (NMS) c = wp[p] * c;
                    v = v - c;
                }
                if constexpr (kPre)
                    acc.add_sel(p, v);
                else
                    acc.add(p, v);
            }
            float m1 = acc.min1, m2 = acc.min2;
            uint32_t meta;
            if constexpr (DVB)
                meta = acc.idx | (dvb_signs<D>(cv, acc, cv.row_edge + e0, i, skip) << 5);
            else
                meta = acc.idx | (acc.signs<D>(D) << 5);
            if constexpr (D == 1) {  // the lone message is +0.0
                m1 = 0.0f;
                m2 = 0.0f;
                meta = 0;
            }
            stu(st, i, make_uint3(__float_as_uint(m1), __float_as_uint(m2), meta));
        }
    });
}

// Variable update of the z variables of block-column bc of one frame:
// belief = L + sum of the rebuilt check messages in ascending block-row
// order; hard decisions packed by wave ballot into 64-bit words (fh may be
// null).  fs: frame's state, fl: the block-column's LLRs, fb: its beliefs.
// Needs blockDim.x % 64 == 0.
template <int MAXD, bool NMS, bool DVB = false>
__device__ __forceinline__ void vn_state_item(const CodeView &cv, uint32_t bc, const uint32_t *__restrict__ fs,
                                              const float *__restrict__ fl, float *__restrict__ fb,
                                              uint64_t *__restrict__ fh, const float *__restrict__ w_cur) {
    const uint3 *fst = reinterpret_cast<const uint3 *>(fs);
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t c0 = uni(ldc(cv.col_ptr + bc));
    const uint32_t dv = uni(ldc(cv.col_ptr + bc + 1)) - c0;
    dispatch_degree<MAXD>(dv, [&](auto DC) {
        constexpr int D = decltype(DC)::value, DA = D > 0 ? D : 1;
        uint2 ci[DA];  // {first state record of the block-row, shift << 5 | position}
        float wq[DA];
#pragma unroll
        for (int q = 0; q < D; ++q) {
            const uint2 r = ldc(cv.col_pack + c0 + q);
            ci[q] = make_uint2(uni(r.x), uni(r.y));
            if constexpr (NMS) wq[q] = __uint_as_float(uni(__float_as_uint(ldc(w_cur + ldc(&cv.col_info[c0 + q].w)))));
        }
        // DVB: per slot the lane below which its check lane does not wrap
        // (its shift, or z for shift 0); whether every shift is 0
        uint32_t thr[DVB ? DA : 1];
        bool uni_all_zero_shift = true;
        if constexpr (DVB) {
#pragma unroll
            for (int q = 0; q < D; ++q) {
                const uint32_t sh = ci[q].y >> 5;
                thr[q] = sh ? sh : cv.z;
                uni_all_zero_shift = uni_all_zero_shift && sh == 0;
            }
        }
        for (uint32_t j0 = 0; j0 < cv.z; j0 += blockDim.x) {
            const uint32_t j = j0 + threadIdx.x;
            const bool valid = j < cv.z;
            float bel = 0.0f;
            if (valid) {
                float m1[DA], m2[DA];
                uint32_t mt[DA];
#pragma unroll
                for (int q = 0; q < D; ++q) {
                    const uint3 o = ldu(fst, ci[q].x + wrap_sub(j, ci[q].y >> 5, cv.z));
                    m1[q] = __uint_as_float(o.x);
                    m2[q] = __uint_as_float(o.y);
                    mt[q] = o.z;
                }
                bel = ldu(fl, j);
                if constexpr (DVB) {
                    // Slots are in ascending standard check order at lane 0
                    // (check a + q*b of block-row a, lane b).  At lane j the
                    // order is that list rotated: slots whose check lane
                    // (j + s) mod z wrapped come first (s = z - shift, so a
                    // slot stays in place while j < shift, or always for
                    // shift 0).  Parity columns other than the last have
                    // only shift-0 slots and keep their order (uniform test).
                    float c[DA];
#pragma unroll
                    for (int q = 0; q < D; ++q) c[q] = recon(m1[q], m2[q], mt[q], ci[q].y & 31u);
                    if (uni_all_zero_shift) {
#pragma unroll
                        for (int q = 0; q < D; ++q) bel = bel + c[q];
                    } else {
                        uint32_t r = 0;
#pragma unroll
                        for (int q = 0; q < D; ++q) r += (j < thr[q]) ? 1u : 0u;
#pragma unroll
                        for (int k = 1; k < D; k <<= 1) {
                            const bool rot = (r & (uint32_t)k) != 0;
                            float t[DA];
#pragma unroll
                            for (int q = 0; q < D; ++q) t[q] = rot ? c[(q + k) % DA] : c[q];
#pragma unroll
                            for (int q = 0; q < D; ++q) c[q] = t[q];
                        }
                        if (bc == cv.wrap_bc) {
                            // the staircase has no edge from check 0 to the last
                            // parity bit (lane z-1, the block-row-0 slot)
                            uint64_t live = 0;
#pragma unroll
                            for (int q = 0; q < D; ++q)
                                live |= (uint64_t)((ci[q].x == 0 && j == cv.z - 1) ? 0u : 1u) << q;
                            const uint64_t full = (D >= 64 ? ~0ull : (1ull << D) - 1);
                            const uint32_t rr = r % (uint32_t)DA;
                            live = ((live >> rr) | (live << ((DA - rr) % 64))) & full;
#pragma unroll
                            for (int q = 0; q < D; ++q)
                                if ((live >> q) & 1ull) bel = bel + c[q];
                        } else {
#pragma unroll
                            for (int q = 0; q < D; ++q) bel = bel + c[q];
                        }
                    }
                } else {
#pragma unroll
                    for (int q = 0; q < D; ++q) {
                        float c = recon(m1[q], m2[q], mt[q], ci[q].y & 31u);
                        if constexpr (NMS) c = wq[q] * c;
                        bel = bel + c;
                    }
                }
                stu(fb, j, bel);
            }
            if (fh) {
                const uint64_t mask = __ballot(valid && bel < 0.0f);
                const uint32_t word_bit = j0 + wave * 64;
                if (lane == 0 && word_bit < cv.z) fh[word_bit >> 6] = mask;
            }
        }
    });
}

// ---------------------------------------------------------------------------
// Syndrome and message extraction of one frame by one workgroup (any size).

// Slot value the syndrome pass leaves in the next running list for a frame
// that stopped.
constexpr uint32_t kRetiredSlot = 0xffffffffu;

// Removes the kRetiredSlot holes from list[0, count) in place, keeping the
// other slots in order, and returns how many remain.  Called by a whole
// workgroup (blockDim.x a multiple of 64, at most 1024; wave_kept holds
// blockDim.x / 64 entries): every write lands at or before the slot it came
// from, so the slots of a batch of chunks can all be loaded before any of
// the batch's writes (which never reach a later batch).  Loading eight chunks
// at once pays one memory latency per batch instead of one per chunk.
__device__ __forceinline__ uint32_t compact_list(uint32_t *__restrict__ list, uint32_t count,
                                                 uint32_t *wave_kept) {
    constexpr int kBatch = 8;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t waves = blockDim.x >> 6;
    uint32_t base = 0;
    for (uint64_t c0 = 0; c0 < count; c0 += kBatch * blockDim.x) {
        uint32_t v[kBatch];
#pragma unroll
        for (int k = 0; k < kBatch; ++k) {
            const uint64_t i = c0 + k * blockDim.x + threadIdx.x;
            v[k] = i < count ? list[i] : kRetiredSlot;
        }
#pragma unroll
        for (int k = 0; k < kBatch; ++k) {
            if (c0 + k * blockDim.x >= count) break;  // uniform over the workgroup
            const bool keep = v[k] != kRetiredSlot;
            const uint64_t kept = __ballot(keep);
            if (lane == 0) wave_kept[wave] = (uint32_t)__popcll(kept);
            __syncthreads();
            uint32_t off = base, total = 0;
            for (uint32_t w = 0; w < waves; ++w) {
                const uint32_t n = wave_kept[w];
                if (w < wave) off += n;
                total += n;
            }
            if (keep) list[off + (uint32_t)__popcll(kept & ((1ull << lane) - 1ull))] = v[k];
            base += total;
            __syncthreads();
        }
    }
    return base;
}

// Bits pos .. pos+63 (cyclically mod z) of a z-bit vector stored in zw
// words whose bits at z and above are zero.  Needs z >= 64.
__device__ __forceinline__ uint64_t cyc64(const uint64_t *v, uint32_t z, uint32_t zw, uint32_t pos) {
    const uint32_t wi = pos >> 6, sh = pos & 63u;
    uint64_t r = v[wi] >> sh;
    if (sh && wi + 1 < zw) r |= v[wi + 1] << (64u - sh);
    const uint32_t avail = z - pos;
    if (avail < 64u) r |= v[0] << avail;
    return r;
}

// Returns nonzero (uniform over the workgroup) when a check is unsatisfied
// by the frame's nb*zw hard-decision words in sh (LDS, complete and visible
// to the whole workgroup).
__device__ __forceinline__ int words_unsat(const CodeView &cv, const uint64_t *sh) {
    const uint32_t lane = threadIdx.x;
    const uint32_t z = cv.z, zw = cv.zw;
    uint64_t bad = 0;
    if (z >= 64) {
        for (uint32_t t = lane; t < cv.mb * zw; t += blockDim.x) {
            const uint32_t br = t / zw, w = t - br * zw;
            uint64_t acc = 0;
            for (uint32_t e = cv.row_ptr[br]; e < cv.row_ptr[br + 1]; ++e) {
                const uint2 re = cv.row_edge[e];
                uint64_t v = cyc64(sh + re.x * zw, z, zw, wrap_add(64u * w, re.y, z));
                if (br == 0 && w == 0 && re.x == cv.wrap_bc) v &= ~1ull;  // no edge at check 0
                acc ^= v;
            }
            if (w == zw - 1 && (z & 63u)) acc &= (1ull << (z & 63u)) - 1;
            bad |= acc;
        }
    } else {
        for (uint32_t c = lane; c < cv.m; c += blockDim.x) {
            const uint32_t br = c / z, i = c - br * z;
            uint32_t par = 0;
            for (uint32_t e = cv.row_ptr[br]; e < cv.row_ptr[br + 1]; ++e) {
                const uint2 re = cv.row_edge[e];
                if (br == 0 && i == 0 && re.x == cv.wrap_bc) continue;
                par ^= (uint32_t)(sh[re.x * zw] >> wrap_add(i, re.y, z)) & 1u;
            }
            bad |= par;
        }
    }
    return __syncthreads_or(bad != 0);
}

// Copies the frame's nb*zw hard-decision words hbf into sh (LDS) and
// returns nonzero (uniform over the workgroup) when a check is unsatisfied.
__device__ __forceinline__ int frame_unsat(const CodeView &cv, uint64_t *sh, const uint64_t *__restrict__ hbf) {
    const uint32_t words = cv.nb * cv.zw;
    for (uint32_t w = threadIdx.x; w < words; w += blockDim.x) sh[w] = hbf[w];
    __syncthreads();
    return words_unsat(cv, sh);
}

// Writes the first k hard decisions (the message of a systematic code) from
// sh as ceil(k/64) packed words to msgf.
__device__ __forceinline__ void frame_message(const CodeView &cv, const uint64_t *sh, uint64_t *__restrict__ msgf) {
    const uint32_t kw = (cv.k + 63) / 64;
    const uint32_t z = cv.z, zw = cv.zw;
    for (uint32_t w = threadIdx.x; w < kw; w += blockDim.x) {
        const uint32_t v0 = 64u * w;
        uint64_t r = 0;
        if (z >= 64) {
            const uint32_t bc = v0 / z, j = v0 - bc * z;
            const uint64_t *blk = sh + bc * zw;
            const uint32_t wi = j >> 6, s = j & 63u;
            r = blk[wi] >> s;
            if (s && wi + 1 < zw) r |= blk[wi + 1] << (64u - s);
            const uint32_t avail = z - j;
            if (avail < 64u && bc + 1 < cv.nb) r |= blk[zw] << avail;
        } else {
            for (uint32_t b = 0; b < 64 && v0 + b < cv.k; ++b) {
                const uint32_t v = v0 + b, bc = v / z, j = v - bc * z;
                r |= ((sh[bc * zw] >> j) & 1ull) << b;
            }
        }
        if (v0 + 64u > cv.k) r &= (1ull << (cv.k - v0)) - 1;
        msgf[w] = r;
    }
}

}  // namespace
}  // namespace gf2l
