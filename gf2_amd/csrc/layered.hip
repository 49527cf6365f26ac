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
#include <vector>

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

// ---------------------------------------------------------------------------
// Staircase kernel (decoding only).  For a QC-IRA code whose parity part is a
// block staircase of unshifted circulants (block-row r has the parity
// block-columns kb+r-1 and kb+r as its last two edges, shift 0), check i of
// every block-row touches parity bit i of its parity columns: those beliefs
// are private to thread i (the workgroup has exactly z threads).  The kb
// information columns live in LDS; the parity beliefs and the check states
// stay in the frame's global rows, each touched only by its owner thread.
// So
//   - the layers are separated by an LDS-only barrier (no wait for the
//     thread's outstanding global loads and stores);
//   - the state and the belief of column kb+r that layer r reads are loaded
//     P layers earlier, P a divisor of mb (the same thread stored them at
//     least two layers before the load is issued, in program order);
//   - column kb+r, updated by layer r, reaches layer r+1 in a register;
//   - the hard decisions of a parity column are taken (ballot) when its
//     belief is final for the iteration, those of the information columns
//     from LDS after the last layer.
// The arithmetic and its order are those of the general kernel with X = kb.
// The vector-memory counter retires loads and stores in issue order, and
// the compiler's wait for a loaded register is exact only where every path
// to it issued the same memory operations: so the loop body has no early
// exit (iterations end at group boundaries, P dividing mb), and every global
// store is unconditional (block-row 0 stores its position D-2 value into the
// unused global row of block-column kb-1, every block-row its column kb+r
// as it stands; the information rows are rewritten at hand-over).

__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// LDS pointers typed as such, so every access is a ds_ instruction (a
// generic pointer held in a struct compiles to flat accesses, which count
// against the vector-memory counter and would wait for the loads ahead).
typedef __attribute__((address_space(3))) float lds_float;
typedef __attribute__((address_space(3))) uint64_t lds_u64;
typedef __attribute__((address_space(3))) uint32_t lds_u32;

struct StairFrame {
    lds_u64 *hb;       // [nb][zw] hard decisions
    const lds_u32 *tab; // [nbe][4] per edge: LDS byte address of its column's
                        // beliefs (information columns), shift, w_t, w_{t-1}
    float *gb;         // the frame's global belief rows [nb][zp]
    uint3 *st;         // the frame's check states [mb][zp]
    const float *llr;  // the frame's channel LLRs [n]
    uint32_t kb, i;    // i = threadIdx.x < z
};

__device__ __forceinline__ float lds_ld(uint32_t addr) { return *(const lds_float *)(size_t)addr; }
__device__ __forceinline__ void lds_st(uint32_t addr, float v) { *(lds_float *)(size_t)addr = v; }
// x * (-1)^s for s in {0, 1}: a sign flip, exact
__device__ __forceinline__ float flip(float x, uint32_t s) { return __uint_as_float(__float_as_uint(x) ^ (s << 31)); }

// Layer br (edges e0..e0+D-1) of iteration t (cin: br > 0; last: br is the
// last block-row; both the caller's constants where it knows them).  (o, bo): this layer's check states and belief
// of column kb+br, loaded ahead.  Code tables and weights come from LDS: no
// scalar memory loads in the layer loop (their counter would also wait for
// the LDS accesses).  In iteration 1 the old message is rebuilt from a zero
// state (read from zero rows, not from the frame's stale ones) with weight 0
// (the table's previous weights are zero then): x = b - (+0), which is b
// exactly (also for -0 and infinities).
//
// The check rule is the general kernel's (CheckAcc) specialised to D <= 8:
// with fewer than nine inputs every input of the reference's min-sum falls in
// the remainder after the full 8-lane chunks, whose sign rule is x < 0, so the
// output sign of position p is the parity of (x < 0) over the other inputs;
// w * (+-m) is computed as +-(w * m) (IEEE multiplication: the same value).
template <int D>
__device__ __forceinline__ void stair_layer(const LayeredArgs &a, const StairFrame &S, uint32_t br, uint32_t e0,
                                            bool cin, bool last, const uint3 o, const float bo, float &carry) {
    static_assert(D >= 2 && D <= 8, "the staircase layer's sign rule needs 2 <= D <= 8");
    const CodeView &cv = a.cv;
    uint32_t cb[D], sh[D];
    float wt[D], wp[D];
#pragma unroll
    for (int p = 0; p < D; ++p) {
        const lds_u32 *r = S.tab + 4 * (e0 + p);
        cb[p] = r[0];
        sh[p] = r[1];
        wt[p] = __uint_as_float(r[2]);
        wp[p] = __uint_as_float(r[3]);
    }
    const float m1o = __uint_as_float(o.x), m2o = __uint_as_float(o.y);
    const uint32_t mo = o.z, io = mo & 31u;
    float x[D];
    uint32_t ad[D];
#pragma unroll
    for (int p = 0; p < D; ++p) {
        float bv;
        ad[p] = cb[p] + 4u * wrap_add(S.i, sh[p], cv.z);
        if (p == D - 1) {
            bv = bo;
        } else if (p == D - 2) {
            // read either way, no branch among the reads (the table gives
            // parity edges the address of column 0: in bounds, unused)
            const float lv = lds_ld(ad[p]);
            bv = cin ? carry : lv;
        } else {
            bv = lds_ld(ad[p]);
        }
        const float mag = (uint32_t)p == io ? m2o : m1o;
        x[p] = bv - flip(wp[p] * mag, (mo >> (5 + p)) & 1u);
    }
    float m1 = __builtin_inff(), m2 = __builtin_inff();
    uint32_t idx = 0, lt = 0;
#pragma unroll
    for (int p = 0; p < D; ++p) {
        // The strict-comparison update (a < m1: m2 = m1, m1 = a; else a < m2:
        // m2 = a) as min / median: with m1 <= m2 the new m2 is the median of
        // (m1, a, m2) and the new m1 is min(m1, a), the same values for any
        // non-NaN a (|x| carries no -0); idx keeps the strict comparison, so
        // the first of equal magnitudes keeps it.
        const float av = fabsf(x[p]);
        idx = av < m1 ? (uint32_t)p : idx;
        m2 = __builtin_amdgcn_fmed3f(m1, av, m2);
        m1 = __builtin_fminf(m1, av);
        lt |= (x[p] < 0.0f ? 1u : 0u) << p;
    }
    const uint32_t sg = ((__popc(lt) & 1u) ? (1u << D) - 1u : 0u) ^ lt;  // output sign bits
    const uint32_t meta = idx | (sg << 5);
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
#pragma unroll
    for (int p = 0; p < D; ++p) {
        const float mag = (uint32_t)p == idx ? m2 : m1;
        const float v = x[p] + flip(wt[p] * mag, (sg >> p) & 1u);
        if (p == D - 1) {
            carry = v;  // column kb+br: to layer br+1, final now for the last block-row
            stu(S.gb + (S.kb + br) * cv.zp, S.i, v);
            if (last) {
                const uint64_t m = __ballot(v < 0.0f);
                if (lane == 0) S.hb[(cv.nb - 1) * cv.zw + wave] = m;
            }
        } else if (p == D - 2) {
            // column kb+br-1, final for this iteration (block-row 0: an
            // information column; the global row of column kb-1 takes the
            // store, unused until hand-over rewrites it)
            stu(S.gb + (S.kb + br - 1) * cv.zp, S.i, v);
            if (cin) {
                const uint64_t m = __ballot(v < 0.0f);
                if (lane == 0) S.hb[(S.kb + br - 1) * cv.zw + wave] = m;
            } else {
                lds_st(ad[p], v);
            }
        } else {
            lds_st(ad[p], v);
        }
    }
    stu(S.st + br * cv.zp, S.i, make_uint3(__float_as_uint(m1), __float_as_uint(m2), meta));
}

// Parity checks over the hard-decision words, the code's edge tables read
// from LDS copies (rp, re); uniform over the workgroup.
__device__ __forceinline__ uint64_t cyc64_lds(const lds_u64 *v, uint32_t z, uint32_t zw, uint32_t pos) {
    const uint32_t wi = pos >> 6, sh = pos & 63u;
    uint64_t r = v[wi] >> sh;
    if (sh && wi + 1 < zw) r |= v[wi + 1] << (64u - sh);
    const uint32_t avail = z - pos;
    if (avail < 64u) r |= v[0] << avail;
    return r;
}

// The staircase kernel runs z threads (zw waves): the block size comes from
// the code view (a scalar), not from the dispatch packet (a vector load,
// whose wait would also wait for every load ahead).
__device__ __forceinline__ uint32_t stair_unsat(const CodeView &cv, const lds_u64 *sh, const lds_u32 *rp,
                                                const lds_u64 *re, lds_u32 *s_bad) {
    const uint32_t z = cv.z, zw = cv.zw, nthr = z;
    uint64_t bad = 0;
    if (z >= 64) {
        for (uint32_t u = threadIdx.x; u < cv.mb * zw; u += nthr) {
            // once a lane of the wave has found an unsatisfied check the
            // frame is unsatisfied: the wave's other words are not needed
            // (the ballot is uniform over the wave, so is the exit)
            if (__ballot(bad != 0)) break;
            const uint32_t br = u / zw, w = u - br * zw;
            uint64_t acc = 0;
            for (uint32_t e = rp[br]; e < rp[br + 1]; ++e) {
                const uint64_t r = re[e];  // block-column | shift << 32
                acc ^= cyc64_lds(sh + (uint32_t)r * zw, z, zw, wrap_add(64u * w, (uint32_t)(r >> 32), z));
            }
            if (w == zw - 1 && (z & 63u)) acc &= (1ull << (z & 63u)) - 1;
            bad |= acc;
        }
    } else {
        for (uint32_t c = threadIdx.x; c < cv.m; c += nthr) {
            if (__ballot(bad != 0)) break;  // as above
            const uint32_t br = c / z, i = c - br * z;
            uint32_t par = 0;
            for (uint32_t e = rp[br]; e < rp[br + 1]; ++e) {
                const uint64_t r = re[e];
                par ^= (uint32_t)(sh[(uint32_t)r * zw] >> wrap_add(i, (uint32_t)(r >> 32), z)) & 1u;
            }
            bad |= par;
        }
    }
    const uint64_t any = __ballot(bad != 0);
    if ((threadIdx.x & 63u) == 0) s_bad[threadIdx.x >> 6] = any != 0;
    lds_barrier();
    uint32_t r = 0;
    for (uint32_t w = 0; w < zw; ++w) r |= s_bad[w];
    return uni(r);  // the same in every lane: keeps the iteration state scalar
}

// Bits b .. b+31 of a 256-bit mask held in four words (b < 256; bits past
// 255 read as zero): one extraction per group of layers, so a layer tests a
// constant bit.
__device__ __forceinline__ uint32_t mask_slice(uint64_t m0, uint64_t m1, uint64_t m2, uint64_t m3, uint32_t b) {
    const uint32_t w = b >> 6, o = b & 63u;
    const uint64_t lo = w == 0 ? m0 : w == 1 ? m1 : w == 2 ? m2 : m3;
    const uint64_t hi = w == 0 ? m1 : w == 1 ? m2 : w == 2 ? m3 : 0ull;
    return (uint32_t)((lo >> o) | (o ? hi << (64u - o) : 0ull));
}

template <int DH, int P>
__global__ void __launch_bounds__(512) lnms_stair_kernel(LayeredArgs a) {
    static_assert(P >= 1 && P <= 32, "a group's mask bits fit one 32-bit slice");
    extern __shared__ __align__(16) uint64_t smem[];
    __shared__ uint32_t s_f;
    __shared__ uint32_t s_bad_arr[8];
    const CodeView &cv = a.cv;
    const uint32_t kb = cv.nb - cv.mb, kw = (cv.k + 63) / 64;
    lds_u32 *s_tab = (lds_u32 *)smem;  // [nbe][4], 16-byte records
    lds_u64 *s_hb = (lds_u64 *)(s_tab + 4 * cv.nbe);
    lds_u64 *s_re = s_hb + cv.nb * cv.zw;  // row-order edges: block-column | shift << 32
    lds_float *s_bel = (lds_float *)(s_re + cv.nbe);
    lds_u32 *s_rp = (lds_u32 *)(s_bel + kb * cv.z);
    lds_u32 *s_bad = (lds_u32 *)s_bad_arr;
    const uint32_t i = threadIdx.x;  // < z: the block has z threads
    const uint32_t bel0 = (uint32_t)(size_t)s_bel;  // LDS byte address of the beliefs
    for (uint32_t e = i; e < cv.nbe; e += cv.z) {
        const uint2 r = cv.row_edge[e];
        s_re[e] = (uint64_t)r.x | ((uint64_t)r.y << 32);
        s_tab[4 * e] = r.x < kb ? bel0 + 4u * r.x * cv.z : bel0;
        s_tab[4 * e + 1] = r.x < kb ? r.y : 0u;
    }
    for (uint32_t r = i; r <= cv.mb; r += cv.z) s_rp[r] = cv.row_ptr[r];
    StairFrame S;
    S.hb = s_hb;
    S.tab = s_tab;
    S.kb = kb;
    S.i = i;
    // row-degree and barrier masks in scalar registers (an indexed
    // kernel-argument read would be a scalar memory load inside the layer loop)
    const uint64_t dm0 = a.stair_deg[0], dm1 = a.stair_deg[1], dm2 = a.stair_deg[2], dm3 = a.stair_deg[3];
    const uint64_t bm0 = a.stair_sync[0], bm1 = a.stair_sync[1], bm2 = a.stair_sync[2], bm3 = a.stair_sync[3];
    for (;;) {
        __syncthreads();  // the previous frame is done with s_f and the LDS
        if (threadIdx.x == 0) s_f = atomicAdd(a.ctr, 1u);
        __syncthreads();
        const uint32_t f = uni(s_f);
        if (f >= a.F) break;
        S.gb = a.belief + (size_t)f * cv.nb * cv.zp;
        S.st = reinterpret_cast<uint3 *>(a.state) + (size_t)f * cv.mb * cv.zp;
        S.llr = a.llr + (size_t)f * cv.n;
        uint3 sst[P];
        float sbo[P];
#pragma unroll
        for (int u = 0; u < P; ++u) {  // layers 0..P-1 of iteration 1 read the LLRs
            sbo[u] = S.llr[(kb + u) * cv.z + i];
            sst[u] = make_uint3(0, 0, 0);
        }
        for (uint32_t c = 0; c < kb; ++c)  // information beliefs start at the channel LLRs
            s_bel[c * cv.z + i] = S.llr[c * cv.z + i];
        for (uint32_t e = i; e < cv.nbe; e += cv.z) {  // iteration 1's weights
            s_tab[4 * e + 2] = __float_as_uint(a.w[e]);
            s_tab[4 * e + 3] = 0u;
        }
        // wait for every load above: on entry to the layer loop nothing is
        // outstanding, so the compiler's waits for the slots inside it count
        // the loop's own memory operations only (vmcnt(0), other counters
        // unrestricted)
        __builtin_amdgcn_s_waitcnt(0x0F70);
        __syncthreads();
        float carry = 0.0f;
        int conv = 0;
        uint32_t t = 1, br = 0, e0 = 0;  // br: first block-row of the group, e0: its first edge
        for (;;) {
            // the group's slots come back P layers on: the same block-rows of
            // the next group, or of the next iteration after the last group
            const bool wrap = br + P == cv.mb;
            const uint32_t tp = t + (wrap ? 1u : 0u), bq = wrap ? 0u : br + P;
            const float *src = tp == 1 ? S.llr + (S.kb + bq) * cv.z : S.gb + (S.kb + bq) * cv.zp;
            const uint32_t src_row = tp == 1 ? cv.z : cv.zp;
            // iteration 1 reads zero states (the frame's rows hold another
            // frame's or nothing yet)
            const uint3 *st_src = tp == 1 ? a.zero_state : S.st;
            // the group's degree and barrier bits (bit u: block-row br + u)
            const uint32_t dsl = mask_slice(dm0, dm1, dm2, dm3, br), ssl = mask_slice(bm0, bm1, bm2, bm3, br);
#pragma unroll
            for (int u = 0; u < P; ++u) {
                const uint32_t r = br + u;
                const bool hi = (dsl >> u) & 1u;
                const bool cin = u > 0 || br > 0;      // a constant in all but the first copy
                const bool last = u == P - 1 && wrap;  // false in all but the last copy
                if (hi)
                    stair_layer<DH>(a, S, r, e0, cin, last, sst[u], sbo[u], carry);
                else
                    stair_layer<DH - 1>(a, S, r, e0, cin, last, sst[u], sbo[u], carry);
                e0 += hi ? DH : DH - 1;
                // loaded whether or not it will be used (past iteration T): in
                // bounds, one definition per slot; wave-uniform row bases, so
                // the loads take the scalar-base form
                sbo[u] = ldu(src + u * src_row, i);
                sst[u] = ldu(st_src + (bq + u) * cv.zp, i);
                // block-rows without a common information column need no
                // barrier between them (parity beliefs and states are private)
                if ((ssl >> u) & 1u) lds_barrier();
            }
            br += P;
            if (br == cv.mb) {  // end of iteration t
                // hard decisions of the information columns: the z threads
                // are the zw waves, so wave w packs word w of every column
                // from its own lanes (the last wave's missing lanes give the
                // zero bits above z)
                const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
                for (uint32_t c = 0; c < kb; ++c) {
                    const uint64_t m = __ballot(s_bel[c * cv.z + i] < 0.0f);
                    if (lane == 0) s_hb[c * cv.zw + wave] = m;
                }
                lds_barrier();
                if (!stair_unsat(cv, s_hb, s_rp, s_re, s_bad)) {
                    conv = 1;
                    break;
                }
                if (t == a.T) break;
                ++t;
                br = 0;
                e0 = 0;
                // weights: those of t-1 move to the previous slot, iteration
                // t's come in (each thread its own records; the last reads
                // were before the barriers above)
                for (uint32_t e = i; e < cv.nbe; e += cv.z) {
                    s_tab[4 * e + 3] = s_tab[4 * e + 2];
                    s_tab[4 * e + 2] = __float_as_uint(a.w[(size_t)(t - 1) * cv.nbe + e]);
                }
                lds_barrier();
            }
        }
        const uint32_t its = conv ? t : a.T;
        if (threadIdx.x == 0) atomicAdd(a.ctr + 1, its);
        if (conv || !a.fallback) {
            frame_message(cv, (const uint64_t *)s_hb, a.msg + (size_t)f * kw);
            if (threadIdx.x == 0) {
                a.iters[f] = its;
                a.conv[f] = (uint8_t)conv;
                a.synd[f] = (uint8_t)conv;
                if (a.fb) a.fb[f] = 0;
            }
        } else {
            // hand the frame to the plain continuation: the information
            // beliefs to the frame's global rows (parity beliefs and check
            // states are there), the frame on the running list
            for (uint32_t c = 0; c < kb; ++c)
                S.gb[c * cv.zp + i] = s_bel[c * cv.z + i];
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
    __shared__ float lred[16];
    __shared__ uint32_t s_f;
    const CodeView &cv = a.cv;
    const uint32_t nw = (cv.n + 63) / 64;
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
        float lsum = 0.0f;
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
        // the frame's loss, reduced in a fixed order whichever workgroup took it
        lsum = wave_sum(lsum);
        if ((threadIdx.x & 63u) == 0) lred[threadIdx.x >> 6] = lsum;
        __syncthreads();
        if (threadIdx.x == 0) {
            float s = 0.0f;
            for (uint32_t w = 0; w < blockDim.x / 64; ++w) s += lred[w];
            a.lpart[f] = s;
        }
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

size_t layered_stair_bytes(const CodeView &cv) {
    return (size_t)cv.nbe * 16 + (size_t)cv.nb * cv.zw * sizeof(uint64_t) + (size_t)cv.nbe * sizeof(uint2) +
           (size_t)(cv.nb - cv.mb) * cv.z * sizeof(float) + (size_t)(cv.mb + 1) * sizeof(uint32_t);
}

// Load-ahead distance of the staircase kernel for mb block-rows: the largest
// instantiated divisor of mb with two groups or more; 0 if none.
uint32_t stair_ahead(uint32_t mb) {
    for (uint32_t p : {10u, 6u})
        if (mb % p == 0 && mb >= 2 * p) return p;
    return 0;
}

// Highest row degree of the staircase kernel's instantiations; 0 when the
// code does not have the structure it needs (block staircase of unshifted
// circulants as the last edges of each block-row, information edges before
// them, row degrees dh-1 and dh only, z <= 512, mb <= 256 a multiple of a
// load-ahead distance with two groups or more).
uint32_t layered_stair_degree(const gf2l_code *code, uint64_t *deg_mask, uint64_t *sync_mask) {
    if (code->is_csr) return 0;
    const CodeView &cv = code->view;
    if (cv.wrap_bc != 0xffffffffu || cv.z > 512 || cv.nb <= cv.mb || cv.mb > 256 || !stair_ahead(cv.mb)) return 0;
    const uint32_t kb = cv.nb - cv.mb;
    uint32_t dmin = ~0u, dmax = 0;
    for (uint32_t br = 0; br < cv.mb; ++br) {
        const uint32_t e0 = code->h_row_ptr[br], d = code->h_row_ptr[br + 1] - e0;
        if (d < 2) return 0;
        const uint2 last = code->h_row_edge[e0 + d - 1];
        if (last.x != kb + br || last.y != 0) return 0;
        const uint32_t nin = br > 0 ? 2 : 1;
        if (br > 0) {
            const uint2 prev = code->h_row_edge[e0 + d - 2];
            if (prev.x != kb + br - 1 || prev.y != 0) return 0;
        }
        for (uint32_t p = 0; p + nin < d; ++p)
            if (code->h_row_edge[e0 + p].x >= kb) return 0;
        dmin = std::min(dmin, d);
        dmax = std::max(dmax, d);
    }
    if (dmin + 1 < dmax || (dmax != 6 && dmax != 7)) return 0;
    if (deg_mask) {
        for (int q = 0; q < 4; ++q) deg_mask[q] = 0;
        for (uint32_t br = 0; br < cv.mb; ++br)
            if (code->h_row_ptr[br + 1] - code->h_row_ptr[br] == dmax) deg_mask[br >> 6] |= 1ull << (br & 63u);
    }
    if (sync_mask) {
        // a barrier after block-row r when block-row r+1 touches an
        // information column (in LDS, read and written by other threads) that
        // a block-row since the last barrier touched; always after the last
        // block-row (the hard decisions read them all)
        for (int q = 0; q < 4; ++q) sync_mask[q] = 0;
        std::vector<uint8_t> since(kb, 0);  // columns touched since the last barrier
        auto mark = [&](uint32_t br) {
            for (uint32_t e = code->h_row_ptr[br]; e < code->h_row_ptr[br + 1]; ++e)
                if (code->h_row_edge[e].x < kb) since[code->h_row_edge[e].x] = 1;
        };
        mark(0);
        for (uint32_t br = 0; br < cv.mb; ++br) {
            bool need = br + 1 == cv.mb;
            if (!need)
                for (uint32_t e = code->h_row_ptr[br + 1]; e < code->h_row_ptr[br + 2]; ++e)
                    if (code->h_row_edge[e].x < kb && since[code->h_row_edge[e].x]) need = true;
            if (need) {
                sync_mask[br >> 6] |= 1ull << (br & 63u);
                std::fill(since.begin(), since.end(), 0);
            }
            if (br + 1 < cv.mb) mark(br + 1);
        }
    }
    return dmax;
}

hipError_t launch_layered_stair(const LayeredArgs &a, uint32_t dh, int num_cus, hipStream_t s) {
    const size_t bytes = layered_stair_bytes(a.cv);
    const uint32_t block = a.cv.z;  // one thread per check of a block-row
    const uint32_t per_cu = (uint32_t)std::max<size_t>(1, std::min<size_t>(8, (160 * 1024) / std::max<size_t>(bytes, 1)));
    const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(a.F, (uint32_t)num_cus * per_cu));
    const uint32_t p = stair_ahead(a.cv.mb);
    if (a.cv.z == 0 || a.cv.z > 512 || !p) return hipErrorInvalidValue;
#define GF2L_LSTAIR(DH, P)                                                              \
    do {                                                                                \
        auto k = lnms_stair_kernel<DH, P>;                                              \
        hipError_t e = allow_lds(k, bytes);                                             \
        if (e != hipSuccess) return e;                                                  \
        hipLaunchKernelGGL(k, dim3(grid), dim3(block), bytes, s, a);                    \
    } while (0)
    if (dh == 6 && p == 6)
        GF2L_LSTAIR(6, 6);
    else if (dh == 6 && p == 10)
        GF2L_LSTAIR(6, 10);
    else if (dh == 7 && p == 6)
        GF2L_LSTAIR(7, 6);
    else if (dh == 7 && p == 10)
        GF2L_LSTAIR(7, 10);
    else
        return hipErrorInvalidValue;
#undef GF2L_LSTAIR
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
