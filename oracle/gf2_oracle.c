/*
 * CPU oracle for the batched LDPC belief-propagation decoder.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path.  Loaded by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg; never by
 * the product.  Built with -O2 -ffp-contract=off (no fast-math) so each
 * float operation below is one IEEE-754 binary32 operation, as in the
 * reference Rust code.
 *
 * Restates (paths relative to the reference repository root):
 *   crates/gf2-coding/src/ldpc/core.rs     code model, QC expansion, decoder
 *   crates/gf2-coding/src/llr.rs           Llr, min-sum rules
 *   crates/gf2-kernels-simd/src/llr.rs     AVX2 min-sum lane structure
 *   crates/gf2-core/src/sparse.rs          CSR/CSC construction, GF(2) matvec
 *   crates/gf2-core/src/bitvec.rs          LE u64 bit layout
 */
#include "gf2_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* QC expansion: core.rs:436-447 (circulant row i -> col (i+s) mod Z)  */
/* and core.rs:578-605 (block-row-major walk, -1 skipped, range check). */
int or_qc_to_edges(uint32_t mb, uint32_t nb, uint32_t z, const int32_t *shifts,
                   uint32_t *out_rc, size_t cap, size_t *count) {
    size_t cnt = 0;
    for (uint32_t br = 0; br < mb; ++br) {
        for (uint32_t bc = 0; bc < nb; ++bc) {
            int32_t s = shifts[(size_t)br * nb + bc];
            if (s == -1) continue;
            if (s < 0 || (uint32_t)s >= z) return -1;
            for (uint32_t i = 0; i < z; ++i) {
                if (cnt >= cap) return -2;
                out_rc[2 * cnt] = br * z + i;
                out_rc[2 * cnt + 1] = bc * z + (i + (uint32_t)s) % z;
                ++cnt;
            }
        }
    }
    *count = cnt;
    return 0;
}

/* ------------------------------------------------------------------ */
static int cmp_u32(const void *a, const void *b) {
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return (x > y) - (x < y);
}

/* SpBitMatrix::from_coo (sparse.rs:166-204): bucket by row, sort each row,
 * keep a column iff it occurs an odd number of times.  Then the CSC by the
 * stable counting transpose of sparse.rs:297-331 (rows ascending within a
 * column). */
or_code *or_code_from_edges(uint32_t m, uint32_t n, const uint32_t *rc, size_t nnz_in) {
    for (size_t e = 0; e < nnz_in; ++e)
        if (rc[2 * e] >= m || rc[2 * e + 1] >= n) return NULL;

    or_code *c = (or_code *)calloc(1, sizeof(or_code));
    c->m = m;
    c->n = n;
    uint64_t *cnt = (uint64_t *)calloc((size_t)m + 1, sizeof(uint64_t));
    for (size_t e = 0; e < nnz_in; ++e) cnt[rc[2 * e] + 1]++;
    for (uint32_t r = 0; r < m; ++r) cnt[r + 1] += cnt[r];
    uint32_t *bucket = (uint32_t *)malloc((nnz_in ? nnz_in : 1) * sizeof(uint32_t));
    uint64_t *fill = (uint64_t *)malloc(((size_t)m + 1) * sizeof(uint64_t));
    memcpy(fill, cnt, ((size_t)m + 1) * sizeof(uint64_t));
    for (size_t e = 0; e < nnz_in; ++e) bucket[fill[rc[2 * e]]++] = rc[2 * e + 1];

    c->row_ptr = (uint64_t *)calloc((size_t)m + 1, sizeof(uint64_t));
    c->row_idx = (uint32_t *)malloc((nnz_in ? nnz_in : 1) * sizeof(uint32_t));
    uint64_t out = 0;
    for (uint32_t r = 0; r < m; ++r) {
        uint32_t *row = bucket + cnt[r];
        size_t len = (size_t)(cnt[r + 1] - cnt[r]);
        qsort(row, len, sizeof(uint32_t), cmp_u32);
        size_t i = 0;
        while (i < len) {
            size_t k = 1;
            while (i + k < len && row[i + k] == row[i]) ++k;
            if (k % 2 == 1) c->row_idx[out++] = row[i];
            i += k;
        }
        c->row_ptr[r + 1] = out;
    }
    c->nnz = out;
    free(bucket);
    free(fill);
    free(cnt);

    /* stable transpose */
    c->col_ptr = (uint64_t *)calloc((size_t)n + 1, sizeof(uint64_t));
    c->col_idx = (uint32_t *)malloc((c->nnz ? c->nnz : 1) * sizeof(uint32_t));
    c->csc_to_csr = (uint64_t *)malloc((c->nnz ? c->nnz : 1) * sizeof(uint64_t));
    for (uint64_t e = 0; e < c->nnz; ++e) c->col_ptr[c->row_idx[e] + 1]++;
    for (uint32_t v = 0; v < n; ++v) c->col_ptr[v + 1] += c->col_ptr[v];
    uint64_t *next = (uint64_t *)malloc(((size_t)n + 1) * sizeof(uint64_t));
    memcpy(next, c->col_ptr, ((size_t)n + 1) * sizeof(uint64_t));
    for (uint32_t r = 0; r < m; ++r)
        for (uint64_t e = c->row_ptr[r]; e < c->row_ptr[r + 1]; ++e) {
            uint64_t pos = next[c->row_idx[e]]++;
            c->col_idx[pos] = r;
            c->csc_to_csr[pos] = e;
        }
    free(next);
    return c;
}

void or_code_free(or_code *c) {
    if (!c) return;
    free(c->row_ptr);
    free(c->row_idx);
    free(c->col_ptr);
    free(c->col_idx);
    free(c->csc_to_csr);
    free(c);
}

/* ------------------------------------------------------------------ */
static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* MINPS lane semantics: dst = (a < b) ? a : b, i.e. the second operand
 * whenever either is NaN or both are zero. */
static inline float minps_lane(float a, float b) { return (a < b) ? a : b; }
/* Rust f32::min: NaN operands are ignored. */
static inline float rust_fmin(float a, float b) {
    if (isnan(a)) return b;
    if (isnan(b)) return a;
    return (a < b) ? a : b;
}

/* minsum_avx2_f32, crates/gf2-kernels-simd/src/llr.rs:117-180. */
float or_minsum_avx2(const float *x, size_t n) {
    if (n == 0) return 0.0f; /* llr.rs:123-125 */
    float vmin[8], vsign[8];
    for (int l = 0; l < 8; ++l) { vmin[l] = INFINITY; vsign[l] = 0.0f; }
    size_t chunks = n / 8;
    for (size_t i = 0; i < chunks; ++i) {
        for (int l = 0; l < 8; ++l) {
            float v = x[i * 8 + l];
            float a = u2f(f2u(v) & 0x7fffffffu);             /* andnot(sign_mask, v) */
            vmin[l] = minps_lane(vmin[l], a);                /* _mm256_min_ps(vec_min, abs) */
            vsign[l] = u2f(f2u(vsign[l]) ^ (f2u(v) & 0x80000000u)); /* xor sign bits */
        }
    }
    float min_abs = INFINITY;
    for (int l = 0; l < 8; ++l) min_abs = rust_fmin(min_abs, vmin[l]);
    float sign_product = 1.0f;
    for (int l = 0; l < 8; ++l)
        if (f2u(vsign[l]) & 0x80000000u) sign_product = -sign_product;
    for (size_t i = chunks * 8; i < n; ++i) {
        float v = x[i];
        min_abs = rust_fmin(min_abs, fabsf(v));
        if (v < 0.0f) sign_product = -sign_product;
    }
    return sign_product * min_abs;
}

/* Llr::boxplus_minsum_n_scalar, crates/gf2-coding/src/llr.rs:272-284. */
float or_minsum_scalar(const float *x, size_t n) {
    float sign_product = 1.0f;
    for (size_t i = 0; i < n; ++i) sign_product *= (x[i] >= 0.0f) ? 1.0f : -1.0f;
    float min_abs = INFINITY;
    for (size_t i = 0; i < n; ++i) min_abs = rust_fmin(min_abs, fabsf(x[i]));
    return sign_product * min_abs;
}

/* ------------------------------------------------------------------ */
/* SpBitMatrix::matvec over GF(2) (sparse.rs:367-380); bit i lives at
 * word i>>6, mask 1<<(i&63) (bitvec.rs:40-63). */
uint64_t or_syndrome(const or_code *c, const uint64_t *x, uint64_t *s) {
    size_t sw = ((size_t)c->m + 63) / 64;
    if (s) memset(s, 0, sw * sizeof(uint64_t));
    uint64_t w = 0;
    for (uint32_t r = 0; r < c->m; ++r) {
        unsigned acc = 0;
        for (uint64_t e = c->row_ptr[r]; e < c->row_ptr[r + 1]; ++e) {
            uint32_t v = c->row_idx[e];
            acc ^= (unsigned)((x[v >> 6] >> (v & 63)) & 1u);
        }
        if (acc) {
            ++w;
            if (s) s[r >> 6] |= (uint64_t)1 << (r & 63);
        }
    }
    return w;
}

/* hard_decode (core.rs:859-865) with Llr::hard_decision = value < 0
 * (llr.rs:73-75), packed as a BitVec. */
static void hard_bits(const float *belief, uint32_t n, uint64_t *words) {
    memset(words, 0, (((size_t)n + 63) / 64) * sizeof(uint64_t));
    for (uint32_t v = 0; v < n; ++v)
        if (belief[v] < 0.0f) words[v >> 6] |= (uint64_t)1 << (v & 63);
}

/* Llr::boxplus_n, crates/gf2-coding/src/llr.rs:210-215:
 * 2 * atanh(prod tanh(L_i / 2)), the product folded from 1.0 in input order
 * (Iterator::product), every step a binary32 operation (tanhf / atanhf of
 * the C library, as Rust's f32::tanh / f32::atanh call). */
float or_boxplus_n(const float *x, size_t n) {
    float prod = 1.0f;
    for (size_t i = 0; i < n; ++i) prod = prod * tanhf(x[i] / 2.0f);
    return 2.0f * atanhf(prod);
}

/* decode_iterative on a freshly constructed decoder (core.rs:653-697,
 * 889-938).  Flat arrays replace the nested Vec<Vec<Llr>>; the CSR/CSC
 * position maps replace the reference's linear neighbour scans
 * (core.rs:841-856) and return the same positions.  spa != 0 replaces the
 * min-sum check update by check_node_update_spa (core.rs:762-787). */
static void decode_core(const or_code *c, const float *llr, uint32_t max_it, const float *edge_w, int spa,
                        uint64_t *msg_words, or_result *res, float *trace_c2v, float *trace_bel,
                        int no_early_stop) {
    const uint32_t m = c->m, n = c->n;
    const uint64_t nnz = c->nnz;
    float *c2v = (float *)calloc(nnz ? nnz : 1, sizeof(float)); /* [check][pos] */
    float *v2c = (float *)malloc((nnz ? nnz : 1) * sizeof(float)); /* [var][pos]  */
    float *belief = (float *)calloc(n ? n : 1, sizeof(float));   /* new(): zeros */
    uint64_t *csr_to_csc = (uint64_t *)malloc((nnz ? nnz : 1) * sizeof(uint64_t));
    for (uint64_t e = 0; e < nnz; ++e) csr_to_csc[c->csc_to_csr[e]] = e;
    uint32_t maxdeg = 0;
    for (uint32_t r = 0; r < m; ++r) {
        uint32_t d = (uint32_t)(c->row_ptr[r + 1] - c->row_ptr[r]);
        if (d > maxdeg) maxdeg = d;
    }
    float *temp = (float *)malloc((maxdeg ? maxdeg : 1) * sizeof(float));
    size_t nw = ((size_t)n + 63) / 64;
    uint64_t *hard = (uint64_t *)malloc((nw ? nw : 1) * sizeof(uint64_t));

    /* core.rs:893-904: c2v := 0, v2c[v][*] := L[v] */
    for (uint32_t v = 0; v < n; ++v)
        for (uint64_t e = c->col_ptr[v]; e < c->col_ptr[v + 1]; ++e) v2c[e] = llr[v];

    uint32_t iterations = 0;
    int converged = 0;
    for (uint32_t it = 0; it < max_it; ++it) {
        iterations = it + 1;
        /* check_node_update_minsum, core.rs:790-814 (check_node_update_spa,
         * core.rs:762-787, when spa) */
        for (uint32_t r = 0; r < m; ++r) {
            uint64_t b = c->row_ptr[r], d = c->row_ptr[r + 1] - b;
            for (uint64_t p = 0; p < d; ++p) {
                size_t t = 0;
                for (uint64_t q = 0; q < d; ++q)
                    if (q != p) temp[t++] = v2c[csr_to_csc[b + q]];
                float msg = (t == 0) ? 0.0f : spa ? or_boxplus_n(temp, t) : or_minsum_avx2(temp, t);
                if (edge_w) msg = edge_w[(size_t)it * nnz + b + p] * msg; /* llr.rs:352-355 */
                c2v[b + p] = msg;
            }
        }
        /* variable_node_update, core.rs:819-838 */
        for (uint32_t v = 0; v < n; ++v) {
            float bel = llr[v];
            for (uint64_t e = c->col_ptr[v]; e < c->col_ptr[v + 1]; ++e)
                bel = bel + c2v[c->csc_to_csr[e]];
            belief[v] = bel;
            for (uint64_t e = c->col_ptr[v]; e < c->col_ptr[v + 1]; ++e)
                v2c[e] = bel - c2v[c->csc_to_csr[e]];
        }
        if (trace_c2v) memcpy(trace_c2v + (size_t)it * nnz, c2v, nnz * sizeof(float));
        if (trace_bel) memcpy(trace_bel + (size_t)it * n, belief, (size_t)n * sizeof(float));
        /* hard decision + syndrome, core.rs:919-923 */
        hard_bits(belief, n, hard);
        if (or_syndrome(c, hard, NULL) == 0) {
            converged = 1;
            if (!no_early_stop) break;
        }
    }
    /* core.rs:926-937 */
    hard_bits(belief, n, hard);
    int synd_ok = or_syndrome(c, hard, NULL) == 0;
    uint32_t k = (n > m) ? n - m : 0; /* LdpcCode::k, core.rs:117-119 */
    size_t kw = ((size_t)k + 63) / 64;
    for (size_t w = 0; w < kw; ++w) msg_words[w] = hard[w];
    if (k % 64) msg_words[kw - 1] &= (((uint64_t)1 << (k % 64)) - 1);
    res->iterations = iterations;
    res->converged = (uint8_t)converged;
    res->syndrome_ok = (uint8_t)synd_ok;

    free(c2v); free(v2c); free(belief); free(csr_to_csc); free(temp); free(hard);
}

void or_decode_iterative_ex(const or_code *c, const float *llr, uint32_t max_it,
                            const float *edge_w, uint64_t *msg_words, or_result *res,
                            float *trace_c2v, float *trace_bel, int no_early_stop) {
    decode_core(c, llr, max_it, edge_w, 0, msg_words, res, trace_c2v, trace_bel, no_early_stop);
}

void or_decode_iterative(const or_code *c, const float *llr, uint32_t max_it,
                         const float *edge_w, uint64_t *msg_words, or_result *res,
                         float *trace_c2v, float *trace_bel) {
    decode_core(c, llr, max_it, edge_w, 0, msg_words, res, trace_c2v, trace_bel, 0);
}

void or_decode_spa(const or_code *c, const float *llr, uint32_t max_it, uint64_t *msg_words,
                   or_result *res, float *trace_c2v, float *trace_bel) {
    decode_core(c, llr, max_it, NULL, 1, msg_words, res, trace_c2v, trace_bel, 0);
}

/* ------------------------------------------------------------------ */
/* decode_batch (core.rs:729-756): independent frames, ordered results.
 * The threads split the frame range statically, which keeps results in
 * input order as rayon's indexed collect does. */
typedef struct {
    const or_code *c; const float *llr; size_t f0, f1; uint32_t max_it;
    const float *edge_w; uint64_t *msg; or_result *res; int spa;
} batch_job;

static void *batch_worker(void *arg) {
    batch_job *j = (batch_job *)arg;
    uint32_t k = (j->c->n > j->c->m) ? j->c->n - j->c->m : 0;
    size_t kw = ((size_t)k + 63) / 64;
    for (size_t f = j->f0; f < j->f1; ++f)
        decode_core(j->c, j->llr + f * j->c->n, j->max_it, j->edge_w, j->spa, j->msg + f * kw, j->res + f,
                    NULL, NULL, 0);
    return NULL;
}

static void run_batch(const or_code *c, const float *llr, size_t frames, uint32_t max_it,
                      const float *edge_w, int spa, uint64_t *msg_words, or_result *res, int threads) {
    if (frames == 0) return;
    if (threads < 1) threads = 1;
    if ((size_t)threads > frames) threads = (int)frames;
    pthread_t *tid = (pthread_t *)malloc(sizeof(pthread_t) * threads);
    batch_job *jobs = (batch_job *)malloc(sizeof(batch_job) * threads);
    for (int t = 0; t < threads; ++t) {
        jobs[t] = (batch_job){c, llr, frames * t / threads, frames * (t + 1) / threads,
                              max_it, edge_w, msg_words, res, spa};
        if (threads == 1) batch_worker(&jobs[t]);
        else pthread_create(&tid[t], NULL, batch_worker, &jobs[t]);
    }
    if (threads > 1)
        for (int t = 0; t < threads; ++t) pthread_join(tid[t], NULL);
    free(tid);
    free(jobs);
}

void or_decode_batch(const or_code *c, const float *llr, size_t frames, uint32_t max_it,
                     const float *edge_w, uint64_t *msg_words, or_result *res, int threads) {
    run_batch(c, llr, frames, max_it, edge_w, 0, msg_words, res, threads);
}

void or_decode_batch_spa(const or_code *c, const float *llr, size_t frames, uint32_t max_it,
                         uint64_t *msg_words, or_result *res, int threads) {
    run_batch(c, llr, frames, max_it, NULL, 1, msg_words, res, threads);
}

/* ------------------------------------------------------------------ */
/* Weighted min-sum forward + reverse-mode gradient in double precision.
 * Forward: the flooding schedule of or_decode_iterative, run for exactly T
 * iterations, c2v_t(e) = w_t(e) * sgn_t(e) * min_{e' != e} |v2c_{t-1}(e')|
 * (the alpha * minsum rule of llr.rs:352-355 with one alpha per edge and
 * iteration).  Loss: mean over t and v of softplus(-y_v * belief_t[v]),
 * y = +1 for bit 0 and -1 for bit 1, which is the BCE of
 * P(bit = 1) = sigmoid(-belief) under the LLR sign of llr.rs:9-15.
 * Min-sum is differentiated as a subgradient: only the extrinsic argmin
 * (first on ties) receives gradient, signs are constant. */
static double softplus(double x) { return x > 0 ? x + log1p(exp(-x)) : log1p(exp(x)); }
static double sigm(double x) { return 1.0 / (1.0 + exp(-x)); }

double or_nms_loss_grad(const or_code *c, const float *llr, const uint8_t *bits,
                        uint32_t T, const double *edge_w, double *grad) {
    const uint32_t m = c->m, n = c->n;
    const uint64_t E = c->nnz;
    const size_t TE = (size_t)(T + 1) * E + 1;
    double *v2c = (double *)calloc(TE, sizeof(double));  /* [t][csr e], t=0: LLR */
    double *mag = (double *)calloc(TE, sizeof(double));  /* min |x| over extrinsic */
    double *sgn = (double *)calloc(TE, sizeof(double));  /* extrinsic sign product */
    int64_t *amin = (int64_t *)calloc(TE, sizeof(int64_t));
    double *bel = (double *)calloc((size_t)(T + 1) * n + 1, sizeof(double));
    double loss = 0.0;
    for (uint32_t r = 0; r < m; ++r)
        for (uint64_t e = c->row_ptr[r]; e < c->row_ptr[r + 1]; ++e)
            v2c[e] = (double)llr[c->row_idx[e]];
    for (uint32_t t = 1; t <= T; ++t) {
        const double *vin = v2c + (size_t)(t - 1) * E;
        const double *w = edge_w + (size_t)(t - 1) * E;
        double *mg = mag + (size_t)t * E, *sg = sgn + (size_t)t * E;
        int64_t *am = amin + (size_t)t * E;
        for (uint32_t r = 0; r < m; ++r) {
            uint64_t b = c->row_ptr[r], d = c->row_ptr[r + 1] - b;
            for (uint64_t p = 0; p < d; ++p) {
                double mn = INFINITY, s = 1.0;
                int64_t arg = -1;
                for (uint64_t q = 0; q < d; ++q) {
                    if (q == p) continue;
                    double x = vin[b + q];
                    if (x < 0) s = -s;
                    if (fabs(x) < mn) { mn = fabs(x); arg = (int64_t)(b + q); }
                }
                mg[b + p] = (arg < 0) ? 0.0 : mn;
                sg[b + p] = s;
                am[b + p] = arg;
            }
        }
        double *bt = bel + (size_t)t * n, *vout = v2c + (size_t)t * E;
        for (uint32_t v = 0; v < n; ++v) {
            double s = (double)llr[v];
            for (uint64_t ec = c->col_ptr[v]; ec < c->col_ptr[v + 1]; ++ec) {
                uint64_t e = c->csc_to_csr[ec];
                s += w[e] * sg[e] * mg[e];
            }
            bt[v] = s;
            loss += softplus(-(bits[v] ? -1.0 : 1.0) * s);
            for (uint64_t ec = c->col_ptr[v]; ec < c->col_ptr[v + 1]; ++ec) {
                uint64_t e = c->csc_to_csr[ec];
                vout[e] = s - w[e] * sg[e] * mg[e];
            }
        }
    }
    const double scale = 1.0 / ((double)T * (double)n);
    loss *= scale;
    if (grad) {
        memset(grad, 0, (size_t)T * E * sizeof(double));
        double *g_v2c = (double *)calloc(E + 1, sizeof(double)); /* dL/dv2c_t   */
        double *g_prev = (double *)calloc(E + 1, sizeof(double)); /* dL/dv2c_t-1 */
        double *g_bel = (double *)calloc((size_t)n + 1, sizeof(double));
        for (uint32_t t = T; t >= 1; --t) {
            const double *bt = bel + (size_t)t * n;
            const double *mg = mag + (size_t)t * E, *sg = sgn + (size_t)t * E;
            const int64_t *am = amin + (size_t)t * E;
            const double *vin = v2c + (size_t)(t - 1) * E;
            const double *w = edge_w + (size_t)(t - 1) * E;
            for (uint32_t v = 0; v < n; ++v) {
                double y = bits[v] ? -1.0 : 1.0;
                double g = -y * sigm(-y * bt[v]) * scale; /* d softplus(-y s) / ds */
                for (uint64_t ec = c->col_ptr[v]; ec < c->col_ptr[v + 1]; ++ec)
                    g += g_v2c[c->csc_to_csr[ec]];
                g_bel[v] = g;
            }
            memset(g_prev, 0, E * sizeof(double));
            for (uint32_t v = 0; v < n; ++v)
                for (uint64_t ec = c->col_ptr[v]; ec < c->col_ptr[v + 1]; ++ec) {
                    uint64_t e = c->csc_to_csr[ec];
                    /* belief_t[v] = L + sum c2v ; v2c_t(e) = belief_t[v] - c2v_t(e) */
                    double g_c2v = g_bel[v] - g_v2c[e];
                    grad[(size_t)(t - 1) * E + e] += g_c2v * sg[e] * mg[e];
                    int64_t a = am[e];
                    if (a >= 0 && t >= 2) /* v2c_0 is the channel LLR: no parameters */
                        g_prev[a] += g_c2v * w[e] * sg[e] * ((vin[a] < 0) ? -1.0 : 1.0);
                }
            double *tmp = g_v2c; g_v2c = g_prev; g_prev = tmp;
        }
        free(g_v2c); free(g_prev); free(g_bel);
    }
    free(v2c); free(mag); free(sgn); free(amin); free(bel);
    return loss;
}

/* ------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11): 10 rounds of the 4x32 S-box with multipliers
 * 0xD2511F53, 0xCD9E8D57 and Weyl key increments 0x9E3779B9, 0xBB67AE85. */
void or_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* ------------------------------------------------------------------ */
/* Frame generator restatement (one frame, bit by bit), the checker of the
 * device generator: Philox4x32-10 information bits, staircase QC-IRA
 * encoding, BPSK/AWGN with LLR = (float)(2 y / sigma^2) in double
 * (crates/gf2-coding/src/channel.rs:51-57, 100-102, 153-158). */
void or_framegen(const int32_t *shifts, uint32_t mb, uint32_t nb, uint32_t z, uint64_t seed,
                 uint64_t g, float ebn0_lo, float ebn0_hi, float *llr, uint64_t *cw_words) {
    const uint32_t kb = nb - mb, zw = (z + 63) / 64, n = nb * z;
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint8_t *bit = (uint8_t *)calloc(n, 1);
    uint32_t o[4];
    for (uint32_t bc = 0; bc < kb; ++bc)
        for (uint32_t j = 0; j < z; ++j) {
            const uint32_t ctr[4] = {bc * zw + j / 64, (uint32_t)g, (uint32_t)(g >> 32), 0};
            or_philox4x32(ctr, key, o);
            const uint64_t w = (uint64_t)o[0] | ((uint64_t)o[1] << 32);
            bit[bc * z + j] = (uint8_t)((w >> (j % 64)) & 1u);
        }
    for (uint32_t br = 0; br < mb; ++br)
        for (uint32_t i = 0; i < z; ++i) {
            uint8_t sacc = 0;
            for (uint32_t bc = 0; bc < kb; ++bc) {
                const int32_t sh = shifts[(size_t)br * nb + bc];
                if (sh >= 0) sacc ^= bit[bc * z + (i + (uint32_t)sh) % z];
            }
            const uint8_t prev = br ? bit[(kb + br - 1) * z + i] : 0;
            bit[(kb + br) * z + i] = sacc ^ prev;
        }
    const uint32_t c2[4] = {0, (uint32_t)g, (uint32_t)(g >> 32), 2};
    or_philox4x32(c2, key, o);
    const double rate = (double)(kb * z) / (double)n;
    const double u = ((double)o[0] + 0.5) * (1.0 / 4294967296.0);
    const double ebn0 = (double)ebn0_lo + ((double)ebn0_hi - (double)ebn0_lo) * u;
    const double s2 = 1.0 / (2.0 * rate * pow(10.0, ebn0 / 10.0));
    const double sigma = sqrt(s2);
    for (uint32_t v = 0; v < n; ++v) {
        const uint32_t c1[4] = {v / 2, (uint32_t)g, (uint32_t)(g >> 32), 1};
        or_philox4x32(c1, key, o);
        const double u1 = ((double)o[0] + 0.5) * (1.0 / 4294967296.0);
        const double u2 = ((double)o[1] + 0.5) * (1.0 / 4294967296.0);
        const double r = sqrt(-2.0 * log(u1));
        const double th = 6.283185307179586476925286766559 * u2;
        const double nz = (v % 2 == 0) ? r * cos(th) : r * sin(th);
        const double x = bit[v] ? -1.0 : 1.0;
        llr[v] = (float)(2.0 * (x + sigma * nz) / s2);
    }
    if (cw_words) {
        memset(cw_words, 0, ((n + 63) / 64) * sizeof(uint64_t));
        for (uint32_t v = 0; v < n; ++v)
            if (bit[v]) cw_words[v >> 6] |= (uint64_t)1 << (v & 63);
    }
    free(bit);
}

/* All-zero codeword frame g of an n-bit, rate k/n code: the Eb/N0 and noise
 * streams of or_framegen with every bit 0 (x = +1), as the reference FER
 * example transmits (crates/gf2-coding/examples/ldpc_awgn.rs:121-175). */
void or_framegen_zero(uint32_t n, uint32_t k, uint64_t seed, uint64_t g, float ebn0_lo, float ebn0_hi,
                      float *llr) {
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    const uint32_t c2[4] = {0, (uint32_t)g, (uint32_t)(g >> 32), 2};
    or_philox4x32(c2, key, o);
    const double rate = (double)k / (double)n;
    const double u = ((double)o[0] + 0.5) * (1.0 / 4294967296.0);
    const double ebn0 = (double)ebn0_lo + ((double)ebn0_hi - (double)ebn0_lo) * u;
    const double s2 = 1.0 / (2.0 * rate * pow(10.0, ebn0 / 10.0));
    const double sigma = sqrt(s2);
    for (uint32_t v = 0; v < n; ++v) {
        const uint32_t c1[4] = {v / 2, (uint32_t)g, (uint32_t)(g >> 32), 1};
        or_philox4x32(c1, key, o);
        const double u1 = ((double)o[0] + 0.5) * (1.0 / 4294967296.0);
        const double u2 = ((double)o[1] + 0.5) * (1.0 / 4294967296.0);
        const double r = sqrt(-2.0 * log(u1));
        const double th = 6.283185307179586476925286766559 * u2;
        const double nz = (v % 2 == 0) ? r * cos(th) : r * sin(th);
        llr[v] = (float)(2.0 * (1.0 + sigma * nz) / s2);
    }
}
