/*
 * CPU oracle of the layered learned decoder (TEST INFRASTRUCTURE ONLY: the
 * checker of gf2l_lnms_*; loaded by tests/ and never by the product).
 *
 * The reference has no learned decoder; ROADMAP.md:93 plans one.  Its only
 * anchor is the alpha * min-sum rule, Llr::boxplus_normalized_minsum_n
 * (crates/gf2-coding/src/llr.rs:352-355), here with one alpha per base-graph
 * edge and iteration.  The min-sum itself is the reference's
 * (or_minsum_avx2, crates/gf2-kernels-simd/src/llr.rs:117-180), and the
 * plain continuation is the reference's flooding decoder
 * (crates/gf2-coding/src/ldpc/core.rs:790-938) started from the learned
 * decoder's state instead of from zero messages.
 *
 * Layered (row-sequential) schedule over a quasi-cyclic code of mb x nb
 * circulant blocks of lift z, block-rows in ascending order; each block-row
 * is one layer whose z checks touch disjoint variables:
 *   x_p     = b[v_p] - w_{t-1}(e_p) * m_{t-1}(e_p, i)   (no subtraction at t = 1)
 *   m_t(p)  = minsum over the other inputs x_q of the check, in row order
 *   b[v_p]  = x_p + w_t(e_p) * m_t(p)
 * beliefs start at the channel LLRs; after each iteration the hard decision
 * b < 0 is checked against every parity check and a frame that satisfies
 * them stops.  A frame still failing after T iterations continues with
 * plain flooding min-sum for up to fb_it iterations from that state: the
 * first check update subtracts w_T * m_T from the layered beliefs, later
 * ones the plain messages; beliefs are L + the messages in ascending check
 * order.  Every float operation is one IEEE binary32 operation (built with
 * -ffp-contract=off).
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "gf2_oracle.h"

typedef struct {
    uint32_t mb, nb, z, nbe;
    uint32_t *row_ptr; /* mb+1 block edges per block-row, block-columns ascending */
    uint32_t *bc;      /* nbe */
    uint32_t *sh;      /* nbe */
    uint32_t max_dc;
} qc_view;

static void qc_build(qc_view *q, const int32_t *shifts, uint32_t mb, uint32_t nb, uint32_t z) {
    q->mb = mb;
    q->nb = nb;
    q->z = z;
    q->nbe = 0;
    for (uint32_t i = 0; i < mb * nb; ++i) q->nbe += shifts[i] >= 0;
    q->row_ptr = (uint32_t *)calloc(mb + 1, sizeof(uint32_t));
    q->bc = (uint32_t *)malloc((q->nbe + 1) * sizeof(uint32_t));
    q->sh = (uint32_t *)malloc((q->nbe + 1) * sizeof(uint32_t));
    q->max_dc = 0;
    uint32_t e = 0;
    for (uint32_t br = 0; br < mb; ++br) {
        for (uint32_t bc = 0; bc < nb; ++bc)
            if (shifts[br * nb + bc] >= 0) {
                q->bc[e] = bc;
                q->sh[e] = (uint32_t)shifts[br * nb + bc];
                ++e;
            }
        q->row_ptr[br + 1] = e;
        if (e - q->row_ptr[br] > q->max_dc) q->max_dc = e - q->row_ptr[br];
    }
}

static void qc_free(qc_view *q) {
    free(q->row_ptr);
    free(q->bc);
    free(q->sh);
}

/* variable of block edge e at lane i (CirculantMatrix::to_edges, core.rs:436-447) */
static inline uint32_t var_of(const qc_view *q, uint32_t e, uint32_t i) {
    return q->bc[e] * q->z + (i + q->sh[e]) % q->z;
}

/* extrinsic min-sum of position p among d inputs (core.rs:790-814: the other
 * d-1 inputs in row order; a degree-1 check sends +0.0, core.rs:804-806) */
static float ext_minsum(const float *x, uint32_t d, uint32_t p, float *tmp) {
    uint32_t t = 0;
    for (uint32_t q = 0; q < d; ++q)
        if (q != p) tmp[t++] = x[q];
    return t ? or_minsum_avx2(tmp, t) : 0.0f;
}

/* zero syndrome of the hard decisions b < 0 (core.rs:129-138, llr.rs:73-75) */
static int satisfied(const qc_view *q, const float *b) {
    for (uint32_t br = 0; br < q->mb; ++br)
        for (uint32_t i = 0; i < q->z; ++i) {
            unsigned par = 0;
            for (uint32_t e = q->row_ptr[br]; e < q->row_ptr[br + 1]; ++e) par ^= b[var_of(q, e, i)] < 0.0f;
            if (par) return 0;
        }
    return 1;
}

static void message_words(const float *b, uint32_t k, uint64_t *msg) {
    const size_t kw = ((size_t)k + 63) / 64;
    memset(msg, 0, kw * sizeof(uint64_t));
    for (uint32_t v = 0; v < k; ++v)
        if (b[v] < 0.0f) msg[v >> 6] |= (uint64_t)1 << (v & 63);
}

void or_layered_decode(const int32_t *shifts, uint32_t mb, uint32_t nb, uint32_t z, const float *llr, uint32_t T,
                       const float *w, uint32_t fb_it, uint64_t *msg_words, or_result *res, uint8_t *fell_back,
                       float *trace_bel) {
    qc_view q;
    qc_build(&q, shifts, mb, nb, z);
    const uint32_t n = nb * z, k = (nb - mb) * z;
    const size_t E = (size_t)q.nbe * z;
    float *b = (float *)malloc((size_t)n * sizeof(float));
    float *m = (float *)calloc(E + 1, sizeof(float)); /* [e][i] unweighted min-sum of the last update */
    float x[64], tmp[64];
    memcpy(b, llr, (size_t)n * sizeof(float));
    uint32_t it = 0;
    int conv = 0;
    *fell_back = 0;
    for (uint32_t t = 1; t <= T && !conv; ++t) {
        const float *wp = t >= 2 ? w + (size_t)(t - 2) * q.nbe : NULL, *wt = w + (size_t)(t - 1) * q.nbe;
        for (uint32_t br = 0; br < mb; ++br) {
            const uint32_t e0 = q.row_ptr[br], d = q.row_ptr[br + 1] - e0;
            for (uint32_t i = 0; i < z; ++i) {
                for (uint32_t p = 0; p < d; ++p) {
                    const float bv = b[var_of(&q, e0 + p, i)];
                    x[p] = wp ? bv - wp[e0 + p] * m[(size_t)(e0 + p) * z + i] : bv;
                }
                for (uint32_t p = 0; p < d; ++p) {
                    const float mp = ext_minsum(x, d, p, tmp);
                    m[(size_t)(e0 + p) * z + i] = mp;
                    b[var_of(&q, e0 + p, i)] = x[p] + wt[e0 + p] * mp;
                }
            }
        }
        it = t;
        if (trace_bel) memcpy(trace_bel + (size_t)(t - 1) * n, b, (size_t)n * sizeof(float));
        conv = satisfied(&q, b);
    }
    if (!conv && fb_it > 0) {
        *fell_back = 1;
        float *c2v = (float *)malloc((E + 1) * sizeof(float));
        const float *wT = w + (size_t)(T - 1) * q.nbe;
        for (uint32_t e = 0; e < q.nbe; ++e)
            for (uint32_t i = 0; i < z; ++i) c2v[(size_t)e * z + i] = wT[e] * m[(size_t)e * z + i];
        /* block edges of each block-column in ascending block-row order */
        for (uint32_t u = 1; u <= fb_it && !conv; ++u) {
            for (uint32_t br = 0; br < mb; ++br) {
                const uint32_t e0 = q.row_ptr[br], d = q.row_ptr[br + 1] - e0;
                for (uint32_t i = 0; i < z; ++i) {
                    for (uint32_t p = 0; p < d; ++p)
                        x[p] = b[var_of(&q, e0 + p, i)] - c2v[(size_t)(e0 + p) * z + i];
                    for (uint32_t p = 0; p < d; ++p) m[(size_t)(e0 + p) * z + i] = ext_minsum(x, d, p, tmp);
                }
            }
            memcpy(c2v, m, E * sizeof(float));
            for (uint32_t v = 0; v < n; ++v) b[v] = llr[v];
            for (uint32_t br = 0; br < mb; ++br)
                for (uint32_t e = q.row_ptr[br]; e < q.row_ptr[br + 1]; ++e)
                    for (uint32_t i = 0; i < z; ++i) {
                        const uint32_t v = var_of(&q, e, i);
                        b[v] = b[v] + c2v[(size_t)e * z + i];
                    }
            it = T + u;
            if (trace_bel) memcpy(trace_bel + (size_t)(T + u - 1) * n, b, (size_t)n * sizeof(float));
            conv = satisfied(&q, b);
        }
        free(c2v);
    }
    message_words(b, k, msg_words);
    res->iterations = it;
    res->converged = (uint8_t)conv;
    res->syndrome_ok = (uint8_t)satisfied(&q, b);
    free(b);
    free(m);
    qc_free(&q);
}

/* Frames [f0, f1) of a batch, one thread per slice (results in input order). */
typedef struct {
    const int32_t *shifts;
    uint32_t mb, nb, z, T, fb_it;
    const float *llr, *w;
    size_t f0, f1;
    uint64_t *msg;
    or_result *res;
    uint8_t *fb;
} layered_job;

static void *layered_worker(void *arg) {
    layered_job *j = (layered_job *)arg;
    const size_t n = (size_t)j->nb * j->z, kw = ((size_t)(j->nb - j->mb) * j->z + 63) / 64;
    for (size_t f = j->f0; f < j->f1; ++f)
        or_layered_decode(j->shifts, j->mb, j->nb, j->z, j->llr + f * n, j->T, j->w, j->fb_it, j->msg + f * kw,
                          j->res + f, j->fb + f, NULL);
    return NULL;
}

void or_layered_decode_batch(const int32_t *shifts, uint32_t mb, uint32_t nb, uint32_t z, const float *llr,
                             size_t frames, uint32_t T, const float *w, uint32_t fb_it, uint64_t *msg_words,
                             or_result *res, uint8_t *fell_back, int threads) {
    if (frames == 0) return;
    if (threads < 1) threads = 1;
    if ((size_t)threads > frames) threads = (int)frames;
    pthread_t tid[64];
    layered_job jobs[64];
    if (threads > 64) threads = 64;
    for (int t = 0; t < threads; ++t) {
        jobs[t] = (layered_job){shifts, mb, nb, z, T, fb_it, llr, w, frames * t / threads,
                                frames * (t + 1) / threads, msg_words, res, fell_back};
        if (threads == 1)
            layered_worker(&jobs[t]);
        else
            pthread_create(&tid[t], NULL, layered_worker, &jobs[t]);
    }
    if (threads > 1)
        for (int t = 0; t < threads; ++t) pthread_join(tid[t], NULL);
}

/* ---- training oracle (double precision) -------------------------------
 * T layered iterations without early stop; loss = sum over t and v of
 * softplus(-y_v * b_t[v]) / (T n), y = +1 for bit 0, -1 for bit 1 (BCE of
 * P(bit = 1) = sigmoid(-b)), b_t the beliefs after iteration t.  Min-sum is
 * differentiated as a subgradient: the message to p depends on the extrinsic
 * argmin a(p) only (first on ties), d|x_a|/dx_a = sign(x_a).  Reverse mode
 * over the layers in the opposite order:
 *   gc_p = G[v_p] + H(e_p, i)           dL/d(weighted message of t)
 *   dL/dw_t(e_p) += gc_p * m_p
 *   gx_p = G[v_p]; gx_{a(p)} += gc_p * w_t(e_p) * sgn_p * sign(x_{a(p)})
 *   G[v_p] = gx_p;  H(e_p, i) = -gx_p   (x_p subtracted the previous message)
 * with G = dL/db, G += the BCE derivative at the end of each iteration. */
static double softplus_d(double x) { return x > 0 ? x + log1p(exp(-x)) : log1p(exp(x)); }

double or_layered_loss_grad(const int32_t *shifts, uint32_t mb, uint32_t nb, uint32_t z, const float *llr,
                            const uint8_t *bits, uint32_t T, const double *w, double *grad) {
    qc_view q;
    qc_build(&q, shifts, mb, nb, z);
    const uint32_t n = nb * z;
    const size_t E = (size_t)q.nbe * z;
    /* tape: inputs x and min-sum outputs of every (t, edge, lane); beliefs per t */
    double *X = (double *)calloc((size_t)T * E + 1, sizeof(double));
    double *M = (double *)calloc((size_t)T * E + 1, sizeof(double));
    int32_t *A = (int32_t *)calloc((size_t)T * E + 1, sizeof(int32_t)); /* extrinsic argmin position */
    double *B = (double *)calloc((size_t)T * n + 1, sizeof(double));
    double *b = (double *)malloc((size_t)n * sizeof(double));
    for (uint32_t v = 0; v < n; ++v) b[v] = (double)llr[v];
    double loss = 0.0;
    for (uint32_t t = 1; t <= T; ++t) {
        const double *wp = t >= 2 ? w + (size_t)(t - 2) * q.nbe : NULL, *wt = w + (size_t)(t - 1) * q.nbe;
        double *Xt = X + (size_t)(t - 1) * E, *Mt = M + (size_t)(t - 1) * E;
        const double *Mp = t >= 2 ? M + (size_t)(t - 2) * E : NULL;
        int32_t *At = A + (size_t)(t - 1) * E;
        for (uint32_t br = 0; br < mb; ++br) {
            const uint32_t e0 = q.row_ptr[br], d = q.row_ptr[br + 1] - e0;
            for (uint32_t i = 0; i < z; ++i) {
                for (uint32_t p = 0; p < d; ++p) {
                    const size_t ei = (size_t)(e0 + p) * z + i;
                    const double bv = b[var_of(&q, e0 + p, i)];
                    Xt[ei] = wp ? bv - wp[e0 + p] * Mp[ei] : bv;
                }
                for (uint32_t p = 0; p < d; ++p) {
                    double mn = INFINITY, s = 1.0;
                    int32_t arg = -1;
                    for (uint32_t r = 0; r < d; ++r) {
                        if (r == p) continue;
                        const double xr = Xt[(size_t)(e0 + r) * z + i];
                        if (xr < 0) s = -s;
                        if (fabs(xr) < mn) {
                            mn = fabs(xr);
                            arg = (int32_t)r;
                        }
                    }
                    const size_t ei = (size_t)(e0 + p) * z + i;
                    Mt[ei] = arg < 0 ? 0.0 : s * mn;
                    At[ei] = arg;
                }
                for (uint32_t p = 0; p < d; ++p) {
                    const size_t ei = (size_t)(e0 + p) * z + i;
                    b[var_of(&q, e0 + p, i)] = Xt[ei] + wt[e0 + p] * Mt[ei];
                }
            }
        }
        memcpy(B + (size_t)(t - 1) * n, b, (size_t)n * sizeof(double));
        for (uint32_t v = 0; v < n; ++v) loss += softplus_d(-(bits[v] ? -1.0 : 1.0) * b[v]);
    }
    const double scale = 1.0 / ((double)T * (double)n);
    loss *= scale;
    if (grad) {
        memset(grad, 0, (size_t)T * q.nbe * sizeof(double));
        double *G = (double *)calloc((size_t)n + 1, sizeof(double));
        double *H = (double *)calloc(E + 1, sizeof(double));
        double gx[64];
        for (uint32_t t = T; t >= 1; --t) {
            const double *Bt = B + (size_t)(t - 1) * n;
            for (uint32_t v = 0; v < n; ++v) {
                const double y = bits[v] ? -1.0 : 1.0;
                G[v] += -y * scale / (1.0 + exp(y * Bt[v]));
            }
            const double *wt = w + (size_t)(t - 1) * q.nbe;
            const double *Xt = X + (size_t)(t - 1) * E, *Mt = M + (size_t)(t - 1) * E;
            const int32_t *At = A + (size_t)(t - 1) * E;
            for (uint32_t br = mb; br-- > 0;) {
                const uint32_t e0 = q.row_ptr[br], d = q.row_ptr[br + 1] - e0;
                for (uint32_t i = 0; i < z; ++i) {
                    for (uint32_t p = 0; p < d; ++p) gx[p] = G[var_of(&q, e0 + p, i)];
                    for (uint32_t p = 0; p < d; ++p) {
                        const size_t ei = (size_t)(e0 + p) * z + i;
                        const double gc = G[var_of(&q, e0 + p, i)] + H[ei];
                        grad[(size_t)(t - 1) * q.nbe + e0 + p] += gc * Mt[ei];
                        const int32_t a = At[ei];
                        if (a >= 0) {
                            /* m_p = sgn_p * |x_a|: d m_p / d x_a = sgn_p * sign(x_a) */
                            const double xa = Xt[(size_t)(e0 + a) * z + i];
                            double sgn_p = 1.0;
                            for (uint32_t r = 0; r < d; ++r)
                                if (r != p && Xt[(size_t)(e0 + r) * z + i] < 0) sgn_p = -sgn_p;
                            gx[a] += gc * wt[e0 + p] * sgn_p * (xa < 0 ? -1.0 : 1.0);
                        }
                    }
                    for (uint32_t p = 0; p < d; ++p) {
                        const size_t ei = (size_t)(e0 + p) * z + i;
                        G[var_of(&q, e0 + p, i)] = gx[p];
                        H[ei] = t >= 2 ? -gx[p] : 0.0;
                    }
                }
            }
        }
        free(G);
        free(H);
    }
    free(X);
    free(M);
    free(A);
    free(B);
    free(b);
    qc_free(&q);
    return loss;
}
