/* CPU test of the oracle (oracle/gf2_oracle.c, oracle/ref_port.cpp) built
 * with AddressSanitizer and UndefinedBehaviorSanitizer by tests/native/
 * Makefile and run by tests/test_native.py.  It drives every exported oracle
 * function on small inputs (edge lists with duplicates and empty rows, the
 * decoders with traces, several threads, the training oracle, the frame
 * generators) and checks a few invariants: thread count and the structural
 * port do not change results, an all-zero codeword with positive LLRs stops
 * at iteration 1, generated codewords have zero syndrome.  Exit status 0 when
 * every expectation holds; a sanitizer report aborts the run. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gf2_oracle.h"

static int failures = 0;
#define EXPECT(cond, what)                                              \
    do {                                                                \
        if (!(cond)) {                                                  \
            fprintf(stderr, "FAIL line %d: %s\n", __LINE__, what);      \
            ++failures;                                                 \
        }                                                               \
    } while (0)

/* staircase QC-IRA base matrix: mb x nb, info columns of weight 3 */
static void qc_ira(uint32_t mb, uint32_t nb, uint32_t z, int32_t *shifts) {
    const uint32_t kb = nb - mb;
    for (uint32_t i = 0; i < mb * nb; ++i) shifts[i] = -1;
    for (uint32_t c = 0; c < kb; ++c)
        for (uint32_t r = 0; r < 3; ++r) shifts[((c + r) % mb) * nb + c] = (int32_t)((7 * c + 3 * r + 1) % z);
    for (uint32_t j = 0; j < mb; ++j) {
        shifts[j * nb + kb + j] = 0;
        if (j + 1 < mb) shifts[(j + 1) * nb + kb + j] = 0;
    }
}

static float urand(uint32_t *s) {
    *s = *s * 1664525u + 1013904223u;
    return (float)(*s >> 8) / 16777216.0f;
}

int main(void) {
    enum { MB = 4, NB = 8, Z = 16, M = MB * Z, N = NB * Z, K = N - M, KW = (K + 63) / 64, F = 12, T = 3 };
    int32_t shifts[MB * NB];
    qc_ira(MB, NB, Z, shifts);
    size_t cap = MB * NB * Z, count = 0;
    uint32_t *rc = malloc(2 * cap * sizeof(uint32_t));
    EXPECT(or_qc_to_edges(MB, NB, Z, shifts, rc, cap, &count) == 0, "qc_to_edges");
    EXPECT(or_qc_to_edges(MB, NB, Z, shifts, rc, 3, &count) == -2, "qc_to_edges capacity");
    or_qc_to_edges(MB, NB, Z, shifts, rc, cap, &count);
    int32_t bad[MB * NB];
    memcpy(bad, shifts, sizeof(bad));
    bad[0] = Z;
    size_t c2 = 0;
    EXPECT(or_qc_to_edges(MB, NB, Z, bad, rc, cap, &c2) == -1, "invalid shift");
    or_code *code = or_code_from_edges(M, N, rc, count);
    EXPECT(code && code->nnz == count, "code from edges");
    /* duplicates cancel, out-of-range index refused */
    const uint32_t dup[] = {0, 1, 0, 1, 0, 1, 2, 3};
    or_code *cd = or_code_from_edges(4, 4, dup, 4);
    EXPECT(cd && cd->nnz == 2, "duplicates cancel in pairs");
    or_code_free(cd);
    const uint32_t oob[] = {4, 0};
    EXPECT(or_code_from_edges(4, 4, oob, 1) == NULL, "out-of-range row");
    or_code *empty = or_code_from_edges(0, 0, NULL, 0);
    EXPECT(empty && empty->nnz == 0, "empty code");
    or_code_free(empty);

    /* min-sum on every length up to 40 with mixed signs and -0.0 */
    float x[40];
    uint32_t seed = 1;
    for (size_t n = 1; n <= 40; ++n) {
        for (size_t i = 0; i < n; ++i) x[i] = (urand(&seed) - 0.5f) * 8.0f;
        if (n > 3) x[2] = -0.0f;
        const float a = or_minsum_avx2(x, n), s = or_minsum_scalar(x, n);
        EXPECT(fabsf(a) == fabsf(s), "avx2 and scalar magnitudes agree");
    }

    /* decoders: threads and the structural port give the same results */
    float *llr = malloc((size_t)F * N * sizeof(float));
    uint64_t *cw = malloc((size_t)F * ((N + 63) / 64) * sizeof(uint64_t));
    for (uint32_t f = 0; f < F; ++f)
        or_framegen(shifts, MB, NB, Z, 99, f, 1.0f, 3.0f, llr + (size_t)f * N, cw + (size_t)f * ((N + 63) / 64));
    for (uint32_t f = 0; f < F; ++f)
        EXPECT(or_syndrome(code, cw + (size_t)f * ((N + 63) / 64), NULL) == 0, "generated codeword");
    uint64_t m1[F * KW], m2[F * KW], m3[F * KW];
    or_result r1[F], r2[F], r3[F];
    or_decode_batch(code, llr, F, 20, NULL, m1, r1, 1);
    or_decode_batch(code, llr, F, 20, NULL, m2, r2, 4);
    void *port = port_code_new(M, N, rc, count);
    port_decode_batch(port, llr, F, 20, m3, r3, 3);
    EXPECT(memcmp(m1, m2, sizeof(m1)) == 0 && memcmp(m1, m3, sizeof(m1)) == 0, "messages agree");
    for (uint32_t f = 0; f < F; ++f)
        EXPECT(r1[f].iterations == r2[f].iterations && r1[f].iterations == r3[f].iterations &&
                   r1[f].converged == r3[f].converged, "iterations agree");
    port_code_free(port);
    or_decode_batch(code, llr, 0, 20, NULL, m1, r1, 4); /* empty batch */

    /* traces and weights; all-zero codeword with +LLR stops at iteration 1 */
    float *tc = malloc((size_t)T * code->nnz * sizeof(float)), *tb = malloc((size_t)T * N * sizeof(float));
    float *w = malloc((size_t)T * code->nnz * sizeof(float));
    for (size_t i = 0; i < (size_t)T * code->nnz; ++i) w[i] = 0.75f;
    uint64_t msg[KW];
    or_result res;
    or_decode_iterative_ex(code, llr, T, w, msg, &res, tc, tb, 1);
    float pos[N];
    for (int i = 0; i < N; ++i) pos[i] = 4.0f;
    or_decode_iterative(code, pos, 10, NULL, msg, &res, NULL, NULL);
    EXPECT(res.iterations == 1 && res.converged && res.syndrome_ok, "all-zero codeword");
    or_decode_iterative(code, pos, 0, NULL, msg, &res, NULL, NULL);
    EXPECT(res.iterations == 0 && !res.converged, "max_it = 0");

    /* training oracle */
    uint8_t bits[N];
    for (int v = 0; v < N; ++v) bits[v] = (uint8_t)((cw[v >> 6] >> (v & 63)) & 1u);
    double *wd = malloc((size_t)T * code->nnz * sizeof(double)), *g = malloc((size_t)T * code->nnz * sizeof(double));
    for (size_t i = 0; i < (size_t)T * code->nnz; ++i) wd[i] = 0.8;
    const double loss = or_nms_loss_grad(code, llr, bits, T, wd, g);
    EXPECT(isfinite(loss) && loss > 0.0, "training loss");

    /* generators and Philox */
    const uint32_t ctr[4] = {0, 0, 0, 0}, key[2] = {0, 0};
    uint32_t out[4];
    or_philox4x32(ctr, key, out);
    EXPECT(out[0] == 0x6627e8d5u, "Philox4x32-10 zero-key known answer");
    or_framegen_zero(N, K, 5, 3, 2.0f, 2.0f, pos);

    free(wd);
    free(g);
    free(w);
    free(tc);
    free(tb);
    free(llr);
    free(cw);
    free(rc);
    or_code_free(code);
    if (failures) {
        fprintf(stderr, "%d expectation(s) failed\n", failures);
        return 1;
    }
    printf("oracle_sanitize_test: all expectations hold\n");
    return 0;
}
