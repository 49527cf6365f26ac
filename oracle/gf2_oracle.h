/*
 * CPU oracle for the batched LDPC belief-propagation decoder.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * decoder (erankavija/gf2, crate gf2-coding) used as the checker for the HIP
 * product path.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product never links or calls it.
 *
 * Every function cites the reference file:line it restates.  Which of them
 * are pinned by what (reference known-answer tests, closed forms, brute
 * force) is listed in tests/test_oracle.py; every function here has at least
 * one such check.
 */
#ifndef GF2_ORACLE_H
#define GF2_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Sparse parity-check matrix in the reference's dual CSR/CSC form
 * (crates/gf2-core/src/sparse.rs:422-425). */
typedef struct or_code {
    uint32_t m, n;
    uint64_t nnz;
    uint64_t *row_ptr; /* m+1 */
    uint32_t *row_idx; /* nnz, ascending within a row            */
    uint64_t *col_ptr; /* n+1 */
    uint32_t *col_idx; /* nnz, ascending check index (stable)    */
    uint64_t *csc_to_csr; /* nnz: CSC edge -> CSR edge of same (c,v) */
} or_code;

/* QuasiCyclicLdpc::to_edges (crates/gf2-coding/src/ldpc/core.rs:578-605)
 * with CirculantMatrix::to_edges (core.rs:436-447).  shifts is mb*nb
 * row-major, -1 = zero block.  Writes 2*count uint32 (row, col) pairs into
 * out_rc (capacity cap pairs).  Returns 0, or -1 on an invalid shift
 * (reference panics, core.rs:589-596), -2 if cap is too small. */
int or_qc_to_edges(uint32_t mb, uint32_t nb, uint32_t z, const int32_t *shifts,
                   uint32_t *out_rc, size_t cap, size_t *count);

/* LdpcCode::from_edges -> SpBitMatrixDual::from_coo
 * (core.rs:96-104, sparse.rs:166-204, 297-331, 441-445).
 * Duplicates cancel in pairs (XOR), rows sorted ascending, CSC by stable
 * transpose.  Returns NULL on an out-of-range index (reference panics). */
or_code *or_code_from_edges(uint32_t m, uint32_t n, const uint32_t *rc, size_t nnz);
void or_code_free(or_code *c);

/* minsum_avx2_f32 (crates/gf2-kernels-simd/src/llr.rs:117-180): 8-lane
 * abs/min/sign-bit XOR over full chunks, horizontal f32::min, scalar
 * remainder with val<0.  Emulated lane by lane, including MINPS operand
 * order for NaN. */
float or_minsum_avx2(const float *x, size_t n);
/* Llr::boxplus_minsum_n_scalar (crates/gf2-coding/src/llr.rs:272-284). */
float or_minsum_scalar(const float *x, size_t n);

/* GF(2) syndrome s = H x (sparse.rs:367-380).  x, s are LE u64 words
 * (bitvec.rs:40-63).  Returns the popcount of s (core.rs:135-138). */
uint64_t or_syndrome(const or_code *c, const uint64_t *x_words, uint64_t *s_words);

/* Per-frame result, mirroring DecoderResult (crates/gf2-coding/src/traits.rs:116-156). */
typedef struct or_result {
    uint32_t iterations;
    uint8_t converged;
    uint8_t syndrome_ok;
} or_result;

/* LdpcDecoder::decode_iterative on a fresh decoder (core.rs:653-697,
 * 790-865, 889-938).
 *   llr         n channel LLRs
 *   max_it      iteration cap
 *   edge_w      optional normalized-min-sum weights, one per CSR edge per
 *               iteration: edge_w[t*nnz + e].  NULL = plain min-sum.  The
 *               check message becomes w * minsum(...) (llr.rs:352-355).
 *   msg_words   ceil(k/64) words, first k hard bits, tail zero
 *   trace_c2v   optional [max_it][nnz] check->variable messages in CSR order
 *   trace_bel   optional [max_it][n] beliefs
 * Iterations that were not run leave their trace rows untouched. */
void or_decode_iterative(const or_code *c, const float *llr, uint32_t max_it,
                         const float *edge_w, uint64_t *msg_words, or_result *res,
                         float *trace_c2v, float *trace_bel);

/* As or_decode_iterative; with no_early_stop != 0 all max_it iterations run
 * (the unrolled forward pass of the learned decoder), converged still
 * reports whether any iteration met the syndrome. */
void or_decode_iterative_ex(const or_code *c, const float *llr, uint32_t max_it,
                            const float *edge_w, uint64_t *msg_words, or_result *res,
                            float *trace_c2v, float *trace_bel, int no_early_stop);

/* LdpcDecoder::decode_batch (core.rs:729-756): frames decoded independently,
 * results in input order.  llr is frames*n, frame-major.  threads<=1 runs
 * serially. */
void or_decode_batch(const or_code *c, const float *llr, size_t frames, uint32_t max_it,
                     const float *edge_w, uint64_t *msg_words, or_result *res, int threads);

/* Sum-product check rule: Llr::boxplus_n (llr.rs:210-215) for n >= 1 inputs,
 * and the decoder with check_node_update_spa (core.rs:762-787) in place of
 * the min-sum update (dead code in the reference; same loop otherwise). */
float or_boxplus_n(const float *x, size_t n);
void or_decode_spa(const or_code *c, const float *llr, uint32_t max_it, uint64_t *msg_words,
                   or_result *res, float *trace_c2v, float *trace_bel);
void or_decode_batch_spa(const or_code *c, const float *llr, size_t frames, uint32_t max_it,
                         uint64_t *msg_words, or_result *res, int threads);

/* ---- normalized min-sum training oracle (double precision) ------------
 * Forward pass of the weighted min-sum decoder with T unrolled iterations
 * (no early stop), and the gradient of
 *     loss = sum_t sum_v BCE(sigmoid(-belief_t[v]), bit[v]) / (T*n)
 * with respect to the per-edge weights edge_w[t*nnz+e].  Pure double
 * arithmetic.  Returns the loss.  grad may be NULL. */
double or_nms_loss_grad(const or_code *c, const float *llr, const uint8_t *bits,
                        uint32_t T, const double *edge_w, double *grad);

/* ---- channel / frame restatements used to pin the frame generator ---- */
/* Philox4x32-10 counter-based generator (Salmon et al., SC'11). */
void or_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Restatement of the device frame generator for one frame with global id g
 * (see gf2l_frames_generate): writes n LLRs and ceil(n/64) codeword words. */
void or_framegen(const int32_t *shifts, uint32_t mb, uint32_t nb, uint32_t z, uint64_t seed,
                 uint64_t g, float ebn0_lo, float ebn0_hi, float *llr, uint64_t *cw_words);
/* All-zero codeword frame g of an n-bit, rate k/n code (same Eb/N0 and noise
 * streams as or_framegen with every bit 0). */
void or_framegen_zero(uint32_t n, uint32_t k, uint64_t seed, uint64_t g, float ebn0_lo, float ebn0_hi,
                      float *llr);

/* ---- layered learned decoder (layered_oracle.c) ------------------------
 * Layered weighted min-sum over a quasi-cyclic code (block-rows in order, one
 * weight per block edge and iteration, w[t*nbe + e], block edges in
 * block-row-major order), T iterations with early stop, then for a frame
 * still failing up to fb_it plain flooding iterations continued from the
 * layered state.  fell_back = 1 when the continuation ran.  trace_bel
 * (optional, [T + fb_it][n]) receives the beliefs after each layered and
 * each continuation iteration (rows of iterations not run are untouched). */
void or_layered_decode(const int32_t *shifts, uint32_t mb, uint32_t nb, uint32_t z, const float *llr, uint32_t T,
                       const float *w, uint32_t fb_it, uint64_t *msg_words, or_result *res, uint8_t *fell_back,
                       float *trace_bel);
/* or_layered_decode over `frames` frames (llr frame-major), `threads` threads. */
void or_layered_decode_batch(const int32_t *shifts, uint32_t mb, uint32_t nb, uint32_t z, const float *llr,
                             size_t frames, uint32_t T, const float *w, uint32_t fb_it, uint64_t *msg_words,
                             or_result *res, uint8_t *fell_back, int threads);
/* Loss (mean BCE of the beliefs after each of T layered iterations, no early
 * stop) and its gradient w.r.t. w[t*nbe + e], double precision. */
double or_layered_loss_grad(const int32_t *shifts, uint32_t mb, uint32_t nb, uint32_t z, const float *llr,
                            const uint8_t *bits, uint32_t T, const double *w, double *grad);

/* ---- structural restatement of the rayon path (ref_port.cpp) ---------- */
void *port_code_new(uint32_t m, uint32_t n, const uint32_t *rc, size_t nnz);
void port_code_free(void *code);
void port_decode_batch(const void *code, const float *llr, size_t frames, uint32_t max_it,
                       uint64_t *msg_words, or_result *res, int threads);

#ifdef __cplusplus
}
#endif
#endif
