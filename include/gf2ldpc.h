/*
 * gf2ldpc.h -- C ABI of the MI355X (gfx950) batched LDPC belief-propagation
 * decoder.
 *
 * This is the drop-in boundary for the reference's rayon batch path,
 *   LdpcDecoder::decode_batch(code: &LdpcCode, llr_blocks: &[Vec<Llr>],
 *                             max_iterations: usize) -> Vec<DecoderResult>
 *   (crates/gf2-coding/src/ldpc/core.rs:729-756),
 * and of the code constructor it depends on,
 *   LdpcCode::from_quasi_cyclic(&QuasiCyclicLdpc)   (core.rs:168-173).
 *
 * Conventions
 *   - Plain pointers and sizes only.  The library never keeps a caller
 *     pointer beyond the call that received it.
 *   - Every entry point returns GF2L_OK (0) or a negative GF2L_ERR_* code;
 *     gf2l_last_error() then holds a message.  Where the reference panics
 *     (invalid shift, LLR length mismatch) these calls return
 *     GF2L_ERR_INVALID with the reference's panic text.  No C++ exception
 *     crosses this boundary: a failed host allocation returns
 *     GF2L_ERR_NOMEM, any other exception GF2L_ERR_INTERNAL.
 *   - A context owns one HIP stream and its device buffers.  It is not
 *     thread-safe: use one context per calling thread (the reference gives
 *     each rayon task its own decoder, core.rs:737-743).
 *   - Results are deterministic and independent of batch size and order.
 *
 * Bit layout of decoded words: bit i of a frame's message is bit (i & 63) of
 * word (i >> 6), little-endian u64 words, tail bits zero -- the BitVec layout
 * of crates/gf2-core/src/bitvec.rs:40-63, so BitVec::from_words(words, k)
 * (bitvec.rs:204) rebuilds DecoderResult::decoded_bits directly.
 */
#ifndef GF2LDPC_H
#define GF2LDPC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GF2L_OK 0
#define GF2L_ERR_INVALID (-1)     /* bad argument; reference would panic       */
#define GF2L_ERR_HIP (-2)         /* HIP runtime error                           */
#define GF2L_ERR_NOMEM (-3)       /* device or host allocation failed            */
#define GF2L_ERR_UNSUPPORTED (-4) /* shape outside what the kernels handle       */
#define GF2L_ERR_INTERNAL (-5)    /* unexpected C++ exception inside the library */

typedef struct gf2l_ctx gf2l_ctx;
typedef struct gf2l_code gf2l_code;

/* ---- context ---------------------------------------------------------- */

/* Creates a context on HIP device `device`.  No reference counterpart: the
 * reference has no device; this replaces rayon's global pool
 * (crates/gf2-core/src/compute/cpu.rs:39-45). */
int gf2l_create(int device, gf2l_ctx **out);
void gf2l_destroy(gf2l_ctx *ctx);
/* Message of the last failing call on this context ("" if none). */
const char *gf2l_last_error(const gf2l_ctx *ctx);
/* Waits for all work queued on the context's stream. */
int gf2l_synchronize(gf2l_ctx *ctx);
/* The context's hipStream_t, for callers that order their own work on it. */
void *gf2l_stream(gf2l_ctx *ctx);

/* Per-kernel timing: HIP events are recorded around kernel launches on the
 * context's stream.  on = GF2L_PROFILE_ALL (or any value other than 0 and
 * GF2L_PROFILE_LAYERED): every check / variable / syndrome launch of the HBM
 * paths and every layered learned-decoder launch; the HBM paths then read the
 * running count back after every iteration (frames processed per launch).
 * on = GF2L_PROFILE_LAYERED: the layered kernel only, so the plain
 * continuation after it runs as it does unprofiled.  0: off.  Enabling (or
 * re-enabling) resets the counters.  No reference counterpart: the reference
 * has no in-code timing (profiling was done with perf,
 * crates/gf2-coding/docs/LDPC_PERFORMANCE.md:80-96). */
#define GF2L_PROFILE_ALL 1
#define GF2L_PROFILE_LAYERED 2
int gf2l_profile_enable(gf2l_ctx *ctx, int on);
/* ms[3] and launches[3] for the check, variable and syndrome kernels;
 * frame_iterations = frames processed summed over iterations. */
int gf2l_profile_read(gf2l_ctx *ctx, double *ms, uint64_t *launches, uint64_t *frame_iterations);
/* One kernel's figures: summed ms, launches, and frame-iterations processed
 * (for the check / variable / syndrome kernels the HBM paths' frames summed
 * over iterations, for the layered learned decoder's kernel its iterations
 * summed over frames). */
#define GF2L_KERNEL_CHECK 0
#define GF2L_KERNEL_VARIABLE 1
#define GF2L_KERNEL_SYNDROME 2
#define GF2L_KERNEL_LAYERED 3
int gf2l_profile_read_kernel(gf2l_ctx *ctx, int kernel, double *ms, uint64_t *launches,
                             uint64_t *frame_iterations);

/* Decoder path selection (tests force each path; default automatic).
 *   GF2L_PATH_AUTO    LDS when a frame's state fits, else STATE when every
 *                     check degree is <= 27, else STREAM
 *   GF2L_PATH_LDS     one workgroup per frame, all iterations in one launch
 *   GF2L_PATH_STREAM  per-iteration kernels over HBM, one stored message per
 *                     edge (check degree <= 32)
 *   GF2L_PATH_STATE   per-iteration kernels over HBM, 12-byte compressed
 *                     state per check from which messages are rebuilt
 * A forced path that cannot run a code makes decode calls return
 * GF2L_ERR_UNSUPPORTED.  All paths give identical results. */
#define GF2L_PATH_AUTO 0
#define GF2L_PATH_LDS 1
#define GF2L_PATH_STREAM 2
#define GF2L_PATH_STATE 3
int gf2l_set_path(gf2l_ctx *ctx, int path);
/* The path a decode of `code` on this context takes under its current path
 * setting and check rule: GF2L_PATH_LDS, _STREAM or _STATE for a
 * quasi-cyclic or DVB code, GF2L_PATH_CSR (reported here only) for a general
 * sparse code.  GF2L_ERR_UNSUPPORTED when the forced path or rule cannot run
 * the code (the error a decode call would return).  No reference
 * counterpart: the reference has one CPU path. */
#define GF2L_PATH_CSR 4
int gf2l_code_path(gf2l_ctx *ctx, const gf2l_code *code, int *path);

/* Check-node rule of the plain decoder (gf2l_decode_batch[_device],
 * gf2l_decode_trace):
 *   GF2L_RULE_MINSUM  (default) LdpcDecoder::check_node_update_minsum
 *                     (core.rs:790-814), bit-exact with the reference
 *   GF2L_RULE_SPA     LdpcDecoder::check_node_update_spa (core.rs:762-787,
 *                     compiled but never called by the reference decoder):
 *                     the message to an edge is Llr::boxplus_n of the others
 *                     (llr.rs:210-215), 2 atanh(prod tanh(L/2)) in binary32.
 *                     The device's tanh/atanh differ from the host C
 *                     library's in the last bits, so results match the
 *                     reference within a tolerance, not bit for bit.  Runs on
 *                     the STREAM path of quasi-cyclic codes only (a forced
 *                     LDS or STATE path, a DVB or general sparse code, the
 *                     learned, mixed-code and no-allocation decoders return
 *                     GF2L_ERR_UNSUPPORTED under it). */
#define GF2L_RULE_MINSUM 0
#define GF2L_RULE_SPA 1
int gf2l_set_check_rule(gf2l_ctx *ctx, int rule);

/* Launch-shape tunables of the HBM paths (results never depend on them).
 *   GF2L_TUNE_WG_PER_CU  persistent workgroups per compute unit (1..32); 0 (default):
 *                        5 with counter-fed items (GF2L_TUNE_CHUNK), else 3
 *   GF2L_TUNE_XCD_MAP    1 (default): each frame's block-rows / block-columns
 *                        run on one XCD so its beliefs and check states are
 *                        re-read from that XCD's L2; 0: frame-major over the grid
 *   GF2L_TUNE_GRAPH_ITERS iterations per host check for running frames (1..64,
 *                        default 4): captured per graph replay by the mixed-code
 *                        decoder, enqueued back to back by the single-code
 *                        decoder (iterations after the last frame retired do
 *                        nothing)
 *   GF2L_TUNE_CHUNK      consecutive items a workgroup of the check-state
 *                        kernels takes per shared-counter fetch (1..64);
 *                        0: fixed-stride item assignment; -1 (default):
 *                        8 when every variable degree is at most 8, else 0
 *   GF2L_TUNE_LAYERED_COLUMNS  block-columns of beliefs the general layered
 *                        kernel keeps in LDS per workgroup (the rest in
 *                        global memory); 0 (default): as many as fit in a
 *                        third of the LDS (three frames per compute unit).
 *                        A value above 0 also selects the general kernel
 *                        when GF2L_TUNE_LAYERED_KERNEL is 0.
 *   GF2L_TUNE_LAYERED_KERNEL  kernel of the layered learned decoder:
 *                        0 (default): the staircase kernel where the code
 *                        allows it (QC-IRA code whose parity part is a block
 *                        staircase of unshifted circulants, z <= 512, a
 *                        block-row count of at most 256 that is a multiple
 *                        of 10 (at least 20) or of 6 (at least 12), row
 *                        degrees dh - 1 and dh with dh 6 or 7, the
 *                        information beliefs of one frame in LDS), else the
 *                        general kernel; 1: the general
 *                        kernel; 2: the staircase kernel (decoding fails
 *                        with GF2L_ERR_UNSUPPORTED where it does not apply).
 *                        Training always runs the general kernel.
 *   GF2L_TUNE_LDS_GROUPS thread groups per workgroup of the LDS path (one
 *                        frame per workgroup): group g updates block-rows
 *                        and block-columns g, g + groups, ... with min(z, 512)
 *                        lanes each; at most 512 / min(z, 512) (larger
 *                        values are clamped).  0 (default): the count with
 *                        the fewest check plus variable passes.
 *   GF2L_TUNE_HBM_GRAPH  1: the plain iterations of the single-code HBM paths
 *                        after the first of the final phase are replayed from
 *                        a graph of GF2L_TUNE_GRAPH_ITERS iterations (rounded
 *                        down to an even count), captured once per code,
 *                        batch size and buffers and kept by the context;
 *                        0 (default): launched one by one.  Never while
 *                        profiling or tracing, never by the no-allocation
 *                        entry point. */
#define GF2L_TUNE_WG_PER_CU 1
#define GF2L_TUNE_XCD_MAP 2
#define GF2L_TUNE_GRAPH_ITERS 3
#define GF2L_TUNE_CHUNK 4
#define GF2L_TUNE_LAYERED_COLUMNS 5
#define GF2L_TUNE_LAYERED_KERNEL 6
#define GF2L_TUNE_LDS_GROUPS 7
#define GF2L_TUNE_HBM_GRAPH 8
int gf2l_set_tuning(gf2l_ctx *ctx, int key, int value);

/* ---- codes ------------------------------------------------------------ */

/* Quasi-cyclic code from an mb x nb base matrix of circulant shifts
 * (row-major, -1 = zero block, else 0 <= s < z) and lift z.
 * Replaces QuasiCyclicLdpc::new + LdpcCode::from_quasi_cyclic
 * (core.rs:519-540, 168-173): block (br, bc) with shift s connects check
 * br*z + i to variable bc*z + (i + s) mod z (CirculantMatrix::to_edges,
 * core.rs:436-447).  Errors: an out-of-range shift returns GF2L_ERR_INVALID
 * with the text of core.rs:589-596. */
int gf2l_code_qc(gf2l_ctx *ctx, uint32_t mb, uint32_t nb, uint32_t z, const int32_t *shifts,
                 gf2l_code **out);
/* General sparse code from (row, col) pairs: LdpcCode::from_edges
 * (core.rs:96-104) -> SpBitMatrixDual::from_coo (crates/gf2-core/src/
 * sparse.rs:166-204, 297-331): duplicates cancel in pairs (XOR), rows sorted,
 * columns by stable transpose.  An out-of-range index returns
 * GF2L_ERR_INVALID with the text of sparse.rs:170-171.  Decoded by a CSR/CSC
 * path with the same arithmetic (plain min-sum only: no weights, no traces).
 * Check degree <= 32. */
int gf2l_code_edges(gf2l_ctx *ctx, uint32_t m, uint32_t n, const uint32_t *rc_pairs, size_t nnz,
                    gf2l_code **out);
/* n = nb*z (LdpcCode::n), m = mb*z (LdpcCode::m), k = n - m (LdpcCode::k,
 * core.rs:107-119), nnz = number of Tanner-graph edges. */
int gf2l_code_info(const gf2l_code *code, uint32_t *n, uint32_t *m, uint32_t *k, uint64_t *nnz);
void gf2l_code_destroy(gf2l_code *code);

/* Seeded quasi-cyclic IRA base matrix (no reference counterpart: the
 * project's synthetic code family).  Info block-columns 0..nb-mb-1 get the
 * given column weights; block rows are chosen to balance check degrees and
 * shifts are drawn uniformly in [0, z), redrawn to avoid length-4 cycles
 * where possible.  The parity part is the block dual-diagonal staircase:
 * block-column nb-mb+j has identity blocks at block-rows j and j+1 (the last
 * one weight 1).  Writes mb*nb shifts, row-major. */
int gf2l_qc_ira_shifts(uint64_t seed, uint32_t mb, uint32_t nb, uint32_t z,
                       const uint32_t *info_col_weight, int32_t *shifts_out);

/* ---- code ingestion: DVB-T2 tables and GF2DATA files ------------------ */

/* DVB-T2 code parameters and ETSI EN 302 755 address tables, as
 * DvbParams::for_code (crates/gf2-coding/src/ldpc/dvb_t2/params.rs:56-86)
 * and the tables of dvb_t2/dvb_t2_matrices.rs.  frame: 0 short (N=16200),
 * 1 normal (N=64800); rate: 0 = 1/2, 1 = 3/5, 2 = 2/3, 3 = 3/4, 4 = 4/5,
 * 5 = 5/6 (the order of CodeRate).  gf2l_dvb_t2_table writes `rows` row
 * lengths and `entries` addresses (row after row); either pointer may be
 * NULL.  GF2L_ERR_INVALID for an unknown frame or rate. */
int gf2l_dvb_t2_params(int frame, int rate, uint32_t *n, uint32_t *k, uint32_t *q, uint32_t *rows,
                       uint32_t *entries);
int gf2l_dvb_t2_table(int frame, int rate, uint32_t *row_len, uint32_t *addresses);

/* DVB-style IRA code from an address table: build_dvb_edges
 * (dvb_t2/builder.rs:74-118) + LdpcCode::from_edges.  Information bit
 * i*z + j of table row i and address p connects to check (p + j*q) mod m
 * (m = n - k); check p connects to parity bits k+p and, for p > 0, k+p-1.
 * Validation as validate_table (builder.rs:15-41): rows == k / z, every
 * address < m, no single-entry placeholder table -- GF2L_ERR_INVALID with
 * the reference's panic text otherwise.  Any table of this shape works
 * (DVB-S2 uses the same construction).  When m = q*z (and no address
 * repeats within a row) the code is decoded by the state kernels in a
 * renumbered quasi-cyclic form (check a + q*b -> lane b of block-row a,
 * parity bit likewise), bit-exact with the standard numbering; otherwise by
 * the CSR path.  DVB codes take plain min-sum only. */
int gf2l_code_dvb_table(gf2l_ctx *ctx, uint32_t n, uint32_t k, uint32_t q, uint32_t z, uint32_t rows,
                        const uint32_t *row_len, const uint32_t *addresses, gf2l_code **out);
/* LdpcCode::dvb_t2_short / LdpcCode::dvb_t2_normal (crates/gf2-coding/src/
 * ldpc/core.rs:197-255): frame and rate as gf2l_dvb_t2_params. */
int gf2l_code_dvb_t2(gf2l_ctx *ctx, int frame, int rate, gf2l_code **out);

/* Code from a GF2DATA file image (`bytes`, `len` bytes, e.g. a whole file
 * read into memory): crates/gf2-core/src/io/mod.rs:6-98, header.rs:33-86.
 * Type 3 (SpBitMatrix, COO pairs; entries kept once, sparse.rs:151-200,
 * 232-259) or type 4 (SpBitMatrixDual, CSR + CSC; the CSR rows are used,
 * sparse.rs:397-449); the code is LdpcCode::from_edges over the matrix
 * entries (rows = checks, cols = variables).  Errors (GF2L_ERR_INVALID)
 * carry the reference's IoError text: "Invalid magic bytes", "Unsupported
 * format version: v", "Unknown type tag: t", data-length and metadata
 * mismatches; a compressed payload (flag bit 0, which the reference does not
 * implement either) returns GF2L_ERR_UNSUPPORTED.  A checksum trailer (flag
 * bit 1) is ignored, as by the reference reader. */
int gf2l_code_gf2data(gf2l_ctx *ctx, const void *bytes, size_t len, gf2l_code **out);

/* ---- device memory (for callers without their own HIP allocator) ------ */

/* hipMalloc / hipFree on the context's device; copies are synchronous on the
 * context's stream.  No reference counterpart. */
int gf2l_device_alloc(gf2l_ctx *ctx, size_t bytes, void **out);
int gf2l_device_free(gf2l_ctx *ctx, void *ptr);
int gf2l_copy_to_device(gf2l_ctx *ctx, void *dst, const void *src, size_t bytes);
int gf2l_copy_to_host(gf2l_ctx *ctx, void *dst, const void *src, size_t bytes);

/* ---- frame generation ----------------------------------------------------- */

/* Generates `frames` frames of a staircase QC-IRA code on the device (frame
 * ids first_frame .. first_frame+frames-1): Philox4x32-10 information bits,
 * systematic staircase encoding, BPSK (bit 0 -> +1) over AWGN with a
 * per-frame Eb/N0 drawn uniformly in [ebn0_lo_db, ebn0_hi_db] (equal bounds
 * give a fixed Eb/N0), sigma^2 = 1/(2 R 10^(EbN0/10)) and LLR = 2y/sigma^2
 * computed in double -- the channel of crates/gf2-coding/src/channel.rs:51-57,
 * 100-102, 153-158 with a counter-based noise stream.  Writes d_llr
 * (frames*n floats, frame-major) and, when d_codeword_words is not NULL, the
 * codewords (frames*ceil(n/64) LE words; the first k bits are the message).
 * Queued on the context's stream.  GF2L_ERR_UNSUPPORTED unless the code's
 * parity part is the dual-diagonal staircase and z >= 64. */
int gf2l_frames_generate(gf2l_ctx *ctx, const gf2l_code *code, uint64_t seed, uint64_t first_frame,
                         size_t frames, float ebn0_lo_db, float ebn0_hi_db, float *d_llr,
                         uint64_t *d_codeword_words);

/* All-zero codeword frames of ANY code (QC, DVB-T2, general sparse): the
 * reference FER example's method (crates/gf2-coding/examples/
 * ldpc_awgn.rs:121-175 sends all-zero codewords; a linear code over a
 * symmetric channel with this decoder has the same FER for every codeword).
 * Same Eb/N0 and noise streams as gf2l_frames_generate with every bit 0:
 * LLR = 2(1 + sigma n)/sigma^2, sigma^2 = 1/(2 (k/n) 10^(EbN0/10)).  Writes
 * frames*n floats to d_llr, queued on the context's stream. */
int gf2l_frames_generate_zero(gf2l_ctx *ctx, const gf2l_code *code, uint64_t seed, uint64_t first_frame,
                              size_t frames, float ebn0_lo_db, float ebn0_hi_db, float *d_llr);

/* ---- decoding --------------------------------------------------------- */

/* LdpcDecoder::decode_batch (core.rs:729-756) on host buffers.
 *   llr          frames * n channel LLRs, frame-major (frame f at llr + f*n)
 *   max_it       iteration cap (max_iterations)
 *   msg_words    frames * ceil(k/64) words: first k hard decisions per frame
 *   iterations   frames entries  (DecoderResult::iterations)
 *   converged    frames entries  (DecoderResult::converged)
 *   syndrome_ok  frames entries  (DecoderResult::syndrome_check_passed)
 * frames == 0 is valid and does nothing (core.rs:1297-1309).  Flooding
 * min-sum with syndrome early stop, results bit-identical to the reference
 * decode_iterative for finite LLRs. */
int gf2l_decode_batch(gf2l_ctx *ctx, const gf2l_code *code, const float *llr, size_t frames,
                      uint32_t max_it, uint64_t *msg_words, uint32_t *iterations,
                      uint8_t *converged, uint8_t *syndrome_ok);

/* Same semantics on device-resident buffers (hipMalloc'd by the caller),
 * queued on the context's stream; returns after queueing (call
 * gf2l_synchronize before reading the outputs).  No reference counterpart
 * beyond decode_batch itself: it lets a GPU-resident pipeline (frame
 * generator, FER sweep) skip the PCIe round trip. */
int gf2l_decode_batch_device(gf2l_ctx *ctx, const gf2l_code *code, const float *d_llr,
                             size_t frames, uint32_t max_it, uint64_t *d_msg_words,
                             uint32_t *d_iterations, uint8_t *d_converged,
                             uint8_t *d_syndrome_ok);

/* gf2l_decode_batch_device without any allocation.  No reference
 * counterpart: it is for callers that cannot allocate device memory (code
 * under stream capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, semantics and results as
 * gf2l_decode_batch_device; the working buffers (check states, beliefs, hard
 * decisions, running lists and counters, permuted DVB LLRs) come from one
 * static device array of GF2L_NOALLOC_SCRATCH_BYTES per GPU, which that GPU
 * holds from the first call on it until the process ends.
 *   - Quasi-cyclic codes (DVB-T2 included) with check degree <= 27 that run
 *     on the LDS or STATE path; others return GF2L_ERR_UNSUPPORTED.
 *   - A batch whose buffers exceed the array returns GF2L_ERR_INVALID and a
 *     message with the largest batch of that code that fits (e.g. about 90
 *     frames of an N=64800 rate-1/2 code; N=2304 codes run in LDS and need
 *     none).
 *   - Every call on one GPU shares that GPU's array: calls that overlap on
 *     one GPU (several contexts or threads) race; serialize them.
 * The host reads the running count back between iterations (a synchronous
 * copy into pageable memory), so the call itself cannot be captured. */
#define GF2L_NOALLOC_SCRATCH_BYTES (64u << 20)
int gf2l_decode_batch_device_noalloc(gf2l_ctx *ctx, const gf2l_code *code, const float *d_llr,
                                     size_t frames, uint32_t max_it, uint64_t *d_msg_words,
                                     uint32_t *d_iterations, uint8_t *d_converged,
                                     uint8_t *d_syndrome_ok);

/* Per-iteration state dump for parity testing (host buffers).  Decodes like
 * gf2l_decode_batch and, after each iteration t < trace_iters, copies
 *   c2v_out[t][f][e*z + i]   check->variable message of block edge e
 *                            (block-row-major, block-columns ascending)
 *                            into check (row-of-e)*z + i
 *   belief_out[t][f][v]      posterior LLR of variable v
 * A frame that stopped early keeps its final values in later rows. */
int gf2l_decode_trace(gf2l_ctx *ctx, const gf2l_code *code, const float *llr, size_t frames,
                      uint32_t max_it, uint32_t trace_iters, float *c2v_out, float *belief_out,
                      uint64_t *msg_words, uint32_t *iterations, uint8_t *converged,
                      uint8_t *syndrome_ok);

/* ---- learned normalized min-sum decoder ----------------------------------
 * Weighted min-sum with one weight per base-graph edge e and iteration t:
 *   c2v_t(e) = w_t(e) * minsum(...)
 * -- the reference's alpha * min-sum rule (Llr::boxplus_normalized_minsum_n,
 * crates/gf2-coding/src/llr.rs:352-355) with a learned alpha per edge and
 * iteration; all other steps are the flooding schedule of decode_iterative
 * (core.rs:889-938).  Weights are stored [iteration][edge] with edges in
 * block-row-major order (block-columns ascending within a block-row).  The
 * reference has no learned decoder (ROADMAP.md:93 plans one); these calls
 * have no reference counterpart beyond that rule. */
typedef struct gf2l_nms gf2l_nms;

/* Creates a learned decoder of `iterations` weighted iterations for `code`
 * (which must outlive it); weights from init_weights (host, iterations*nbe)
 * or all equal to init_value. */
int gf2l_nms_create(gf2l_ctx *ctx, const gf2l_code *code, uint32_t iterations,
                    const float *init_weights, float init_value, gf2l_nms **out);
/* A learned decoder with the layered (row-sequential) schedule: the
 * block-rows of the code are processed in ascending order and each updates
 * the beliefs of its variables in place,
 *   x = b - w_{t-1}(e) m_{t-1},  m_t = sign * min over the other inputs,
 *   b = x + w_t(e) m_t
 * (the reference's alpha * min-sum, llr.rs:352-355, per base edge and
 * iteration, on a layered instead of a flooding schedule; the reference has
 * no layered decoder).  Iterations stop early when the hard decisions
 * satisfy every check.  Its fallback always continues from its state
 * (GF2L_FALLBACK_CONTINUE).  The same calls train and decode it as a
 * flooding decoder (gradient, train_step, decode_device, weights, Adam).
 * Check degree <= 27; the beliefs of as many block-columns as fit stay in
 * LDS (GF2L_TUNE_LAYERED_COLUMNS). */
int gf2l_nms_create_layered(gf2l_ctx *ctx, const gf2l_code *code, uint32_t iterations,
                            const float *init_weights, float init_value, gf2l_nms **out);
/* What the plain fallback of gf2l_nms_decode_device starts from:
 *   GF2L_FALLBACK_RESTART   the channel LLRs, zero messages (default of a
 *                           flooding decoder)
 *   GF2L_FALLBACK_CONTINUE  the learned decoder's beliefs and last messages:
 *                           the first plain check update subtracts
 *                           w_T(e) m_T, later ones the plain messages (the
 *                           only mode of a layered decoder) */
#define GF2L_FALLBACK_RESTART 0
#define GF2L_FALLBACK_CONTINUE 1
int gf2l_nms_set_fallback(gf2l_ctx *ctx, gf2l_nms *nms, int mode);
/* Block-columns of beliefs a layered decoder keeps in LDS when it runs on
 * this context (0 for a flooding decoder). */
int gf2l_nms_layered_columns(gf2l_ctx *ctx, const gf2l_nms *nms, uint32_t *columns);
void gf2l_nms_destroy(gf2l_nms *nms);
int gf2l_nms_info(const gf2l_nms *nms, uint32_t *iterations, uint32_t *weights_per_iteration);
int gf2l_nms_get_weights(gf2l_ctx *ctx, const gf2l_nms *nms, float *weights);
int gf2l_nms_set_weights(gf2l_ctx *ctx, gf2l_nms *nms, const float *weights);
/* Adam state for checkpoint / resume (host buffers of iterations*nbe). */
int gf2l_nms_get_adam(gf2l_ctx *ctx, const gf2l_nms *nms, float *m, float *v, uint64_t *step);
int gf2l_nms_set_adam(gf2l_ctx *ctx, gf2l_nms *nms, const float *m, const float *v, uint64_t step);

/* Loss and gradient on a batch (device LLRs d_llr [frames][n], device
 * codewords d_cw [frames][ceil(n/64)] as written by gf2l_frames_generate):
 * loss = mean over frames, iterations 1..T and variables of
 * softplus(-y*belief), y = +1 for bit 0 / -1 for bit 1 (BCE of
 * P(bit=1) = sigmoid(-belief)); grad (host, iterations*nbe) = d loss / d w.
 * Check degree <= 27. */
int gf2l_nms_gradient(gf2l_ctx *ctx, gf2l_nms *nms, const float *d_llr, const uint64_t *d_cw,
                      size_t frames, float *grad, float *loss);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8, learning rate lr) on the
 * gradient of that loss; *loss (if not NULL) receives the batch loss. */
int gf2l_nms_train_step(gf2l_ctx *ctx, gf2l_nms *nms, const float *d_llr, const uint64_t *d_cw,
                        size_t frames, float lr, float *loss);
/* Decodes like gf2l_decode_batch_device with the learned rule for the
 * decoder's iterations; frames whose syndrome still fails are decoded again
 * from their LLRs by plain min-sum for up to fallback_max_it iterations
 * (0: no fallback).  d_iterations counts weighted plus fallback iterations;
 * d_fallback (may be NULL) marks re-decoded frames. */
int gf2l_nms_decode_device(gf2l_ctx *ctx, const gf2l_nms *nms, const float *d_llr, size_t frames,
                           uint32_t fallback_max_it, uint64_t *d_msg_words, uint32_t *d_iterations,
                           uint8_t *d_converged, uint8_t *d_syndrome_ok, uint8_t *d_fallback);

/* ---- mixed-code batches ----------------------------------------------- */

typedef struct gf2l_codeset gf2l_codeset;

/* A set of 1..16 quasi-cyclic codes (each check degree <= 27; the codes must
 * outlive the set) whose frames can be decoded together.  No reference
 * counterpart: the reference decodes one code per decode_batch call
 * (core.rs:729-756); a caller with frames of several codes calls it once per
 * code.  The set owns a workspace and a captured iteration graph, so calls on
 * one set must not overlap. */
int gf2l_codeset_create(gf2l_ctx *ctx, const gf2l_code *const *codes, uint32_t count, gf2l_codeset **out);
void gf2l_codeset_destroy(gf2l_codeset *set);
/* Plain min-sum decode of `frames` frames, frame f of code code_ids[f]
 * (host array, index into the set), each with exactly the results of
 * gf2l_decode_batch_device on its own code: per-frame early stop, at most
 * max_it iterations.  Device buffers, frames packed back to back in input
 * order: d_llr holds n(code_ids[f]) floats per frame, d_msg_words
 * ceil(k/64) words per frame; d_iterations, d_converged, d_syndrome_ok one
 * entry per frame.  The iteration loop runs as replays of a hipGraph of
 * GF2L_TUNE_GRAPH_ITERS iterations.  Errors: a code id outside the set
 * returns GF2L_ERR_INVALID. */
int gf2l_decode_mixed_device(gf2l_ctx *ctx, gf2l_codeset *set, const uint16_t *code_ids, size_t frames,
                             const float *d_llr, uint32_t max_it, uint64_t *d_msg_words,
                             uint32_t *d_iterations, uint8_t *d_converged, uint8_t *d_syndrome_ok);
/* Learned decoding of a mixed batch: decoders[c] is a flooding learned
 * decoder (gf2l_nms_create) of code c of the set, all with the same number T
 * of weighted iterations and the same fallback mode.  Every frame gets
 * exactly the results of gf2l_nms_decode_device with its code's decoder: T
 * weighted iterations with early stop, then, for frames still failing the
 * syndrome, plain min-sum for up to fallback_max_it iterations, restarted
 * from the LLRs or (gf2l_nms_set_fallback continue) continued from the
 * weighted state;
 * d_iterations counts weighted plus fallback iterations and d_fallback (may
 * be NULL) marks re-decoded frames.  The weighted rule is the reference's
 * alpha * min-sum (llr.rs:352-355) with one alpha per base edge and
 * iteration; the fallback is the reference decoder (core.rs:889-938). */
int gf2l_decode_mixed_nms_device(gf2l_ctx *ctx, gf2l_codeset *set, const gf2l_nms *const *decoders,
                                 const uint16_t *code_ids, size_t frames, const float *d_llr,
                                 uint32_t fallback_max_it, uint64_t *d_msg_words, uint32_t *d_iterations,
                                 uint8_t *d_converged, uint8_t *d_syndrome_ok, uint8_t *d_fallback);

/* ---- SDR-style decoder (GNU Radio blocks) ---------------------------- */

/* The C FFI planned in crates/gf2-coding/docs/SDR_INTEGRATION.md:39-57 for a
 * GNU Radio DVB-T2 LDPC decoder block.  A handle owns its own context on HIP
 * device 0, one DVB-T2 code and its staging buffers; one handle per thread.
 *   create   frame_size 0 = short (16200), 1 = normal (64800); code_rate
 *            0 = 1/2, 1 = 3/5, 2 = 2/3, 3 = 3/4, 4 = 4/5, 5 = 5/6;
 *            NULL on a bad argument or when no HIP device is usable.
 *   decode   one frame: `len` must equal n; writes the k information bits,
 *            one per byte (0 or 1, GNU Radio's unpacked-bit stream), and
 *            returns the iteration count when the syndrome is satisfied,
 *            -1 when it is not (the output still holds the hard decision),
 *            -2 on an error (gf2_ldpc_last_error says which).
 *   decode_frames  `frames` frames at once (the block's work() with several
 *            items): outputs k bytes per frame, per-frame iteration counts
 *            (-1 = not converged) when `iterations` is not NULL; returns the
 *            number of converged frames or -2.
 * Results equal gf2l_decode_batch on the same code. */
typedef struct gf2_ldpc_decoder gf2_ldpc_decoder;
gf2_ldpc_decoder *gf2_ldpc_decoder_create(uint32_t frame_size, uint32_t code_rate, uint32_t max_iterations);
int32_t gf2_ldpc_decode(gf2_ldpc_decoder *decoder, const float *llrs, size_t len, uint8_t *output);
int32_t gf2_ldpc_decode_frames(gf2_ldpc_decoder *decoder, const float *llrs, size_t frames, uint8_t *output,
                               int32_t *iterations);
int gf2_ldpc_decoder_info(const gf2_ldpc_decoder *decoder, uint32_t *n, uint32_t *k);
const char *gf2_ldpc_last_error(const gf2_ldpc_decoder *decoder);
void gf2_ldpc_decoder_destroy(gf2_ldpc_decoder *decoder);

#ifdef __cplusplus
}
#endif
#endif /* GF2LDPC_H */
