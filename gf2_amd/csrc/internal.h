// Internal types shared by the host code and the HIP kernels of libgf2ldpc.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <exception>
#include <new>
#include <string>
#include <vector>

namespace gf2l {

// Largest check degree of the explicit-message kernels (sign masks are 64-bit
// with one spare bit for the chunk arithmetic).
constexpr uint32_t kMaxCheckDegree = 32;
// Largest check degree of the compressed-state kernels: argmin (5 bits) and
// one output sign bit per edge share a 32-bit word.
constexpr uint32_t kMaxStateDegree = 27;

// Device view of a quasi-cyclic code: base-matrix block edges in two orders.
//   row order  e = 0..nbe-1: block-rows ascending, block-columns ascending
//              within a row -- the order of SpBitMatrix::row_iter for every
//              lifted check of the block-row.
//   col order  block-columns ascending, block-rows ascending within a
//              column -- the order of the CSC col_iter for every lifted
//              variable, which fixes the belief summation order.
// Explicit message arrays are indexed [e][i], i = check index inside the
// block-row; compressed check state and padded beliefs use a row stride zp.
struct CodeView {
    uint32_t mb, nb, z, n, m, k, nbe;
    uint32_t max_dc, max_dv;
    uint32_t zw;                 // ceil(z/64): 64-bit words per block-column of hard bits
    uint32_t zp;                 // z rounded up to 32 (128-byte rows)
    const uint32_t *row_ptr;     // mb+1, into row-order edges
    const uint2 *row_edge;       // nbe: {bc, shift}
    const uint32_t *col_ptr;     // nb+1, into col-order list
    const uint2 *col_edge;       // nbe: {row-order edge id, shift}
    const uint4 *col_info;       // nbe: {block-row, position in row, shift, row-order edge id}
    const uint2 *col_pack;       // nbe: {block-row * zp, shift << 5 | position} (z < 2^27)
    // DVB-T2/S2 codes in quasi-cyclic order (check c -> (c mod q, c div q),
    // parity bit k+t likewise): the staircase block of block-row 0 that has no
    // edge at lane 0 (~0u for every other code), kb = k/z, q.
    uint32_t wrap_bc = 0xffffffffu;
    uint32_t dvb_kb = 0, dvb_q = 0;
};

// Device view of a general sparse code (LdpcCode::from_edges): CSR rows with
// ascending columns and, per variable, its CSR edges in ascending check
// order (the stable transpose of SpBitMatrix::transpose).
struct CsrView {
    uint32_t m, n, k, nnz;
    const uint32_t *row_ptr;   // m+1
    const uint32_t *col_idx;   // nnz
    const uint32_t *col_ptr;   // n+1
    const uint32_t *csc_edge;  // nnz: CSR edge index
};

// Decoder paths (values of gf2l_set_path).
enum Path : int { kAuto = 0, kLds = 1, kStream = 2, kState = 3 };

// Per-batch device state of the two HBM paths.
struct Workspace {
    size_t cap_frames = 0;
    size_t cap_c2v = 0, cap_state = 0, cap_bel = 0, cap_hb = 0;
    float *c2v = nullptr;        // explicit path: [frames][nbe*z]
    uint32_t *state = nullptr;   // state path: [frames][mb][zp][3] min1, min2, meta
    float *belief = nullptr;     // explicit: [frames][n]; state: [frames][nb][zp]
    uint64_t *hb = nullptr;      // [frames][nb][zw] hard decisions
    uint32_t *lists = nullptr;   // [2][frames] running frame lists
    uint32_t *counts = nullptr;  // [0..1] running counts; item counters from kSchedOffset
    uint32_t *h_count = nullptr; // pinned host mirror of kProfileSlots running counts
    float *llr_perm = nullptr;   // DVB codes: [frames][n] LLRs in quasi-cyclic order
    size_t cap_llr = 0;
};

// Iterations the HBM paths enqueue between two readbacks of the running
// count at most (GF2L_TUNE_GRAPH_ITERS is clamped to it), and the number of
// per-iteration event / count slots kept for them.
constexpr int kProfileSlots = 64;

// Accumulated per-kernel timings of the HBM paths (HIP events recorded
// around each launch on the context's stream; four per iteration, for the
// iterations enqueued between two readbacks).
struct Profile {
    bool on = false;
    bool all = true;                  // every launch; false: the layered kernel only
    hipEvent_t ev[4 * kProfileSlots] = {};
    hipEvent_t lev[2] = {};           // around the layered kernel
    double ms[4] = {0, 0, 0, 0};      // check, variable, syndrome, layered kernels
    uint64_t launches[4] = {0, 0, 0, 0};
    uint64_t frame_iters = 0;         // sum over HBM-path launches of frames processed
    uint64_t layered_frame_iters = 0; // layered iterations run, summed over frames
};

// Launch-shape tunables of the HBM paths (set by gf2l_set_tuning or the
// autotuner).
struct Tuning {
    int wg_per_cu = 0;   // persistent workgroups per compute unit (0: 5 with counter-fed items, else 3)
    int xcd_map = 1;     // 1: frames pinned to XCDs (L2 reuse), 0: frame-major over the grid
    int graph_iters = 4; // iterations per graph replay of the mixed-code decoder
    int chunk = -1;      // items per counter fetch of the check-state kernels (0: fixed stride, -1: automatic)
    int layered_cols = 0; // block-columns of beliefs the general layered kernel keeps in LDS (0: automatic)
    int layered_kernel = 0; // 0: staircase kernel where it applies, 1: general kernel, 2: staircase kernel
    int lds_groups = 0;  // thread groups per workgroup of the LDS path (0: automatic)
    int hbm_graph = 0;   // 1: plain iterations of the HBM paths replayed from a captured graph
};

// Word offset of the item counters in Workspace::counts and its size (two
// kernels x 8 XCDs, 64 words apart).
constexpr uint32_t kSchedOffset = 64;
// Word of Workspace::counts holding the iteration number of the next
// syndrome pass (IterArgs::itc).
constexpr uint32_t kItcWord = 2;
constexpr uint32_t kCountsWords = kSchedOffset + 16 * 64;

// Arguments of one iteration of an HBM path.
struct IterArgs {
    const float *llr;
    uint32_t frames;
    uint32_t it;                     // iteration within the current phase (1 = first)
    uint32_t cur;                    // parity of the running-frame list to read
    uint32_t it_report;              // iteration count reported for frames retiring now
    int last;                        // retire frames that still fail the syndrome
    int fb_phase;                    // value written to fb[] for retiring frames
    const float *w_prev, *w_cur;     // per-block-edge weights (NULL = plain min-sum)
    int spa;                         // sum-product check rule (STREAM path, no weights)
    // device iteration word (may be NULL): a direct launch sets it to
    // it_report + 1 for the next iteration; an iteration replayed from a
    // captured graph (itc_graph) takes its number from it and advances it,
    // with last = (number >= itc_last), fb_phase = (number > itc_fb_from)
    uint32_t *itc;
    int itc_graph;
    uint32_t itc_last, itc_fb_from;
    uint64_t *msg;
    uint32_t *iters;
    uint8_t *conv, *synd, *fb;
};

// Arguments of the layered learned decoder's forward kernel (layered.hip).
struct LayeredArgs {
    CodeView cv;
    const float *llr;        // [F][n] channel LLRs
    const float *w;          // [T][nbe] weights, block edges in row order
    uint32_t F, T;           // frames, layered iterations
    uint32_t X;              // block-columns whose beliefs live in LDS (0..X-1)
    int fallback;            // frames failing after T: 1 -> continuation list, 0 -> retire
    uint32_t *state;         // [F][mb][zp][3] compressed check states
    const uint3 *zero_state; // staircase kernel: [mb][zp] zero states, read in place of iteration 1's
    float *belief;           // [F][nb][zp] beliefs of block-columns >= X (all at hand-over)
    uint32_t *ctr;           // [0] next frame, [1] layered frame-iterations run (zeroed by the host)
    uint32_t *list, *count;  // running list of the continuation and its length
    uint64_t *msg;
    uint32_t *iters;
    uint8_t *conv, *synd, *fb;
    uint32_t *tape_state;    // training: [T][F][mb][zp] uint4 (min1, min2, meta, idx2 | argmin input signs)
    float *tape_bel;         // training: [T][F][nb][zp] beliefs after each iteration
    uint64_t stair_deg[4];   // staircase kernel: bit r set when block-row r has the higher of its two degrees
    uint64_t stair_sync[4];  // staircase kernel: bit r set when a barrier must follow block-row r
};

// Arguments of the layered backward kernel (layered.hip).
struct LayeredBwdArgs {
    CodeView cv;
    const float *w;           // [T][nbe]
    uint32_t F, T, X;
    const uint32_t *tape_state;
    const float *tape_bel;
    const uint64_t *cw;       // [F][ceil(n/64)] codewords
    float *g_glob;            // [F][nb][zp] dL/db of block-columns >= X
    float *H;                 // [F][nbe][z] dL/d(message) carried to the previous iteration
    float *gpart;             // [T][nbe][F] per-frame weight gradients
    float scale;              // 1 / (T n F)
    float *lpart;             // [F] loss of each frame
    uint32_t *ctr;            // [0] next frame (zeroed by the host)
};

}  // namespace gf2l

struct gf2l_code {
    bool is_csr = false;         // built by gf2l_code_edges (general sparse path)
    gf2l::CsrView csr{};
    gf2l::CodeView view;         // device pointers
    std::vector<int32_t> shifts; // host copy of the base matrix
    std::vector<uint32_t> h_row_ptr, h_col_ptr;
    std::vector<uint2> h_row_edge, h_col_edge;
    std::vector<uint4> h_col_info;
    void *d_tables = nullptr;    // one allocation backing the view's tables
    int device = 0;
};

struct gf2l_nms;

struct gf2l_ctx {
    int device = 0;
    gf2l::Profile prof;
    hipStream_t stream = nullptr;
    int num_cus = 0;
    int path = 0;
    int rule = 0;                // GF2L_RULE_*: check update of the plain decoder
    gf2l::Tuning tune;
    size_t lds_limit = 0;        // bytes of dynamic LDS one workgroup may use
    std::string err;
    gf2l::Workspace ws;
    // staging for the host-pointer entry point
    void *d_io = nullptr;
    size_t d_io_bytes = 0;
    // plain iterations of the HBM paths captured as a graph, and what they
    // were captured for (code, sizes, buffers, launch shape)
    hipGraphExec_t hbm_exec = nullptr;
    std::vector<uintptr_t> hbm_key;
};

namespace gf2l {

// Records `msg` as the context's last error and returns `code`.
inline int fail(gf2l_ctx *ctx, int code, const std::string &msg) {
    if (ctx) {
        try {
            ctx->err = msg;
        } catch (...) {
            ctx->err.clear();
        }
    }
    return code;
}

// Runs f() and turns any C++ exception into an error code, so that none
// crosses the C ABI: std::bad_alloc -> GF2L_ERR_NOMEM, anything else ->
// GF2L_ERR_INTERNAL.  Every extern "C" entry point that allocates host
// memory (vectors, strings, new) runs its body through this.
template <typename F>
int guard(gf2l_ctx *ctx, F &&f) noexcept {
    auto note = [ctx](const char *msg) noexcept {
        if (!ctx) return;
        try {
            ctx->err = msg;
        } catch (...) {
            ctx->err.clear();
        }
    };
    try {
        return f();
    } catch (const std::bad_alloc &) {
        note("host allocation failed");
        return -3;  // GF2L_ERR_NOMEM
    } catch (const std::exception &e) {
        note(e.what());
        return -5;  // GF2L_ERR_INTERNAL
    } catch (...) {
        note("unknown C++ exception");
        return -5;
    }
}

// Returns from the enclosing function with a GF2L error when a HIP call fails.
#define GF2L_HIP(ctx, expr)                                                                        \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return ::gf2l::fail((ctx), e_ == hipErrorOutOfMemory ? -3 : -2,                        \
                                std::string(#expr) + ": " + hipGetErrorString(e_));               \
    } while (0)

// Decoder driver (gf2ldpc_api.hip): plain min-sum for max_it iterations, or
// with weights d_w ([max_it][nbe]) max_it weighted iterations followed by up
// to fb_it plain iterations for frames still failing the syndrome, restarted
// from their LLRs or, with fb_continue, continued from the weighted state
// (the first plain check update subtracts the last weighted messages).
// Device buffers; may synchronize the context's stream.
int decode_device(gf2l_ctx *ctx, const gf2l_code *code, const float *d_llr, size_t frames,
                  uint32_t max_it, const float *d_w, uint32_t fb_it, bool fb_continue, uint64_t *d_msg,
                  uint32_t *d_iters, uint8_t *d_conv, uint8_t *d_synd, uint8_t *d_fb,
                  float *h_trace_c2v, float *h_trace_bel, uint32_t trace_iters);

// Iteration loop of the HBM paths (STATE / STREAM) over a workspace the
// caller has prepared (gf2ldpc_api.hip); allocates nothing.  Used by
// decode_device after growing the context's workspace and by the
// no-allocation entry point over its static scratch array.
int run_hbm(gf2l_ctx *ctx, const CodeView &cv, int path, Workspace &ws, const float *d_llr, uint32_t F,
            uint32_t max_it, const float *d_w, uint32_t fb_it, bool fb_continue, uint64_t *d_msg,
            uint32_t *d_iters, uint8_t *d_conv, uint8_t *d_synd, uint8_t *d_fb, float *h_trace_c2v,
            float *h_trace_bel, uint32_t trace_iters, float *d_tc, float *d_tb, bool graphs = true);
// Decoder path (GF2L_PATH_*) the context would run a quasi-cyclic code on,
// or -1 when a forced path cannot run it.
int resolve_path(const gf2l_ctx *ctx, const CodeView &cv);

// Uploads the tables of a quasi-cyclic code whose block-row lists
// (h_row_ptr, h_row_edge) are filled and sets its view.  dvb_q > 0: DVB
// order (each block-column's edges by ascending standard check index of its
// lane 0; block-column kb+q-1's edge in block-row 0 has no lane-0 edge).
int build_qc_view(gf2l_ctx *ctx, gf2l_code *c, uint32_t mb, uint32_t nb, uint32_t z, uint32_t dvb_q);
// LLRs [frames][n] in the standard's order -> quasi-cyclic order of a DVB code.
hipError_t launch_dvb_permute(const CodeView &cv, const float *llr, float *out, uint32_t frames, hipStream_t s);

// Grows the context's workspace to at least these element counts per buffer
// (c2v floats, state words, beliefs, hard-decision words, permuted LLRs) and
// `frames` running-list entries (gf2ldpc_api.hip).
int ensure_buffers(gf2l_ctx *ctx, size_t frames, size_t c2v, size_t state, size_t bel, size_t hb, size_t llr);

// Weights [iterations][nbe] (device), weighted iterations and code of a
// learned decoder, and whether it is a flooding decoder whose fallback
// restarts from the LLRs (nms.hip).
const float *nms_weights(const gf2l_nms *q, uint32_t *iterations, const gf2l_code **code, bool *layered,
                         bool *fb_continue);

// Layered learned decoder (layered.hip): LDS bytes of a workgroup holding
// X block-columns of beliefs, the largest X that fits lds_limit, and the
// forward (tape: the training forward, no early stop, no outputs) and
// backward launchers.
size_t layered_lds_bytes(const CodeView &cv, uint32_t X);
uint32_t layered_lds_columns(const CodeView &cv, size_t lds_limit);
hipError_t launch_layered_fwd(const LayeredArgs &a, bool tape, int num_cus, hipStream_t s);
hipError_t launch_layered_bwd(const LayeredBwdArgs &a, int num_cus, hipStream_t s);
// Staircase kernel (decoding only, X = kb information columns in LDS): its
// LDS bytes, the row degree of its instantiation for a code (0: the code
// lacks the structure; deg_mask and sync_mask, when given, receive
// LayeredArgs::stair_deg and stair_sync) and its launcher.
size_t layered_stair_bytes(const CodeView &cv);
uint32_t layered_stair_degree(const gf2l_code *code, uint64_t *deg_mask = nullptr, uint64_t *sync_mask = nullptr);
hipError_t launch_layered_stair(const LayeredArgs &a, uint32_t dh, int num_cus, hipStream_t s);

// Plain min-sum on a general sparse code (csr.hip).
int decode_csr(gf2l_ctx *ctx, const gf2l_code *code, const float *d_llr, uint32_t F, uint32_t max_it,
               uint64_t *d_msg, uint32_t *d_iters, uint8_t *d_conv, uint8_t *d_synd);

// Thread layout of the LDS path's workgroup: `groups` groups of `lanes`
// threads; group g updates block-rows / block-columns g, g + groups, ...
struct LdsShape {
    uint32_t groups, lanes;
};
// Float offset of the code tables in the LDS path's dynamic LDS: after the
// messages, the beliefs and the LLRs, 8-byte aligned.
__host__ __device__ inline uint32_t lds_table_offset(const CodeView &cv) {
    const uint32_t f = cv.nbe * cv.z + 2u * cv.n;
    return (f + 1u) & ~1u;
}

// Kernel launchers (gf2ldpc_kernels.hip).  All enqueue on `s`.
// Dynamic LDS of the LDS path (frame state and code tables).
size_t lds_path_bytes(const CodeView &cv);
// Phase 0: max_it iterations, weighted by w[it-1][e] if w != NULL; then, for
// frames still failing the syndrome, fb_it plain iterations from scratch or,
// with fb_continue, from the phase-0 state.  groups: thread groups per
// workgroup (0: automatic, else at most 512 / min(z, 512)).
hipError_t launch_lds_decode(const CodeView &cv, const float *llr, uint32_t frames, uint32_t max_it,
                             const float *w, uint32_t fb_it, bool fb_continue, uint64_t *msg, uint32_t *iters,
                             uint8_t *conv, uint8_t *synd, uint8_t *fb, float *trace_c2v,
                             float *trace_bel, uint32_t trace_iters, int groups, hipStream_t s);
hipError_t launch_hbm_init(const CodeView &cv, Workspace &ws, int path, const float *llr,
                           uint32_t frames, hipStream_t s);
hipError_t launch_hbm_iteration(const CodeView &cv, Workspace &ws, int path, const IterArgs &a,
                                int num_cus, const Tuning &tune, hipStream_t s, hipEvent_t *ev);
// One forward iteration of the state kernels on explicit buffers (no
// syndrome, no hard decisions): used by the training forward pass.
hipError_t launch_state_forward(const CodeView &cv, const uint32_t *list, const uint32_t *count,
                                uint32_t frames, bool first, const uint32_t *state_in,
                                uint32_t *state_out, const float *belief_in, float *belief_out,
                                const float *llr, const float *w_prev, const float *w_cur,
                                int num_cus, const Tuning &tune, hipStream_t s);
// Frame generator (framegen.hip).
hipError_t launch_framegen(const CodeView &cv, uint64_t seed, uint64_t first_frame, uint32_t frames,
                           float ebn0_lo, float ebn0_hi, float *llr, uint64_t *cw_words, int num_cus,
                           hipStream_t s);
// All-zero codeword frames of an n-bit, rate k/n code (framegen.hip).
hipError_t launch_framegen_zero(uint32_t n, uint32_t k, uint64_t seed, uint64_t first_frame, uint32_t frames,
                                float ebn0_lo, float ebn0_hi, float *llr, int num_cus, hipStream_t s);
// Expands the state path's compressed checks and padded beliefs into the
// explicit trace layout ([f][e*z+i] and [f][n]) for all `frames` frames.
hipError_t launch_state_expand(const CodeView &cv, const Workspace &ws, uint32_t frames,
                               const float *w, float *c2v_out, float *bel_out, hipStream_t s);

}  // namespace gf2l
