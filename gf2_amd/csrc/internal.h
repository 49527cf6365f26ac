// Internal types shared by the host code and the HIP kernels of libgf2ldpc.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace gf2l {

// Largest check degree the kernels accept; per-check sign bits live in one
// 64-bit mask.
constexpr uint32_t kMaxCheckDegree = 32;

// Device view of a quasi-cyclic code: base-matrix block edges in two orders.
//   row order  e = 0..nbe-1: block-rows ascending, block-columns ascending
//              within a row -- the order of SpBitMatrix::row_iter for every
//              lifted check of the block-row.
//   col order  block-columns ascending, block-rows ascending within a
//              column -- the order of the CSC col_iter for every lifted
//              variable, which fixes the belief summation order.
// Message arrays are indexed [e][i], i = check index inside the block-row.
struct CodeView {
    uint32_t mb, nb, z, n, m, k, nbe;
    uint32_t max_dc, max_dv;
    uint32_t zw;                 // ceil(z/64): 64-bit words per block-column of hard bits
    const uint32_t *row_ptr;     // mb+1, into row-order edges
    const uint2 *row_edge;       // nbe: {bc, shift}
    const uint32_t *col_ptr;     // nb+1, into col-order list
    const uint2 *col_edge;       // nbe: {row-order edge id, shift}
};

// Per-batch device state of the streaming path.
struct Workspace {
    size_t cap_frames = 0;
    size_t cap_e = 0, cap_n = 0, cap_hb = 0, cap_kw = 0;
    float *c2v = nullptr;        // [frames][nbe*z]
    float *belief = nullptr;     // [frames][n]
    uint64_t *hb = nullptr;      // [frames][nb][zw]
    uint32_t *lists = nullptr;   // [2][frames] active frame lists
    uint32_t *counts = nullptr;  // [2]
    uint32_t *h_count = nullptr; // pinned host mirror of one count
};

// Accumulated per-kernel timings of the streaming path (HIP events recorded
// around each launch on the context's stream).
struct Profile {
    bool on = false;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    double ms[3] = {0, 0, 0};         // check, variable, syndrome kernels
    uint64_t launches[3] = {0, 0, 0};
    uint64_t frame_iters = 0;         // sum over launches of frames processed
};

}  // namespace gf2l

struct gf2l_code {
    gf2l::CodeView view;         // device pointers
    std::vector<int32_t> shifts; // host copy of the base matrix
    std::vector<uint32_t> h_row_ptr, h_col_ptr;
    std::vector<uint2> h_row_edge, h_col_edge;
    void *d_tables = nullptr;    // one allocation backing the view's tables
    int device = 0;
};

struct gf2l_ctx {
    int device = 0;
    gf2l::Profile prof;
    hipStream_t stream = nullptr;
    int num_cus = 0;
    int path = 0;
    size_t lds_limit = 0;        // bytes of dynamic LDS one workgroup may use
    std::string err;
    gf2l::Workspace ws;
    // staging for the host-pointer entry point
    void *d_io = nullptr;
    size_t d_io_bytes = 0;
};

namespace gf2l {

// Kernel launchers (gf2ldpc_kernels.hip).  All enqueue on `s`.
size_t lds_path_bytes(const CodeView &cv);
hipError_t launch_lds_decode(const CodeView &cv, const float *llr, uint32_t frames, uint32_t max_it,
                             const float *w, uint32_t w_iters, uint64_t *msg, uint32_t *iters,
                             uint8_t *conv, uint8_t *synd, float *trace_c2v, float *trace_bel,
                             uint32_t trace_iters, hipStream_t s);
hipError_t launch_stream_init(const CodeView &cv, Workspace &ws, const float *llr, uint32_t frames,
                              hipStream_t s);
hipError_t launch_stream_iteration(const CodeView &cv, Workspace &ws, const float *llr,
                                   uint32_t frames, uint32_t it, uint32_t max_it, const float *w_it,
                                   uint64_t *msg, uint32_t *iters, uint8_t *conv, uint8_t *synd,
                                   int num_cus, hipStream_t s, hipEvent_t *ev);

}  // namespace gf2l
