// Code ingestion: DVB-T2 codes built from the ETSI address tables and
// parity-check matrices read from GF2DATA files.
//
//   DVB-T2   LdpcCode::dvb_t2_short / dvb_t2_normal (crates/gf2-coding/src/
//            ldpc/core.rs:197-255) -> DvbParams::for_code (dvb_t2/params.rs:
//            56-86) -> build_dvb_edges (dvb_t2/builder.rs:15-118) ->
//            LdpcCode::from_edges.  Info bit i*360 + j of table row i and
//            address p connects to check (p + j*q) mod m; the parity part is
//            the bit-level staircase (check p -> parity bits p and p-1, no
//            wrap at p = 0).  Renumbered (check a + q*b -> lane b of
//            block-row a) these codes are quasi-cyclic and run on the state
//            kernels; a table outside that form runs on the CSR path.
//   Sizes and indices from the caller or the file are validated by the
//   host-only code of host_codes.cpp before anything is allocated.
//   GF2DATA  crates/gf2-core/src/io/mod.rs:6-98 (32-byte header, JSON
//            metadata, little-endian payload), header.rs:33-86,
//            sparse.rs:151-200 (type 3, COO pairs, set semantics on load)
//            and sparse.rs:397-449 (type 4, CSR + CSC).  A code is
//            LdpcCode::from_edges over the matrix's (row, col) entries, the
//            only way the reference turns a loaded matrix into a code.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "dvb_t2_tables.h"
#include "gf2ldpc.h"
#include "host_codes.h"
#include "internal.h"

namespace {

constexpr uint32_t kDvbZ = 360;

struct DvbShape {
    uint32_t n, k, q;
};

// DvbParams::for_code (params.rs:59-73): [frame][rate]
constexpr DvbShape kDvbShapes[2][6] = {
    {{16200, 7200, 25}, {16200, 9720, 18}, {16200, 10800, 15},
     {16200, 11880, 12}, {16200, 12600, 10}, {16200, 13320, 8}},
    {{64800, 32400, 90}, {64800, 38880, 72}, {64800, 43200, 60},
     {64800, 48600, 45}, {64800, 51840, 36}, {64800, 54000, 30}},
};

bool dvb_index_ok(int frame, int rate) { return frame >= 0 && frame < 2 && rate >= 0 && rate < 6; }

int invalid(gf2l_ctx *ctx, const std::string &msg) { return gf2l::fail(ctx, GF2L_ERR_INVALID, msg); }

}  // namespace

extern "C" int gf2l_dvb_t2_params(int frame, int rate, uint32_t *n, uint32_t *k, uint32_t *q, uint32_t *rows,
                                  uint32_t *entries) {
    if (!dvb_index_ok(frame, rate)) return GF2L_ERR_INVALID;
    const DvbShape &s = kDvbShapes[frame][rate];
    const gf2l::dvb_tables::Table &t = gf2l::dvb_tables::kTables[frame][rate];
    uint32_t total = 0;
    for (uint32_t r = 0; r < t.rows; ++r) total += t.row_len[r];
    if (n) *n = s.n;
    if (k) *k = s.k;
    if (q) *q = s.q;
    if (rows) *rows = t.rows;
    if (entries) *entries = total;
    return GF2L_OK;
}

extern "C" int gf2l_dvb_t2_table(int frame, int rate, uint32_t *row_len, uint32_t *addresses) {
    if (!dvb_index_ok(frame, rate)) return GF2L_ERR_INVALID;
    const gf2l::dvb_tables::Table &t = gf2l::dvb_tables::kTables[frame][rate];
    uint32_t at = 0;
    for (uint32_t r = 0; r < t.rows; ++r) {
        if (row_len) row_len[r] = t.row_len[r];
        for (uint32_t i = 0; i < t.row_len[r]; ++i, ++at)
            if (addresses) addresses[at] = t.addr[at];
    }
    return GF2L_OK;
}

extern "C" int gf2l_code_dvb_table(gf2l_ctx *ctx, uint32_t n, uint32_t k, uint32_t q, uint32_t z, uint32_t rows,
                                   const uint32_t *row_len, const uint32_t *addresses, gf2l_code **out) {
    if (!ctx || !out) return GF2L_ERR_INVALID;
    *out = nullptr;
    return gf2l::guard(ctx, [&]() -> int {
        // validate_table (builder.rs:15-41) and the size limits, before any
        // allocation sized from the table
        size_t total = 0;
        std::string err;
        const int vrc = gf2l::host::check_dvb_table(n, k, z, rows, row_len, addresses, total, err);
        if (vrc != GF2L_OK) return gf2l::fail(ctx, vrc, err);
        const uint32_t m = n - k;
        // Quasi-cyclic form (the decoder's fast path): check c = a + q*b and
        // parity bit k + a + q*b become lane b of block-row / parity block-column
        // a.  Table address p of row i is then a circulant in block (p mod q, i)
        // with shift (z - p div q) mod z; the staircase gives identity blocks
        // (a, kb+a) and (a, kb+a-1), and block (0, kb+q-1) with shift z-1 that
        // lacks its lane-0 edge (check 0 has one parity bit).  Needs m = q*z,
        // k = rows*z, and no address repeated within a row (a repeat cancels in
        // from_coo); otherwise the code runs on the CSR path.
        bool qc = m == (uint64_t)q * z && k == (uint64_t)rows * z && q >= 2 && z >= 2;
        for (uint32_t r = 0, at = 0; qc && r < rows; at += row_len[r], ++r) {
            std::vector<uint32_t> v(addresses + at, addresses + at + row_len[r]);
            std::sort(v.begin(), v.end());
            qc = std::adjacent_find(v.begin(), v.end()) == v.end();
        }
        if (qc) {
            const uint32_t kb = rows, nb = kb + q;
            std::vector<std::vector<uint2>> per_row(q);
            for (uint32_t r = 0, at = 0; r < rows; ++r)
                for (uint32_t i = 0; i < row_len[r]; ++i, ++at) {
                    const uint32_t p = addresses[at];
                    per_row[p % q].push_back(make_uint2(r, (z - p / q) % z));
                }
            gf2l_code *c = new gf2l_code;
            c->device = ctx->device;
            c->h_row_ptr.assign(q + 1, 0);
            uint32_t max_dc = 0;
            for (uint32_t a = 0; a < q; ++a) {
                auto &v = per_row[a];
                std::stable_sort(v.begin(), v.end(), [](const uint2 &x, const uint2 &y) {
                    return x.x != y.x ? x.x < y.x : x.y < y.y;
                });
                if (a == 0) {
                    v.push_back(make_uint2(kb + q - 1, z - 1));
                    v.push_back(make_uint2(kb, 0));
                } else {
                    v.push_back(make_uint2(kb + a - 1, 0));
                    v.push_back(make_uint2(kb + a, 0));
                }
                c->h_row_edge.insert(c->h_row_edge.end(), v.begin(), v.end());
                c->h_row_ptr[a + 1] = (uint32_t)c->h_row_edge.size();
                max_dc = std::max<uint32_t>(max_dc, (uint32_t)v.size());
            }
            if (max_dc <= gf2l::kMaxStateDegree) {
                if (hipSetDevice(ctx->device) != hipSuccess) {
                    delete c;
                    return gf2l::fail(ctx, GF2L_ERR_HIP, "hipSetDevice failed");
                }
                const int rc = gf2l::build_qc_view(ctx, c, q, nb, z, q);
                if (rc == GF2L_OK) *out = c;
                return rc;
            }
            delete c;
        }
        // build_dvb_edges (builder.rs:74-118): table edges, then the staircase
        std::vector<uint32_t> rc;
        rc.reserve(2 * (total * z + 2ull * m));
        for (uint32_t r = 0, at = 0; r < rows; ++r)
            for (uint32_t i = 0; i < row_len[r]; ++i, ++at)
                for (uint32_t j = 0; j < z; ++j) {
                    rc.push_back((uint32_t)((addresses[at] + (uint64_t)j * q) % m));
                    rc.push_back(r * z + j);
                }
        for (uint32_t p = 0; p < m; ++p) {
            rc.push_back(p);
            rc.push_back(k + p);
            if (p > 0) {
                rc.push_back(p);
                rc.push_back(k + p - 1);
            }
        }
        return gf2l_code_edges(ctx, m, n, rc.data(), rc.size() / 2, out);
    });
}

extern "C" int gf2l_code_dvb_t2(gf2l_ctx *ctx, int frame, int rate, gf2l_code **out) {
    if (!ctx || !out) return GF2L_ERR_INVALID;
    *out = nullptr;
    return gf2l::guard(ctx, [&]() -> int {
        if (!dvb_index_ok(frame, rate))
            return invalid(ctx, "DVB-T2 frame size must be 0 (short) or 1 (normal) and rate 0..5");
        const DvbShape &s = kDvbShapes[frame][rate];
        const gf2l::dvb_tables::Table &t = gf2l::dvb_tables::kTables[frame][rate];
        std::vector<uint32_t> len(t.rows), addr;
        for (uint32_t r = 0, at = 0; r < t.rows; ++r) {
            len[r] = t.row_len[r];
            for (uint32_t i = 0; i < len[r]; ++i, ++at) addr.push_back(t.addr[at]);
        }
        return gf2l_code_dvb_table(ctx, s.n, s.k, s.q, kDvbZ, t.rows, len.data(), addr.data(), out);
    });
}

extern "C" int gf2l_code_gf2data(gf2l_ctx *ctx, const void *bytes, size_t len, gf2l_code **out) {
    if (!ctx || !out) return GF2L_ERR_INVALID;
    *out = nullptr;
    return gf2l::guard(ctx, [&]() -> int {
        uint32_t rows = 0, cols = 0;
        std::vector<uint32_t> rc;
        std::string err;
        const int prc = gf2l::host::parse_gf2data(bytes, len, rows, cols, rc, err);
        if (prc != GF2L_OK) return gf2l::fail(ctx, prc, err);
        return gf2l_code_edges(ctx, rows, cols, rc.empty() ? nullptr : rc.data(), rc.size() / 2, out);
    });
}
