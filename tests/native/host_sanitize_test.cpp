// CPU test of the library's host-side input handling (gf2_amd/csrc/
// host_codes.cpp), built with AddressSanitizer and UndefinedBehaviorSanitizer
// by tests/native/Makefile and run by tests/test_native.py.  It feeds hostile
// sizes (rows or cols = 2^32 - 1 in GF2DATA headers and edge lists), truncated
// and mutated files, and checks that every call returns an error code instead
// of allocating from or indexing with the hostile value.  Exit status 0 when
// every expectation holds; a sanitizer report aborts the run.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "gf2ldpc.h"
#include "host_codes.h"

using gf2l::host::build_csr;
using gf2l::host::check_dvb_table;
using gf2l::host::Csr;
using gf2l::host::parse_gf2data;

static int failures = 0;

#define EXPECT(cond, what)                                                   \
    do {                                                                     \
        if (!(cond)) {                                                       \
            std::fprintf(stderr, "FAIL line %d: %s\n", __LINE__, what);      \
            ++failures;                                                      \
        }                                                                    \
    } while (0)

static void put32(std::vector<uint8_t> &b, size_t at, uint32_t v) {
    for (int i = 0; i < 4; ++i) b[at + i] = (uint8_t)(v >> (8 * i));
}
static void put64(std::vector<uint8_t> &b, size_t at, uint64_t v) {
    for (int i = 0; i < 8; ++i) b[at + i] = (uint8_t)(v >> (8 * i));
}

// A GF2DATA image: 32-byte header, JSON metadata, payload of u32 words.
static std::vector<uint8_t> gf2data(uint8_t type, const std::string &meta, const std::vector<uint32_t> &payload,
                                    int64_t data_len_override = -1) {
    std::vector<uint8_t> b(32 + meta.size() + 4 * payload.size(), 0);
    std::memcpy(b.data(), "GF2DATA\0", 8);
    b[8] = 1;  // version 1, little-endian u16
    b[10] = type;
    put32(b, 20, (uint32_t)meta.size());
    put64(b, 24, data_len_override >= 0 ? (uint64_t)data_len_override : 4ull * payload.size());
    std::memcpy(b.data() + 32, meta.data(), meta.size());
    for (size_t i = 0; i < payload.size(); ++i) put32(b, 32 + meta.size() + 4 * i, payload[i]);
    return b;
}

static std::string coo_meta(uint64_t rows, uint64_t cols, uint64_t nnz) {
    return "{\"type\":\"SpBitMatrix\",\"rows\":" + std::to_string(rows) + ",\"cols\":" + std::to_string(cols) +
           ",\"nnz\":" + std::to_string(nnz) + ",\"format\":\"coo\"}";
}
static std::string dual_meta(uint64_t rows, uint64_t cols, uint64_t nnz) {
    return "{\"type\":\"SpBitMatrixDual\",\"rows\":" + std::to_string(rows) + ",\"cols\":" + std::to_string(cols) +
           ",\"nnz\":" + std::to_string(nnz) + "}";
}

static void test_build_csr() {
    Csr c;
    std::string err;
    // from_coo: duplicates cancel in pairs, rows sorted (sparse.rs:166-204)
    const uint32_t rc[] = {1, 2, 0, 3, 1, 0, 1, 2, 1, 2, 0, 1};
    EXPECT(build_csr(2, 4, rc, 6, c, err) == GF2L_OK, "small from_coo");
    const std::vector<uint32_t> rp = {0, 2, 4}, ci = {1, 3, 0, 2};
    EXPECT(c.row_ptr == rp && c.col_idx == ci, "from_coo rows and columns");
    const std::vector<uint32_t> cp = {0, 1, 2, 3, 4}, ce = {2, 0, 3, 1};
    EXPECT(c.col_ptr == cp && c.csc_edge == ce, "stable transpose");
    EXPECT(c.max_dc == 2, "max check degree");
    // an empty matrix and an empty edge list
    EXPECT(build_csr(0, 0, nullptr, 0, c, err) == GF2L_OK && c.row_ptr.size() == 1, "empty matrix");
    // hostile sizes: no allocation from m or n, an error code instead
    const uint32_t one[] = {0, 0};
    EXPECT(build_csr(0xffffffffu, 0xffffffffu, one, 1, c, err) == GF2L_ERR_UNSUPPORTED, "m = n = 2^32-1");
    EXPECT(build_csr(4, 0xffffffffu, one, 1, c, err) == GF2L_ERR_UNSUPPORTED, "n = 2^32-1");
    EXPECT(build_csr(0xffffffffu, 4, one, 1, c, err) == GF2L_ERR_UNSUPPORTED, "m = 2^32-1");
    EXPECT(build_csr((1u << 28) + 1, 4, one, 1, c, err) == GF2L_ERR_UNSUPPORTED, "m = 2^28+1");
    // out-of-range entries: the reference's panic texts
    const uint32_t bad_r[] = {5, 0}, bad_c[] = {0, 9};
    EXPECT(build_csr(4, 4, bad_r, 1, c, err) == GF2L_ERR_INVALID && err.find("row index 5") == 0, "row range");
    EXPECT(build_csr(4, 4, bad_c, 1, c, err) == GF2L_ERR_INVALID && err.find("col index 9") == 0, "col range");
    EXPECT(build_csr(4, 4, nullptr, 3, c, err) == GF2L_ERR_INVALID, "null edge list");
}

static void test_gf2data() {
    uint32_t rows = 0, cols = 0;
    std::vector<uint32_t> rc;
    std::string err;
    // a valid type-3 file: set semantics on load
    auto ok3 = gf2data(3, coo_meta(2, 3, 3), {0, 1, 1, 2, 0, 1});
    EXPECT(parse_gf2data(ok3.data(), ok3.size(), rows, cols, rc, err) == GF2L_OK, "valid type 3");
    EXPECT(rows == 2 && cols == 3 && rc == std::vector<uint32_t>({0, 1, 1, 2}), "type 3 entries kept once");
    // a valid type-4 file (CSR + CSC of [[0,1,0],[0,0,1]])
    auto ok4 = gf2data(4, dual_meta(2, 3, 2), {0, 1, 2, 1, 2, 0, 0, 1, 2, 0, 1});
    EXPECT(parse_gf2data(ok4.data(), ok4.size(), rows, cols, rc, err) == GF2L_OK, "valid type 4");
    EXPECT(rc == std::vector<uint32_t>({0, 1, 1, 2}), "type 4 CSR entries");
    // hostile headers: rows and / or cols = 2^32-1 with one valid entry
    const uint64_t big = 0xffffffffull;
    auto h1 = gf2data(3, coo_meta(big, big, 1), {0, 0});
    EXPECT(parse_gf2data(h1.data(), h1.size(), rows, cols, rc, err) == GF2L_ERR_UNSUPPORTED, "type 3 rows=cols=2^32-1");
    auto h2 = gf2data(3, coo_meta(4, big, 1), {0, 0});
    EXPECT(parse_gf2data(h2.data(), h2.size(), rows, cols, rc, err) == GF2L_ERR_UNSUPPORTED, "type 3 cols=2^32-1");
    auto h3 = gf2data(3, coo_meta(big, 4, 1), {0, 0});
    EXPECT(parse_gf2data(h3.data(), h3.size(), rows, cols, rc, err) == GF2L_ERR_UNSUPPORTED, "type 3 rows=2^32-1");
    auto h4 = gf2data(3, coo_meta(big + 1, 4, 1), {0, 0});
    EXPECT(parse_gf2data(h4.data(), h4.size(), rows, cols, rc, err) == GF2L_ERR_UNSUPPORTED, "rows = 2^32");
    // type 4 with rows = 2^32-1: the payload cannot be in the buffer
    auto h5 = gf2data(4, dual_meta(big, 4, 1), {0, 1, 0, 0, 1, 0, 0});
    EXPECT(parse_gf2data(h5.data(), h5.size(), rows, cols, rc, err) == GF2L_ERR_INVALID, "type 4 rows=2^32-1");
    auto h6 = gf2data(4, dual_meta(2, big, 2), {0, 1, 2, 1, 2, 0, 0, 1, 2, 0, 1},
                      4ll * (2 + 1 + 2 + (int64_t)big + 1 + 2));
    EXPECT(parse_gf2data(h6.data(), h6.size(), rows, cols, rc, err) == GF2L_ERR_INVALID, "type 4 cols=2^32-1");
    // nnz = 2^32-1 with a short payload
    auto h7 = gf2data(3, coo_meta(4, 4, big), {0, 0}, (int64_t)(8 * big));
    EXPECT(parse_gf2data(h7.data(), h7.size(), rows, cols, rc, err) == GF2L_ERR_INVALID, "nnz=2^32-1 truncated");
    // metadata length past the end, truncated header, bad magic / version / tag
    auto m1 = ok3;
    put32(m1, 20, 0xffffffffu);
    EXPECT(parse_gf2data(m1.data(), m1.size(), rows, cols, rc, err) == GF2L_ERR_INVALID, "meta past end");
    EXPECT(parse_gf2data(ok3.data(), 31, rows, cols, rc, err) == GF2L_ERR_INVALID, "short header");
    EXPECT(parse_gf2data(ok3.data(), ok3.size() - 1, rows, cols, rc, err) == GF2L_ERR_INVALID, "short payload");
    auto m2 = ok3;
    m2[0] = 'X';
    EXPECT(parse_gf2data(m2.data(), m2.size(), rows, cols, rc, err) == GF2L_ERR_INVALID &&
               err == "Invalid magic bytes", "magic");
    auto m3 = ok3;
    m3[8] = 2;
    EXPECT(parse_gf2data(m3.data(), m3.size(), rows, cols, rc, err) == GF2L_ERR_INVALID &&
               err == "Unsupported format version: 2", "version");
    auto m4 = ok3;
    m4[10] = 9;
    EXPECT(parse_gf2data(m4.data(), m4.size(), rows, cols, rc, err) == GF2L_ERR_INVALID &&
               err == "Unknown type tag: 9", "type tag");
    // type 4 with offsets that run past nnz
    auto m5 = gf2data(4, dual_meta(2, 3, 2), {0, 7, 2, 1, 2, 0, 0, 1, 2, 0, 1});
    EXPECT(parse_gf2data(m5.data(), m5.size(), rows, cols, rc, err) == GF2L_ERR_INVALID, "offsets past nnz");
    EXPECT(parse_gf2data(nullptr, 100, rows, cols, rc, err) == GF2L_ERR_INVALID, "null buffer");
    // mutated files: any result is acceptable, a memory error is not
    std::mt19937 rng(12345);
    const std::vector<std::vector<uint8_t>> seeds = {ok3, ok4};
    for (int t = 0; t < 20000; ++t) {
        auto b = seeds[t & 1];
        const int flips = 1 + (int)(rng() % 4);
        for (int f = 0; f < flips; ++f) b[rng() % b.size()] ^= (uint8_t)(1u << (rng() % 8));
        const size_t len = (t % 7 == 0) ? rng() % (b.size() + 1) : b.size();
        const int r = parse_gf2data(b.data(), len, rows, cols, rc, err);
        EXPECT(r == GF2L_OK || r == GF2L_ERR_INVALID || r == GF2L_ERR_UNSUPPORTED, "mutated file result");
        if (r == GF2L_OK) {
            Csr c;
            std::string e2;
            const int r2 = build_csr(rows, cols, rc.data(), rc.size() / 2, c, e2);
            EXPECT(r2 == GF2L_OK || r2 == GF2L_ERR_UNSUPPORTED, "mutated file -> code");
        }
    }
}

static void test_dvb_table() {
    size_t total = 0;
    std::string err;
    const uint32_t len2[] = {2, 2}, addr[] = {0, 3, 1, 2};
    EXPECT(check_dvb_table(16, 8, 4, 2, len2, addr, total, err) == GF2L_OK && total == 4, "valid table");
    const uint32_t addr_bad[] = {0, 3, 1, 8};
    EXPECT(check_dvb_table(16, 8, 4, 2, len2, addr_bad, total, err) == GF2L_ERR_INVALID &&
               err.find("Invalid parity index 8") == 0, "address range");
    EXPECT(check_dvb_table(16, 8, 4, 3, len2, addr, total, err) == GF2L_ERR_INVALID, "row count");
    EXPECT(check_dvb_table(0xffffffffu, 8, 4, 2, len2, addr, total, err) == GF2L_ERR_UNSUPPORTED, "n = 2^32-1");
    EXPECT(check_dvb_table(16, 16, 4, 2, len2, addr, total, err) == GF2L_ERR_INVALID, "k = n");
    EXPECT(check_dvb_table(16, 8, 0, 2, len2, addr, total, err) == GF2L_ERR_INVALID, "z = 0");
    const uint32_t one[] = {1}, a0[] = {0};
    EXPECT(check_dvb_table(16, 8, 8, 1, one, a0, total, err) == GF2L_ERR_INVALID &&
               err.find("placeholder") != std::string::npos, "placeholder table");
    EXPECT(check_dvb_table(16, 8, 4, 2, nullptr, addr, total, err) == GF2L_ERR_INVALID, "null table");
}

int main() {
    test_build_csr();
    test_gf2data();
    test_dvb_table();
    if (failures) {
        std::fprintf(stderr, "%d expectation(s) failed\n", failures);
        return 1;
    }
    std::printf("host_sanitize_test: all expectations hold\n");
    return 0;
}
