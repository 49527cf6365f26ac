// Host-only validation and construction of code inputs: the from_coo CSR /
// CSC build behind gf2l_code_edges, the GF2DATA reader behind
// gf2l_code_gf2data and the DVB table checks of gf2l_code_dvb_table.  Plain
// C++17 with no HIP dependency, so the same source also builds on the CPU
// with AddressSanitizer and UndefinedBehaviorSanitizer (tests/native).
//
// Every size read from a caller or a file is bounded before anything is
// allocated from it: at most 2^28 checks or variables per code (the limit of
// the device kernels) and at most 2^32 - 1 edges.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace gf2l {
namespace host {

constexpr uint64_t kMaxNodes = 1ull << 28;         // checks or variables of one code
constexpr uint64_t kMaxEdges = 0xffffffffull;      // Tanner-graph edges of one code

// A parity-check matrix in the reference's dual form (sparse.rs:422-425):
// CSR rows with ascending columns, and per variable its CSR edge indices in
// ascending check order (the stable transpose of sparse.rs:297-331).
struct Csr {
    uint32_t m = 0, n = 0, max_dc = 0;
    std::vector<uint32_t> row_ptr;   // m+1
    std::vector<uint32_t> col_idx;   // nnz
    std::vector<uint32_t> col_ptr;   // n+1
    std::vector<uint32_t> csc_edge;  // nnz
};

// SpBitMatrix::from_coo (crates/gf2-core/src/sparse.rs:166-204) + transpose:
// bounds-checked (row, col) pairs, rows sorted, duplicates cancel in pairs.
// Returns GF2L_OK, GF2L_ERR_INVALID (out-of-range index, with the
// reference's panic text) or GF2L_ERR_UNSUPPORTED (beyond the size limits);
// `err` holds the message.
int build_csr(uint32_t m, uint32_t n, const uint32_t *rc_pairs, size_t nnz, Csr &out, std::string &err);

// GF2DATA image (crates/gf2-core/src/io/mod.rs:6-98, header.rs:33-86) of a
// type-3 (COO, set semantics on load: sparse.rs:232-259) or type-4 (CSR +
// CSC, the CSR rows are used: sparse.rs:397-449) sparse matrix -> its
// dimensions and (row, col) pairs.  Errors carry the reference's IoError
// texts; GF2L_ERR_UNSUPPORTED for compressed payloads and for dimensions
// beyond the size limits.
int parse_gf2data(const void *bytes, size_t len, uint32_t &rows, uint32_t &cols, std::vector<uint32_t> &rc,
                  std::string &err);

// validate_table (crates/gf2-coding/src/ldpc/dvb_t2/builder.rs:15-41) plus
// the size limits: 0 < k < n <= 2^28, z > 0, rows == k / z, every address
// < n - k, and at most kMaxEdges edges.  `total` receives the number of
// table entries.
int check_dvb_table(uint32_t n, uint32_t k, uint32_t z, uint32_t rows, const uint32_t *row_len,
                    const uint32_t *addresses, size_t &total, std::string &err);

}  // namespace host
}  // namespace gf2l
