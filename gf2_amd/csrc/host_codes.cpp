// Host-only code construction and input validation (no HIP): see
// host_codes.h.
#include "host_codes.h"

#include <algorithm>
#include <cstring>

#include "gf2ldpc.h"

namespace gf2l {
namespace host {
namespace {

uint32_t rd32(const uint8_t *p) { return p[0] | (p[1] << 8) | (p[2] << 16) | ((uint32_t)p[3] << 24); }
uint64_t rd64(const uint8_t *p) { return rd32(p) | ((uint64_t)rd32(p + 4) << 32); }

int invalid(std::string &err, const std::string &msg) {
    err = msg;
    return GF2L_ERR_INVALID;
}
int unsupported(std::string &err, const std::string &msg) {
    err = msg;
    return GF2L_ERR_UNSUPPORTED;
}

// Minimal reader of the flat JSON object the GF2DATA writer emits: one
// number or string value per key.
struct Meta {
    std::string text;
    bool find(const char *key, size_t &pos) const {
        const std::string k = std::string("\"") + key + "\"";
        size_t at = text.find(k);
        if (at == std::string::npos) return false;
        at += k.size();
        while (at < text.size() && (text[at] == ' ' || text[at] == '\t' || text[at] == '\n' || text[at] == '\r')) ++at;
        if (at >= text.size() || text[at] != ':') return false;
        ++at;
        while (at < text.size() && (text[at] == ' ' || text[at] == '\t' || text[at] == '\n' || text[at] == '\r')) ++at;
        pos = at;
        return at < text.size();
    }
    bool number(const char *key, uint64_t &v) const {
        size_t p;
        if (!find(key, p) || text[p] < '0' || text[p] > '9') return false;
        v = 0;
        for (; p < text.size() && text[p] >= '0' && text[p] <= '9'; ++p) {
            if (v > (UINT64_MAX - 9) / 10) return false;
            v = v * 10 + (uint64_t)(text[p] - '0');
        }
        return true;
    }
    bool string(const char *key, std::string &v) const {
        size_t p;
        if (!find(key, p) || text[p] != '"') return false;
        const size_t e = text.find('"', p + 1);
        if (e == std::string::npos) return false;
        v = text.substr(p + 1, e - p - 1);
        return true;
    }
};

// Groups the column indices of `nnz` (row, col) pairs by row with a counting
// sort: on return row r's columns are cols[begin[r] .. begin[r+1]) in input
// order.  Indices must already be checked against `rows`.
void group_by_row(uint32_t rows, size_t nnz, const uint32_t *rc, std::vector<uint32_t> &begin,
                  std::vector<uint32_t> &cols) {
    begin.assign((size_t)rows + 1, 0);
    for (size_t e = 0; e < nnz; ++e) begin[(size_t)rc[2 * e] + 1]++;
    for (size_t r = 0; r < rows; ++r) begin[r + 1] += begin[r];
    std::vector<uint32_t> at(begin.begin(), begin.end() - 1);
    cols.resize(nnz);
    for (size_t e = 0; e < nnz; ++e) cols[at[rc[2 * e]]++] = rc[2 * e + 1];
}

}  // namespace

int build_csr(uint32_t m, uint32_t n, const uint32_t *rc, size_t nnz, Csr &out, std::string &err) {
    if (nnz && !rc) return invalid(err, "null edge list");
    // the reference asserts every index first (sparse.rs:170-171)
    for (size_t e = 0; e < nnz; ++e) {
        const uint32_t r = rc[2 * e], c = rc[2 * e + 1];
        if (r >= m)
            return invalid(err, "row index " + std::to_string(r) + " out of bounds (rows=" + std::to_string(m) + ")");
        if (c >= n)
            return invalid(err, "col index " + std::to_string(c) + " out of bounds (cols=" + std::to_string(n) + ")");
    }
    if (m > kMaxNodes || n > kMaxNodes)
        return unsupported(err, "code too large (more than 2^28 checks or variables)");
    if (nnz > kMaxEdges) return unsupported(err, "more than 2^32 - 1 edges");
    std::vector<uint32_t> begin, cols;
    group_by_row(m, nnz, rc, begin, cols);
    // rows sorted; a pair of equal entries cancels (GF(2) addition)
    out = Csr{};
    out.m = m;
    out.n = n;
    out.row_ptr.assign((size_t)m + 1, 0);
    out.col_idx.reserve(nnz);
    for (size_t r = 0; r < m; ++r) {
        auto b = cols.begin() + begin[r], e = cols.begin() + begin[r + 1];
        std::sort(b, e);
        while (b != e) {
            auto k = b + 1;
            while (k != e && *k == *b) ++k;
            if ((k - b) % 2 == 1) out.col_idx.push_back(*b);
            b = k;
        }
        out.row_ptr[r + 1] = (uint32_t)out.col_idx.size();
        out.max_dc = std::max(out.max_dc, out.row_ptr[r + 1] - out.row_ptr[r]);
    }
    // stable transpose (sparse.rs:297-331): each column's checks ascending
    const size_t E = out.col_idx.size();
    out.col_ptr.assign((size_t)n + 1, 0);
    out.csc_edge.resize(E);
    for (size_t e = 0; e < E; ++e) out.col_ptr[(size_t)out.col_idx[e] + 1]++;
    for (size_t v = 0; v < n; ++v) out.col_ptr[v + 1] += out.col_ptr[v];
    std::vector<uint32_t> next(out.col_ptr.begin(), out.col_ptr.end() - 1);
    for (size_t r = 0; r < m; ++r)
        for (uint32_t e = out.row_ptr[r]; e < out.row_ptr[r + 1]; ++e) out.csc_edge[next[out.col_idx[e]]++] = e;
    return GF2L_OK;
}

int parse_gf2data(const void *bytes, size_t len, uint32_t &rows_out, uint32_t &cols_out, std::vector<uint32_t> &rc,
                  std::string &err) {
    rc.clear();
    if (!bytes) return invalid(err, "null GF2DATA buffer");
    const uint8_t *b = static_cast<const uint8_t *>(bytes);
    // Header::read_from (header.rs:33-86)
    if (len < 32) return invalid(err, "I/O error: file shorter than the 32-byte header");
    if (std::memcmp(b, "GF2DATA\0", 8) != 0) return invalid(err, "Invalid magic bytes");
    const uint32_t version = b[8] | (b[9] << 8);
    if (version != 1) return invalid(err, "Unsupported format version: " + std::to_string(version));
    const uint8_t type = b[10], flags = b[11];
    if (type < 1 || type > 4) return invalid(err, "Unknown type tag: " + std::to_string(type));
    if (type != 3 && type != 4)
        return invalid(err, "Invalid data: a code needs a sparse matrix (type 3 or 4), got type " +
                                std::to_string(type));
    if (flags & 0x01) return unsupported(err, "Invalid data: compressed payloads are not supported");
    const uint32_t meta_len = rd32(b + 20);
    const uint64_t data_len = rd64(b + 24);
    if ((uint64_t)len - 32 < meta_len) return invalid(err, "I/O error: metadata runs past the end of the buffer");
    Meta meta{std::string(reinterpret_cast<const char *>(b + 32), meta_len)};
    std::string tname;
    uint64_t rows = 0, cols = 0, nnz = 0;
    if (!meta.string("type", tname) || !meta.number("rows", rows) || !meta.number("cols", cols) ||
        !meta.number("nnz", nnz))
        return invalid(err, "Invalid data: Failed to parse metadata");
    const char *want = type == 3 ? "SpBitMatrix" : "SpBitMatrixDual";
    if (tname != want) return invalid(err, std::string("Invalid data: Expected ") + want + " type, got " + tname);
    if (rows > 0xFFFFFFFFull || cols > 0xFFFFFFFFull || nnz > 0xFFFFFFFFull)
        return unsupported(err, "matrix dimensions above 2^32");
    const uint64_t expected = type == 3 ? nnz * 8 : (rows + 1 + nnz + cols + 1 + nnz) * 4;
    if (type == 3) {
        std::string fmt;
        if (!meta.string("format", fmt) || fmt != "coo")
            return invalid(err, "Invalid data: Expected COO format, got " + fmt);
    }
    if (data_len != expected)
        return invalid(err, "Invalid data: Data length mismatch: header says " + std::to_string(data_len) +
                                " bytes, expected " + std::to_string(expected));
    if ((uint64_t)len - 32 - meta_len < data_len)
        return invalid(err, "I/O error: payload runs past the end of the buffer");
    // the payload is in the buffer; now the decoder's own size limits, before
    // anything is allocated from the header's dimensions
    if (rows > kMaxNodes || cols > kMaxNodes)
        return unsupported(err, "code too large (more than 2^28 checks or variables)");
    const uint8_t *pl = b + 32 + meta_len;
    if (type == 3) {
        // SpBitMatrix::from_coo_deduplicated (sparse.rs:232-259): set
        // semantics, so each distinct entry is kept once.
        std::vector<uint32_t> pairs(2 * nnz);
        for (uint64_t e = 0; e < nnz; ++e) {
            const uint32_t r = rd32(pl + 8 * e), c = rd32(pl + 8 * e + 4);
            if (r >= rows)
                return invalid(err, "row index " + std::to_string(r) + " out of bounds (rows=" + std::to_string(rows) + ")");
            if (c >= cols)
                return invalid(err, "col index " + std::to_string(c) + " out of bounds (cols=" + std::to_string(cols) + ")");
            pairs[2 * e] = r;
            pairs[2 * e + 1] = c;
        }
        std::vector<uint32_t> begin, by_row;
        group_by_row((uint32_t)rows, nnz, pairs.data(), begin, by_row);
        rc.reserve(2 * nnz);
        for (size_t r = 0; r < rows; ++r) {
            auto s = by_row.begin() + begin[r], e = by_row.begin() + begin[r + 1];
            std::sort(s, e);
            e = std::unique(s, e);
            for (auto it = s; it != e; ++it) {
                rc.push_back((uint32_t)r);
                rc.push_back(*it);
            }
        }
    } else {
        // SpBitMatrixDual::from_csr_csc keeps the CSR rows as stored; the code
        // is built from those (row, col) entries.  The CSC half only has to be
        // present.
        if (rd32(pl) != 0 || rd32(pl + 4 * rows) != nnz)
            return invalid(err, "Invalid data: row offsets do not span nnz entries");
        const uint8_t *ri = pl + 4 * (rows + 1);
        rc.reserve(2 * nnz);
        for (uint64_t r = 0; r < rows; ++r) {
            const uint32_t e0 = rd32(pl + 4 * r), e1 = rd32(pl + 4 * (r + 1));
            if (e1 < e0 || e1 > nnz) return invalid(err, "Invalid data: row offsets decrease");
            for (uint32_t e = e0; e < e1; ++e) {
                const uint32_t c = rd32(ri + 4ull * e);
                if (c >= cols)
                    return invalid(err, "col index " + std::to_string(c) + " out of bounds (cols=" + std::to_string(cols) + ")");
                rc.push_back((uint32_t)r);
                rc.push_back(c);
            }
        }
    }
    rows_out = (uint32_t)rows;
    cols_out = (uint32_t)cols;
    return GF2L_OK;
}

int check_dvb_table(uint32_t n, uint32_t k, uint32_t z, uint32_t rows, const uint32_t *row_len,
                    const uint32_t *addresses, size_t &total, std::string &err) {
    total = 0;
    if (!row_len || !addresses || rows == 0) return invalid(err, "null DVB table");
    if (z == 0 || k >= n) return invalid(err, "DVB parameters need 0 < k < n and z > 0");
    if (n > kMaxNodes) return unsupported(err, "code too large (more than 2^28 checks or variables)");
    const uint32_t m = n - k;
    // validate_table (builder.rs:15-41)
    if (!(rows > 1 || row_len[0] > 1)) return invalid(err, "DVB-T2 table not yet implemented (placeholder detected)");
    if (rows != k / z)
        return invalid(err, "Table must have " + std::to_string(k / z) + " rows (info blocks), found " +
                                std::to_string(rows));
    for (uint32_t r = 0; r < rows; ++r) total += row_len[r];
    size_t at = 0;
    for (uint32_t r = 0; r < rows; ++r)
        for (uint32_t i = 0; i < row_len[r]; ++i, ++at)
            if (addresses[at] >= m)
                return invalid(err, "Invalid parity index " + std::to_string(addresses[at]) + " in table row " +
                                        std::to_string(r) + " (must be < " + std::to_string(m) + ")");
    if ((uint64_t)total * z + 2ull * m > kMaxEdges) return unsupported(err, "more than 2^32 - 1 edges");
    return GF2L_OK;
}

}  // namespace host
}  // namespace gf2l
