// Structural C++ restatement of the reference's rayon batch decoder, used as
// the timed CPU baseline ("port") and cross-checked bit for bit against the
// flat-array oracle in gf2_oracle.c.
//
// TEST INFRASTRUCTURE ONLY (loaded by tests/ and bench.py's cpu_baseline leg).
//
// It keeps the reference's data structures and per-frame costs so the timing
// reflects the reference path, not a re-optimised CPU decoder:
//   * LdpcDecoder::decode_batch (crates/gf2-coding/src/ldpc/core.rs:729-756):
//     one fresh decoder per frame, built from a deep copy of the code
//     (`code.clone()`, core.rs:740), frames pulled by a pool of worker threads
//     (rayon work stealing), results stored in input order.
//   * LdpcDecoder::new (core.rs:653-697): Vec<Vec<usize>> neighbour lists,
//     Vec<Vec<Llr>> message storage.
//   * check_node_update_minsum (core.rs:790-814) with the linear
//     find_check_position scan (core.rs:841-846) and a per-edge temp vector.
//   * Llr::boxplus_minsum_n (crates/gf2-coding/src/llr.rs:246-269): copies its
//     input into a fresh Vec<f32> (llr.rs:261) and calls the AVX2 kernel
//     minsum_avx2_f32 (crates/gf2-kernels-simd/src/llr.rs:117-180).
//   * variable_node_update (core.rs:819-838) with the check_to_var_message
//     scan (core.rs:849-856).
//   * hard_decode -> BitVec (core.rs:859-865), is_valid_codeword -> CSR matvec
//     producing a BitVec and count_ones (core.rs:129-138, sparse.rs:367-380).
#include <immintrin.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "gf2_oracle.h"

namespace {

struct BitVecP {  // crates/gf2-core/src/bitvec.rs:40-63
    std::vector<uint64_t> data;
    size_t len = 0;
    void push_bit(bool b) {
        if ((len & 63) == 0) data.push_back(0);
        if (b) data[len >> 6] |= uint64_t(1) << (len & 63);
        ++len;
    }
    bool get(size_t i) const { return (data[i >> 6] >> (i & 63)) & 1u; }
    size_t count_ones() const {
        size_t c = 0;
        for (uint64_t w : data) c += (size_t)__builtin_popcountll(w);
        return c;
    }
};

struct SpMat {  // SpBitMatrix{indptr, indices}, sparse.rs:67-72
    size_t rows = 0, cols = 0;
    std::vector<size_t> indptr, indices;
};

struct CodeP {  // LdpcCode{h: SpBitMatrixDual{csr, csc}, n, m}, core.rs:60-70
    SpMat csr, csc;
    size_t n = 0, m = 0;
    size_t k() const { return n > m ? n - m : 0; }
    BitVecP syndrome(const BitVecP &x) const {  // SpBitMatrix::matvec
        BitVecP y;
        for (size_t r = 0; r < csr.rows; ++r) {
            bool acc = false;
            for (size_t p = csr.indptr[r]; p < csr.indptr[r + 1]; ++p) acc ^= x.get(csr.indices[p]);
            y.push_bit(acc);
        }
        return y;
    }
    bool is_valid_codeword(const BitVecP &x) const { return syndrome(x).count_ones() == 0; }
};

__attribute__((target("avx2"))) float minsum_avx2_f32(const float *inputs, size_t n) {
    if (n == 0) return 0.0f;
    const __m256 sign_mask = _mm256_set1_ps(-0.0f);
    __m256 vec_min = _mm256_set1_ps(INFINITY);
    __m256 vec_sign = _mm256_setzero_ps();
    size_t chunks = n / 8;
    for (size_t i = 0; i < chunks; ++i) {
        __m256 vals = _mm256_loadu_ps(inputs + i * 8);
        __m256 abs_vals = _mm256_andnot_ps(sign_mask, vals);
        vec_min = _mm256_min_ps(vec_min, abs_vals);
        vec_sign = _mm256_xor_ps(vec_sign, _mm256_and_ps(vals, sign_mask));
    }
    float temp[8];
    _mm256_storeu_ps(temp, vec_min);
    float min_abs = INFINITY;
    for (float v : temp) min_abs = std::isnan(v) ? min_abs : (std::isnan(min_abs) ? v : std::min(min_abs, v));
    _mm256_storeu_ps(temp, vec_sign);
    float sign_product = 1.0f;
    for (float s : temp) {
        uint32_t u;
        std::memcpy(&u, &s, 4);
        if (u & 0x80000000u) sign_product = -sign_product;
    }
    for (size_t i = chunks * 8; i < n; ++i) {
        float v = std::fabs(inputs[i]);
        if (!std::isnan(v)) min_abs = std::isnan(min_abs) ? v : std::min(min_abs, v);
        if (inputs[i] < 0.0f) sign_product = -sign_product;
    }
    return sign_product * min_abs;
}

float boxplus_minsum_n(const std::vector<float> &llrs) {
    std::vector<float> f32_vals(llrs.begin(), llrs.end());  // llr.rs:261
    return minsum_avx2_f32(f32_vals.data(), f32_vals.size());
}

struct DecoderP {
    CodeP code;
    std::vector<float> beliefs;
    std::vector<std::vector<float>> check_to_var, var_to_check;
    std::vector<std::vector<size_t>> check_neighbors, var_neighbors;
    std::vector<float> temp_inputs;

    explicit DecoderP(CodeP c) : code(std::move(c)) {
        const size_t n = code.n, m = code.m;
        check_neighbors.resize(m);
        for (size_t r = 0; r < m; ++r)
            check_neighbors[r].assign(code.csr.indices.begin() + code.csr.indptr[r],
                                      code.csr.indices.begin() + code.csr.indptr[r + 1]);
        var_neighbors.resize(n);
        for (size_t v = 0; v < n; ++v)
            var_neighbors[v].assign(code.csc.indices.begin() + code.csc.indptr[v],
                                    code.csc.indices.begin() + code.csc.indptr[v + 1]);
        size_t maxd = 0;
        for (auto &nb : check_neighbors) maxd = std::max(maxd, nb.size());
        check_to_var.resize(m);
        for (size_t r = 0; r < m; ++r) check_to_var[r].assign(check_neighbors[r].size(), 0.0f);
        var_to_check.resize(n);
        for (size_t v = 0; v < n; ++v) var_to_check[v].assign(var_neighbors[v].size(), 0.0f);
        beliefs.assign(n, 0.0f);
        temp_inputs.reserve(maxd);
    }

    size_t find_check_position(size_t var, size_t check) const {
        const auto &nb = var_neighbors[var];
        return (size_t)(std::find(nb.begin(), nb.end(), check) - nb.begin());
    }
    float check_to_var_message(size_t var, size_t pos) const {
        size_t check = var_neighbors[var][pos];
        const auto &nb = check_neighbors[check];
        size_t cpos = (size_t)(std::find(nb.begin(), nb.end(), var) - nb.begin());
        return check_to_var[check][cpos];
    }
    void check_node_update_minsum() {
        for (size_t check = 0; check < check_neighbors.size(); ++check) {
            const auto &nb = check_neighbors[check];
            for (size_t pos = 0; pos < nb.size(); ++pos) {
                temp_inputs.clear();
                for (size_t op = 0; op < nb.size(); ++op) {
                    if (op == pos) continue;
                    size_t other = nb[op];
                    temp_inputs.push_back(var_to_check[other][find_check_position(other, check)]);
                }
                check_to_var[check][pos] = temp_inputs.empty() ? 0.0f : boxplus_minsum_n(temp_inputs);
            }
        }
    }
    void variable_node_update(const float *llr) {
        for (size_t var = 0; var < code.n; ++var) {
            float belief = llr[var];
            for (size_t pos = 0; pos < var_neighbors[var].size(); ++pos)
                belief = belief + check_to_var_message(var, pos);
            beliefs[var] = belief;
            for (size_t pos = 0; pos < var_neighbors[var].size(); ++pos)
                var_to_check[var][pos] = belief - check_to_var_message(var, pos);
        }
    }
    BitVecP hard_decode() const {
        BitVecP d;
        for (float b : beliefs) d.push_bit(b < 0.0f);
        return d;
    }
    void decode_iterative(const float *llr, uint32_t max_it, uint64_t *msg, or_result *res) {
        for (auto &cm : check_to_var) std::fill(cm.begin(), cm.end(), 0.0f);
        for (size_t v = 0; v < code.n; ++v)
            std::fill(var_to_check[v].begin(), var_to_check[v].end(), llr[v]);
        uint32_t iterations = 0;
        bool converged = false;
        for (uint32_t it = 0; it < max_it; ++it) {
            iterations = it + 1;
            check_node_update_minsum();
            variable_node_update(llr);
            if (code.is_valid_codeword(hard_decode())) {
                converged = true;
                break;
            }
        }
        BitVecP cw = hard_decode();
        bool synd = code.is_valid_codeword(cw);
        BitVecP message;
        for (size_t i = 0; i < code.k(); ++i) message.push_bit(cw.get(i));
        std::copy(message.data.begin(), message.data.end(), msg);
        res->iterations = iterations;
        res->converged = converged;
        res->syndrome_ok = synd;
    }
};

}  // namespace

extern "C" {

void *port_code_new(uint32_t m, uint32_t n, const uint32_t *rc, size_t nnz) {
    or_code *oc = or_code_from_edges(m, n, rc, nnz);  // from_coo ordering
    if (!oc) return nullptr;
    CodeP *c = new CodeP;
    c->m = m;
    c->n = n;
    c->csr.rows = m;
    c->csr.cols = n;
    c->csr.indptr.assign(oc->row_ptr, oc->row_ptr + m + 1);
    c->csr.indices.assign(oc->row_idx, oc->row_idx + oc->nnz);
    c->csc.rows = n;
    c->csc.cols = m;
    c->csc.indptr.assign(oc->col_ptr, oc->col_ptr + n + 1);
    c->csc.indices.assign(oc->col_idx, oc->col_idx + oc->nnz);
    or_code_free(oc);
    return c;
}

void port_code_free(void *c) { delete static_cast<CodeP *>(c); }

// decode_batch: `threads` workers pull frame indices from a shared counter;
// each frame gets DecoderP(code copy) as in core.rs:740.
void port_decode_batch(const void *code, const float *llr, size_t frames, uint32_t max_it,
                       uint64_t *msg_words, or_result *res, int threads) {
    const CodeP &c = *static_cast<const CodeP *>(code);
    const size_t kw = (c.k() + 63) / 64;
    std::atomic<size_t> next{0};
    auto worker = [&]() {
        for (;;) {
            size_t f = next.fetch_add(1);
            if (f >= frames) break;
            DecoderP dec(c);
            dec.decode_iterative(llr + f * c.n, max_it, msg_words + f * kw, res + f);
        }
    };
    if (threads < 1) threads = 1;
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; ++t) pool.emplace_back(worker);
    worker();
    for (auto &t : pool) t.join();
}

}  // extern "C"
