// SDR-style decoder handle: the C FFI planned for GNU Radio blocks in
// crates/gf2-coding/docs/SDR_INTEGRATION.md:39-57 (gf2_ldpc_decoder_create /
// gf2_ldpc_decode / gf2_ldpc_decoder_destroy), over a DVB-T2 code and this
// library's batch decoder.  A handle owns its own context (HIP device 0,
// one stream), the code and the host staging of one call; it is not
// thread-safe.
#include <cstdint>
#include <new>
#include <string>
#include <vector>

#include "gf2ldpc.h"

struct gf2_ldpc_decoder {
    gf2l_ctx *ctx = nullptr;
    gf2l_code *code = nullptr;
    uint32_t n = 0, k = 0, max_it = 0;
    std::vector<uint64_t> words;
    std::vector<uint32_t> iters;
    std::vector<uint8_t> conv, synd;
    std::string err;
};

namespace {

int32_t sdr_fail(gf2_ldpc_decoder *d, const char *msg) noexcept {
    try {
        d->err = msg;
    } catch (...) {
        d->err.clear();
    }
    return -2;
}

// Runs `frames` frames through gf2l_decode_batch and unpacks the first k hard
// decisions of each into bytes (0 or 1).  No exception leaves: a failed host
// allocation of the staging vectors is an error return.
int32_t sdr_decode(gf2_ldpc_decoder *d, const float *llrs, size_t frames, uint8_t *output,
                   int32_t *iterations) noexcept {
    const size_t kw = (d->k + 63) / 64;
    try {
        d->words.resize(frames * kw);
        d->iters.resize(frames);
        d->conv.resize(frames);
        d->synd.resize(frames);
    } catch (...) {
        return sdr_fail(d, "host allocation failed");
    }
    const int rc = gf2l_decode_batch(d->ctx, d->code, llrs, frames, d->max_it, d->words.data(), d->iters.data(),
                                     d->conv.data(), d->synd.data());
    if (rc != GF2L_OK) return sdr_fail(d, gf2l_last_error(d->ctx));
    int32_t ok = 0;
    for (size_t f = 0; f < frames; ++f) {
        const uint64_t *w = d->words.data() + f * kw;
        uint8_t *o = output + f * d->k;
        for (uint32_t i = 0; i < d->k; ++i) o[i] = (uint8_t)((w[i >> 6] >> (i & 63)) & 1u);
        if (iterations) iterations[f] = d->conv[f] ? (int32_t)d->iters[f] : -1;
        ok += d->conv[f] ? 1 : 0;
    }
    d->err.clear();
    return ok;
}

}  // namespace

extern "C" {

gf2_ldpc_decoder *gf2_ldpc_decoder_create(uint32_t frame_size, uint32_t code_rate, uint32_t max_iterations) {
    if (frame_size > 1 || code_rate > 5) return nullptr;
    gf2_ldpc_decoder *d = new (std::nothrow) gf2_ldpc_decoder;
    if (!d) return nullptr;
    if (gf2l_create(0, &d->ctx) != GF2L_OK ||
        gf2l_code_dvb_t2(d->ctx, (int)frame_size, (int)code_rate, &d->code) != GF2L_OK) {
        if (d->ctx) gf2l_destroy(d->ctx);
        delete d;
        return nullptr;
    }
    uint32_t m = 0;
    uint64_t nnz = 0;
    gf2l_code_info(d->code, &d->n, &m, &d->k, &nnz);
    d->max_it = max_iterations;
    return d;
}

int32_t gf2_ldpc_decode(gf2_ldpc_decoder *decoder, const float *llrs, size_t len, uint8_t *output) {
    if (!decoder) return -2;
    if (!llrs || !output) return sdr_fail(decoder, "null LLR or output buffer");
    if (len != decoder->n) return sdr_fail(decoder, "LLR length must equal n");
    int32_t it = -1;
    const int32_t rc = sdr_decode(decoder, llrs, 1, output, &it);
    return rc < 0 ? rc : it;
}

int32_t gf2_ldpc_decode_frames(gf2_ldpc_decoder *decoder, const float *llrs, size_t frames, uint8_t *output,
                               int32_t *iterations) {
    if (!decoder) return -2;
    if (frames == 0) return 0;
    if (!llrs || !output) return sdr_fail(decoder, "null LLR or output buffer");
    return sdr_decode(decoder, llrs, frames, output, iterations);
}

int gf2_ldpc_decoder_info(const gf2_ldpc_decoder *decoder, uint32_t *n, uint32_t *k) {
    if (!decoder) return GF2L_ERR_INVALID;
    if (n) *n = decoder->n;
    if (k) *k = decoder->k;
    return GF2L_OK;
}

const char *gf2_ldpc_last_error(const gf2_ldpc_decoder *decoder) {
    return decoder ? decoder->err.c_str() : "null decoder";
}

void gf2_ldpc_decoder_destroy(gf2_ldpc_decoder *decoder) {
    if (!decoder) return;
    gf2l_code_destroy(decoder->code);
    gf2l_destroy(decoder->ctx);
    delete decoder;
}

}  // extern "C"
