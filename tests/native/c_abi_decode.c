/* A plain C caller of the library's C ABI (include/gf2ldpc.h), compiled with
 * gcc as a C11 program and linked against gf2_amd/libgf2ldpc.so: what a
 * foreign-language binding (the Rust FFI crate, a GNU Radio block) sees.
 * Run by tests/test_gpu_native.py, which checks its outputs against the
 * oracle.
 *
 *   c_abi_decode <mb> <nb> <z> <shifts.i32> <llr.f32> <frames> <max_it> <out.bin>
 *
 * Builds the code with gf2l_code_qc from the base matrix file (mb*nb int32,
 * row-major, -1 = zero block), decodes the frames (frames*n float32,
 * frame-major) with gf2l_decode_batch and writes, in order: the message words
 * (frames*ceil(k/64) u64), iterations (frames u32), converged and syndrome_ok
 * (frames bytes each).  Before that it checks the error contract on the same
 * context: an out-of-range shift returns GF2L_ERR_INVALID with the
 * reference's panic text, an empty batch returns GF2L_OK, a null context
 * returns GF2L_ERR_INVALID.  Exit status 0 on success. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gf2ldpc.h"

static void *read_file(const char *path, size_t bytes) {
    FILE *f = fopen(path, "rb");
    if (!f) return NULL;
    void *p = malloc(bytes ? bytes : 1);
    const size_t got = fread(p, 1, bytes, f);
    fclose(f);
    if (got != bytes) {
        free(p);
        return NULL;
    }
    return p;
}

int main(int argc, char **argv) {
    if (argc != 9) {
        fprintf(stderr, "usage: %s mb nb z shifts llr frames max_it out\n", argv[0]);
        return 2;
    }
    const uint32_t mb = (uint32_t)atoi(argv[1]), nb = (uint32_t)atoi(argv[2]), z = (uint32_t)atoi(argv[3]);
    const size_t frames = (size_t)atol(argv[6]);
    const uint32_t max_it = (uint32_t)atoi(argv[7]);
    int32_t *shifts = read_file(argv[4], (size_t)mb * nb * sizeof(int32_t));
    if (!shifts) {
        fprintf(stderr, "cannot read %s\n", argv[4]);
        return 2;
    }
    if (gf2l_decode_batch(NULL, NULL, NULL, 0, 0, NULL, NULL, NULL, NULL) != GF2L_ERR_INVALID) {
        fprintf(stderr, "null context not rejected\n");
        return 1;
    }
    gf2l_ctx *ctx = NULL;
    int rc = gf2l_create(0, &ctx);
    if (rc != GF2L_OK) {
        fprintf(stderr, "gf2l_create: %d\n", rc);
        return 1;
    }
    /* the reference's panic text for a bad shift (core.rs:589-596) */
    gf2l_code *code = NULL;
    int32_t saved = shifts[0];
    shifts[0] = (int32_t)z;
    rc = gf2l_code_qc(ctx, mb, nb, z, shifts, &code);
    char want[160];
    snprintf(want, sizeof want, "Shift value %u at position (0,0) must be -1 or in range [0,%u)", z, z);
    if (rc != GF2L_ERR_INVALID || code != NULL || strcmp(gf2l_last_error(ctx), want) != 0) {
        fprintf(stderr, "bad shift: rc %d, message '%s'\n", rc, gf2l_last_error(ctx));
        return 1;
    }
    shifts[0] = saved;
    rc = gf2l_code_qc(ctx, mb, nb, z, shifts, &code);
    if (rc != GF2L_OK) {
        fprintf(stderr, "gf2l_code_qc: %d %s\n", rc, gf2l_last_error(ctx));
        return 1;
    }
    uint32_t n = 0, m = 0, k = 0;
    uint64_t nnz = 0;
    gf2l_code_info(code, &n, &m, &k, &nnz);
    const size_t kw = (k + 63) / 64;
    float *llr = read_file(argv[5], frames * n * sizeof(float));
    if (!llr) {
        fprintf(stderr, "cannot read %s\n", argv[5]);
        return 2;
    }
    /* empty batch: nothing to do (core.rs:1297-1309) */
    if (gf2l_decode_batch(ctx, code, llr, 0, max_it, NULL, NULL, NULL, NULL) != GF2L_OK) {
        fprintf(stderr, "empty batch rejected\n");
        return 1;
    }
    uint64_t *msg = calloc(frames * kw + 1, sizeof(uint64_t));
    uint32_t *iters = calloc(frames + 1, sizeof(uint32_t));
    uint8_t *conv = calloc(frames + 1, 1), *synd = calloc(frames + 1, 1);
    rc = gf2l_decode_batch(ctx, code, llr, frames, max_it, msg, iters, conv, synd);
    if (rc != GF2L_OK) {
        fprintf(stderr, "gf2l_decode_batch: %d %s\n", rc, gf2l_last_error(ctx));
        return 1;
    }
    FILE *out = fopen(argv[8], "wb");
    if (!out) return 2;
    fwrite(msg, sizeof(uint64_t), frames * kw, out);
    fwrite(iters, sizeof(uint32_t), frames, out);
    fwrite(conv, 1, frames, out);
    fwrite(synd, 1, frames, out);
    fclose(out);
    gf2l_code_destroy(code);
    gf2l_destroy(ctx);
    free(msg);
    free(iters);
    free(conv);
    free(synd);
    free(llr);
    free(shifts);
    printf("c_abi_decode: %zu frames of n=%u k=%u decoded\n", frames, n, k);
    return 0;
}
