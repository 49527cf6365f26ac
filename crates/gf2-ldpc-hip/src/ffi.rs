//! Raw bindings of libgf2ldpc's C ABI (`include/gf2ldpc.h`), one `extern`
//! item per declaration, in the header's order.  Status codes: `GF2L_OK` (0)
//! or a negative `GF2L_ERR_*`; `gf2l_last_error` holds the message.
#![allow(non_camel_case_types)]

use std::os::raw::{c_char, c_int, c_void};

/// Opaque library handles.
#[repr(C)]
pub struct gf2l_ctx {
    _private: [u8; 0],
}
#[repr(C)]
pub struct gf2l_code {
    _private: [u8; 0],
}
#[repr(C)]
pub struct gf2l_nms {
    _private: [u8; 0],
}
#[repr(C)]
pub struct gf2l_codeset {
    _private: [u8; 0],
}
#[repr(C)]
pub struct gf2_ldpc_decoder {
    _private: [u8; 0],
}

pub const GF2L_OK: c_int = 0;
pub const GF2L_ERR_INVALID: c_int = -1;
pub const GF2L_ERR_HIP: c_int = -2;
pub const GF2L_ERR_NOMEM: c_int = -3;
pub const GF2L_ERR_UNSUPPORTED: c_int = -4;
pub const GF2L_ERR_INTERNAL: c_int = -5;

pub const GF2L_KERNEL_CHECK: c_int = 0;
pub const GF2L_KERNEL_VARIABLE: c_int = 1;
pub const GF2L_KERNEL_SYNDROME: c_int = 2;
pub const GF2L_KERNEL_LAYERED: c_int = 3;

pub const GF2L_PROFILE_ALL: c_int = 1;
pub const GF2L_PROFILE_LAYERED: c_int = 2;

pub const GF2L_PATH_AUTO: c_int = 0;
pub const GF2L_PATH_LDS: c_int = 1;
pub const GF2L_PATH_STREAM: c_int = 2;
pub const GF2L_PATH_STATE: c_int = 3;
pub const GF2L_PATH_CSR: c_int = 4;

pub const GF2L_RULE_MINSUM: c_int = 0;
pub const GF2L_RULE_SPA: c_int = 1;

pub const GF2L_TUNE_WG_PER_CU: c_int = 1;
pub const GF2L_TUNE_XCD_MAP: c_int = 2;
pub const GF2L_TUNE_GRAPH_ITERS: c_int = 3;
pub const GF2L_TUNE_CHUNK: c_int = 4;
pub const GF2L_TUNE_LAYERED_COLUMNS: c_int = 5;
pub const GF2L_TUNE_LAYERED_KERNEL: c_int = 6;
pub const GF2L_TUNE_LDS_GROUPS: c_int = 7;
pub const GF2L_TUNE_HBM_GRAPH: c_int = 8;

pub const GF2L_NOALLOC_SCRATCH_BYTES: u32 = 64 << 20;

pub const GF2L_FALLBACK_RESTART: c_int = 0;
pub const GF2L_FALLBACK_CONTINUE: c_int = 1;

extern "C" {
    pub fn gf2l_create(device: c_int, out: *mut *mut gf2l_ctx) -> c_int;
    pub fn gf2l_destroy(ctx: *mut gf2l_ctx);
    pub fn gf2l_last_error(ctx: *const gf2l_ctx) -> *const c_char;
    pub fn gf2l_synchronize(ctx: *mut gf2l_ctx) -> c_int;
    pub fn gf2l_stream(ctx: *mut gf2l_ctx) -> *mut c_void;
    pub fn gf2l_profile_enable(ctx: *mut gf2l_ctx, on: c_int) -> c_int;
    pub fn gf2l_profile_read(ctx: *mut gf2l_ctx, ms: *mut f64, launches: *mut u64, frame_iterations: *mut u64)
        -> c_int;
    pub fn gf2l_profile_read_kernel(
        ctx: *mut gf2l_ctx,
        kernel: c_int,
        ms: *mut f64,
        launches: *mut u64,
        frame_iterations: *mut u64,
    ) -> c_int;
    pub fn gf2l_set_path(ctx: *mut gf2l_ctx, path: c_int) -> c_int;
    pub fn gf2l_code_path(ctx: *mut gf2l_ctx, code: *const gf2l_code, path: *mut c_int) -> c_int;
    pub fn gf2l_set_check_rule(ctx: *mut gf2l_ctx, rule: c_int) -> c_int;
    pub fn gf2l_set_tuning(ctx: *mut gf2l_ctx, key: c_int, value: c_int) -> c_int;

    pub fn gf2l_code_qc(
        ctx: *mut gf2l_ctx,
        mb: u32,
        nb: u32,
        z: u32,
        shifts: *const i32,
        out: *mut *mut gf2l_code,
    ) -> c_int;
    pub fn gf2l_code_edges(
        ctx: *mut gf2l_ctx,
        m: u32,
        n: u32,
        rc_pairs: *const u32,
        nnz: usize,
        out: *mut *mut gf2l_code,
    ) -> c_int;
    pub fn gf2l_code_info(code: *const gf2l_code, n: *mut u32, m: *mut u32, k: *mut u32, nnz: *mut u64) -> c_int;
    pub fn gf2l_code_destroy(code: *mut gf2l_code);
    pub fn gf2l_qc_ira_shifts(
        seed: u64,
        mb: u32,
        nb: u32,
        z: u32,
        info_col_weight: *const u32,
        shifts_out: *mut i32,
    ) -> c_int;
    pub fn gf2l_dvb_t2_params(
        frame: c_int,
        rate: c_int,
        n: *mut u32,
        k: *mut u32,
        q: *mut u32,
        rows: *mut u32,
        entries: *mut u32,
    ) -> c_int;
    pub fn gf2l_dvb_t2_table(frame: c_int, rate: c_int, row_len: *mut u32, addresses: *mut u32) -> c_int;
    pub fn gf2l_code_dvb_table(
        ctx: *mut gf2l_ctx,
        n: u32,
        k: u32,
        q: u32,
        z: u32,
        rows: u32,
        row_len: *const u32,
        addresses: *const u32,
        out: *mut *mut gf2l_code,
    ) -> c_int;
    pub fn gf2l_code_dvb_t2(ctx: *mut gf2l_ctx, frame: c_int, rate: c_int, out: *mut *mut gf2l_code) -> c_int;
    pub fn gf2l_code_gf2data(ctx: *mut gf2l_ctx, bytes: *const c_void, len: usize, out: *mut *mut gf2l_code)
        -> c_int;

    pub fn gf2l_device_alloc(ctx: *mut gf2l_ctx, bytes: usize, out: *mut *mut c_void) -> c_int;
    pub fn gf2l_device_free(ctx: *mut gf2l_ctx, ptr: *mut c_void) -> c_int;
    pub fn gf2l_copy_to_device(ctx: *mut gf2l_ctx, dst: *mut c_void, src: *const c_void, bytes: usize) -> c_int;
    pub fn gf2l_copy_to_host(ctx: *mut gf2l_ctx, dst: *mut c_void, src: *const c_void, bytes: usize) -> c_int;
    pub fn gf2l_frames_generate(
        ctx: *mut gf2l_ctx,
        code: *const gf2l_code,
        seed: u64,
        first_frame: u64,
        frames: usize,
        ebn0_lo_db: f32,
        ebn0_hi_db: f32,
        d_llr: *mut f32,
        d_codeword_words: *mut u64,
    ) -> c_int;
    pub fn gf2l_frames_generate_zero(
        ctx: *mut gf2l_ctx,
        code: *const gf2l_code,
        seed: u64,
        first_frame: u64,
        frames: usize,
        ebn0_lo_db: f32,
        ebn0_hi_db: f32,
        d_llr: *mut f32,
    ) -> c_int;

    pub fn gf2l_decode_batch(
        ctx: *mut gf2l_ctx,
        code: *const gf2l_code,
        llr: *const f32,
        frames: usize,
        max_it: u32,
        msg_words: *mut u64,
        iterations: *mut u32,
        converged: *mut u8,
        syndrome_ok: *mut u8,
    ) -> c_int;
    pub fn gf2l_decode_batch_device(
        ctx: *mut gf2l_ctx,
        code: *const gf2l_code,
        d_llr: *const f32,
        frames: usize,
        max_it: u32,
        d_msg_words: *mut u64,
        d_iterations: *mut u32,
        d_converged: *mut u8,
        d_syndrome_ok: *mut u8,
    ) -> c_int;
    pub fn gf2l_decode_batch_device_noalloc(
        ctx: *mut gf2l_ctx,
        code: *const gf2l_code,
        d_llr: *const f32,
        frames: usize,
        max_it: u32,
        d_msg_words: *mut u64,
        d_iterations: *mut u32,
        d_converged: *mut u8,
        d_syndrome_ok: *mut u8,
    ) -> c_int;
    pub fn gf2l_decode_trace(
        ctx: *mut gf2l_ctx,
        code: *const gf2l_code,
        llr: *const f32,
        frames: usize,
        max_it: u32,
        trace_iters: u32,
        c2v_out: *mut f32,
        belief_out: *mut f32,
        msg_words: *mut u64,
        iterations: *mut u32,
        converged: *mut u8,
        syndrome_ok: *mut u8,
    ) -> c_int;

    pub fn gf2l_nms_create(
        ctx: *mut gf2l_ctx,
        code: *const gf2l_code,
        iterations: u32,
        init_weights: *const f32,
        init_value: f32,
        out: *mut *mut gf2l_nms,
    ) -> c_int;
    pub fn gf2l_nms_create_layered(
        ctx: *mut gf2l_ctx,
        code: *const gf2l_code,
        iterations: u32,
        init_weights: *const f32,
        init_value: f32,
        out: *mut *mut gf2l_nms,
    ) -> c_int;
    pub fn gf2l_nms_set_fallback(ctx: *mut gf2l_ctx, nms: *mut gf2l_nms, mode: c_int) -> c_int;
    pub fn gf2l_nms_layered_columns(ctx: *mut gf2l_ctx, nms: *const gf2l_nms, columns: *mut u32) -> c_int;
    pub fn gf2l_nms_destroy(nms: *mut gf2l_nms);
    pub fn gf2l_nms_info(nms: *const gf2l_nms, iterations: *mut u32, weights_per_iteration: *mut u32) -> c_int;
    pub fn gf2l_nms_get_weights(ctx: *mut gf2l_ctx, nms: *const gf2l_nms, weights: *mut f32) -> c_int;
    pub fn gf2l_nms_set_weights(ctx: *mut gf2l_ctx, nms: *mut gf2l_nms, weights: *const f32) -> c_int;
    pub fn gf2l_nms_get_adam(ctx: *mut gf2l_ctx, nms: *const gf2l_nms, m: *mut f32, v: *mut f32, step: *mut u64)
        -> c_int;
    pub fn gf2l_nms_set_adam(ctx: *mut gf2l_ctx, nms: *mut gf2l_nms, m: *const f32, v: *const f32, step: u64)
        -> c_int;
    pub fn gf2l_nms_gradient(
        ctx: *mut gf2l_ctx,
        nms: *mut gf2l_nms,
        d_llr: *const f32,
        d_cw: *const u64,
        frames: usize,
        grad: *mut f32,
        loss: *mut f32,
    ) -> c_int;
    pub fn gf2l_nms_train_step(
        ctx: *mut gf2l_ctx,
        nms: *mut gf2l_nms,
        d_llr: *const f32,
        d_cw: *const u64,
        frames: usize,
        lr: f32,
        loss: *mut f32,
    ) -> c_int;
    pub fn gf2l_nms_decode_device(
        ctx: *mut gf2l_ctx,
        nms: *const gf2l_nms,
        d_llr: *const f32,
        frames: usize,
        fallback_max_it: u32,
        d_msg_words: *mut u64,
        d_iterations: *mut u32,
        d_converged: *mut u8,
        d_syndrome_ok: *mut u8,
        d_fallback: *mut u8,
    ) -> c_int;

    pub fn gf2l_codeset_create(
        ctx: *mut gf2l_ctx,
        codes: *const *const gf2l_code,
        count: u32,
        out: *mut *mut gf2l_codeset,
    ) -> c_int;
    pub fn gf2l_codeset_destroy(set: *mut gf2l_codeset);
    pub fn gf2l_decode_mixed_device(
        ctx: *mut gf2l_ctx,
        set: *mut gf2l_codeset,
        code_ids: *const u16,
        frames: usize,
        d_llr: *const f32,
        max_it: u32,
        d_msg_words: *mut u64,
        d_iterations: *mut u32,
        d_converged: *mut u8,
        d_syndrome_ok: *mut u8,
    ) -> c_int;
    pub fn gf2l_decode_mixed_nms_device(
        ctx: *mut gf2l_ctx,
        set: *mut gf2l_codeset,
        decoders: *const *const gf2l_nms,
        code_ids: *const u16,
        frames: usize,
        d_llr: *const f32,
        fallback_max_it: u32,
        d_msg_words: *mut u64,
        d_iterations: *mut u32,
        d_converged: *mut u8,
        d_syndrome_ok: *mut u8,
        d_fallback: *mut u8,
    ) -> c_int;

    pub fn gf2_ldpc_decoder_create(frame_size: u32, code_rate: u32, max_iterations: u32) -> *mut gf2_ldpc_decoder;
    pub fn gf2_ldpc_decode(decoder: *mut gf2_ldpc_decoder, llrs: *const f32, len: usize, output: *mut u8) -> i32;
    pub fn gf2_ldpc_decode_frames(
        decoder: *mut gf2_ldpc_decoder,
        llrs: *const f32,
        frames: usize,
        output: *mut u8,
        iterations: *mut i32,
    ) -> i32;
    pub fn gf2_ldpc_decoder_info(decoder: *const gf2_ldpc_decoder, n: *mut u32, k: *mut u32) -> c_int;
    pub fn gf2_ldpc_last_error(decoder: *const gf2_ldpc_decoder) -> *const c_char;
    pub fn gf2_ldpc_decoder_destroy(decoder: *mut gf2_ldpc_decoder);
}
