//! gf2-coding's batch LDPC decoder on one AMD MI355X.
//!
//! The drop-in is [`HipLdpcDecoder::decode_batch`]: the signature, argument
//! meaning, results and panics of `LdpcDecoder::decode_batch`
//! (crates/gf2-coding/src/ldpc/core.rs:729-756), with the code given as a
//! [`HipCode`] (the parity-check matrix uploaded once to the GPU) instead of
//! an `LdpcCode` (whose matrix gf2-coding does not expose outside the crate).
//! Every frame is decoded by the same min-sum flooding rule with syndrome
//! early stop, bit for bit: messages, iteration counts and both flags equal
//! `LdpcDecoder::decode_iterative`'s (core.rs:888-938).
//!
//! [`HipLdpcDecoder`] also implements `SoftDecoder` and
//! `IterativeSoftDecoder` (crates/gf2-coding/src/traits.rs:242-347) for
//! single frames; batches are where the GPU pays off.
//!
//! ```ignore
//! use gf2_coding::{Llr, QuasiCyclicLdpc};
//! use gf2_ldpc_hip::{Context, HipCode, HipLdpcDecoder};
//!
//! let ctx = Context::new(0)?;
//! let code = HipCode::from_quasi_cyclic(&ctx, &qc)?;      // qc: QuasiCyclicLdpc
//! let results = HipLdpcDecoder::decode_batch(&code, &llr_blocks, 50);
//! ```
pub mod ffi;

use std::ffi::CStr;
use std::fmt;
use std::ptr;

use gf2_coding::bch::CodeRate;
use gf2_coding::llr::Llr;
use gf2_coding::traits::{DecoderResult, IterativeSoftDecoder, SoftDecoder};
use gf2_coding::QuasiCyclicLdpc;
use gf2_core::BitVec;

/// A libgf2ldpc status other than `GF2L_OK`, with the library's message.
#[derive(Debug, Clone, PartialEq, Eq)]
pub struct Error {
    pub code: i32,
    pub message: String,
}

impl fmt::Display for Error {
    fn fmt(&self, f: &mut fmt::Formatter<'_>) -> fmt::Result {
        write!(f, "gf2ldpc error {}: {}", self.code, self.message)
    }
}

impl std::error::Error for Error {}

/// One HIP device, its stream and the library's device buffers.
pub struct Context {
    raw: *mut ffi::gf2l_ctx,
}

// The library serialises nothing internally: a context is used from one
// thread at a time (Send, not Sync).
unsafe impl Send for Context {}

impl Context {
    pub fn new(device: i32) -> Result<Self, Error> {
        let mut raw = ptr::null_mut();
        let rc = unsafe { ffi::gf2l_create(device, &mut raw) };
        if rc != ffi::GF2L_OK {
            return Err(Error { code: rc, message: format!("no HIP device {device}") });
        }
        Ok(Self { raw })
    }

    fn check(&self, rc: i32) -> Result<(), Error> {
        if rc == ffi::GF2L_OK {
            return Ok(());
        }
        let msg = unsafe {
            let p = ffi::gf2l_last_error(self.raw);
            if p.is_null() {
                String::new()
            } else {
                CStr::from_ptr(p).to_string_lossy().into_owned()
            }
        };
        Err(Error { code: rc, message: msg })
    }

    /// Launch-shape tunable (`GF2L_TUNE_*`); results never depend on it.
    pub fn set_tuning(&self, key: i32, value: i32) -> Result<(), Error> {
        self.check(unsafe { ffi::gf2l_set_tuning(self.raw, key, value) })
    }
}

impl Drop for Context {
    fn drop(&mut self) {
        unsafe { ffi::gf2l_destroy(self.raw) }
    }
}

/// A parity-check matrix resident on the GPU: the device counterpart of
/// `LdpcCode` (core.rs:60-255).
pub struct HipCode<'c> {
    ctx: &'c Context,
    raw: *mut ffi::gf2l_code,
    n: usize,
    m: usize,
    k: usize,
}

impl<'c> HipCode<'c> {
    fn wrap(ctx: &'c Context, rc: i32, raw: *mut ffi::gf2l_code) -> Result<Self, Error> {
        ctx.check(rc)?;
        let (mut n, mut m, mut k, mut nnz) = (0u32, 0u32, 0u32, 0u64);
        ctx.check(unsafe { ffi::gf2l_code_info(raw, &mut n, &mut m, &mut k, &mut nnz) })?;
        Ok(Self { ctx, raw, n: n as usize, m: m as usize, k: k as usize })
    }

    /// `LdpcCode::from_quasi_cyclic` (core.rs:168-173).  The base matrix is
    /// read back from `QuasiCyclicLdpc::to_edges` (core.rs:578-600): row i of
    /// a circulant with shift s has its one at column (i + s) mod Z
    /// (`CirculantMatrix::to_edges`, core.rs:436-446), so the edge of row 0
    /// of each block gives its shift.
    pub fn from_quasi_cyclic(ctx: &'c Context, qc: &QuasiCyclicLdpc) -> Result<Self, Error> {
        let (mb, nb, z) = (qc.base_rows(), qc.base_cols(), qc.expansion_factor());
        let mut shifts = vec![-1i32; mb * nb];
        for (r, c) in qc.to_edges() {
            if r % z == 0 {
                shifts[(r / z) * nb + c / z] = (c % z) as i32;
            }
        }
        let mut raw = ptr::null_mut();
        let rc = unsafe { ffi::gf2l_code_qc(ctx.raw, mb as u32, nb as u32, z as u32, shifts.as_ptr(), &mut raw) };
        Self::wrap(ctx, rc, raw)
    }

    /// `LdpcCode::from_edges` (core.rs:96-105): (check, variable) pairs with
    /// `SpBitMatrixDual::from_coo`'s semantics (a pair given twice cancels).
    pub fn from_edges(ctx: &'c Context, m: usize, n: usize, edges: &[(usize, usize)]) -> Result<Self, Error> {
        let mut rc_pairs = Vec::with_capacity(2 * edges.len());
        for &(r, c) in edges {
            rc_pairs.push(r as u32);
            rc_pairs.push(c as u32);
        }
        let mut raw = ptr::null_mut();
        let rc = unsafe {
            ffi::gf2l_code_edges(ctx.raw, m as u32, n as u32, rc_pairs.as_ptr(), edges.len(), &mut raw)
        };
        Self::wrap(ctx, rc, raw)
    }

    /// `LdpcCode::dvb_t2_short` / `dvb_t2_normal` (core.rs:197-255).
    pub fn dvb_t2(ctx: &'c Context, normal: bool, rate: CodeRate) -> Result<Self, Error> {
        let r = match rate {
            CodeRate::Rate1_2 => 0,
            CodeRate::Rate3_5 => 1,
            CodeRate::Rate2_3 => 2,
            CodeRate::Rate3_4 => 3,
            CodeRate::Rate4_5 => 4,
            CodeRate::Rate5_6 => 5,
        };
        let mut raw = ptr::null_mut();
        let rc = unsafe { ffi::gf2l_code_dvb_t2(ctx.raw, normal as i32, r, &mut raw) };
        Self::wrap(ctx, rc, raw)
    }

    pub fn n(&self) -> usize {
        self.n
    }
    pub fn m(&self) -> usize {
        self.m
    }
    pub fn k(&self) -> usize {
        self.k
    }
}

impl Drop for HipCode<'_> {
    fn drop(&mut self) {
        unsafe { ffi::gf2l_code_destroy(self.raw) }
    }
}

/// The GPU decoder of one code.
pub struct HipLdpcDecoder<'a, 'c> {
    code: &'a HipCode<'c>,
    last_iterations: usize,
}

impl<'a, 'c> HipLdpcDecoder<'a, 'c> {
    pub fn new(code: &'a HipCode<'c>) -> Self {
        Self { code, last_iterations: 0 }
    }

    /// `LdpcDecoder::decode_batch` (core.rs:729-756): one `DecoderResult` per
    /// LLR block, in order.  Panics, as the reference does, when a block's
    /// length is not n ("LLR length must equal n", core.rs:890); also when the
    /// GPU call fails (the library's message).
    pub fn decode_batch(code: &HipCode, llr_blocks: &[Vec<Llr>], max_iterations: usize) -> Vec<DecoderResult> {
        match Self::try_decode_batch(code, llr_blocks, max_iterations) {
            Ok(r) => r,
            Err(e) => panic!("{e}"),
        }
    }

    /// `decode_batch` returning the library's error instead of panicking
    /// (the LLR length check still panics with the reference's text).
    pub fn try_decode_batch(
        code: &HipCode,
        llr_blocks: &[Vec<Llr>],
        max_iterations: usize,
    ) -> Result<Vec<DecoderResult>, Error> {
        let (n, k) = (code.n, code.k);
        let frames = llr_blocks.len();
        if frames == 0 {
            return Ok(Vec::new());
        }
        let mut llr = Vec::with_capacity(frames * n);
        for block in llr_blocks {
            assert_eq!(block.len(), n, "LLR length must equal n");
            llr.extend(block.iter().map(|l| l.value()));
        }
        let kw = (k + 63) / 64;
        let mut msg = vec![0u64; frames * kw];
        let mut iters = vec![0u32; frames];
        let mut conv = vec![0u8; frames];
        let mut synd = vec![0u8; frames];
        let max_it = u32::try_from(max_iterations).unwrap_or(u32::MAX);
        code.ctx.check(unsafe {
            ffi::gf2l_decode_batch(
                code.ctx.raw,
                code.raw,
                llr.as_ptr(),
                frames,
                max_it,
                msg.as_mut_ptr(),
                iters.as_mut_ptr(),
                conv.as_mut_ptr(),
                synd.as_mut_ptr(),
            )
        })?;
        Ok((0..frames)
            .map(|f| {
                let bits = BitVec::from_words(msg[f * kw..(f + 1) * kw].to_vec(), k);
                DecoderResult::new(bits, iters[f] as usize, conv[f] != 0, synd[f] != 0)
            })
            .collect())
    }
}

impl SoftDecoder for HipLdpcDecoder<'_, '_> {
    fn k(&self) -> usize {
        self.code.k
    }

    fn n(&self) -> usize {
        self.code.n
    }

    /// Hard decisions of the input LLRs, as the reference's non-iterative
    /// interface (core.rs:877-885); no GPU work.
    fn decode_soft(&self, llrs: &[Llr]) -> BitVec {
        assert_eq!(llrs.len(), self.code.n);
        let mut decoded = BitVec::with_capacity(self.code.n);
        for &llr in llrs {
            decoded.push_bit(llr.hard_decision());
        }
        decoded
    }
}

impl IterativeSoftDecoder for HipLdpcDecoder<'_, '_> {
    /// `LdpcDecoder::decode_iterative` (core.rs:888-938) on the GPU: a batch
    /// of one frame.
    fn decode_iterative(&mut self, llrs: &[Llr], max_iterations: usize) -> DecoderResult {
        assert_eq!(llrs.len(), self.code.n, "LLR length must equal n");
        let mut r = Self::decode_batch(self.code, &[llrs.to_vec()], max_iterations);
        let result = r.pop().expect("one result per frame");
        self.last_iterations = result.iterations;
        result
    }

    fn last_iteration_count(&self) -> usize {
        self.last_iterations
    }

    /// The GPU decoder keeps no state between frames.
    fn reset(&mut self) {
        self.last_iterations = 0;
    }
}
