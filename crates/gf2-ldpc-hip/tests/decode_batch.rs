//! The GPU batch decoder against gf2-coding's own `LdpcDecoder::decode_batch`
//! on the same LLRs: identical messages, iteration counts and flags.
//! Needs a HIP device: `cargo test -p gf2-ldpc-hip -- --ignored`.
use gf2_coding::channel::{AwgnChannel, BpskModulator};
use gf2_coding::ldpc::{LdpcCode, LdpcDecoder, QuasiCyclicLdpc};
use gf2_coding::llr::Llr;
use gf2_ldpc_hip::{Context, HipCode, HipLdpcDecoder};
use rand::rngs::StdRng;
use rand::SeedableRng;

/// A 4x8 base graph with a dual-diagonal parity part, Z = 16.
fn small_qc() -> QuasiCyclicLdpc {
    let base = vec![
        vec![3, 0, 7, -1, 0, -1, -1, -1],
        vec![-1, 5, 1, 12, 0, 0, -1, -1],
        vec![9, -1, 14, 2, -1, 0, 0, -1],
        vec![4, 11, -1, 6, -1, -1, 0, 0],
    ];
    QuasiCyclicLdpc::new(base, 16)
}

/// All-zero codewords over BPSK/AWGN (the method of examples/ldpc_awgn.rs).
fn noisy_zero_frames(n: usize, frames: usize, ebn0_db: f64, seed: u64) -> Vec<Vec<Llr>> {
    let channel = AwgnChannel::from_eb_n0_db(ebn0_db, 0.5);
    let mut rng = StdRng::seed_from_u64(seed);
    (0..frames)
        .map(|_| {
            let tx = BpskModulator::modulate_bits(&vec![false; n]);
            channel.to_llrs(&channel.transmit_symbols(&tx, &mut rng))
        })
        .collect()
}

#[test]
#[ignore = "needs a HIP device"]
fn decode_batch_matches_the_cpu_decoder() {
    let qc = small_qc();
    let cpu_code = LdpcCode::from_quasi_cyclic(&qc);
    let ctx = Context::new(0).expect("HIP device 0");
    let gpu_code = HipCode::from_quasi_cyclic(&ctx, &qc).expect("code upload");
    assert_eq!((gpu_code.n(), gpu_code.m(), gpu_code.k()), (cpu_code.n(), cpu_code.m(), cpu_code.k()));
    for (ebn0, seed) in [(1.0, 1u64), (2.5, 2), (4.0, 3)] {
        let llrs = noisy_zero_frames(cpu_code.n(), 64, ebn0, seed);
        let cpu = LdpcDecoder::decode_batch(&cpu_code, &llrs, 50);
        let gpu = HipLdpcDecoder::decode_batch(&gpu_code, &llrs, 50);
        assert_eq!(cpu.len(), gpu.len());
        for (f, (a, b)) in cpu.iter().zip(gpu.iter()).enumerate() {
            assert_eq!(a.decoded_bits, b.decoded_bits, "frame {f} at {ebn0} dB: message");
            assert_eq!(a.iterations, b.iterations, "frame {f} at {ebn0} dB: iterations");
            assert_eq!(a.converged, b.converged, "frame {f} at {ebn0} dB: converged");
            assert_eq!(a.syndrome_check_passed, b.syndrome_check_passed, "frame {f} at {ebn0} dB: syndrome");
        }
    }
}

#[test]
#[ignore = "needs a HIP device"]
fn empty_batch_and_wrong_length() {
    let qc = small_qc();
    let ctx = Context::new(0).expect("HIP device 0");
    let code = HipCode::from_quasi_cyclic(&ctx, &qc).expect("code upload");
    assert!(HipLdpcDecoder::decode_batch(&code, &[], 50).is_empty());
    let short = vec![vec![Llr::new(1.0); code.n() - 1]];
    let r = std::panic::catch_unwind(std::panic::AssertUnwindSafe(|| {
        HipLdpcDecoder::decode_batch(&code, &short, 50)
    }));
    assert!(r.is_err(), "an LLR block of the wrong length panics");
}
