"""CPU checks of the host-side harnesses: FER error counting, code lookup and
the autotuner's cache and candidate logic (no GPU calls)."""
import json

import numpy as np
import pytest

from gf2_amd import _lib, autotune, codes, fer


def test_count_errors_against_unpacked_bits():
    rng = np.random.default_rng(0)
    F, k, n = 6, 130, 200
    cw = rng.integers(0, 2 ** 63, size=(F, (n + 63) // 64), dtype=np.uint64)
    msg = cw[:, :(k + 63) // 64].copy()
    msg[1, 0] ^= np.uint64(1 << 5)             # one bit error
    msg[2, 2] ^= np.uint64(1 << 1)             # bit 129 < k: error
    msg[3, 2] ^= np.uint64(1 << 10)            # bit 138 >= k: ignored
    conv = np.array([1, 1, 1, 1, 0, 1], dtype=bool)
    synd = np.array([1, 1, 1, 1, 1, 0], dtype=bool)
    fe, be = fer.count_errors(msg, cw, conv, synd, k)
    # frames 1, 2 (bit errors), 4 (not converged), 5 (syndrome failed)
    assert fe == 4 and be == 2


def test_spec_lookup_and_ranges():
    assert fer.spec_by_name("n2304").n == 2304
    assert fer.spec_by_name("n16200-k9720").k == 9720
    with pytest.raises(ValueError):
        fer.spec_by_name("n999")
    assert fer._parse_range("1.0:1.5:0.25") == [1.0, 1.25, 1.5]
    assert fer._parse_range("1.1,1.3") == [1.1, 1.3]


def test_tuning_cache_roundtrip(tmp_path):
    path = tmp_path / "t.json"
    c = autotune.TuningCache(str(path))
    assert c.get("x") is None
    c.put("single:n64800:F4096", {"wg_per_cu": 3, "xcd_map": 1}, 74.4, 8)
    c.save()
    d = json.loads(path.read_text())
    assert d["single:n64800:F4096"]["config"] == {"wg_per_cu": 3, "xcd_map": 1}
    assert autotune.TuningCache(str(path)).get("single:n64800:F4096")["wg_per_cu"] == 3


def test_candidates_and_buckets():
    cands = autotune.candidates(["wg_per_cu", "xcd_map"])
    assert len(cands) == len(autotune.CANDIDATES["wg_per_cu"]) * 2
    assert {"wg_per_cu": 3, "xcd_map": 0} in cands
    assert autotune.batch_bucket(4096) == 4096 and autotune.batch_bucket(4097) == 8192
    assert autotune.batch_bucket(300) == 512 and autotune.batch_bucket(1) == 1


def test_apply_rejects_unknown_knob():
    class Dummy:
        def set_tuning(self, k, v):
            pass

        def set_path(self, p):
            pass
    autotune.apply(Dummy(), {"wg_per_cu": 2, "chunk": 8, "path": _lib.PATH_STATE})
    with pytest.raises(ValueError):
        autotune.apply(Dummy(), {"frames_per_wg": 2})


def test_simulation_result_csv_matches_reference_format():
    """simulation.rs:276-312 (test_csv_export, test_simulation_result_complete)
    and the column layout of results_to_csv (simulation.rs:209-229)."""
    from gf2_amd.fer import FerPoint, SimulationResult, results_to_csv
    r = SimulationResult(eb_n0_db=3.0, ber=0.01, num_bits=10000, num_errors=100)
    csv = results_to_csv([r], True)
    assert csv == "eb_n0_db,ber,num_bits,num_errors\n3,0.01,10000,100\n"
    assert r.is_complete(50) and not r.is_complete(200)
    p = FerPoint(ebn0_db=1.5, frames=4096, frame_errors=12, bit_errors=345, fer=12 / 4096,
                 ber=345 / (4096 * 32400), mean_iterations=24.0, decode_seconds=1.0, info_gbps=1.0)
    row = SimulationResult.from_point(p, 32400).to_csv_row().split(",")
    assert row[0] == "1.5" and row[2] == str(4096 * 32400) and row[3] == "345"
    assert float(row[1]) == p.ber and float(row[4]) == p.fer and row[5:] == ["4096", "12"]
    assert "e" not in row[1]  # Rust's f64 Display never uses an exponent
    both = results_to_csv([r, SimulationResult.from_point(p, 32400)], True)
    assert both.splitlines()[0] == "eb_n0_db,ber,num_bits,num_errors,fer,num_frames,num_frame_errors"
    assert results_to_csv([r], False) == "3,0.01,10000,100\n"
