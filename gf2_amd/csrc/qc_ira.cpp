// Seeded quasi-cyclic IRA base-matrix generator (the project's synthetic code
// family; the reference builds QC codes from a caller-given base matrix,
// crates/gf2-coding/src/ldpc/core.rs:485-606, and this produces such a
// matrix).
//
// Layout of the mb x nb base matrix (kb = nb - mb info block-columns first,
// so the first k = kb*z code bits are the information bits, as
// decode_iterative assumes at core.rs:931-936):
//   info block-column c  weight info_col_weight[c]; its block-rows are the
//                        ones with the fewest edges so far (random
//                        tie-break), so check degrees stay balanced; shifts
//                        uniform in [0, z), redrawn up to 64 times while the
//                        new block closes a length-4 cycle
//   parity column kb+j   identity blocks at block-rows j and j+1 (the last
//                        parity column has weight 1): the dual-diagonal
//                        staircase, encodable by running XOR.
#include <algorithm>
#include <cstdint>
#include <vector>

#include "gf2ldpc.h"
#include "internal.h"

namespace {

struct SplitMix64 {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
};

// Would shift s at (a, c) close a 4-cycle a-c-b-d with existing blocks?
bool closes_4cycle(const std::vector<int32_t> &sh, uint32_t mb, uint32_t nb, uint32_t z, uint32_t a,
                   uint32_t c, int64_t s) {
    for (uint32_t b = 0; b < mb; ++b) {
        if (b == a) continue;
        const int32_t sbc = sh[(size_t)b * nb + c];
        if (sbc < 0) continue;
        for (uint32_t d = 0; d < nb; ++d) {
            if (d == c) continue;
            const int32_t sad = sh[(size_t)a * nb + d], sbd = sh[(size_t)b * nb + d];
            if (sad < 0 || sbd < 0) continue;
            int64_t t = (s - sbc + sbd - sad) % (int64_t)z;
            if (t == 0) return true;
        }
    }
    return false;
}

}  // namespace

extern "C" int gf2l_qc_ira_shifts(uint64_t seed, uint32_t mb, uint32_t nb, uint32_t z,
                                  const uint32_t *info_col_weight, int32_t *shifts_out) {
    return gf2l::guard(nullptr, [&]() -> int {
        if (!info_col_weight || !shifts_out || mb == 0 || z == 0 || nb <= mb) return GF2L_ERR_INVALID;
        const uint32_t kb = nb - mb;
        for (uint32_t c = 0; c < kb; ++c)
            if (info_col_weight[c] == 0 || info_col_weight[c] > mb) return GF2L_ERR_INVALID;
        std::vector<int32_t> sh((size_t)mb * nb, -1);
        std::vector<uint32_t> deg(mb, 0);
        for (uint32_t j = 0; j < mb; ++j) {
            sh[(size_t)j * nb + kb + j] = 0;
            ++deg[j];
            if (j + 1 < mb) {
                sh[(size_t)(j + 1) * nb + kb + j] = 0;
                ++deg[j + 1];
            }
        }
        SplitMix64 rng{seed};
        std::vector<std::pair<uint64_t, uint32_t>> order(mb);
        for (uint32_t c = 0; c < kb; ++c) {
            for (uint32_t r = 0; r < mb; ++r) order[r] = {((uint64_t)deg[r] << 40) | (rng.next() >> 24), r};
            std::sort(order.begin(), order.end());
            std::vector<uint32_t> rows;
            for (uint32_t t = 0; t < info_col_weight[c]; ++t) rows.push_back(order[t].second);
            std::sort(rows.begin(), rows.end());
            for (uint32_t r : rows) {
                int64_t s = 0;
                for (int attempt = 0; attempt < 64; ++attempt) {
                    s = (int64_t)(rng.next() % z);
                    if (!closes_4cycle(sh, mb, nb, z, r, c, s)) break;
                }
                sh[(size_t)r * nb + c] = (int32_t)s;
                ++deg[r];
            }
        }
        std::copy(sh.begin(), sh.end(), shifts_out);
        return GF2L_OK;
    });
}
