// born.cpp — seeded stream, Born table and shot drawing of terminal sampling
// (src/sample.rs:85-189). Host code; the statevector comes from the device projection.
#include "born.h"

#include <cmath>
#include <cstring>
#include <map>

namespace wqc {

namespace {

inline uint32_t rotl32(uint32_t v, int s) { return (v << s) | (v >> (32 - s)); }

inline void quarter(uint32_t* x, int a, int b, int c, int d) {
    x[a] += x[b];
    x[d] = rotl32(x[d] ^ x[a], 16);
    x[c] += x[d];
    x[b] = rotl32(x[b] ^ x[c], 12);
    x[a] += x[b];
    x[d] = rotl32(x[d] ^ x[a], 8);
    x[c] += x[d];
    x[b] = rotl32(x[b] ^ x[c], 7);
}

// snap_amplitude_component (src/engine.rs:30-32): round half away from zero on the 1e-6 grid.
inline double snap6(double v) { return std::round(v * 1e6) / 1e6; }

}  // namespace

void chacha_block(const uint32_t key[8], uint64_t counter, uint64_t nonce, int rounds,
                  uint32_t out[16]) {
    const uint32_t init[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u,
                               key[0], key[1], key[2], key[3], key[4], key[5], key[6], key[7],
                               (uint32_t)counter, (uint32_t)(counter >> 32),
                               (uint32_t)nonce, (uint32_t)(nonce >> 32)};
    uint32_t x[16];
    std::memcpy(x, init, sizeof x);
    for (int r = 0; r < rounds; r += 2) {
        quarter(x, 0, 4, 8, 12);
        quarter(x, 1, 5, 9, 13);
        quarter(x, 2, 6, 10, 14);
        quarter(x, 3, 7, 11, 15);
        quarter(x, 0, 5, 10, 15);
        quarter(x, 1, 6, 11, 12);
        quarter(x, 2, 7, 8, 13);
        quarter(x, 3, 4, 9, 14);
    }
    for (int i = 0; i < 16; ++i) out[i] = x[i] + init[i];
}

ChaCha12Stream::ChaCha12Stream(uint64_t seed) {
    // rand_core 0.6 SeedableRng::seed_from_u64: a PCG32 (XSH-RR) step per 4-byte key word.
    uint64_t st = seed;
    for (int i = 0; i < 8; ++i) {
        st = st * 6364136223846793005ull + 11634580027462260723ull;
        const uint32_t xs = (uint32_t)(((st >> 18) ^ st) >> 27);
        const unsigned rot = (unsigned)(st >> 59);
        key_[i] = (xs >> rot) | (xs << ((32u - rot) & 31u));
    }
}

void ChaCha12Stream::refill() {
    for (int b = 0; b < 4; ++b) chacha_block(key_, block_ + (uint64_t)b, 0, 12, words_ + 16 * b);
    block_ += 4;
    pos_ = 0;
}

uint32_t ChaCha12Stream::next_u32() {
    if (pos_ >= 64) refill();
    return words_[pos_++];
}

uint64_t ChaCha12Stream::next_u64() {
    // rand_core BlockRng::next_u64: two consecutive words, low first; a pair that straddles a
    // refill takes the last word of the old buffer and the first of the new one.
    if (pos_ + 1 < 64) {
        const uint64_t lo = words_[pos_], hi = words_[pos_ + 1];
        pos_ += 2;
        return (hi << 32) | lo;
    }
    if (pos_ >= 64) {
        refill();
        pos_ = 2;
        return ((uint64_t)words_[1] << 32) | words_[0];
    }
    const uint64_t lo = words_[63];
    refill();
    pos_ = 1;
    return ((uint64_t)words_[0] << 32) | lo;
}

double ChaCha12Stream::next_f64() { return (double)(next_u64() >> 11) * (1.0 / 9007199254740992.0); }

std::vector<BornEntry> born_table(const double* sv, int n, const int32_t* mq, const int32_t* mc,
                                  size_t n_measures, size_t classical_bits) {
    (void)classical_bits;
    std::map<uint64_t, double> acc;
    const uint64_t dim = (uint64_t)1 << n;
    for (uint64_t basis = 0; basis < dim; ++basis) {
        const double re = snap6(sv[2 * basis]), im = snap6(sv[2 * basis + 1]);
        const double p = re * re + im * im;
        if (p == 0.0) continue;
        uint64_t key = 0;
        for (size_t k = 0; k < n_measures; ++k) {
            const uint64_t bit = (basis >> mq[k]) & 1u;
            const uint64_t m = (uint64_t)1 << mc[k];
            key = bit ? (key | m) : (key & ~m);
        }
        acc[key] += p;
    }
    std::vector<BornEntry> out;
    out.reserve(acc.size());
    for (const auto& kv : acc) out.push_back(BornEntry{kv.first, kv.second});
    return out;
}

std::vector<std::pair<uint64_t, uint64_t>> draw_counts(const std::vector<BornEntry>& table,
                                                       uint64_t shots, uint64_t seed) {
    std::vector<std::pair<uint64_t, uint64_t>> out;
    if (shots == 0 || table.empty()) return out;
    double total = 0.0;
    for (const auto& e : table) total += e.p;
    std::vector<double> cum(table.size());
    double run = 0.0;
    for (size_t i = 0; i < table.size(); ++i) {
        run += table[i].p / total;
        cum[i] = run;
    }
    std::vector<uint64_t> hits(table.size(), 0);
    ChaCha12Stream rng(seed);
    for (uint64_t s = 0; s < shots; ++s) {
        const double r = rng.next_f64();
        size_t i = 0;
        while (i < cum.size() && !(r <= cum[i])) ++i;
        if (i == cum.size()) i = cum.size() - 1;
        ++hits[i];
    }
    for (size_t i = 0; i < table.size(); ++i)
        if (hits[i]) out.emplace_back(table[i].key, hits[i]);
    return out;
}

}  // namespace wqc
