// born.h — host side of terminal sampling: the reference's seeded random stream, the Born
// table of a projected statevector and shot drawing (src/sample.rs:85-189).
#pragma once

#include <cstddef>
#include <cstdint>
#include <utility>
#include <vector>

namespace wqc {

// rand 0.8 StdRng seeded with StdRng::seed_from_u64 (src/sample.rs:177): ChaCha with 12
// rounds, key from rand_core 0.6's PCG32 expansion of the u64, stream 0, 64-bit block
// counter, buffered four blocks at a time as rand_chacha does; next_f64 is rand's Standard
// f64, the top 53 bits of next_u64 times 2^-53.
class ChaCha12Stream {
public:
    explicit ChaCha12Stream(uint64_t seed);
    uint32_t next_u32();
    uint64_t next_u64();
    double next_f64();

private:
    void refill();
    uint32_t key_[8];
    uint64_t block_ = 0;
    uint32_t words_[64];
    unsigned pos_ = 64;
};

// One 64-byte ChaCha block (RFC 8439 layout with a 64-bit counter and 64-bit nonce).
void chacha_block(const uint32_t key[8], uint64_t counter, uint64_t nonce, int rounds,
                  uint32_t out[16]);

// Outcome key: bit c holds classical bit c, so the numeric order of keys is the
// lexicographic order of the reference's Qiskit bitstrings (cbit 0 rightmost,
// src/sample.rs:144-155) for registers up to 64 bits.
struct BornEntry {
    uint64_t key;
    double p;
};

// compute_outcome_probabilities (src/sample.rs:85-131): amplitudes snapped to the 1e-6
// grid, p = re^2 + im^2, zero entries skipped, summed per key in basis order; returned in
// key order. sv holds 2^n interleaved (re, im) pairs; measure k maps qubit mq[k] to cbit
// mc[k].
std::vector<BornEntry> born_table(const double* sv, int n, const int32_t* mq, const int32_t* mc,
                                  size_t n_measures, size_t classical_bits);

// sample_counts_from_probabilities (src/sample.rs:158-189): cumulative p / total in key order,
// one next_f64 per shot, the first entry with r <= cumulative (else the last one).
std::vector<std::pair<uint64_t, uint64_t>> draw_counts(const std::vector<BornEntry>& table,
                                                       uint64_t shots, uint64_t seed);

}  // namespace wqc
