// rand 0.8's StdRng and Uniform<f64>, as GsInit::Random draws its starting
// phases (crates/optics-inverse/src/gs.rs:449-458: StdRng::seed_from_u64,
// Uniform::new(0.0, span), lo + sample).  The crates are not vendored with
// the reference; Cargo.lock pins rand 0.8.8, rand_chacha 0.3.1 and
// rand_core 0.6.4, whose published algorithms this restates:
//  - rand_core 0.6 SeedableRng::seed_from_u64: eight PCG32 outputs
//    (multiplier 6364136223846793005, increment 11634580027462260723, state
//    advanced before each output, XSH-RR output) fill the 32-byte seed,
//    each as four little-endian bytes;
//  - rand_chacha 0.3 ChaCha12Rng: the ChaCha block function with 12 rounds
//    (six double rounds), the seed as the eight key words, a 64-bit block
//    counter from 0 in words 12-13 and a zero stream id in words 14-15;
//    the output words of consecutive blocks in order, and next_u64 takes
//    two of them, low word first;
//  - rand 0.8 UniformFloat<f64>: new() shrinks the scale by one ulp until
//    scale * (1 - 2^-52) + low < high; sample() maps the top 52 bits of a
//    u64 to [1, 2) through the exponent, subtracts 1 and returns
//    value * scale + low, two roundings (no FMA: the library is built with
//    -ffp-contract=off).
// Host code only: the starting phases are made once per run, on the CPU,
// as in the reference.
#pragma once

#include <cstdint>
#include <cstring>

namespace och {

class StdRng {
public:
    explicit StdRng(uint64_t seed) {
        for (int i = 0; i < 8; ++i) {
            seed = seed * 6364136223846793005ull + 11634580027462260723ull;
            const uint32_t xs = (uint32_t)(((seed >> 18) ^ seed) >> 27);
            const uint32_t rot = (uint32_t)(seed >> 59);
            key_[i] = (xs >> rot) | (xs << ((32u - rot) & 31u));
        }
    }

    uint64_t next_u64() {
        if (pos_ >= 16) refill();
        const uint64_t lo = words_[pos_], hi = words_[pos_ + 1];
        pos_ += 2;
        return lo | (hi << 32);
    }

private:
    static uint32_t rotl(uint32_t v, int c) { return (v << c) | (v >> (32 - c)); }
    static void quarter(uint32_t* x, int a, int b, int c, int d) {
        x[a] += x[b]; x[d] = rotl(x[d] ^ x[a], 16);
        x[c] += x[d]; x[b] = rotl(x[b] ^ x[c], 12);
        x[a] += x[b]; x[d] = rotl(x[d] ^ x[a], 8);
        x[c] += x[d]; x[b] = rotl(x[b] ^ x[c], 7);
    }
    void refill() {
        uint32_t in[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u};
        for (int i = 0; i < 8; ++i) in[4 + i] = key_[i];
        in[12] = (uint32_t)counter_;
        in[13] = (uint32_t)(counter_ >> 32);
        in[14] = in[15] = 0u;
        uint32_t x[16];
        std::memcpy(x, in, sizeof x);
        for (int r = 0; r < 6; ++r) {
            quarter(x, 0, 4, 8, 12); quarter(x, 1, 5, 9, 13);
            quarter(x, 2, 6, 10, 14); quarter(x, 3, 7, 11, 15);
            quarter(x, 0, 5, 10, 15); quarter(x, 1, 6, 11, 12);
            quarter(x, 2, 7, 8, 13); quarter(x, 3, 4, 9, 14);
        }
        for (int i = 0; i < 16; ++i) words_[i] = x[i] + in[i];
        ++counter_;
        pos_ = 0;
    }

    uint32_t key_[8];
    uint32_t words_[16] = {};
    uint64_t counter_ = 0;
    int pos_ = 16;
};

// Uniform::new(low, high) for f64 (finite low < high), then sample().
class UniformF64 {
public:
    UniformF64(double low, double high) : low_(low), scale_(high - low) {
        const double max_rand = 1.0 - 0x1p-52;
        for (;;) {
            const double top = scale_ * max_rand;
            if (!(top + low_ >= high)) break;
            uint64_t b;
            std::memcpy(&b, &scale_, 8);
            --b;
            std::memcpy(&scale_, &b, 8);
        }
    }
    double sample(StdRng& rng) const {
        const uint64_t bits = (rng.next_u64() >> 12) | 0x3ff0000000000000ull;
        double v;
        std::memcpy(&v, &bits, 8);
        const double v01 = v - 1.0;
        const double prod = v01 * scale_;
        return prod + low_;
    }

private:
    double low_, scale_;
};

// GsInit::Random's starting phases (gs.rs:449-458) for slot bounds lo, hi.
inline void gs_random_phase(uint64_t seed, double lo, double hi, size_t n, double* out) {
    const double tau = 6.283185307179586;
    double span = hi - lo;
    span = span < 0.0 ? 0.0 : (span > tau ? tau : span);
    if (span > 0.0) {
        StdRng rng(seed);
        const UniformF64 u(0.0, span);
        for (size_t k = 0; k < n; ++k) out[k] = lo + u.sample(rng);
    } else {
        for (size_t k = 0; k < n; ++k) out[k] = lo;
    }
}

}  // namespace och
