// Device random numbers.
//
// philox4x32_10: the Philox-4x32 counter-based generator with 10 rounds, from Salmon, Moraes,
// Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3" (SC'11): round function
// (L0,R0,L1,R1) -> (hi(M1*R1)^L0^k0, lo(M1*R1), hi(M0*R0)^L1^k1, lo(M0*R0)) ... written here
// with the paper's multipliers and Weyl key increments.
//
// The validation helpers turn raw std::mt19937_64 outputs (drawn on the host) into the values
// libstdc++ derives from them:
//   canonical_from_u64 — std::generate_canonical<double,53> for a 64-bit engine:
//                        double(x) * 2^-64, clamped below 1 (one engine call per uniform);
//   lemire_pick        — uniform_int_distribution over [0, range-1] for a 64-bit engine, Lemire's
//                        multiply-and-reject method (D. Lemire, "Fast Random Integer Generation in
//                        an Interval", ACM TOMACS 2019), consuming more draws only on rejection.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace qa {

struct Philox4 {
    uint32_t x, y, z, w;
};

__host__ __device__ __forceinline__ uint32_t mulhi32(uint32_t a, uint32_t b) {
    return static_cast<uint32_t>((static_cast<uint64_t>(a) * b) >> 32);
}

// A Philox-4x32 evaluation in progress: counter words and the round key.
struct PhiloxRun {
    uint32_t c0, c1, c2, c3, k0, k1;

    __host__ __device__ __forceinline__ void start(uint32_t a, uint32_t b, uint32_t c, uint32_t d, uint32_t key0,
                                                   uint32_t key1) {
        c0 = a;
        c1 = b;
        c2 = c;
        c3 = d;
        k0 = key0;
        k1 = key1;
    }
    // Apply R rounds (a full evaluation is 10; evaluations may be spread over several calls).
    template <int R>
    __host__ __device__ __forceinline__ void rounds() {
        constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
        constexpr uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            // one 32 x 32 -> 64-bit product per multiplier (a single v_mad_u64_u32 on the device)
            const uint64_t p0 = static_cast<uint64_t>(M0) * c0, p1 = static_cast<uint64_t>(M1) * c2;
            const uint32_t hi0 = static_cast<uint32_t>(p0 >> 32), lo0 = static_cast<uint32_t>(p0);
            const uint32_t hi1 = static_cast<uint32_t>(p1 >> 32), lo1 = static_cast<uint32_t>(p1);
            const uint32_t n0 = hi1 ^ c1 ^ k0;
            const uint32_t n2 = hi0 ^ c3 ^ k1;
            c0 = n0;
            c1 = lo1;
            c2 = n2;
            c3 = lo0;
            k0 += W0;
            k1 += W1;
        }
    }
    __host__ __device__ __forceinline__ Philox4 result() const { return Philox4{c0, c1, c2, c3}; }
};

__host__ __device__ __forceinline__ Philox4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                                  uint32_t k0, uint32_t k1) {
    PhiloxRun r;
    r.start(c0, c1, c2, c3, k0, k1);
    r.rounds<10>();
    return r.result();
}

// Uniform float in [0, 1) with 24 random bits.
__device__ __forceinline__ float u01(uint32_t x) { return static_cast<float>(x >> 8) * (1.0f / 16777216.0f); }

// Uniform integer in [0, m) from a 32-bit draw (multiply-high; bias < m / 2^32).
__host__ __device__ __forceinline__ uint32_t below(uint32_t x, uint32_t m) { return mulhi32(x, m); }

__host__ __device__ __forceinline__ double canonical_from_u64(uint64_t x) {
    double r = static_cast<double>(x) * 5.421010862427522170037264e-20;  // 2^-64, exact scaling
    if (r >= 1.0) r = 0.99999999999999988897769753748;                    // nextafter(1, 0)
    return r;
}

// Returns the pick and advances *cursor by the number of draws consumed.
__host__ __device__ __forceinline__ uint64_t lemire_pick(const uint64_t *stream, int64_t *cursor,
                                                         int64_t cap, uint64_t range, bool *ok) {
    if (*cursor >= cap) { *ok = false; return 0; }
    uint64_t x = stream[(*cursor)++];
    unsigned __int128 prod = static_cast<unsigned __int128>(x) * range;
    uint64_t low = static_cast<uint64_t>(prod);
    if (low < range) {
        const uint64_t threshold = (0 - range) % range;
        while (low < threshold) {
            if (*cursor >= cap) { *ok = false; return 0; }
            x = stream[(*cursor)++];
            prod = static_cast<unsigned __int128>(x) * range;
            low = static_cast<uint64_t>(prod);
        }
    }
    return static_cast<uint64_t>(prod >> 64);
}

}  // namespace qa
