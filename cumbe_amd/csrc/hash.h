// Order-independent 64-bit hash of a set of maximal bicliques.
//
// The reference never materialises bicliques (it only counts them,
// src/mbe_cuMBE.cuh:295-300), so this engine defines the hash it reports:
//   term_U(a) = mix64(a ^ SALT_U),  term_V(b) = mix64(b ^ SALT_V)   (file ids)
//   S_U = sum term_U over A,  S_V = sum term_V over B                (mod 2^64)
//   h(A, B) = mix64(S_U ^ rotl64(S_V, 29) ^ (|A| K1 + |B| K2))
//   H = sum of h over all maximal bicliques                           (mod 2^64)
// mix64 is the SplitMix64 finaliser (Steele, Lea, Flood, OOPSLA 2014). The
// sums make h a function of the sorted sets, H independent of enumeration
// order and of which side the engine enumerates, and partial results from
// any partition of the work add up exactly.
#pragma once
#include <cstdint>

#if defined(__HIPCC__)
#define MBE_HD __host__ __device__ __forceinline__
#else
#define MBE_HD inline
#endif

namespace mbe {

constexpr uint64_t kSaltU = 0x9E3779B97F4A7C15ULL;
constexpr uint64_t kSaltV = 0xC2B2AE3D27D4EB4FULL;
constexpr uint64_t kK1 = 0x165667B19E3779F9ULL;
constexpr uint64_t kK2 = 0xD6E8FEB86659FD93ULL;

MBE_HD uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

// Inverse of mix64 (each step of the finaliser is invertible: x ^ (x >> s)
// by repeated shifts, the odd multipliers by their inverses mod 2^64, found
// by Newton's iteration y <- y (2 - c y), which doubles the correct low bits).
constexpr uint64_t inv_odd64(uint64_t c) {
    uint64_t y = c;  // correct to 3 bits for odd c
    for (int i = 0; i < 5; i++) y *= 2 - c * y;
    return y;
}
inline uint64_t unmix64(uint64_t z) {
    z = z ^ (z >> 31) ^ (z >> 62);
    z *= inv_odd64(0x94d049bb133111ebULL);
    z = z ^ (z >> 27) ^ (z >> 54);
    z *= inv_odd64(0xbf58476d1ce4e5b9ULL);
    return z ^ (z >> 30) ^ (z >> 60);
}

MBE_HD uint64_t rotl64(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }

// h(A, B) from the side sums and sizes.
MBE_HD uint64_t biclique_hash(uint64_t s_u, uint64_t s_v, uint64_t n_a, uint64_t n_b) {
    return mix64(s_u ^ rotl64(s_v, 29) ^ (n_a * kK1 + n_b * kK2));
}

}  // namespace mbe
