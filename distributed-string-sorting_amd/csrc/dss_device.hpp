// Device-side helpers shared by the kernels of libdss_gpu (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace dss {

constexpr uint32_t kNone = 0xFFFFFFFFu;

// ---------------------------------------------------------------- RNG / shuffle
// Counter-based generator and Feistel shuffle of the seeded input generators
// (identical to the host specification so host and device inputs agree).
__host__ __device__ inline uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__host__ __device__ inline uint64_t rng(uint64_t seed, uint64_t stream, uint64_t i) {
    return mix64(mix64(seed ^ (stream * 0xD1B54A32D192ED03ull)) + i);
}

__host__ __device__ inline uint64_t shuffle_index(uint64_t x, uint64_t n, uint64_t seed) {
    if (n <= 1) return x;
    unsigned bits = 0;
    while ((1ull << bits) < n) ++bits;
    if (bits & 1) ++bits;
    if (bits < 2) bits = 2;
    const unsigned half = bits / 2;
    const uint64_t mask = (1ull << half) - 1;
    do {
        uint64_t L = x >> half, R = x & mask;
        for (unsigned r = 0; r < 4; ++r) {
            uint64_t F = mix64(seed ^ ((uint64_t)(r + 1) * 0xA24BAED4963EE407ull) ^ R) & mask;
            uint64_t t = R;
            R = L ^ F;
            L = t;
        }
        x = (L << half) | R;
    } while (x >= n);
    return x;
}

// ---------------------------------------------------------------- key loads
// 8 characters of a string starting at `depth`, big endian (first char in the
// most significant byte), zero after the end of the string: the reference's
// get_uint64 (src/strings/stringset.hpp:213-239).  Two aligned 8-byte loads
// and a funnel shift; the arena is padded so the second word is readable.
__device__ inline uint64_t load_word_le(const uint8_t* __restrict__ chars, uint64_t a) {
    const uint64_t* w = reinterpret_cast<const uint64_t*>(chars + (a & ~7ull));
    const unsigned sh = (unsigned)(a & 7) * 8;
    const uint64_t lo = w[0];
    if (sh == 0) return lo;
    const uint64_t hi = w[1];
    return (lo >> sh) | (hi << (64 - sh));
}

__device__ inline uint64_t load_key8(const uint8_t* __restrict__ chars, uint64_t off,
                                     uint32_t len, uint32_t depth) {
    if (depth >= len) return 0;
    uint64_t v = __builtin_bswap64(load_word_le(chars, off + depth));
    const uint32_t rem = len - depth;
    if (rem < 8) v &= ~0ull << (8 * (8 - rem));
    return v;
}

// Four consecutive keys (32 characters) from `depth`, from five aligned
// 8-byte loads; the arena padding keeps the fifth word readable.
__device__ inline void load_keys4(const uint8_t* __restrict__ chars, uint64_t off, uint32_t len,
                                  uint32_t depth, uint64_t k[4]) {
    if (depth >= len) {
        k[0] = k[1] = k[2] = k[3] = 0;
        return;
    }
    const uint64_t a = off + depth;
    const uint64_t* w = reinterpret_cast<const uint64_t*>(chars + (a & ~7ull));
    const unsigned sh = (unsigned)(a & 7) * 8;
    uint64_t x[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) x[i] = w[i];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        uint64_t v = sh ? (x[j] >> sh) | (x[j + 1] << (64 - sh)) : x[j];
        v = __builtin_bswap64(v);
        const uint32_t d = depth + 8 * j;
        if (d >= len) v = 0;
        else if (len - d < 8) v &= ~0ull << (8 * (8 - (len - d)));
        k[j] = v;
    }
}

// Number of equal leading characters of two big-endian 8-char keys that are
// known to differ.
__device__ inline uint32_t key_lcp(uint64_t a, uint64_t b) {
    return (uint32_t)(__builtin_clzll(a ^ b) >> 3);
}

// Characters before the end of string inside a key (0..8).
__device__ inline uint32_t key_strlen(uint64_t k) {
    return k == 0 ? 0u : 8u - ((uint32_t)__builtin_ctzll(k) >> 3);
}

// Compares two string views (offset, length) from character `from` on, the
// way scmp does on NUL-terminated strings (src/strings/stringtools.hpp:108-
// 115); returns <0, 0, >0 and the LCP (total, from 0) in *lcp_out.  32
// characters per step (five aligned loads per string and step), so long
// common prefixes take few dependent steps.
__device__ inline int view_cmp32(const uint8_t* __restrict__ chars, uint64_t ao, uint32_t al, uint64_t bo,
                                 uint32_t bl, uint32_t* lcp_out, uint32_t from = 0) {
    for (uint32_t d = from;; d += 32) {
        uint64_t x[4], y[4];
        load_keys4(chars, ao, al, d, x);
        load_keys4(chars, bo, bl, d, y);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (x[j] != y[j]) {
                if (lcp_out) *lcp_out = d + 8 * j + key_lcp(x[j], y[j]);
                return x[j] < y[j] ? -1 : 1;
            }
            if ((x[j] & 0xFF) == 0) {
                if (lcp_out) *lcp_out = d + 8 * j + key_strlen(x[j]);
                return 0;
            }
        }
    }
}

// kN consecutive keys (8 kN characters) from `depth`; words past the
// string's end are not read (the arena padding covers only the five words
// of load_keys4), keys past it are zero.
template <int kN>
__device__ inline void load_keys_n(const uint8_t* __restrict__ chars, uint64_t off, uint32_t len, uint32_t depth,
                                   uint64_t k[kN]) {
    if (depth >= len) {
#pragma unroll
        for (int j = 0; j < kN; ++j) k[j] = 0;
        return;
    }
    const uint64_t a = off + depth, wa = a & ~7ull, end = off + len;
    const uint64_t* w = reinterpret_cast<const uint64_t*>(chars + wa);
    const unsigned sh = (unsigned)(a & 7) * 8;
    uint64_t x[kN + 1];
#pragma unroll
    for (int i = 0; i <= kN; ++i) x[i] = wa + 8 * (uint64_t)i < end ? w[i] : 0ull;
#pragma unroll
    for (int j = 0; j < kN; ++j) {
        uint64_t v = sh ? (x[j] >> sh) | (x[j + 1] << (64 - sh)) : x[j];
        v = __builtin_bswap64(v);
        const uint32_t d = depth + 8 * j;
        if (d >= len) v = 0;
        else if (len - d < 8) v &= ~0ull << (8 * (8 - (len - d)));
        k[j] = v;
    }
}

// ---------------------------------------------------------------- wave helpers
__device__ inline uint32_t lane_id() { return __lane_id(); }

__device__ inline uint64_t lanemask_lt() {
    const uint32_t l = lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// Comparison of two strings by a whole wave: lane l compares the 8-char keys
// at 8l of a 512-character window, a ballot finds the first difference or
// end.  Every lane of the wave must call it with the same arguments (the
// loop is controlled by the ballot).
__device__ inline int wave_cmp(const uint8_t* __restrict__ chars, uint64_t ao, uint32_t al, uint64_t bo,
                               uint32_t bl) {
    const uint32_t lane = lane_id();
    for (uint32_t base = 0;; base += 512) {
        const uint32_t pos = base + 8 * lane;
        const uint64_t a = load_key8(chars, ao, al, pos);
        const uint64_t b = load_key8(chars, bo, bl, pos);
        const uint64_t m = __ballot(a != b || (a & 0xFF) == 0);
        if (m) {
            const int f = __ffsll((unsigned long long)m) - 1;
            const uint64_t fa = __shfl((unsigned long long)a, f), fb = __shfl((unsigned long long)b, f);
            return fa < fb ? -1 : (fa > fb ? 1 : 0);
        }
    }
}

// LCP of two strings by a whole wave (same window scheme as wave_cmp; every
// lane of the wave must call it with the same arguments), for strings known
// to agree on their first `from` characters.
__device__ inline uint32_t wave_lcp(const uint8_t* __restrict__ chars, uint64_t ao, uint32_t al, uint64_t bo,
                                    uint32_t bl, uint32_t from = 0) {
    const uint32_t lane = lane_id();
    for (uint32_t base = from;; base += 512) {
        const uint32_t pos = base + 8 * lane;
        const uint64_t a = load_key8(chars, ao, al, pos);
        const uint64_t b = load_key8(chars, bo, bl, pos);
        const uint64_t m = __ballot(a != b || (a & 0xFF) == 0);
        if (m) {
            const int f = __ffsll((unsigned long long)m) - 1;
            const uint64_t fa = __shfl((unsigned long long)a, f), fb = __shfl((unsigned long long)b, f);
            return base + 8 * (uint32_t)f + (fa != fb ? key_lcp(fa, fb) : key_strlen(fa));
        }
    }
}

}  // namespace dss
