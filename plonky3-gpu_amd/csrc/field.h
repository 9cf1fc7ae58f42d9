// Device field arithmetic for gfx950 (32-bit integer VALU; no MFMA -- this is
// exact modular arithmetic, not a floating-point contraction).
//
// BabyBear  p = 2^31 - 2^27 + 1.  Buffers hold canonical values in [0, p).
//   Multiplication uses Montgomery reduction with R = 2^32; a "Montgomery
//   constant" c_m = c * R mod p multiplies a canonical x to the canonical
//   x * c (mont(x, c_m) = x * c * R * R^-1).  Twiddle and coset tables are
//   therefore stored in Montgomery form and the data stays canonical.
// Goldilocks p = 2^64 - 2^32 + 1.  Canonical u64 values; products are reduced
//   with 2^64 = 2^32 - 1 and 2^96 = -1 (mod p).
//
// The reference's scalar forms are baby_bear.hpp:75-88 and goldilocks.hpp:62-189.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace p3h {

struct BabyBear {
    using T = uint32_t;
    static constexpr uint32_t P = 0x78000001u;
    static constexpr uint32_t NEG_PINV = 0x77FFFFFFu;  // -p^-1 mod 2^32
    static constexpr uint32_t R_MOD_P = 0x0FFFFFFEu;   // 2^32 mod p
    static constexpr uint32_t R2_MOD_P = 0x45DDDDE3u;  // 2^64 mod p

    // [0, 2p) -> [0, p): for x < p, x - p wraps above 2^32 - p > x.
    __device__ __forceinline__ static uint32_t reduce2p(uint32_t x) { return min(x, x - P); }
    __device__ __forceinline__ static uint32_t add(uint32_t a, uint32_t b) { return reduce2p(a + b); }
    __device__ __forceinline__ static uint32_t sub(uint32_t a, uint32_t b) {
        uint32_t d = a - b;
        return min(d, d + P);
    }
    // a * b * 2^-32 mod p in [0, 2p).  Valid when a * b < 2^32 * p, which holds
    // for a < 2p and b < p (t + m*p then stays below 2^64).
    __device__ __forceinline__ static uint32_t mont_lazy(uint32_t a, uint32_t b) {
        uint64_t t = uint64_t(a) * b;
        uint32_t m = uint32_t(t) * NEG_PINV;
        return uint32_t((t + uint64_t(m) * P) >> 32);
    }
    __device__ __forceinline__ static uint32_t mont(uint32_t a, uint32_t b) { return reduce2p(mont_lazy(a, b)); }
    // canonical x times a constant held in Montgomery form
    __device__ __forceinline__ static uint32_t mul_c(uint32_t x, uint32_t c_mont) { return mont(x, c_mont); }
    __device__ __forceinline__ static uint32_t to_mont(uint32_t x) { return mont(x, R2_MOD_P); }
    __device__ __forceinline__ static uint32_t from_mont(uint32_t x) { return mont(x, 1u); }
};

struct Goldilocks {
    using T = uint64_t;
    static constexpr uint64_t P = 0xFFFFFFFF00000001ull;
    static constexpr uint64_t EPS = 0xFFFFFFFFull;  // 2^64 mod p

    __device__ __forceinline__ static uint64_t canon(uint64_t x) { return x >= P ? x - P : x; }
    __device__ __forceinline__ static uint64_t add(uint64_t a, uint64_t b) {
        uint64_t s = a + b;
        if (s < a) return s + EPS;  // wrapped: a + b - 2^64 + (2^64 mod p); below p
        return canon(s);
    }
    __device__ __forceinline__ static uint64_t sub(uint64_t a, uint64_t b) {
        uint64_t d = a - b;
        return a < b ? d + P : d;  // d + p wraps back into [0, p)
    }
    // (hi, lo) = 128-bit value; returns a value below 2^64 congruent to it mod p
    // (not necessarily canonical).  Neither correction can wrap again: after a
    // borrow t0 >= 2^64 - 2^33, and t0 + t1 - 2^64 + EPS < 2^64 - 1.
    __device__ __forceinline__ static uint64_t reduce128_lazy(uint64_t hi, uint64_t lo) {
        uint64_t hi_hi = hi >> 32, hi_lo = hi & EPS;
        uint64_t t0 = lo - hi_hi;
        if (lo < hi_hi) t0 -= EPS;                   // borrow: -2^64 == -(2^32 - 1)
        uint64_t t1 = (hi_lo << 32) - hi_lo;         // hi_lo * (2^32 - 1), no overflow
        uint64_t t2 = t0 + t1;
        if (t2 < t0) t2 += EPS;                      // carry: +2^64 == +(2^32 - 1)
        return t2;
    }
    // (hi, lo) = 128-bit value; returns it mod p, canonical.
    __device__ __forceinline__ static uint64_t reduce128(uint64_t hi, uint64_t lo) {
        return canon(reduce128_lazy(hi, lo));
    }
    // the 128-bit product from four 32 x 32 -> 64 partial products (each one
    // v_mad_u64_u32 with the running carry as its addend):
    // a b = a1 b1 2^64 + (a0 b1 + a1 b0) 2^32 + a0 b0; m0 and m1 stay below
    // (2^32 - 1)^2 + 2^32 - 1 < 2^64.
    __device__ __forceinline__ static void mul128(uint64_t a, uint64_t b, uint64_t& hi, uint64_t& lo) {
        const uint32_t a0 = uint32_t(a), a1 = uint32_t(a >> 32), b0 = uint32_t(b), b1 = uint32_t(b >> 32);
        const uint64_t p00 = uint64_t(a0) * b0;
        const uint64_t m0 = uint64_t(a0) * b1 + (p00 >> 32);
        const uint64_t m1 = uint64_t(a1) * b0 + uint32_t(m0);
        hi = uint64_t(a1) * b1 + (m0 >> 32) + (m1 >> 32);
        lo = (m1 << 32) | uint32_t(p00);
    }
    __device__ __forceinline__ static uint64_t mul(uint64_t a, uint64_t b) {
        uint64_t hi, lo;
        mul128(a, b, hi, lo);
        return reduce128(hi, lo);
    }
    // any a, b < 2^64 -> a*b mod p as a value below 2^64 (not necessarily canonical)
    __device__ __forceinline__ static uint64_t mul_lazy(uint64_t a, uint64_t b) {
        uint64_t hi, lo;
        mul128(a, b, hi, lo);
        return reduce128_lazy(hi, lo);
    }
    // x * 2^i mod p, canonical, for 1 <= i < 32: the product's high word is
    // below 2^i, so only the hi * (2^32 - 1) term of the reduction remains.
    __device__ __forceinline__ static uint64_t mul_pow2(uint64_t x, uint32_t i) {
        const uint64_t lo = x << i, hi = x >> (64 - i);
        const uint64_t t1 = (hi << 32) - hi;
        uint64_t t2 = lo + t1;
        if (t2 < lo) t2 += EPS;
        return canon(t2);
    }
    __device__ __forceinline__ static uint64_t mul_c(uint64_t x, uint64_t c) { return mul(x, c); }
};

}  // namespace p3h
