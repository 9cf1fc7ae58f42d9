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
    // no power-of-two roots of unity to exploit in the NTT
    static constexpr bool POW2_ROOTS = false;
    static constexpr int LOG2_ROOT64_FWD = 0, LOG2_ROOT64_INV = 0;

    // [0, 2p) -> [0, p): x - p, or x itself when that subtraction borrows.
    // The borrow and the select go through VCC (v_sub_co_u32, v_cndmask_b32):
    // on gfx950 that pair issues at about half the cycles of the
    // v_sub_u32 + v_min_u32 form min(x, x - p), whose v_min_u32 issues at the
    // rate of a multiply.
    __device__ __forceinline__ static uint32_t reduce2p(uint32_t x) {
        uint32_t t;
        asm("v_sub_co_u32 %0, vcc, %1, %2\n\t"
            "v_cndmask_b32 %0, %0, %1, vcc"
            : "=&v"(t)
            : "v"(x), "v"(P)
            : "vcc");
        return t;
    }
    __device__ __forceinline__ static uint32_t add(uint32_t a, uint32_t b) { return reduce2p(a + b); }
    // a - b, plus p when the subtraction borrows
    __device__ __forceinline__ static uint32_t sub(uint32_t a, uint32_t b) {
        uint32_t d, e;
        asm("v_sub_co_u32 %0, vcc, %2, %3\n\t"
            "v_add_u32 %1, %0, %4\n\t"
            "v_cndmask_b32 %0, %0, %1, vcc"
            : "=&v"(d), "=&v"(e)
            : "v"(a), "v"(b), "v"(P)
            : "vcc");
        return d;
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
    // -c of a canonical value (or of its multiply-constant form)
    __device__ __forceinline__ static uint32_t neg_c(uint32_t c) { return c ? P - c : 0u; }
    __device__ __forceinline__ static uint32_t to_mont(uint32_t x) { return mont(x, R2_MOD_P); }
    __device__ __forceinline__ static uint32_t from_mont(uint32_t x) { return mont(x, 1u); }

    // NTT butterflies (ntt_impl.h) on canonical words; w in multiply-constant form.
    // dit: (a, b) -> (a + w b, a - w b); dif: (a, b) -> (a + b, (a - b) w);
    // the *1 forms are the same with w = 1.
    __device__ __forceinline__ static void dit(uint32_t& a, uint32_t& b, uint32_t w) {
        const uint32_t t = mul_c(b, w), x = a;
        a = add(x, t);
        b = sub(x, t);
    }
    __device__ __forceinline__ static void dif(uint32_t& a, uint32_t& b, uint32_t w) {
        const uint32_t x = a, y = b;
        a = add(x, y);
        b = mul_c(sub(x, y), w);
    }
    __device__ __forceinline__ static void bf1(uint32_t& a, uint32_t& b) {
        const uint32_t x = a, y = b;
        a = add(x, y);
        b = sub(x, y);
    }
    // NTT words are canonical throughout
    __device__ __forceinline__ static uint32_t mul_ntt(uint32_t x, uint32_t c) { return mul_c(x, c); }
    __device__ __forceinline__ static uint32_t store_form(uint32_t x) { return x; }
};

struct Goldilocks {
    using T = uint64_t;
    static constexpr uint64_t P = 0xFFFFFFFF00000001ull;
    static constexpr uint64_t EPS = 0xFFFFFFFFull;  // 2^64 mod p
    // 2 has order 192 (2^96 = -1), so every 64th root of unity is a power of
    // two.  With generator 7 (omega_{2^32} = 7^((p-1)/2^32)) the NTT's
    // omega_64 is 2^39 and omega_64^-1 = 2^153; omega_{2^(j+1)}^u is then
    // 2^(39 u 2^(5-j) mod 192) for j <= 5.
    static constexpr bool POW2_ROOTS = true;
    static constexpr int LOG2_ROOT64_FWD = 39, LOG2_ROOT64_INV = 153;

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
    // -c of a canonical value
    __device__ __forceinline__ static uint64_t neg_c(uint64_t c) { return c ? P - c : 0ull; }

    // ---- NTT words: lazily reduced (any value below 2^64 congruent mod p) ----
    // The carry chains are written in gfx950 assembly, one block per
    // operation with VCC as the carry: the compiler's own lowering of
    // __builtin_addc / __builtin_subc chains folds a borrow into the next
    // addition as an added 0xFFFFFFFF and then takes that addition's carry,
    // which is wrong after a borrow (seen on canon_lz(sub_lz(a, b))).  Every
    // instruction here is a 32-bit VALU op (v_add_co / v_addc / v_sub_co /
    // v_subb / v_cndmask); no 64-bit compare or (x, 0) register pair.
    __device__ __forceinline__ static uint64_t pack(uint32_t lo, uint32_t hi) { return (uint64_t(hi) << 32) | lo; }
    // canonical form: x + (2^32 - 1) carries out of 64 bits exactly when x >= p,
    // and is then x - p
    __device__ __forceinline__ static uint64_t canon_lz(uint64_t x) {
        uint32_t l, h;
        asm("v_add_co_u32 %0, vcc, -1, %2\n\t"
            "v_addc_co_u32 %1, vcc, 0, %3, vcc\n\t"
            "v_cndmask_b32 %0, %2, %0, vcc\n\t"
            "v_cndmask_b32 %1, %3, %1, vcc"
            : "=&v"(l), "=&v"(h)
            : "v"(uint32_t(x)), "v"(uint32_t(x >> 32))
            : "vcc");
        return pack(l, h);
    }
    // a lazy, b canonical: a + b, lazy.  A carry out of 2^64 is worth
    // 2^32 - 1; the wrapped sum is below p, so the correction cannot wrap.
    __device__ __forceinline__ static uint64_t add_lz(uint64_t a, uint64_t b) {
        uint32_t l, h, m;
        asm("v_add_co_u32 %0, vcc, %3, %5\n\t"
            "v_addc_co_u32 %1, vcc, %4, %6, vcc\n\t"
            "v_cndmask_b32_e64 %2, 0, -1, vcc\n\t"
            "v_add_co_u32 %0, vcc, %0, %2\n\t"
            "v_addc_co_u32 %1, vcc, 0, %1, vcc"
            : "=&v"(l), "=&v"(h), "=&v"(m)
            : "v"(uint32_t(a)), "v"(uint32_t(a >> 32)), "v"(uint32_t(b)), "v"(uint32_t(b >> 32))
            : "vcc");
        return pack(l, h);
    }
    // a lazy, b canonical: a - b, lazy.  A borrow (a < b) leaves a - b + 2^64,
    // which is at least 2^64 - p + 1 = 2^32, so subtracting 2^32 - 1 cannot wrap.
    __device__ __forceinline__ static uint64_t sub_lz(uint64_t a, uint64_t b) {
        uint32_t l, h, m;
        asm("v_sub_co_u32 %0, vcc, %3, %5\n\t"
            "v_subb_co_u32 %1, vcc, %4, %6, vcc\n\t"
            "v_cndmask_b32_e64 %2, 0, -1, vcc\n\t"
            "v_sub_co_u32 %0, vcc, %0, %2\n\t"
            "v_subbrev_co_u32 %1, vcc, 0, %1, vcc"
            : "=&v"(l), "=&v"(h), "=&v"(m)
            : "v"(uint32_t(a)), "v"(uint32_t(a >> 32)), "v"(uint32_t(b)), "v"(uint32_t(b >> 32))
            : "vcc");
        return pack(l, h);
    }
    // (hi, lo) = 128-bit value r3 r2 r1 r0 (32-bit limbs): with 2^64 = 2^32 - 1
    // and 2^96 = -1 (mod p), x = (r1 r0) - r3 + r2 (2^32 - 1), lazy.  A borrow
    // of the subtraction adds 2^64 - (2^32 - 1) (the difference is then at
    // least 2^64 - 2^32 + 1, so it cannot wrap); a carry of the addition adds
    // 2^32 - 1 (the wrapped sum is below 2^64 - 2^33 + 2, so it cannot wrap).
    // Both corrections take their condition from VCC: the same value as
    // reduce128_lazy without its 64-bit compares (each needs r3 moved into an
    // (r3, 0) register pair) or the move that packs lo; u = r2 (2^32 - 1) is
    // one v_mad_u64_u32.
    __device__ __forceinline__ static uint64_t reduce128_lz(uint64_t hi, uint32_t r0, uint32_t r1) {
        const uint64_t u = uint64_t(uint32_t(hi)) * 0xFFFFFFFFull;
        uint32_t tl, th, m;
        asm("v_sub_co_u32 %0, vcc, %3, %5\n\t"          // t = (r1 r0) - r3
            "v_subbrev_co_u32 %1, vcc, 0, %4, vcc\n\t"
            "v_cndmask_b32_e64 %2, 0, -1, vcc\n\t"       // borrow: t += 0xFFFFFFFF00000001
            "v_addc_co_u32 %0, vcc, 0, %0, vcc\n\t"
            "v_addc_co_u32 %1, vcc, %2, %1, vcc\n\t"
            "v_add_co_u32 %0, vcc, %0, %6\n\t"           // t += u
            "v_addc_co_u32 %1, vcc, %1, %7, vcc\n\t"
            "v_cndmask_b32_e64 %2, 0, -1, vcc\n\t"       // carry: t += 2^32 - 1
            "v_add_co_u32 %0, vcc, %0, %2\n\t"
            "v_addc_co_u32 %1, vcc, 0, %1, vcc"
            : "=&v"(tl), "=&v"(th), "=&v"(m)
            : "v"(r0), "v"(r1), "v"(uint32_t(hi >> 32)), "v"(uint32_t(u)), "v"(uint32_t(u >> 32))
            : "vcc");
        return pack(tl, th);
    }
    // any a, b < 2^64 -> a b mod p below 2^64 (lazy): mul128's partial
    // products, whose carry addends the compiler places in v_mad_u64_u32, with
    // the reduction of reduce128_lz.  Every 32-bit limb sum as an assembly
    // carry chain instead (no addends) issued slower in the NTT and the hash.
    __device__ __forceinline__ static uint64_t mul_lazy_vcc(uint64_t a, uint64_t b) {
        const uint32_t a0 = uint32_t(a), a1 = uint32_t(a >> 32), b0 = uint32_t(b), b1 = uint32_t(b >> 32);
        const uint64_t p00 = uint64_t(a0) * b0;
        const uint64_t m0 = uint64_t(a0) * b1 + (p00 >> 32);
        const uint64_t m1 = uint64_t(a1) * b0 + uint32_t(m0);
        const uint64_t hi = uint64_t(a1) * b1 + (m0 >> 32) + (m1 >> 32);
        return reduce128_lz(hi, uint32_t(p00), uint32_t(m1));
    }

    // x lazy, 0 <= SH < 96: x * 2^SH mod p, lazy.  With x * 2^(SH mod 32) =
    // q2 q1 q0 (32-bit limbs; q2 < 2^31) and 2^64 = 2^32 - 1, 2^96 = -1:
    //   SH < 32: (q1 q0) + q2 (2^32 - 1)
    //   SH < 64: (q0 0) + q1 (2^32 - 1) - q2
    //   SH < 96: q0 (2^32 - 1) - (q2 q1)
    // u = q (2^32 - 1) is q 2^32 - q: (q - [q > 0], -q).  A carry out of 2^64
    // adds 2^32 - 1 and a borrow subtracts it; in each case the corrected word
    // stays inside [0, 2^64) (the wrapped sum is below 2^64 - 2^32, the wrapped
    // difference at least 2^63), so one correction suffices.
    template <int SH>
    __device__ __forceinline__ static uint64_t mul_pow2_lz(uint64_t x) {
        static_assert(SH >= 0 && SH < 96, "shift below 96");
        if constexpr (SH == 0) {
            return x;
        } else {
            constexpr int S1 = SH & 31;
            const uint32_t x0 = uint32_t(x), x1 = uint32_t(x >> 32);
            uint32_t q0, q1, q2;
            if constexpr (S1 == 0) {
                q0 = x0;
                q1 = x1;
                q2 = 0;
            } else {
                q0 = x0 << S1;
                q1 = (x1 << S1) | (x0 >> (32 - S1));
                q2 = x1 >> (32 - S1);
            }
            uint32_t tl, th, ul, uh, m;
            if constexpr (SH < 32) {
                asm("v_sub_co_u32 %2, vcc, 0, %7\n\t"          // u = q2 2^32 - q2
                    "v_subbrev_co_u32 %3, vcc, 0, %7, vcc\n\t"
                    "v_add_co_u32 %0, vcc, %5, %2\n\t"         // t = (q1 q0) + u
                    "v_addc_co_u32 %1, vcc, %6, %3, vcc\n\t"
                    "v_cndmask_b32_e64 %4, 0, -1, vcc\n\t"     // carry: t += 2^32 - 1
                    "v_add_co_u32 %0, vcc, %0, %4\n\t"
                    "v_addc_co_u32 %1, vcc, 0, %1, vcc"
                    : "=&v"(tl), "=&v"(th), "=&v"(ul), "=&v"(uh), "=&v"(m)
                    : "v"(q0), "v"(q1), "v"(q2)
                    : "vcc");
            } else if constexpr (SH < 64 && S1 == 0) {
                asm("v_sub_co_u32 %2, vcc, 0, %6\n\t"          // u = q1 2^32 - q1
                    "v_subbrev_co_u32 %3, vcc, 0, %6, vcc\n\t"
                    "v_add_co_u32 %1, vcc, %3, %5\n\t"         // t = (q0 0) + u
                    "v_cndmask_b32_e64 %4, 0, -1, vcc\n\t"     // carry: t += 2^32 - 1
                    "v_add_co_u32 %0, vcc, %2, %4\n\t"
                    "v_addc_co_u32 %1, vcc, 0, %1, vcc"
                    : "=&v"(tl), "=&v"(th), "=&v"(ul), "=&v"(uh), "=&v"(m)
                    : "v"(q0), "v"(q1)
                    : "vcc");
            } else if constexpr (SH < 64) {
                asm("v_sub_co_u32 %2, vcc, 0, %6\n\t"          // u = q1 2^32 - q1
                    "v_subbrev_co_u32 %3, vcc, 0, %6, vcc\n\t"
                    "v_add_co_u32 %1, vcc, %3, %5\n\t"         // t = (q0 0) + u
                    "v_cndmask_b32_e64 %4, 0, -1, vcc\n\t"     // carry: t += 2^32 - 1
                    "v_add_co_u32 %0, vcc, %2, %4\n\t"
                    "v_addc_co_u32 %1, vcc, 0, %1, vcc\n\t"
                    "v_sub_co_u32 %0, vcc, %0, %7\n\t"         // t -= q2
                    "v_subbrev_co_u32 %1, vcc, 0, %1, vcc\n\t"
                    "v_cndmask_b32_e64 %4, 0, -1, vcc\n\t"     // borrow: t -= 2^32 - 1
                    "v_sub_co_u32 %0, vcc, %0, %4\n\t"
                    "v_subbrev_co_u32 %1, vcc, 0, %1, vcc"
                    : "=&v"(tl), "=&v"(th), "=&v"(ul), "=&v"(uh), "=&v"(m)
                    : "v"(q0), "v"(q1), "v"(q2)
                    : "vcc");
            } else if constexpr (S1 == 0) {
                asm("v_sub_co_u32 %2, vcc, 0, %5\n\t"          // u = q0 2^32 - q0
                    "v_subbrev_co_u32 %3, vcc, 0, %5, vcc\n\t"
                    "v_sub_co_u32 %0, vcc, %2, %6\n\t"         // t = u - q1
                    "v_subbrev_co_u32 %1, vcc, 0, %3, vcc\n\t"
                    "v_cndmask_b32_e64 %4, 0, -1, vcc\n\t"     // borrow: t -= 2^32 - 1
                    "v_sub_co_u32 %0, vcc, %0, %4\n\t"
                    "v_subbrev_co_u32 %1, vcc, 0, %1, vcc"
                    : "=&v"(tl), "=&v"(th), "=&v"(ul), "=&v"(uh), "=&v"(m)
                    : "v"(q0), "v"(q1)
                    : "vcc");
            } else {
                asm("v_sub_co_u32 %2, vcc, 0, %5\n\t"          // u = q0 2^32 - q0
                    "v_subbrev_co_u32 %3, vcc, 0, %5, vcc\n\t"
                    "v_sub_co_u32 %0, vcc, %2, %6\n\t"         // t = u - (q2 q1)
                    "v_subb_co_u32 %1, vcc, %3, %7, vcc\n\t"
                    "v_cndmask_b32_e64 %4, 0, -1, vcc\n\t"     // borrow: t -= 2^32 - 1
                    "v_sub_co_u32 %0, vcc, %0, %4\n\t"
                    "v_subbrev_co_u32 %1, vcc, 0, %1, vcc"
                    : "=&v"(tl), "=&v"(th), "=&v"(ul), "=&v"(uh), "=&v"(m)
                    : "v"(q0), "v"(q1), "v"(q2)
                    : "vcc");
            }
            return pack(tl, th);
        }
    }

    // Butterflies whose twiddle is 2^E (0 <= E < 192, 2^96 = -1), lazy words
    // as dit / dif.  A twiddle -2^(E-96) swaps the sum and the difference; in
    // the DIF form the difference is then taken the other way round, which
    // needs the other operand canonical.
    template <int E>
    __device__ __forceinline__ static void dit_p2(uint64_t& a, uint64_t& b) {
        if constexpr (E == 0) {
            bf1(a, b);
        } else {
            const uint64_t t = canon_lz(mul_pow2_lz<E % 96>(b)), x = a;
            if constexpr (E < 96) {
                a = add_lz(x, t);
                b = sub_lz(x, t);
            } else {
                a = sub_lz(x, t);
                b = add_lz(x, t);
            }
        }
    }
    template <int E>
    __device__ __forceinline__ static void dif_p2(uint64_t& a, uint64_t& b) {
        if constexpr (E == 0) {
            bf1(a, b);
        } else {
            const uint64_t y = canon_lz(b), x = a;
            a = add_lz(x, y);
            if constexpr (E < 96) b = mul_pow2_lz<E>(sub_lz(x, y));
            else b = mul_pow2_lz<E - 96>(sub_lz(y, canon_lz(x)));
        }
    }

    // NTT butterflies (ntt_impl.h) on lazy words, results lazy; w canonical.
    // The operand added to or subtracted from a is made canonical first, as
    // add_lz / sub_lz require.  Words are made canonical when stored to global
    // memory (store_form).
    __device__ __forceinline__ static void dit(uint64_t& a, uint64_t& b, uint64_t w) {
        const uint64_t t = canon_lz(mul_lazy_vcc(b, w)), x = a;
        a = add_lz(x, t);
        b = sub_lz(x, t);
    }
    __device__ __forceinline__ static void dif(uint64_t& a, uint64_t& b, uint64_t w) {
        const uint64_t y = canon_lz(b), x = a;
        a = add_lz(x, y);
        b = mul_lazy_vcc(sub_lz(x, y), w);
    }
    __device__ __forceinline__ static void bf1(uint64_t& a, uint64_t& b) {
        const uint64_t y = canon_lz(b), x = a;
        a = add_lz(x, y);
        b = sub_lz(x, y);
    }
    __device__ __forceinline__ static uint64_t mul_ntt(uint64_t x, uint64_t c) { return mul_lazy_vcc(x, c); }
    __device__ __forceinline__ static uint64_t store_form(uint64_t x) { return canon_lz(x); }
};

// Global-memory form of an NTT word (canonical), by word type.
__device__ __forceinline__ uint32_t ntt_store_form(uint32_t x) { return BabyBear::store_form(x); }
__device__ __forceinline__ uint64_t ntt_store_form(uint64_t x) { return Goldilocks::store_form(x); }

}  // namespace p3h
