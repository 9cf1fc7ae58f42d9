// Poseidon2 permutation on the device: one lane owns one whole state (16
// BabyBear or 8 Goldilocks elements in VGPRs); round constants arrive in the
// kernel-argument block and are read with wave-uniform indices, so they stay in
// scalar registers.  BabyBear state is kept in Montgomery form for the whole
// permutation (constants are pre-converted on the host); Goldilocks is
// canonical.
//
// The permutation is the reference's generic Poseidon2
// (poseidon2.hpp:108-183, external.hpp:49-61/121-162/243-275,
// generic.hpp:58-142): MDS-light, rf/2 full rounds, rp partial rounds with the
// internal layer s_i <- 2^i * s_i + sum(s), rf/2 full rounds; S-box x^7.
#pragma once

#include <type_traits>

#include "field.h"

namespace p3h {

constexpr int P2_MAX_HALF_FULL = 4;   // up to 8 full rounds
constexpr int P2_MAX_PARTIAL = 32;

template <class T, int W>
struct P2Params {
    T init[P2_MAX_HALF_FULL][W];
    T term[P2_MAX_HALF_FULL][W];
    T internal[P2_MAX_PARTIAL];
    T diag[W];        // 2^i in the state domain
    uint32_t half_full;
    uint32_t partial;
};

// State-domain arithmetic.
struct BBDom {
    using F = BabyBear;
    using T = uint32_t;
    static constexpr int W = 16, RATE = 8, DIG = 8;
    __device__ __forceinline__ static T add(T a, T b) { return F::add(a, b); }
    __device__ __forceinline__ static T mul(T a, T b) { return F::mont(a, b); }
    // a product that only feeds further products
    __device__ __forceinline__ static T mul_inner(T a, T b) { return F::mont(a, b); }
    // s * diag_i, diag_i = 2^i in the state domain
    __device__ __forceinline__ static T mul_diag(T s, uint32_t, T diag_i) { return F::mont(s, diag_i); }
    __device__ __forceinline__ static T to_dom(T x) { return F::to_mont(x); }
    __device__ __forceinline__ static T from_dom(T x) { return F::from_mont(x); }
};

struct GLDom {
    using F = Goldilocks;
    using T = uint64_t;
    static constexpr int W = 8, RATE = 4, DIG = 4;
    __device__ __forceinline__ static T add(T a, T b) { return F::add(a, b); }
    __device__ __forceinline__ static T mul(T a, T b) { return F::mul(a, b); }
    // a product that only feeds further products: any value below 2^64 will do
    __device__ __forceinline__ static T mul_inner(T a, T b) { return F::mul_lazy(a, b); }
    // s * 2^i by shifting (diag_i is 2^i)
    __device__ __forceinline__ static T mul_diag(T s, uint32_t i, T) { return F::mul_pow2(s, i); }
    __device__ __forceinline__ static T to_dom(T x) { return x; }
    __device__ __forceinline__ static T from_dom(T x) { return x; }
};

// Goldilocks width-8 permutation with exact-integer linear layers.
//
// State words are any 64-bit values congruent to the field elements; only the
// permutation's output is canonical.
//   S-box: four Goldilocks::mul_lazy products, which accept any 64-bit inputs.
//   Round constant: one 64-bit add; a carry adds 2^64 mod p = 2^32 - 1, and a
//     second carry cannot occur because the constant is canonical (at most
//     2^64 - 2^32), so the wrapped sum is at most 2^64 - 2^32 - 1.
//   Linear layers (MDS-light, internal layer): every output is a combination
//     of the inputs with small non-negative integer coefficients.  The low and
//     high 32-bit halves are combined separately in 64-bit registers, where
//     each term is one 64-bit add or shift-add with no carry to handle
//     (MDS-light: coefficient sum 21, so below 21 * 2^32 < 2^37; internal
//     layer: below 2^7 * 2^32 + 8 * 2^32 < 2^40), and each output is folded
//     once: H * 2^32 + L (mod p).
// This replaces a canonical modular add (compare, select, canonical
// subtraction) per term of the linear layers.
struct Poseidon2GL {
    static constexpr uint64_t EPS = 0xFFFFFFFFull;  // 2^64 mod p

    // H * 2^32 + L (mod p) as a value below 2^64, for H < 2^56 and any L < 2^64:
    // H 2^32 + L = h_hi 2^64 + (h_lo + l_hi) 2^32 + l_lo = k 2^64 + T with the
    // carry of h_lo + l_hi moved into k < 2^24 + 1; 2^64 == 2^32 - 1, so the
    // value is T + k (2^32 - 1), whose wrap-around (at most once: the wrapped
    // sum is below k 2^32 < 2^57) adds 2^32 - 1 again.
    __device__ __forceinline__ static uint64_t fold(uint64_t L, uint64_t H) {
        const uint32_t l_lo = uint32_t(L), l_hi = uint32_t(L >> 32);
        const uint32_t h_lo = uint32_t(H), h_hi = uint32_t(H >> 32);
        const uint32_t t_hi = h_lo + l_hi;
        const uint32_t k = h_hi + (t_hi < h_lo ? 1u : 0u);
        const uint64_t T = (uint64_t(t_hi) << 32) | l_lo;
        const uint64_t m = (uint64_t(k) << 32) - k;
        const uint64_t r = T + m;
        return r < T ? r + EPS : r;
    }
    __device__ __forceinline__ static uint64_t add_rc(uint64_t x, uint64_t rc) {
        const uint64_t t = x + rc;
        return t < x ? t + EPS : t;
    }
    __device__ __forceinline__ static uint64_t sbox(uint64_t x) {
        const uint64_t x2 = Goldilocks::mul_lazy_vcc(x, x);
        const uint64_t x3 = Goldilocks::mul_lazy_vcc(x2, x);
        const uint64_t x4 = Goldilocks::mul_lazy_vcc(x2, x2);
        return Goldilocks::mul_lazy_vcc(x3, x4);
    }
    // [[2,3,1,1],[1,2,3,1],[1,1,2,3],[3,1,1,2]] x over the integers
    __device__ __forceinline__ static void mat4(uint64_t* x) {
        const uint64_t t01 = x[0] + x[1];
        const uint64_t t23 = x[2] + x[3];
        const uint64_t t0123 = t01 + t23;
        const uint64_t t01123 = t0123 + x[1];
        const uint64_t t01233 = t0123 + x[3];
        const uint64_t y3 = t01233 + (x[0] << 1);
        const uint64_t y1 = t01123 + (x[2] << 1);
        const uint64_t y0 = t01123 + t01;
        const uint64_t y2 = t01233 + t23;
        x[0] = y0;
        x[1] = y1;
        x[2] = y2;
        x[3] = y3;
    }
    // mat4 on both halves of the state, then s_i + (s_i + s_{i+-4}).  With
    // rc, the next round's constants are added as their 32-bit halves to the
    // halves before the fold (each half then stays below 22 * 2^32), in place
    // of a carry-checked 64-bit addition per lane.
    __device__ __forceinline__ static void mds_light(uint64_t* s, const uint64_t* rc = nullptr) {
        uint64_t lo[8], hi[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            lo[i] = s[i] & 0xFFFFFFFFull;
            hi[i] = s[i] >> 32;
        }
        mat4(lo);
        mat4(lo + 4);
        mat4(hi);
        mat4(hi + 4);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint64_t sl = lo[k] + lo[k + 4], sh = hi[k] + hi[k + 4];
            uint64_t l0 = lo[k] + sl, h0 = hi[k] + sh, l1 = lo[k + 4] + sl, h1 = hi[k + 4] + sh;
            if (rc) {
                l0 += rc[k] & EPS;
                h0 += rc[k] >> 32;
                l1 += rc[k + 4] & EPS;
                h1 += rc[k + 4] >> 32;
            }
            s[k] = fold(l0, h0);
            s[k + 4] = fold(l1, h1);
        }
    }
    // s_0 <- s_0 + sum(s), s_i <- 2^i s_i + sum(s)
    __device__ __forceinline__ static void internal_layer(uint64_t* s) {
        uint64_t lo[8], hi[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            lo[i] = s[i] & 0xFFFFFFFFull;
            hi[i] = s[i] >> 32;
        }
        uint64_t sl = lo[0], sh = hi[0];
#pragma unroll
        for (int i = 1; i < 8; ++i) {
            sl += lo[i];
            sh += hi[i];
        }
        s[0] = fold(lo[0] + sl, hi[0] + sh);
#pragma unroll
        for (int i = 1; i < 8; ++i) s[i] = fold((lo[i] << i) + sl, (hi[i] << i) + sh);
    }
    // The 22 partial rounds on split accumulators: lane i is a pair (L_i, H_i)
    // of 64-bit words with value H_i 2^32 + L_i, so the internal layer is
    // L_i <- 2^i L_i + sum(L), H_i <- 2^i H_i + sum(H) with no fold.  Lane 0 is
    // folded before its S-box every round and split again after it; lane i >= 1
    // is folded (and split) only after the rounds whose bit i is set in
    // FOLD22, and every lane once at the end.  The schedule keeps every word
    // below 2^64 and every folded H below 2^56 (fold's range) in the worst
    // case, which tests/test_gl_fold_schedule.py checks by magnitude recursion
    // over the table read from this header: 83 folds in the 22 rounds against
    // 176 with a fold of every lane in every round.  Each round's constant is
    // added to lane 0 as its 32-bit halves just before the fold.
    static constexpr uint8_t FOLD22[22] = {0x00, 0x00, 0xc0, 0xbe, 0xc0, 0xfe, 0x00, 0x00, 0xc0, 0xbe, 0xc0,
                                           0xfe, 0x00, 0x00, 0xc0, 0xbe, 0xc0, 0xfe, 0x00, 0x00, 0xc0, 0x00};
    // a b + c from a 32-bit a: one v_mad_u64_u32.  b is one of the opaque
    // multipliers below, so the compiler cannot turn a product by a power of
    // two into a shift of a zero-extended word, which costs two moves to build
    // the (a, 0) register pair and a 64-bit shift-add.
    __device__ __forceinline__ static uint64_t mad32(uint32_t a, uint32_t b, uint64_t c) {
        return uint64_t(a) * uint64_t(b) + c;
    }
    // 2^i, i < 8, in VGPRs the compiler cannot see through
    struct Pow2 {
        uint32_t m[8];
        __device__ __forceinline__ Pow2() {
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                m[i] = 1u << i;
                asm("" : "+v"(m[i]));
            }
        }
    };
    // Lanes whose (L, H) are 32-bit halves on entry to partial round R: every
    // lane in round 0 (the split of the MDS-light output), lane 0 in every
    // round (the split of its S-box output), and the lanes folded at the end
    // of round R - 1.  Their terms go through mad32.
    template <int R>
    static constexpr uint8_t fresh_lanes() { return R == 0 ? 0xFF : uint8_t(FOLD22[R - 1] | 1u); }

    template <int R>
    __device__ __forceinline__ static void partial_rounds22(uint64_t* L, uint64_t* H, const uint64_t* internal,
                                                            const Pow2& m) {
        if constexpr (R < 22) {
            constexpr uint8_t FRESH = fresh_lanes<R>();
            const uint64_t rc = internal[R];
            const uint64_t x = sbox(fold(L[0] + (rc & EPS), H[0] + (rc >> 32)));
            const uint32_t x_lo = uint32_t(x), x_hi = uint32_t(x >> 32);
            // sums of the 8 lanes: 64-bit accumulators first, then the halves
            uint64_t sl = 0, sh = 0;
            bool started = false;
#pragma unroll
            for (int i = 1; i < 8; ++i) {
                if (!((FRESH >> i) & 1)) {
                    sl = started ? sl + L[i] : L[i];
                    sh = started ? sh + H[i] : H[i];
                    started = true;
                }
            }
            if (!started) {
                sl = x_lo;
                sh = x_hi;
            } else {
                sl = mad32(x_lo, m.m[0], sl);
                sh = mad32(x_hi, m.m[0], sh);
            }
#pragma unroll
            for (int i = 1; i < 8; ++i) {
                if ((FRESH >> i) & 1) {
                    sl = mad32(uint32_t(L[i]), m.m[0], sl);
                    sh = mad32(uint32_t(H[i]), m.m[0], sh);
                }
            }
            L[0] = mad32(x_lo, m.m[0], sl);
            H[0] = mad32(x_hi, m.m[0], sh);
#pragma unroll
            for (int i = 1; i < 8; ++i) {
                if ((FRESH >> i) & 1) {
                    L[i] = mad32(uint32_t(L[i]), m.m[i], sl);
                    H[i] = mad32(uint32_t(H[i]), m.m[i], sh);
                } else {
                    L[i] = (L[i] << i) + sl;
                    H[i] = (H[i] << i) + sh;
                }
            }
#pragma unroll
            for (int i = 1; i < 8; ++i) {
                if ((FOLD22[R] >> i) & 1) {
                    const uint64_t v = fold(L[i], H[i]);
                    L[i] = v & EPS;
                    H[i] = v >> 32;
                }
            }
            partial_rounds22<R + 1>(L, H, internal, m);
        }
    }
    // A full round's constants are added inside the MDS-light before it (the
    // initial one for init round 0), and term round 0's in the folds that end
    // the 22 partial rounds.  RELOAD = true reads the constants at an offset
    // z that is zero but opaque to the compiler, once per call, so that a
    // caller looping over permutations re-reads them with scalar loads
    // instead of keeping them in scalar registers across the loop (more than
    // it has), spilled to VGPR lanes and read back with v_readlane_b32.
    template <bool RELOAD = false, class P>
    __device__ __forceinline__ static void permute(uint64_t* s, const P& p) {
        const uint32_t hf = p.half_full;
        uint32_t z = 0;
        if constexpr (RELOAD) asm volatile("" : "+s"(z));
        mds_light(s, hf ? p.init[z] : nullptr);
        for (uint32_t r = 0; r < hf; ++r) {
#pragma unroll
            for (int i = 0; i < 8; ++i) s[i] = sbox(s[i]);
            mds_light(s, r + 1 < hf ? p.init[r + 1] : nullptr);
        }
        if (p.partial == 22) {
            uint64_t L[8], H[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                L[i] = s[i] & EPS;
                H[i] = s[i] >> 32;
            }
            partial_rounds22<0>(L, H, p.internal + z, Pow2());
            if (hf) {
#pragma unroll
                for (int i = 0; i < 8; ++i) s[i] = fold(L[i] + (p.term[z][i] & EPS), H[i] + (p.term[z][i] >> 32));
            } else {
#pragma unroll
                for (int i = 0; i < 8; ++i) s[i] = fold(L[i], H[i]);
            }
        } else {
            for (uint32_t r = 0; r < p.partial; ++r) {
                s[0] = sbox(add_rc(s[0], p.internal[r]));
                internal_layer(s);
            }
            if (hf) {
#pragma unroll
                for (int i = 0; i < 8; ++i) s[i] = add_rc(s[i], p.term[0][i]);
            }
        }
        for (uint32_t r = 0; r < hf; ++r) {
#pragma unroll
            for (int i = 0; i < 8; ++i) s[i] = sbox(s[i]);
            mds_light(s, r + 1 < hf ? p.term[r + 1] : nullptr);
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) s[i] = Goldilocks::canon(s[i]);
    }
};

template <class D>
struct Poseidon2 {
    using T = typename D::T;
    static constexpr int W = D::W;

    __device__ __forceinline__ static T sbox(T x) {
        T x2 = D::mul_inner(x, x);
        T x3 = D::mul_inner(x2, x);
        T x4 = D::mul_inner(x2, x2);
        return D::mul(x3, x4);
    }
    // [[2,3,1,1],[1,2,3,1],[1,1,2,3],[3,1,1,2]] x
    __device__ __forceinline__ static void mat4(T* x) {
        T t01 = D::add(x[0], x[1]);
        T t23 = D::add(x[2], x[3]);
        T t0123 = D::add(t01, t23);
        T t01123 = D::add(t0123, x[1]);
        T t01233 = D::add(t0123, x[3]);
        T y3 = D::add(t01233, D::add(x[0], x[0]));
        T y1 = D::add(t01123, D::add(x[2], x[2]));
        T y0 = D::add(t01123, t01);
        T y2 = D::add(t01233, t23);
        x[0] = y0;
        x[1] = y1;
        x[2] = y2;
        x[3] = y3;
    }
    __device__ __forceinline__ static void mds_light(T* s) {
#pragma unroll
        for (int i = 0; i < W; i += 4) mat4(s + i);
        T sums[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            sums[k] = s[k];
#pragma unroll
            for (int j = k + 4; j < W; j += 4) sums[k] = D::add(sums[k], s[j]);
        }
#pragma unroll
        for (int i = 0; i < W; ++i) s[i] = D::add(s[i], sums[i & 3]);
    }
    __device__ __forceinline__ static void full_round(T* s, const T* rc) {
#pragma unroll
        for (int i = 0; i < W; ++i) s[i] = sbox(D::add(s[i], rc[i]));
        mds_light(s);
    }
    __device__ __forceinline__ static void internal_layer(T* s, const T* diag) {
        T sum = s[0];
#pragma unroll
        for (int i = 1; i < W; ++i) sum = D::add(sum, s[i]);
        s[0] = D::add(s[0], sum);  // diag[0] = 1
#pragma unroll
        for (int i = 1; i < W; ++i) s[i] = D::add(D::mul_diag(s[i], uint32_t(i), diag[i]), sum);
    }
    // RELOAD: see Poseidon2GL::permute (the Goldilocks engine only)
    template <bool RELOAD = false>
    __device__ __forceinline__ static void permute(T* s, const P2Params<T, W>& p) {
        if constexpr (std::is_same<D, GLDom>::value) {
            Poseidon2GL::permute<RELOAD>(s, p);
            return;
        }
        mds_light(s);
        for (uint32_t r = 0; r < p.half_full; ++r) full_round(s, p.init[r]);
        for (uint32_t r = 0; r < p.partial; ++r) {
            s[0] = sbox(D::add(s[0], p.internal[r]));
            internal_layer(s, p.diag);
        }
        for (uint32_t r = 0; r < p.half_full; ++r) full_round(s, p.term[r]);
    }
};

}  // namespace p3h
