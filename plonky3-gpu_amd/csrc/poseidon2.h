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
    __device__ __forceinline__ static void permute(T* s, const P2Params<T, W>& p) {
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
