// BabyBear Poseidon2 (width 16) evaluated in IEEE binary64, exactly.
//
// Every state word is a double holding an integer v with v == x (mod p),
// p = 2^31 - 2^27 + 1.  Integers below 2^53 are exact in binary64, so the
// linear layers (MDS-light, the internal sum and the 2^i diagonal) need no
// modular reduction at all; only products are reduced.  All operations are
// full-rate FP64 VALU instructions on gfx950 (v_fma_f64 / v_mul_f64 /
// v_add_f64).  The library is compiled with -ffp-contract=off, so every fused
// multiply-add below is an explicit fma and nothing else is fused.
//
// Reduction of an integer double v with |v| < 2^81:
//   q = round(v / p) via t = fma(v, 1/p, M) - M, M = 1.5 * 2^52 (valid while
//       |v / p| < 2^51: the sum M + v/p then has unit spacing),
//   r = fma(-q, p, v)  -- exact, because the true value v - q p is an integer of
//       magnitude <= p/2 + |v| 2^-53 + ... < 2^31, hence representable.
// Product a*b (|a*b| < 2^81): h = a*b rounded, l = fma(a, b, -h) exact (two-
// product), then reduce(h) + l.  |l| <= ulp(h)/2 <= 2^28, so the result has
// magnitude < p/2 + 2^29 < 2^31.
//
// Bounds through one permutation (the largest magnitudes, all < 2^53):
//   inputs (canonical or reduced)           <= p
//   S-box output                             <  2^31
//   MDS-light output (coefficient sum 35)    <  35 * 2^31 < 2^37.2
//   S-box input (MDS output + constant)      <  2^37.3, square < 2^74.6 < 2^81
//   internal layer: lanes enter <= 2^37.2, 2^15 * lane + sum < 2^52.3, then every
//   lane is reduced (|r| < 2^31) before the next round
//   end of the permutation: every lane reduced, so a sponge's carried lanes
//   start the next permutation below 2^31.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace p3h {

struct P2ParamsF64 {
    double init[4][16];
    double term[4][16];
    double internal[32];
    uint32_t half_full;
    uint32_t partial;
};

struct BBF64 {
    static constexpr double P = 2013265921.0;
    static constexpr double PINV = 1.0 / 2013265921.0;
    static constexpr double MAGIC = 6755399441055744.0;  // 1.5 * 2^52

    __device__ __forceinline__ static double reduce(double v) {
        double q = __builtin_fma(v, PINV, MAGIC) - MAGIC;
        return __builtin_fma(-q, P, v);
    }
    __device__ __forceinline__ static double mul(double a, double b) {
        double h = a * b;
        double l = __builtin_fma(a, b, -h);
        return reduce(h) + l;
    }
    // v with |v| < 2^51 * p -> canonical [0, p)
    __device__ __forceinline__ static uint32_t canonical(double v) {
        double r = reduce(v);
        r = r < 0.0 ? r + P : r;
        return uint32_t(r);
    }
};

struct Poseidon2F64 {
    __device__ __forceinline__ static double sbox(double x) {
        double x2 = BBF64::mul(x, x);
        double x3 = BBF64::mul(x2, x);
        double x4 = BBF64::mul(x2, x2);
        return BBF64::mul(x3, x4);
    }
    // [[2,3,1,1],[1,2,3,1],[1,1,2,3],[3,1,1,2]] x, exact integer arithmetic
    __device__ __forceinline__ static void mat4(double* x) {
        double t01 = x[0] + x[1];
        double t23 = x[2] + x[3];
        double t0123 = t01 + t23;
        double t01123 = t0123 + x[1];
        double t01233 = t0123 + x[3];
        double y3 = t01233 + (x[0] + x[0]);
        double y1 = t01123 + (x[2] + x[2]);
        double y0 = t01123 + t01;
        double y2 = t01233 + t23;
        x[0] = y0;
        x[1] = y1;
        x[2] = y2;
        x[3] = y3;
    }
    __device__ __forceinline__ static void mds_light(double* s) {
#pragma unroll
        for (int i = 0; i < 16; i += 4) mat4(s + i);
        double sums[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) sums[k] = (s[k] + s[k + 4]) + (s[k + 8] + s[k + 12]);
#pragma unroll
        for (int i = 0; i < 16; ++i) s[i] = s[i] + sums[i & 3];
    }
    __device__ __forceinline__ static void full_round(double* s, const double* rc) {
#pragma unroll
        for (int i = 0; i < 16; ++i) s[i] = sbox(s[i] + rc[i]);
        mds_light(s);
    }
    // s_i <- 2^i s_i + sum(s), then every lane reduced
    __device__ __forceinline__ static void internal_layer(double* s) {
        double sum = 0.0;
#pragma unroll
        for (int i = 0; i < 16; ++i) sum = sum + s[i];
        s[0] = BBF64::reduce(s[0] + sum);
#pragma unroll
        for (int i = 1; i < 16; ++i) s[i] = BBF64::reduce(__builtin_fma(s[i], double(1u << i), sum));
    }
    __device__ __forceinline__ static void permute(double* s, const P2ParamsF64& p) {
        mds_light(s);
        for (uint32_t r = 0; r < p.half_full; ++r) full_round(s, p.init[r]);
        for (uint32_t r = 0; r < p.partial; ++r) {
            s[0] = sbox(s[0] + p.internal[r]);
            internal_layer(s);
        }
        for (uint32_t r = 0; r < p.half_full; ++r) full_round(s, p.term[r]);
#pragma unroll
        for (int i = 0; i < 16; ++i) s[i] = BBF64::reduce(s[i]);
    }
};

}  // namespace p3h
