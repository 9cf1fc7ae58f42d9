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
// Product with a shared multiplier (mul_q, |a*b| < 2^79): with bp = b/p
// rounded, t = fma(a, bp, M) = M + q, q = round(a b / p) up to a rounding of
// |a b / p| 2^-52; -q (p - 1) = fma(t, -(p - 1), M (p - 1)) is exact because
// p - 1 = 15 * 2^27 and 15 q < 2^53; r = fma(a, b, -q (p - 1)) = a b - q p + q
// is exact (below 2^53); (r - t) + (M + c) = a b - q p + c, with every step
// exact.  |a b - q p| <= p (1/2 + |a b / p| 2^-52) < 0.57 p.  Five FP64
// instructions instead of six, bp is computed once per multiplier, and the
// last addition carries a constant c for free.
//
// Bounds through one permutation (the largest magnitudes, all < 2^53):
//   inputs (canonical or reduced)           <= p
//   S-box output (plus a folded constant)    <  2^31 (0.51 p + p/2)
//   MDS-light output (coefficient sum 35)    <  35 * 2^31 < 2^36.2
//   full-round S-box input (MDS output, or   <  2^36.2, square < 2^72.4 < 2^79
//     a reduced value plus a constant)
//   internal layer: lanes enter < 2^36.2, 2^15 * lane + sum < 2^52.3; which
//   lanes are reduced after each partial round is fixed per engine below, with
//   its own bound argument
//   end of the permutation: every lane reduced (lanes 0..7 optionally left
//   for the caller), so a sponge's carried lanes start the next permutation
//   below 2^31.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace p3h {

struct P2ParamsF64 {
    double init[4][16];
    double term[4][16];
    double internal[32];
    // Folded round constants of the one-lane-per-state engine: the S-box of
    // full round r adds c with mc = MAGIC + c, where c is the balanced residue
    // of MDS-light^-1 applied to the constants of the round after it (init
    // round r + 1; after the last init round, the first partial round's
    // constant on lane 0; term round r + 1; nothing after the last term
    // round), so those rounds need no constant addition of their own.
    double mc_init[4][16];
    double mc_term[4][16];
    // 2^i, the internal layer's diagonal: read from the argument block, so its
    // fmas take it from a scalar register and update their lane in place
    double diag[16];
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
    static constexpr double PM1 = 2013265920.0;          // p - 1 = 15 * 2^27
    static constexpr double MAGIC_PM1 = MAGIC * PM1;     // 45 * 2^78, exact
    // The fma that forms -q (p - 1) takes two constants.  A gfx950 VOP3
    // instruction reads at most one scalar operand and no literal, so one
    // constant sits in a VGPR pair (QConsts::mpm1) and the other in an SGPR
    // pair (QConsts::npm1); both are opaque to the compiler, which otherwise
    // copies the VGPR constant into a fresh pair for an accumulating
    // v_fmac_f64 with the other as a literal, one extra move per product.
    struct QConsts {
        double mpm1, npm1;
    };
    __device__ __forceinline__ static QConsts qconsts() {
        QConsts c{MAGIC_PM1, -PM1};
        asm("" : "+v"(c.mpm1));
        asm("" : "+s"(c.npm1));
        return c;
    }
    // a b - q p + c for mc = MAGIC + c, bp = b * PINV; |a b| < 2^79, |c| < 2^51
    __device__ __forceinline__ static double mul_q(double a, double b, double bp, double mc, const QConsts& k) {
        const double t = __builtin_fma(a, bp, MAGIC);     // MAGIC + q
        const double nq = __builtin_fma(t, k.npm1, k.mpm1);  // -q (p - 1), exact
        const double r = __builtin_fma(a, b, nq);         // a b - q p + q, exact
        return (r - t) + mc;
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
    // x^7 + c (mc = MAGIC + c) for |x| < 2^39.5, every product through mul_q:
    // x2 = x x, x3 = x2 x, x6 = x3 x3, x7 = x6 x, so x / p serves three of the
    // four products; 22 FP64 instructions against 24 for sbox() plus a
    // constant addition.  Products: x^2 < 2^79, x2 x and x6 x < 2^70.
    __device__ __forceinline__ static double sbox_q(double x, double mc, const BBF64::QConsts& k) {
        const double xp = x * BBF64::PINV;
        const double x2 = BBF64::mul_q(x, x, xp, BBF64::MAGIC, k);
        const double x3 = BBF64::mul_q(x2, x, xp, BBF64::MAGIC, k);
        const double x3p = x3 * BBF64::PINV;
        const double x6 = BBF64::mul_q(x3, x3, x3p, BBF64::MAGIC, k);
        return BBF64::mul_q(x6, x, xp, mc, k);
    }
    // x^7 for a partial-round lane 0 (|x| < 2^41, so x^2 may pass 2^79): the
    // square in the two-product form, the rest through mul_q (x2 x and x6 x
    // below 2^72); 23 instructions.
    __device__ __forceinline__ static double sbox_p(double x, const BBF64::QConsts& k) {
        const double x2 = BBF64::mul(x, x);
        const double xp = x * BBF64::PINV;
        const double x3 = BBF64::mul_q(x2, x, xp, BBF64::MAGIC, k);
        const double x3p = x3 * BBF64::PINV;
        const double x6 = BBF64::mul_q(x3, x3, x3p, BBF64::MAGIC, k);
        return BBF64::mul_q(x6, x, xp, BBF64::MAGIC, k);
    }
    // [[2,3,1,1],[1,2,3,1],[1,1,2,3],[3,1,1,2]] x, exact integer arithmetic
    __device__ __forceinline__ static void mat4(double* x) {
        double t01 = x[0] + x[1];
        double t23 = x[2] + x[3];
        double t0123 = t01 + t23;
        double t01123 = t0123 + x[1];
        double t01233 = t0123 + x[3];
        double y3 = __builtin_fma(x[0], 2.0, t01233);  // exact: integers below 2^53
        double y1 = __builtin_fma(x[2], 2.0, t01123);
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
    // S-box plus the next round's folded constants, then MDS-light
    __device__ __forceinline__ static void full_round_q(double* s, const double* mc, const BBF64::QConsts& k) {
#pragma unroll
        for (int i = 0; i < 16; ++i) s[i] = sbox_q(s[i], mc[i], k);
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
    // Partial rounds with deferred reductions, for the 13 partial rounds of the
    // BabyBear width-16 permutation: after round r only the lanes whose bit is
    // set in LAZY13[r] are reduced; the others carry exact unreduced integers.
    // Worst-case magnitudes (log2), with S-box outputs < 2^31, reduced values
    // <= p/2 + 1 < 2^30 and round constants < 2^31: lanes enter at < 35 * 2^31
    // (MDS-light output); every round keeps lane 0 + constant < 2^40.9 (so the
    // S-box square stays below 2^81.8, inside reduce's range) and every lane
    // and the row sum below 2^53 (exact); after round 12 every lane is reduced,
    // so the terminal full rounds see inputs < 2^31.  The masks are the fewest
    // reductions (132 of 208) a search over these bounds found.
    static constexpr uint16_t LAZY13[13] = {0xff80, 0xfc71, 0xff81, 0xffff, 0xfc00, 0xf3c0, 0xfc01,
                                            0xfff1, 0xffff, 0xfc00, 0xf3c0, 0xfc01, 0xffff};
    template <uint32_t MASK>
    __device__ __forceinline__ static void internal_layer_lazy(double* s, const double* diag) {
        double sum = 0.0;
#pragma unroll
        for (int i = 0; i < 16; ++i) sum = sum + s[i];
        s[0] = (MASK & 1u) ? BBF64::reduce(s[0] + sum) : s[0] + sum;
#pragma unroll
        for (int i = 1; i < 16; ++i) {
            const double v = __builtin_fma(s[i], diag[i], sum);
            s[i] = ((MASK >> i) & 1u) ? BBF64::reduce(v) : v;
        }
    }
    // Round 0's constant is folded into the last init round (half_full >= 1).
    template <int R>
    __device__ __forceinline__ static void partial_rounds13(double* s, const double* internal, const double* diag,
                                                            const BBF64::QConsts& k) {
        if constexpr (R < 13) {
            s[0] = sbox_p(R == 0 ? s[0] : s[0] + internal[R], k);
            internal_layer_lazy<LAZY13[R]>(s, diag);
            partial_rounds13<R + 1>(s, internal, diag, k);
        }
    }
    // Constants: init round 0 and term round 0 add theirs; every other full
    // round's, and partial round 0's when there are full rounds, arrive folded
    // into the S-boxes of the round before (P2ParamsF64::mc_*).
    // REDUCE_RATE = false leaves lanes 0..7 unreduced (below 2^36.2), for a
    // sponge that overwrites all of them next or canonicalises them.
    // RELOAD = true reads the constants from the argument block at an offset
    // z that is zero but opaque to the compiler, once per call.  A caller that
    // loops over permutations then re-reads them with scalar loads, off the
    // VALU; otherwise the compiler keeps them in scalar registers across the
    // loop (more than the 102 it has), spills them to VGPR lanes and reads
    // them back with one v_readlane_b32 per 32-bit half, in every permutation.
    // Kernels that run two permutations and a sponge per thread (the injecting
    // compression) keep RELOAD = false: there the opaque offset makes the
    // compiler copy the argument block to scratch.
    template <bool REDUCE_RATE = true, bool RELOAD = false>
    __device__ __forceinline__ static void permute(double* s, const P2ParamsF64& p) {
        const BBF64::QConsts k = BBF64::qconsts();
        uint32_t z = 0;
        if constexpr (RELOAD) asm volatile("" : "+s"(z));
        mds_light(s);
        if (p.half_full) {
#pragma unroll
            for (int i = 0; i < 16; ++i) s[i] = s[i] + p.init[z][i];
            for (uint32_t r = 0; r < p.half_full; ++r) full_round_q(s, p.mc_init[r], k);
        }
        if (p.partial == 13 && p.half_full) {
            partial_rounds13<0>(s, p.internal + z, p.diag + z, k);
        } else {
            for (uint32_t r = 0; r < p.partial; ++r) {
                s[0] = sbox_p(r == 0 && p.half_full ? s[0] : s[0] + p.internal[r], k);
                internal_layer(s);
            }
        }
        if (p.half_full) {
#pragma unroll
            for (int i = 0; i < 16; ++i) s[i] = s[i] + p.term[z][i];
            for (uint32_t r = 0; r < p.half_full; ++r) full_round_q(s, p.mc_term[r], k);
        }
#pragma unroll
        for (int i = REDUCE_RATE ? 0 : 8; i < 16; ++i) s[i] = BBF64::reduce(s[i]);
    }
};

// The same permutation with one state spread over a 16-lane DPP row: lane i
// holds s_i, so a wave64 carries 4 states.  The latency of one permutation is
// about 900 dependent instructions instead of 5261, which is what the narrow
// levels of a Merkle tree are bound by.  Cross-lane reads use DPP inside a
// row: quad_perm for the 4x4 circulant, row_ror for sums across the row.  The
// arithmetic is exact integer arithmetic on values congruent to those of
// Poseidon2F64 (the order of exact additions and which lanes are reduced when
// differ), so the canonical output is identical.  Every lane of the row must
// execute it.
struct Poseidon2F64Lanes {
    template <int CTRL>
    __device__ __forceinline__ static double dpp(double v) {
        // every lane of the row is written (full masks, in-row sources), so no old value is needed
        const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, true);
        const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, true);
        return __hiloint2double(hi, lo);
    }
    // quad_perm selectors: lane j of a quad reads lane j+1, j+2, j+3 (mod 4)
    static constexpr int QP1 = 1 | (2 << 2) | (3 << 4) | (0 << 6);
    static constexpr int QP2 = 2 | (3 << 2) | (0 << 4) | (1 << 6);
    static constexpr int QP3 = 3 | (0 << 2) | (1 << 4) | (2 << 6);
    static constexpr int ROR = 0x120;  // row_ror:n = ROR + n

    // MDS-light: y_j = 2 x_j + 3 x_{j+1} + x_{j+2} + x_{j+3} inside each quad,
    // then every lane adds the sum of its column over the 4 quads.
    __device__ __forceinline__ static double mds_light(double x) {
        const double x1 = dpp<QP1>(x), x2 = dpp<QP2>(x), x3 = dpp<QP3>(x);
        const double sum = (x + x1) + (x2 + x3);
        const double y = __builtin_fma(x1, 2.0, sum + x);
        double t = y + dpp<ROR + 4>(y);
        t = t + dpp<ROR + 8>(t);
        return y + t;
    }
    // s_i <- reduce(2^i s_i + sum of the row)
    __device__ __forceinline__ static double internal_layer(double s, double diag) {
        double t = s + dpp<ROR + 8>(s);
        t = t + dpp<ROR + 4>(t);
        t = t + dpp<ROR + 2>(t);
        t = t + dpp<ROR + 1>(t);
        return BBF64::reduce(__builtin_fma(s, diag, t));
    }
    // One lane's round constants, loaded once per kernel: a permutation here is
    // a chain of dependent instructions, and a constant load inside a round
    // would put its memory latency on that chain every round.  Full-round
    // constants are per lane (one state word each), partial-round ones shared.
    struct Consts {
        double init[4], term[4], internal[32];
        uint32_t half_full, partial;
        double diag;
    };
    __device__ __forceinline__ static Consts consts(const P2ParamsF64& p, uint32_t lane) {
        Consts c;
        c.half_full = p.half_full;
        c.partial = p.partial;
        c.diag = double(1u << lane);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            c.init[r] = p.init[r][lane];
            c.term[r] = p.term[r][lane];
        }
#pragma unroll
        for (int r = 0; r < 32; ++r) c.internal[r] = p.internal[r];
        return c;
    }
    // half_full <= 4 and partial <= 32 (the sizes of P2ParamsF64's tables)
    __device__ __forceinline__ static double permute(double s, const Consts& c, uint32_t lane) {
        s = mds_light(s);
#pragma unroll
        for (uint32_t r = 0; r < 4; ++r)
            if (r < c.half_full) s = mds_light(Poseidon2F64::sbox(s + c.init[r]));
#pragma unroll
        for (uint32_t r = 0; r < 32; ++r) {
            if (r < c.partial) {
                const double t = Poseidon2F64::sbox(s + c.internal[r]);
                s = internal_layer(lane == 0 ? t : s, c.diag);
            }
        }
#pragma unroll
        for (uint32_t r = 0; r < 4; ++r)
            if (r < c.half_full) s = mds_light(Poseidon2F64::sbox(s + c.term[r]));
        return BBF64::reduce(s);
    }
};

}  // namespace p3h
