// BabyBear^4 = F_p[a]/(a^4 - 11) on the device, canonical coefficients
// (reference: BinomialExtensionField, extension_field.hpp:32-293).  Shared by
// the FRI fold and the PCS opening kernels.
#pragma once

#include "field.h"

namespace p3h {
namespace ext4 {

using BB = BabyBear;

struct E4 {
    uint32_t c[4];
};

__device__ __forceinline__ E4 e4_add(const E4& a, const E4& b) {
    E4 r;
#pragma unroll
    for (int i = 0; i < 4; ++i) r.c[i] = BB::add(a.c[i], b.c[i]);
    return r;
}
__device__ __forceinline__ E4 e4_sub(const E4& a, const E4& b) {
    E4 r;
#pragma unroll
    for (int i = 0; i < 4; ++i) r.c[i] = BB::sub(a.c[i], b.c[i]);
    return r;
}
// canonical a times b given in Montgomery form -> canonical a*b
__device__ __forceinline__ E4 e4_mul_m(const E4& a, const E4& bm) {
    constexpr uint32_t ELEVEN_M = 0x37FFFFE9u;  // 11 * 2^32 mod p
    uint32_t c0 = BB::mul_c(a.c[0], bm.c[0]);
    uint32_t c1 = BB::add(BB::mul_c(a.c[0], bm.c[1]), BB::mul_c(a.c[1], bm.c[0]));
    uint32_t c2 = BB::add(BB::add(BB::mul_c(a.c[0], bm.c[2]), BB::mul_c(a.c[1], bm.c[1])), BB::mul_c(a.c[2], bm.c[0]));
    uint32_t c3 = BB::add(BB::add(BB::mul_c(a.c[0], bm.c[3]), BB::mul_c(a.c[1], bm.c[2])),
                          BB::add(BB::mul_c(a.c[2], bm.c[1]), BB::mul_c(a.c[3], bm.c[0])));
    uint32_t h0 = BB::add(BB::add(BB::mul_c(a.c[1], bm.c[3]), BB::mul_c(a.c[2], bm.c[2])), BB::mul_c(a.c[3], bm.c[1]));
    uint32_t h1 = BB::add(BB::mul_c(a.c[2], bm.c[3]), BB::mul_c(a.c[3], bm.c[2]));
    uint32_t h2 = BB::mul_c(a.c[3], bm.c[3]);
    E4 r;
    r.c[0] = BB::add(c0, BB::mul_c(h0, ELEVEN_M));
    r.c[1] = BB::add(c1, BB::mul_c(h1, ELEVEN_M));
    r.c[2] = BB::add(c2, BB::mul_c(h2, ELEVEN_M));
    r.c[3] = c3;
    return r;
}
__device__ __forceinline__ E4 e4_to_m(const E4& a) {
    E4 r;
#pragma unroll
    for (int i = 0; i < 4; ++i) r.c[i] = BB::to_mont(a.c[i]);
    return r;
}
__device__ __forceinline__ E4 e4_scale_m(const E4& a, uint32_t sm) {
    E4 r;
#pragma unroll
    for (int i = 0; i < 4; ++i) r.c[i] = BB::mul_c(a.c[i], sm);
    return r;
}
__device__ __forceinline__ uint32_t bb_pow_m(uint32_t base_m, uint64_t e) {
    uint32_t acc = BB::R_MOD_P;
    while (e) {
        if (e & 1) acc = BB::mont(acc, base_m);
        base_m = BB::mont(base_m, base_m);
        e >>= 1;
    }
    return acc;
}
// canonical a -> canonical a^-1 (a != 0), Fermat
__device__ __forceinline__ uint32_t bb_inv(uint32_t a) {
    return BB::from_mont(bb_pow_m(BB::to_mont(a), BB::P - 2));
}
// canonical a -> canonical a^-1 in BabyBear^4: the product t of the three
// Frobenius conjugates satisfies a * t = N(a) in F_p, so a^-1 = t / N(a).
// Frobenius: phi(a)_i = a_i * z^i with z = 11^((p-1)/4).
__device__ __forceinline__ E4 e4_inv(const E4& a) {
    constexpr uint32_t Z1 = 1728404513u;  // 11^((p-1)/4)
    constexpr uint32_t Z2 = 2013265920u;  // z^2 = -1
    constexpr uint32_t Z3 = 284861408u;   // z^3
    // phi^k(a)_i = a_i * z^(i k)
    const uint32_t zpow[4] = {1u, Z1, Z2, Z3};
    E4 t = {{1, 0, 0, 0}};
    for (int k = 1; k < 4; ++k) {
        E4 c;
#pragma unroll
        for (int i = 0; i < 4; ++i) c.c[i] = BB::mul_c(a.c[i], BB::to_mont(zpow[(i * k) & 3]));
        t = e4_mul_m(t, e4_to_m(c));
    }
    E4 n = e4_mul_m(a, e4_to_m(t));  // lies in F_p: n.c[1..3] == 0
    uint32_t ninv = bb_inv(n.c[0]);
    return e4_scale_m(t, BB::to_mont(ninv));
}

}  // namespace ext4
}  // namespace p3h
