// FRI commit-phase folding over BabyBear^4 = F_p[a]/(a^4 - 11).
//
// Reference: TwoAdicFriFolding::fold_row / fold_matrix (fri_folding.hpp:74-222),
// GPU twin fri_fold_arity2_kernel (fri_fold_cuda.hpp:80-109).
// Arity 2: out[i] = (e0 + e1)/2 + beta * (e0 - e1) / (2x) with
// x = omega_{2^log_h}^bitrev(i, log_h - 1).  The factor 1/(2x) comes from one
// table in bit-reversed order, tinv[i] = omega_{2^LH}^-bitrev(i, LH-1) / 2,
// which serves every layer of height <= 2^LH (its first 2^(lh-1) entries are
// exactly the layer-lh values).  Arity 4 and 8 follow the reference's weight
// formula literally (see fold_general_kernel).

#include "fri.h"
#include "ext4.h"

#include <mutex>

namespace p3h {

namespace {

using BB = BabyBear;
using namespace ext4;

// beta_dev (nullable): canonical beta in device memory, used instead of beta_m
__global__ void fold_arity2_kernel(const uint32_t* in, uint32_t* out, uint64_t n_out, E4 beta_m,
                                   const uint32_t* beta_dev, const uint32_t* tinv_m) {
    uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n_out) return;
    if (beta_dev) {
#pragma unroll
        for (int q = 0; q < 4; ++q) beta_m.c[q] = BB::to_mont(beta_dev[q]);
    }
    constexpr uint32_t HALF_M = 0x07FFFFFFu;  // 2^-1 * 2^32 mod p
    const uint4* src = reinterpret_cast<const uint4*>(in) + 2 * i;
    uint4 a = src[0], b = src[1];
    E4 e0 = {{a.x, a.y, a.z, a.w}}, e1 = {{b.x, b.y, b.z, b.w}};
    E4 s = e4_scale_m(e4_add(e0, e1), HALF_M);
    E4 d = e4_scale_m(e4_sub(e0, e1), tinv_m[i]);
    E4 r = e4_add(s, e4_mul_m(d, beta_m));
    reinterpret_cast<uint4*>(out)[i] = make_uint4(r.c[0], r.c[1], r.c[2], r.c[3]);
}

__device__ __forceinline__ uint32_t dbitrev(uint32_t x, uint32_t bits) {
    return bits == 0 ? 0u : (__brev(x) >> (32 - bits));
}

// General arity, reference weights: t_j = x * w^bitrev(j),
// L_j = prod_k (beta - t_k) / ((beta - t_j) * x^(a-1) * w^(-j) * a);
// if beta equals some t_j the row returns e_j.
__global__ void fold_general_kernel(const uint32_t* in, uint32_t* out, uint64_t n_out, uint32_t log_h,
                                    uint32_t log_arity, E4 beta, const uint32_t* beta_dev, uint32_t omega_h_m,
                                    uint32_t w_m, uint32_t arity_inv_m) {
    uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n_out) return;
    if (beta_dev) {
#pragma unroll
        for (int q = 0; q < 4; ++q) beta.c[q] = beta_dev[q];
    }
    const uint32_t arity = 1u << log_arity;
    const uint32_t* ev = in + 4 * uint64_t(i) * arity;
    uint32_t x_m = bb_pow_m(omega_h_m, dbitrev(uint32_t(i), log_h - log_arity));
    uint32_t t[8];
    for (uint32_t j = 0; j < arity; ++j)
        t[j] = BB::from_mont(BB::mont(x_m, bb_pow_m(w_m, dbitrev(j, log_arity))));
    for (uint32_t j = 0; j < arity; ++j) {
        if (beta.c[0] == t[j] && beta.c[1] == 0 && beta.c[2] == 0 && beta.c[3] == 0) {
            for (int q = 0; q < 4; ++q) out[4 * i + q] = ev[4 * j + q];
            return;
        }
    }
    E4 full = {{1, 0, 0, 0}};
    for (uint32_t k = 0; k < arity; ++k) {
        E4 d = beta;
        d.c[0] = BB::sub(d.c[0], t[k]);
        full = e4_mul_m(full, e4_to_m(d));
    }
    full = e4_scale_m(full, arity_inv_m);
    uint32_t x_am1_m = bb_pow_m(x_m, arity - 1);
    uint32_t w_inv_m = bb_pow_m(w_m, arity - 1);
    E4 res = {{0, 0, 0, 0}};
    for (uint32_t j = 0; j < arity; ++j) {
        uint32_t base_m = BB::mont(x_am1_m, bb_pow_m(w_inv_m, j));
        E4 d = beta;
        d.c[0] = BB::sub(d.c[0], t[j]);
        E4 den = e4_scale_m(d, base_m);
        E4 lj = e4_mul_m(full, e4_to_m(e4_inv(den)));
        E4 e = {{ev[4 * j], ev[4 * j + 1], ev[4 * j + 2], ev[4 * j + 3]}};
        res = e4_add(res, e4_mul_m(e, e4_to_m(lj)));
    }
    for (int q = 0; q < 4; ++q) out[4 * i + q] = res.c[q];
}

// Arity-2 table of the allocation-free fold: tinv_m[i] =
// omega_{2^(bits+1)}^-bitrev(i, bits) / 2 in Montgomery form (base_m =
// omega^-1 and half_m = 1/2, both Montgomery), the entries fri_tinv caches.
__global__ void fold_tinv_kernel(uint32_t* tinv_m, uint64_t n, uint32_t bits, uint32_t base_m, uint32_t half_m) {
    uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    tinv_m[i] = BB::mont(bb_pow_m(base_m, dbitrev(uint32_t(i), bits)), half_m);
}

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch_ptr[MAX_DEVICES] = {};

}  // namespace

// Scratch of the allocation-free fold: a file-scope device array, one copy per
// GPU, held until the process ends.  Its name is a host-side stand-in in host
// code; the device address comes from fold_scratch_setup().
constexpr uint64_t FOLD_SCRATCH_WORDS = uint64_t(1) << 22;  // 16 MiB
static __device__ __align__(256) uint32_t scratch[FOLD_SCRATCH_WORDS];

uint64_t fold_scratch_words() { return FOLD_SCRATCH_WORDS; }

int fold_scratch_setup(int device) {
    if (device < 0 || device >= MAX_DEVICES) return int(hipErrorInvalidDevice);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return int(hipSuccess);
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return int(e);
    g_scratch_ptr[device] = static_cast<uint32_t*>(p);
    return int(hipSuccess);
}

uint32_t* fold_scratch_ptr(int device) {
    if (device < 0 || device >= MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_ptr[device];
}

int fold_noalloc_device(const uint32_t* in, uint32_t* out, uint64_t n_out, uint32_t log_h, uint32_t log_arity,
                        const uint32_t* beta, uint32_t* scratch_ptr, uint32_t tinv_base_m, uint32_t half_m,
                        uint32_t omega_h_m, uint32_t w_m, uint32_t arity_inv_m, hipStream_t st) {
    if (n_out == 0) return 0;
    if (log_arity == 1) {
        if (!scratch_ptr || n_out > FOLD_SCRATCH_WORDS || log_h < 1) return int(hipErrorInvalidValue);
        uint32_t blocks = uint32_t((n_out + 255) / 256);
        hipLaunchKernelGGL(fold_tinv_kernel, dim3(blocks), dim3(256), 0, st, scratch_ptr, n_out, log_h - 1,
                           tinv_base_m, half_m);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return int(e);
    }
    return fold_device(in, out, n_out, log_h, log_arity, beta, nullptr, log_arity == 1 ? scratch_ptr : nullptr,
                       omega_h_m, w_m, arity_inv_m, st);
}

int fold_device(const uint32_t* in, uint32_t* out, uint64_t n_out, uint32_t log_h, uint32_t log_arity,
                const uint32_t* beta, const uint32_t* beta_dev, const uint32_t* tinv_m, uint32_t omega_h_m,
                uint32_t w_m, uint32_t arity_inv_m, hipStream_t st) {
    if (n_out == 0) return 0;
    uint32_t blocks = uint32_t((n_out + 255) / 256);
    E4 b = {{0, 0, 0, 0}};
    if (log_arity == 1 && tinv_m) {
        if (!beta_dev)
            for (int q = 0; q < 4; ++q) b.c[q] = uint32_t((uint64_t(beta[q]) << 32) % BabyBear::P);
        hipLaunchKernelGGL(fold_arity2_kernel, dim3(blocks), dim3(256), 0, st, in, out, n_out, b, beta_dev, tinv_m);
    } else {
        if (!beta_dev)
            for (int q = 0; q < 4; ++q) b.c[q] = beta[q];
        hipLaunchKernelGGL(fold_general_kernel, dim3(blocks), dim3(256), 0, st, in, out, n_out, log_h, log_arity, b,
                           beta_dev, omega_h_m, w_m, arity_inv_m);
    }
    return int(hipGetLastError());
}

}  // namespace p3h
