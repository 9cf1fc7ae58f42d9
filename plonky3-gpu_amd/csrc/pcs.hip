// PCS opening on the device.
//
// Opened values (barycentric interpolation on the LDE coset gH, |H| = N):
//   f(z) = (z^N - g^N) * sum_j [h_j / (N g^(N-1))] * lde[j] / (z - g h_j)
// as interpolate_coset_with_precomputation computes it
// (interpolation.hpp:152-185).  The weights u_j = h_j c / (z - g h_j) depend
// on the row only, so one pass builds them and a second pass is a dot product
// of every column with them: each LDE word is read once for all points.
//
// Quotients (two_adic_fri_pcs.hpp:335-381): for matrix m with alpha offset a,
// points z_p and opened values f_p,
//   q[j] += sum_p sum_c alpha^(a + p w + c) (f_p[c] - lde[j][c]) / (z_p - x_j)
//         = sum_p alpha^(a + p w) (S_p - R_j) / (z_p - x_j),
// with S_p = sum_c alpha^c f_p[c] (host) and R_j = sum_c alpha^c lde[j][c]
// (one pass over the row), an exact regrouping of the same field sum.  The
// result goes to position bitrev(j) (the reference bit-reverses the vector).
//
// Sums of Montgomery products are accumulated lazily in 64-bit integers and
// reduced once.

#include "ext4.h"
#include "pcs.h"

namespace p3h {

namespace {

using namespace ext4;

__device__ __forceinline__ uint32_t lde_row_of(uint64_t s, uint32_t L, uint32_t k) {
    // stored row s = r * 2^L + i  ->  LDE row i * 2^k + r
    const uint64_t r = s >> L, i = s & ((1ull << L) - 1);
    return uint32_t((i << k) | r);
}

__device__ __forceinline__ uint32_t reduce64(uint64_t x) { return uint32_t(x % BB::P); }

__global__ void bary_weights_kernel(LdeView v, uint32_t shift, uint32_t omega_m, uint32_t common_m,
                                    const uint32_t* z, uint32_t P, uint32_t* u_m) {
    const uint64_t s = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const uint64_t N = 1ull << (v.L + v.k);
    if (s >= N) return;
    const uint32_t j = lde_row_of(s, v.L, v.k);
    const uint32_t h = BB::from_mont(bb_pow_m(omega_m, j));
    const uint32_t x = BB::from_mont(BB::mont(BB::to_mont(h), BB::to_mont(shift)));
    const uint32_t hc_m = BB::to_mont(BB::mul_c(h, common_m));
    for (uint32_t p = 0; p < P; ++p) {
        E4 d = {{z[4 * p], z[4 * p + 1], z[4 * p + 2], z[4 * p + 3]}};
        d.c[0] = BB::sub(d.c[0], x);
        const E4 u = e4_scale_m(e4_inv(d), hc_m);
        const E4 um = e4_to_m(u);
#pragma unroll
        for (int q = 0; q < 4; ++q) u_m[(s * P + p) * 4 + q] = um.c[q];
    }
}

// blockDim.x columns per block (grid.y column tiles), grid.x row blocks
__global__ void bary_dot_kernel(LdeView v, const uint32_t* u_m, uint32_t P, uint32_t rows_per_block,
                                uint32_t* partial) {
    const uint32_t c = blockIdx.y * blockDim.x + threadIdx.x;
    const uint64_t N = 1ull << (v.L + v.k);
    const uint64_t s0 = uint64_t(blockIdx.x) * rows_per_block;
    const uint64_t s1 = s0 + rows_per_block < N ? s0 + rows_per_block : N;
    if (c >= v.w) return;
    uint64_t acc[PCS_MAX_POINTS][4] = {};
    for (uint64_t s = s0; s < s1; ++s) {
        const uint32_t x = v.data[s * v.w + c];
        const uint32_t* u = u_m + s * P * 4;
#pragma unroll
        for (uint32_t p = 0; p < PCS_MAX_POINTS; ++p) {
            if (p < P) {
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[p][q] += BB::mont_lazy(x, u[4 * p + q]);
            }
        }
    }
    for (uint32_t p = 0; p < P; ++p)
#pragma unroll
        for (int q = 0; q < 4; ++q)
            partial[((uint64_t(blockIdx.x) * P + p) * v.w + c) * 4 + q] = reduce64(acc[p][q]);
}

__global__ void bary_finish_kernel(const uint32_t* partial, uint32_t n_blocks, uint32_t P, uint32_t w,
                                   const uint32_t* van, uint32_t* out) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= P * w) return;
    const uint32_t p = t / w;
    uint64_t acc[4] = {0, 0, 0, 0};
    for (uint32_t b = 0; b < n_blocks; ++b)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[q] += partial[(uint64_t(b) * P * w + t) * 4 + q];
    const E4 sum = {{reduce64(acc[0]), reduce64(acc[1]), reduce64(acc[2]), reduce64(acc[3])}};
    const E4 vz = {{van[4 * p], van[4 * p + 1], van[4 * p + 2], van[4 * p + 3]}};
    const E4 r = e4_mul_m(sum, e4_to_m(vz));
#pragma unroll
    for (int q = 0; q < 4; ++q) out[uint64_t(t) * 4 + q] = r.c[q];
}

__global__ void quotient_kernel(LdeView v, uint32_t shift, uint32_t omega_m, const uint32_t* apow_m,
                                const uint32_t* z, const uint32_t* s_c, const uint32_t* aoff_m, uint32_t P,
                                uint32_t* qv) {
    const uint64_t s = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const uint32_t lN = v.L + v.k;
    if (s >= (1ull << lN)) return;
    const uint32_t j = lde_row_of(s, v.L, v.k);
    // R = sum_c alpha^c * lde[j][c]
    uint64_t acc[4] = {0, 0, 0, 0};
    const uint32_t* row = v.data + s * v.w;
    for (uint32_t c = 0; c < v.w; ++c) {
        const uint32_t x = row[c];
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[q] += BB::mont_lazy(x, apow_m[4 * c + q]);
    }
    const E4 R = {{reduce64(acc[0]), reduce64(acc[1]), reduce64(acc[2]), reduce64(acc[3])}};
    const uint32_t h = BB::from_mont(bb_pow_m(omega_m, j));
    const uint32_t x = BB::from_mont(BB::mont(BB::to_mont(h), BB::to_mont(shift)));
    E4 sum = {{0, 0, 0, 0}};
    for (uint32_t p = 0; p < P; ++p) {
        E4 d = {{z[4 * p], z[4 * p + 1], z[4 * p + 2], z[4 * p + 3]}};
        d.c[0] = BB::sub(d.c[0], x);
        const E4 sp = {{s_c[4 * p], s_c[4 * p + 1], s_c[4 * p + 2], s_c[4 * p + 3]}};
        const E4 ao = {{aoff_m[4 * p], aoff_m[4 * p + 1], aoff_m[4 * p + 2], aoff_m[4 * p + 3]}};
        const E4 t = e4_mul_m(e4_mul_m(e4_sub(sp, R), ao), e4_to_m(e4_inv(d)));
        sum = e4_add(sum, t);
    }
    const uint64_t pos = lN == 0 ? 0 : (__brev(j) >> (32 - lN));
    uint32_t* o = qv + pos * 4;
#pragma unroll
    for (int q = 0; q < 4; ++q) o[q] = BB::add(o[q], sum.c[q]);
}

__global__ void fri_mix_kernel(uint32_t* cur, const uint32_t* next, uint64_t n, E4 beta, const uint32_t* beta_dev,
                               uint32_t log_arity) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    E4 b = beta;
    if (beta_dev) {
#pragma unroll
        for (int q = 0; q < 4; ++q) b.c[q] = beta_dev[q];
    }
    for (uint32_t r = 0; r < log_arity; ++r) b = e4_mul_m(b, e4_to_m(b));  // beta^(2^log_arity)
    const E4 x = {{next[4 * i], next[4 * i + 1], next[4 * i + 2], next[4 * i + 3]}};
    const E4 y = {{cur[4 * i], cur[4 * i + 1], cur[4 * i + 2], cur[4 * i + 3]}};
    const E4 r = e4_add(y, e4_mul_m(x, e4_to_m(b)));
#pragma unroll
    for (int q = 0; q < 4; ++q) cur[4 * i + q] = r.c[q];
}

inline uint32_t blocks_for(uint64_t n, uint32_t t) { return uint32_t((n + t - 1) / t); }

}  // namespace

int pcs_bary_weights_device(const LdeView& v, uint32_t shift, uint32_t omega_m, uint32_t common_m, const uint32_t* z,
                            uint32_t P, uint32_t* u_m, hipStream_t st) {
    if (P == 0 || P > PCS_MAX_POINTS) return int(hipErrorInvalidValue);
    const uint64_t N = 1ull << (v.L + v.k);
    hipLaunchKernelGGL(bary_weights_kernel, dim3(blocks_for(N, 256)), dim3(256), 0, st, v, shift, omega_m, common_m,
                       z, P, u_m);
    return int(hipGetLastError());
}

int pcs_bary_dot_device(const LdeView& v, const uint32_t* u_m, uint32_t P, uint32_t rows_per_block,
                        uint32_t* partial, hipStream_t st) {
    if (P == 0 || P > PCS_MAX_POINTS || rows_per_block == 0) return int(hipErrorInvalidValue);
    const uint64_t N = 1ull << (v.L + v.k);
    const uint32_t threads = v.w < 256 ? ((v.w + 63) / 64) * 64 : 256;
    hipLaunchKernelGGL(bary_dot_kernel, dim3(blocks_for(N, rows_per_block), blocks_for(v.w, threads)), dim3(threads),
                       0, st, v, u_m, P, rows_per_block, partial);
    return int(hipGetLastError());
}

int pcs_bary_finish_device(const uint32_t* partial, uint32_t n_blocks, uint32_t P, uint32_t w, const uint32_t* van,
                           uint32_t* out, hipStream_t st) {
    hipLaunchKernelGGL(bary_finish_kernel, dim3(blocks_for(uint64_t(P) * w, 256)), dim3(256), 0, st, partial,
                       n_blocks, P, w, van, out);
    return int(hipGetLastError());
}

int pcs_quotient_device(const LdeView& v, uint32_t shift, uint32_t omega_m, const uint32_t* apow_m,
                        const uint32_t* z, const uint32_t* s, const uint32_t* aoff_m, uint32_t P, uint32_t* q,
                        hipStream_t st) {
    const uint64_t N = 1ull << (v.L + v.k);
    hipLaunchKernelGGL(quotient_kernel, dim3(blocks_for(N, 256)), dim3(256), 0, st, v, shift, omega_m, apow_m, z, s,
                       aoff_m, P, q);
    return int(hipGetLastError());
}

int fri_mix_device(uint32_t* cur, const uint32_t* next, uint64_t n, const uint32_t* beta_host,
                   const uint32_t* beta_dev, uint32_t log_arity, hipStream_t st) {
    if (n == 0) return 0;
    E4 b = {{0, 0, 0, 0}};
    if (!beta_dev)
        for (int q = 0; q < 4; ++q) b.c[q] = beta_host[q];
    hipLaunchKernelGGL(fri_mix_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, st, cur, next, n, b, beta_dev,
                       log_arity);
    return int(hipGetLastError());
}

}  // namespace p3h
