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

#include <algorithm>

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

// one workgroup per output (p, c): the threads stride over the row blocks'
// partials and reduce in LDS
__global__ __launch_bounds__(256) void bary_finish_kernel(const uint32_t* partial, uint32_t n_blocks, uint32_t P,
                                                          uint32_t w, const uint32_t* van, uint32_t* out) {
    __shared__ uint64_t red[4][256];
    const uint32_t t = blockIdx.x;  // = p * w + c
    const uint32_t p = t / w;
    uint64_t acc[4] = {0, 0, 0, 0};
    for (uint32_t b = threadIdx.x; b < n_blocks; b += 256)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[q] += partial[(uint64_t(b) * P * w + t) * 4 + q];
#pragma unroll
    for (int q = 0; q < 4; ++q) red[q][threadIdx.x] = acc[q];
    __syncthreads();
    for (uint32_t h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h)
#pragma unroll
            for (int q = 0; q < 4; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const E4 sum = {{reduce64(red[0][0]), reduce64(red[1][0]), reduce64(red[2][0]), reduce64(red[3][0])}};
        const E4 vz = {{van[4 * p], van[4 * p + 1], van[4 * p + 2], van[4 * p + 3]}};
        const E4 r = e4_mul_m(sum, e4_to_m(vz));
#pragma unroll
        for (int q = 0; q < 4; ++q) out[uint64_t(t) * 4 + q] = r.c[q];
    }
}

// Quotient, pass 1 (one thread per row): with w_p = aoff_p / (z_p - x_j),
// A_j = sum_p w_p S_p and B_j = sum_p w_p, so that q_j = A_j - R_j B_j.
__global__ void quotient_weights_kernel(LdeView v, uint32_t shift, uint32_t omega_m, const uint32_t* z,
                                        const uint32_t* s_c, const uint32_t* aoff_m, uint32_t P, uint32_t* ab) {
    const uint64_t s = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= (1ull << (v.L + v.k))) return;
    const uint32_t j = lde_row_of(s, v.L, v.k);
    const uint32_t h = BB::from_mont(bb_pow_m(omega_m, j));
    const uint32_t x = BB::from_mont(BB::mont(BB::to_mont(h), BB::to_mont(shift)));
    E4 A = {{0, 0, 0, 0}}, B = {{0, 0, 0, 0}};
    for (uint32_t p = 0; p < P; ++p) {
        E4 d = {{z[4 * p], z[4 * p + 1], z[4 * p + 2], z[4 * p + 3]}};
        d.c[0] = BB::sub(d.c[0], x);
        const E4 ao = {{aoff_m[4 * p], aoff_m[4 * p + 1], aoff_m[4 * p + 2], aoff_m[4 * p + 3]}};
        const E4 w = e4_mul_m(e4_inv(d), ao);
        const E4 sp = {{s_c[4 * p], s_c[4 * p + 1], s_c[4 * p + 2], s_c[4 * p + 3]}};
        A = e4_add(A, e4_mul_m(sp, e4_to_m(w)));
        B = e4_add(B, w);
    }
    const E4 Bm = e4_to_m(B);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        ab[s * 8 + q] = A.c[q];
        ab[s * 8 + 4 + q] = Bm.c[q];
    }
}

// x mod p for x < 2^64: x = hi 2^32 + lo, hi 2^32 = mont(hi, 2^64 mod p).
__device__ __forceinline__ uint32_t reduce_u64(uint64_t x) {
    const uint32_t hi = uint32_t(x >> 32), lo = uint32_t(x);
    uint32_t l = lo;
    l = l >= BB::P ? l - BB::P : l;
    l = l >= BB::P ? l - BB::P : l;
    return BB::add(BB::mont(hi, BB::R2_MOD_P), l);
}

// Quotient, pass 2: 16 lanes per row read the row with 16-byte loads (16
// lanes x 4 columns per step), multiply by alpha^c from an LDS table and sum
// across the 16 lanes; lane 0 of the row stores A_j - R_j B_j over the row's
// A_j (words 0-3 of its 8 in ab), in stored-row order.  Pass 3 adds it at
// bitrev(j): scattered reads and contiguous writes, instead of one scattered
// read-modify-write per row at the end of pass 2.
// The LDS table is coefficient-major (at[q * W + c] = coefficient q of
// alpha^c, W = the width rounded up to a multiple of 4), so a lane's 4 columns
// are one 16-byte read per coefficient and the 16 lanes of a row read 256
// consecutive bytes: no bank conflicts.  It is dynamic LDS of 16 W bytes, so
// narrow matrices leave the CU room for many workgroups.
constexpr uint32_t QR_THREADS = 256;
constexpr uint32_t QR_LDS_COLS = 4096;
__global__ __launch_bounds__(QR_THREADS) void quotient_reduce_kernel(LdeView v, const uint32_t* apow_m,
                                                                     uint32_t* ab) {
    extern __shared__ __align__(16) uint32_t at[];
    const bool in_lds = v.w <= QR_LDS_COLS;
    const uint32_t W = (v.w + 3) & ~3u;
    if (in_lds)
        for (uint32_t t = threadIdx.x; t < v.w * 4; t += QR_THREADS) at[(t & 3) * W + (t >> 2)] = apow_m[t];
    __syncthreads();
    const uint32_t lN = v.L + v.k;
    const uint64_t N = 1ull << lN;
    const uint32_t sub = threadIdx.x & 15;
    const bool vec = (v.w % 4) == 0 && (reinterpret_cast<uintptr_t>(v.data) & 15) == 0;
    for (uint64_t s = (uint64_t(blockIdx.x) * QR_THREADS + threadIdx.x) >> 4; s < N;
         s += (uint64_t(gridDim.x) * QR_THREADS) >> 4) {
        const uint32_t* row = v.data + s * v.w;
        uint64_t acc[4] = {0, 0, 0, 0};
        if (vec && in_lds) {
            for (uint32_t c = sub * 4; c < v.w; c += 64) {
                const uint4 x = *reinterpret_cast<const uint4*>(row + c);
                const uint32_t xs[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const uint4 a = *reinterpret_cast<const uint4*>(at + q * W + c);
                    acc[q] += BB::mont_lazy(xs[0], a.x);
                    acc[q] += BB::mont_lazy(xs[1], a.y);
                    acc[q] += BB::mont_lazy(xs[2], a.z);
                    acc[q] += BB::mont_lazy(xs[3], a.w);
                }
            }
        } else if (in_lds) {
            for (uint32_t c = sub; c < v.w; c += 16) {
                const uint32_t x = row[c];
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[q] += BB::mont_lazy(x, at[q * W + c]);
            }
        } else if (vec) {
            const uint32_t* ap = apow_m;
            for (uint32_t c = sub * 4; c < v.w; c += 64) {
                const uint4 x = *reinterpret_cast<const uint4*>(row + c);
                const uint32_t xs[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
                for (int e = 0; e < 4; ++e)
#pragma unroll
                    for (int q = 0; q < 4; ++q) acc[q] += BB::mont_lazy(xs[e], ap[4 * (c + e) + q]);
            }
        } else {
            const uint32_t* ap = apow_m;
            for (uint32_t c = sub; c < v.w; c += 16) {
                const uint32_t x = row[c];
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[q] += BB::mont_lazy(x, ap[4 * c + q]);
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
            for (int off = 8; off > 0; off >>= 1) acc[q] += __shfl_xor(acc[q], off, 16);
        if (sub == 0) {
            const E4 R = {{reduce_u64(acc[0]), reduce_u64(acc[1]), reduce_u64(acc[2]), reduce_u64(acc[3])}};
            const uint32_t* a = ab + s * 8;
            const E4 A = {{a[0], a[1], a[2], a[3]}}, Bm = {{a[4], a[5], a[6], a[7]}};
            const E4 t = e4_sub(A, e4_mul_m(R, Bm));
            uint32_t* o = ab + s * 8;
#pragma unroll
            for (int q = 0; q < 4; ++q) o[q] = t.c[q];
        }
    }
}

// Quotient, pass 3 (one thread per output position): q[pos] += term of LDE
// row j = bitrev(pos), read from that row's stored position in ab.
__global__ void quotient_scatter_kernel(LdeView v, const uint32_t* ab, uint32_t* qv) {
    const uint32_t lN = v.L + v.k;
    const uint64_t pos = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (pos >= (1ull << lN)) return;
    const uint32_t j = lN == 0 ? 0u : (__brev(uint32_t(pos)) >> (32 - lN));
    const uint64_t s = (uint64_t(j & ((1u << v.k) - 1)) << v.L) | (j >> v.k);  // stored row of LDE row j
    const uint32_t* t = ab + s * 8;
    uint32_t* o = qv + pos * 4;
#pragma unroll
    for (int q = 0; q < 4; ++q) o[q] = BB::add(o[q], t[q]);
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
    if (uint64_t(P) * w == 0) return 0;
    hipLaunchKernelGGL(bary_finish_kernel, dim3(P * w), dim3(256), 0, st, partial, n_blocks, P, w, van, out);
    return int(hipGetLastError());
}

int pcs_quotient_device(const LdeView& v, uint32_t shift, uint32_t omega_m, const uint32_t* apow_m,
                        const uint32_t* z, const uint32_t* s, const uint32_t* aoff_m, uint32_t P, uint32_t* ab,
                        uint32_t* q, hipStream_t st) {
    const uint64_t N = 1ull << (v.L + v.k);
    hipLaunchKernelGGL(quotient_weights_kernel, dim3(blocks_for(N, 256)), dim3(256), 0, st, v, shift, omega_m, z, s,
                       aoff_m, P, ab);
    const uint64_t grid = std::min<uint64_t>(blocks_for(N * 16, QR_THREADS), 1u << 16);
    const uint32_t lds = v.w <= QR_LDS_COLS ? ((v.w + 3) & ~3u) * 16 : 0;
    hipLaunchKernelGGL(quotient_reduce_kernel, dim3(uint32_t(grid)), dim3(QR_THREADS), lds, st, v, apow_m, ab);
    hipLaunchKernelGGL(quotient_scatter_kernel, dim3(blocks_for(N, 256)), dim3(256), 0, st, v, ab, q);
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
