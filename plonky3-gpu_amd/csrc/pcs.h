// PCS opening kernels (reference: TwoAdicFriPcs::open, two_adic_fri_pcs.hpp:271-395,
// and interpolate_coset_with_precomputation, interpolation.hpp:43-185) and the
// FRI input-mixing step (fri_prover.hpp:123-135).  BabyBear base field,
// BabyBear^4 points and results, all canonical unless marked _m (Montgomery).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace p3h {

// An LDE as produced by the coset LDE kernels: 2^(L+k) rows of w words,
// coset-major (stored row r * 2^L + i is LDE row i * 2^k + r), evaluations of
// the polynomial at shift * omega_N^j, N = 2^(L+k).
struct LdeView {
    const uint32_t* data;
    uint32_t L, k, w;
};

constexpr uint32_t PCS_MAX_POINTS = 4;  // points per launch; more run as several launches

// u[s * P + p] = h_j * common / (z_p - shift * h_j) in Montgomery form, for
// stored row s of LDE row j, h_j = omega_N^j, common = 1 / (N shift^(N-1)).
int pcs_bary_weights_device(const LdeView& v, uint32_t shift, uint32_t omega_m, uint32_t common_m, const uint32_t* z,
                            uint32_t P, uint32_t* u_m, hipStream_t st);
// partial[((b * P + p) * w + c) * 4 + q] = sum over the rows of block b of
// u_p(row) * lde[row][c], coefficient q; rows_per_block rows per block.
int pcs_bary_dot_device(const LdeView& v, const uint32_t* u_m, uint32_t P, uint32_t rows_per_block,
                        uint32_t* partial, hipStream_t st);
// out[(p * w + c) * 4 + q] = van_p * sum_b partial[b][p][c]
int pcs_bary_finish_device(const uint32_t* partial, uint32_t n_blocks, uint32_t P, uint32_t w, const uint32_t* van,
                           uint32_t* out, hipStream_t st);
// q[bitrev(j)] += sum_p aoff_p * (s_p - sum_c apow_c * lde[j][c]) / (z_p - shift * omega_N^j)
// ab: scratch of 8 words per LDE row.
int pcs_quotient_device(const LdeView& v, uint32_t shift, uint32_t omega_m, const uint32_t* apow_m,
                        const uint32_t* z, const uint32_t* s, const uint32_t* aoff_m, uint32_t P, uint32_t* ab,
                        uint32_t* q, hipStream_t st);
// cur[i] += beta^(2^log_arity) * next[i] for i < n (BabyBear^4 vectors); beta
// from beta_dev (device) when not null, else from beta_host.
int fri_mix_device(uint32_t* cur, const uint32_t* next, uint64_t n, const uint32_t* beta_host,
                   const uint32_t* beta_dev, uint32_t log_arity, hipStream_t st);

}  // namespace p3h
