/*
 * CPU oracle: a plain C++ restatement of the reference's CPU path for the STARK
 * commit hot path (coset LDE, Poseidon2 sponge/compression, both MMCS trees,
 * FRI commit phase).  TEST INFRASTRUCTURE ONLY: it is loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and only as the
 * checker or the CPU baseline.  The product library never links or calls it.
 *
 * Values crossing this interface are canonical field elements: uint32 in
 * [0, p) for BabyBear (p = 2^31 - 2^27 + 1), uint64 in [0, p) for Goldilocks
 * (p = 2^64 - 2^32 + 1), and BabyBear^4 as 4 consecutive uint32 coefficients.
 *
 * Poseidon2 constants: initial[rf/2][W], terminal[rf/2][W], internal[rp],
 * canonical, the field order of Poseidon2Cuda (src/include/poseidon2_cuda.hpp:37-46).
 * BabyBear: W = 16, rate 8, digest 8.  Goldilocks: W = 8, rate 4, digest 4.
 * The S-box exponent is 7 throughout.
 *
 * Every function returns 0 on success and a negative value on a bad argument.
 */
#ifndef P3_ORACLE_H
#define P3_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* field helpers (used by the tests' closed-form checks) */
uint32_t orc_bb_mul(uint32_t a, uint32_t b);
uint32_t orc_bb_pow(uint32_t a, uint64_t e);
uint32_t orc_bb_two_adic_generator(uint32_t bits);
uint64_t orc_gl_mul(uint64_t a, uint64_t b);
uint64_t orc_gl_pow(uint64_t a, uint64_t e);
uint64_t orc_gl_two_adic_generator(uint32_t bits);
void orc_bb4_mul(const uint32_t* a, const uint32_t* b, uint32_t* out);

/* seeded inputs: std::mt19937_64(seed), value = rng() % p, row-major */
void orc_bb_fill(uint64_t seed, uint32_t* out, size_t n);
void orc_gl_fill(uint64_t seed, uint64_t* out, size_t n);

/* Poseidon2 permutation over n states (in place) */
int orc_bb_permute(const uint32_t* consts, int rf, int rp, uint32_t* states, size_t n);
int orc_gl_permute(const uint64_t* consts, int rf, int rp, uint64_t* states, size_t n);

/* padding-free overwrite sponge and 2-to-1 truncated-permutation compression */
int orc_bb_sponge(const uint32_t* consts, int rf, int rp, const uint32_t* data, size_t len, uint32_t* out8);
int orc_gl_sponge(const uint64_t* consts, int rf, int rp, const uint64_t* data, size_t len, uint64_t* out4);
int orc_bb_compress(const uint32_t* consts, int rf, int rp, const uint32_t* in16, uint32_t* out8);
int orc_gl_compress(const uint64_t* consts, int rf, int rp, const uint64_t* in8, uint64_t* out4);

/* DFT family over row-major h x w matrices; kind 0 dft, 1 idft, 2 coset_dft, 3 coset_idft */
int orc_bb_dft(int kind, const uint32_t* in, size_t log_h, size_t w, uint32_t shift, uint32_t* out);
int orc_gl_dft(int kind, const uint64_t* in, size_t log_h, size_t w, uint64_t shift, uint64_t* out);
/* coset LDE: (2^log_h x w) -> (2^(log_h+added) x w), natural row order */
int orc_bb_lde(const uint32_t* in, size_t log_h, size_t w, size_t added, uint32_t shift, uint32_t* out);
int orc_gl_lde(const uint64_t* in, size_t log_h, size_t w, size_t added, uint64_t shift, uint64_t* out);
/* the same LDE by direct O(n*N) polynomial evaluation, for small sizes */
int orc_bb_lde_brute(const uint32_t* in, size_t log_h, size_t w, size_t added, uint32_t shift, uint32_t* out);
int orc_gl_lde_brute(const uint64_t* in, size_t log_h, size_t w, size_t added, uint64_t shift, uint64_t* out);

/* single-matrix MMCS (odd layers padded with a zero digest); tree_out (optional)
 * receives every layer, leaves first */
int orc_bb_flat_commit(const uint32_t* consts, int rf, int rp, const uint32_t* data, size_t height,
                       size_t width, uint32_t* root8, uint32_t* tree_out);
int orc_gl_flat_commit(const uint64_t* consts, int rf, int rp, const uint64_t* data, size_t height,
                       size_t width, uint64_t* root4, uint64_t* tree_out);
/* LDE + flat commit on `threads` host threads (column slabs for the LDE, row
 * and node ranges for the tree): the same root as orc_*_lde then
 * orc_*_flat_commit; the CPU baseline's multi-core figure */
int orc_bb_lde_commit_mt(const uint32_t* consts, int rf, int rp, const uint32_t* in, size_t log_h, size_t w,
                         size_t added, uint32_t shift, int threads, uint32_t* root8);
int orc_gl_lde_commit_mt(const uint64_t* consts, int rf, int rp, const uint64_t* in, size_t log_h, size_t w,
                         size_t added, uint64_t shift, int threads, uint64_t* root4);
/* the same on `threads` host threads, with openings at n_idx LDE rows:
 * values n_idx x w (natural row order), proof n_idx x (log_h + added) digests,
 * leaf level first */
int orc_bb_lde_commit_open_mt(const uint32_t* consts, int rf, int rp, const uint32_t* in, size_t log_h, size_t w,
                              size_t added, uint32_t shift, int threads, const uint64_t* idx, size_t n_idx,
                              uint32_t* root8, uint32_t* values, uint32_t* proof);
int orc_gl_lde_commit_open_mt(const uint64_t* consts, int rf, int rp, const uint64_t* in, size_t log_h, size_t w,
                              size_t added, uint64_t shift, int threads, const uint64_t* idx, size_t n_idx,
                              uint64_t* root4, uint64_t* values, uint64_t* proof);
/* batched commit on `threads` host threads: LDE of every BabyBear trace (its
 * own added bits, common shift), one multi-height MMCS over the LDEs in trace
 * order with a cap of 2^cap_height digests, openings at n_idx indices of the
 * tallest LDE (values: every LDE's row in trace order; proof: *proof_len
 * digests per index) */
int orc_bb_batch_commit_mt(const uint32_t* consts, int rf, int rp, size_t n, const size_t* log_hs, const size_t* ws,
                           const size_t* adds, const uint32_t* const* traces, uint32_t shift, size_t cap_height,
                           int threads, const uint64_t* idx, size_t n_idx, uint32_t* cap_out, uint32_t* values_out,
                           uint32_t* proof_out, size_t* proof_len);
/* single-matrix opening: row values and log2(height) sibling digests */
int orc_bb_flat_open(const uint32_t* consts, int rf, int rp, const uint32_t* data, size_t height,
                     size_t width, size_t index, uint32_t* row_out, uint32_t* proof_out);

/* multi-height MMCS: commit (cap of 2^cap_height digests) and optional opening
 * at open_index (< 0: none).  values_out: opened rows in original matrix order,
 * concatenated; proof_out: *proof_len digests. */
int orc_bb_mmcs(const uint32_t* consts, int rf, int rp, size_t n_mats, const size_t* heights,
                const size_t* widths, const uint32_t* const* mats, size_t cap_height,
                uint32_t* cap_out, long long open_index, uint32_t* values_out,
                uint32_t* proof_out, size_t* proof_len);
int orc_gl_mmcs(const uint64_t* consts, int rf, int rp, size_t n_mats, const size_t* heights,
                const size_t* widths, const uint64_t* const* mats, size_t cap_height,
                uint64_t* cap_out, long long open_index, uint64_t* values_out,
                uint64_t* proof_out, size_t* proof_len);
/* multi-height verification; returns 1 accept, 0 reject */
int orc_bb_mmcs_verify(const uint32_t* consts, int rf, int rp, size_t n_mats, const size_t* heights,
                       const size_t* widths, size_t cap_height, const uint32_t* cap,
                       size_t index, const uint32_t* values, const uint32_t* proof, size_t proof_len);

/* FRI: fold (any arity 2^1..2^3) and the commit phase with supplied betas */
int orc_bb4_fold(size_t log_height, size_t log_arity, const uint32_t* beta, const uint32_t* cur,
                 size_t len, uint32_t* out);
int orc_bb4_fri_commit(const uint32_t* consts, int rf, int rp, const uint32_t* codeword,
                       size_t log_len, size_t log_blowup, size_t log_final_poly_len,
                       size_t max_log_arity, const uint32_t* betas, uint32_t* roots_out,
                       uint32_t* final_out, size_t* n_rounds);

/* the FRI commit phase with the folds and layer trees on `threads` host threads */
int orc_bb4_fri_commit_mt(const uint32_t* consts, int rf, int rp, const uint32_t* codeword,
                          size_t log_len, size_t log_blowup, size_t log_final_poly_len,
                          size_t max_log_arity, const uint32_t* betas, int threads, uint32_t* roots_out,
                          uint32_t* final_out, size_t* n_rounds);

/* PCS commit: LDE every matrix (added bits = log_blowup, shift per matrix),
 * concatenate rows cyclically, zero-pad the width to a power of two, flat commit */
int orc_bb_challenger_run(const uint32_t* consts, int rf, int rp, const uint64_t* ops, size_t n_ops, uint64_t* out,
                          size_t* n_out);
int orc_bb4_fri_commit_chal(const uint32_t* consts, int rf, int rp, const uint32_t* codeword, size_t log_len,
                            size_t log_blowup, size_t log_final_poly_len, size_t max_log_arity, size_t pow_bits,
                            const uint32_t* pre, size_t n_pre, uint32_t* roots_out, uint32_t* betas_out,
                            uint64_t* witnesses_out, uint32_t* final_out, uint32_t* post_out, size_t* n_rounds);
int orc_bb_pcs_open(const uint32_t* consts, int rf, int rp, size_t n, const size_t* log_hs, const size_t* ws,
                    const uint32_t* shifts, const uint32_t* const* mats, size_t log_blowup, const size_t* n_points,
                    const uint32_t* points, const uint32_t* pre, size_t n_pre, size_t lfpl, size_t mla,
                    size_t num_queries, size_t commit_pow_bits, size_t query_pow_bits, uint32_t* root_out,
                    uint32_t* opened_out, uint32_t* roots_out, uint64_t* cwit_out, uint32_t* final_out,
                    uint64_t* qwit_out, uint64_t* qidx_out, uint32_t* qrows_out, uint32_t* post_out,
                    size_t* n_rounds);
int orc_bb_pcs_commit(const uint32_t* consts, int rf, int rp, size_t n_mats, const size_t* log_heights,
                      const size_t* widths, const uint32_t* shifts, const uint32_t* const* mats,
                      size_t log_blowup, uint32_t* root8);

#ifdef __cplusplus
}
#endif

#endif /* P3_ORACLE_H */
