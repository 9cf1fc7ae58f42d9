/*
 * libp3hip -- MI355X (gfx950) STARK trace-commit hot path behind a C ABI.
 *
 * Scope: coset low-degree extension of BabyBear / Goldilocks columns, the
 * Poseidon2 Merkle commitment over the extended rows (multi-height MMCS and the
 * single-matrix MMCS used by the PCS and FRI), the FRI commit phase over
 * BabyBear^4, batched commits of many traces and a serving queue.
 *
 * Each entry point replaces a piece of the reference's C++ API
 * (dloghin/plonky3-gpu, src/include/), cited beside it.  The reference passes
 * RowMajorMatrix<F> by value and throws std::runtime_error /
 * std::invalid_argument; here every call returns an int status (P3H_OK = 0)
 * and p3h_last_error() holds the message, which the C++ adapters in
 * plonky3-gpu_amd/cpp/p3hip.hpp turn back into the same exception types.
 *
 * Data: canonical field elements only -- uint32 in [0, p) for BabyBear
 * (p = 2^31 - 2^27 + 1), uint64 in [0, p) for Goldilocks (p = 2^64 - 2^32 + 1),
 * row-major.  BabyBear^4 values are 4 consecutive uint32 coefficients.
 * Digests: 8 x uint32 (BabyBear, Poseidon2 width 16, rate 8) or 4 x uint64
 * (Goldilocks, Poseidon2 width 8, rate 4).
 *
 * Threading: a context owns one HIP stream and is not thread-safe (like the
 * reference's NttCuda, ntt_cuda.hpp:301-302); separate contexts may be used
 * from separate threads.  Handles belong to the context that made them.
 */
#ifndef P3HIP_H
#define P3HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct p3h_ctx p3h_ctx;
typedef struct p3h_mat p3h_mat;   /* device-resident matrix */
typedef struct p3h_tree p3h_tree; /* device-resident Merkle tree (prover data) */
typedef struct p3h_fri p3h_fri;   /* device-resident FRI commit-phase data */
typedef struct p3h_batch p3h_batch; /* planned batched commit */
typedef struct p3h_serve p3h_serve; /* serving queue */
typedef struct p3h_chal p3h_chal;   /* device-resident Fiat-Shamir transcript */

enum { P3H_BABYBEAR = 0, P3H_GOLDILOCKS = 1 };

enum {
    P3H_OK = 0,
    P3H_EINVAL = 1,  /* bad argument or shape (reference: std::invalid_argument) */
    P3H_ENOMEM = 2,  /* device allocation failed */
    P3H_EDEVICE = 3, /* HIP runtime error (reference: P3_CUDA_CHECK -> runtime_error) */
    P3H_EBUDGET = 4, /* request exceeds the context's HBM budget */
    P3H_ESTATE = 5   /* missing setup (e.g. Poseidon2 constants not set) */
};

/* ------------------------------------------------------- pinned host memory */
/* Page-locked host buffers (hipHostMalloc).  Matrix uploads and serving
 * submissions from them run at the full PCIe DMA rate, about twice the rate
 * from pageable memory.  No counterpart in the reference, whose matrices live
 * in pageable RowMajorMatrix storage.  Returns P3H_ENOMEM when the allocation
 * fails; p3h_host_free(NULL) is a no-op. */
int p3h_host_alloc(uint64_t bytes, void** out);
int p3h_host_free(void* p);

/* ------------------------------------------------------ synthetic inputs */
/* Host-side fill of n canonical field elements: element i is the i-th output
 * of std::mt19937_64(seed) reduced mod p (uint32 for BabyBear, uint64 for
 * Goldilocks), row-major when read as a matrix.  This is the synthetic input
 * stream of the benchmark configurations; the reference's examples draw their
 * matrices the same way from a 32-bit std::mt19937, rng() % p
 * (examples/merkle_tree_cuda_example.cu:53-66).
 * No device work. */
int p3h_seeded_fill(int field, uint64_t seed, void* out, uint64_t n);

/* ---------------------------------------------------------------- context */
/* hbm_budget_bytes = 0 means "all device memory". */
int p3h_ctx_create(int device, uint64_t hbm_budget_bytes, p3h_ctx** out);
int p3h_ctx_destroy(p3h_ctx* ctx);
const char* p3h_last_error(const p3h_ctx* ctx);
int p3h_ctx_sync(p3h_ctx* ctx);
/* bytes held by live handles, the budget, and bytes parked in the allocator cache */
int p3h_ctx_memory(const p3h_ctx* ctx, uint64_t* in_use, uint64_t* budget, uint64_t* cached);
/* The part of in_use held by the context's table caches (twiddles, coset
 * twists, row scales, fold tables), which live until p3h_ctx_destroy; with
 * no handle alive, in_use equals it. */
int p3h_ctx_table_bytes(const p3h_ctx* ctx, uint64_t* bytes);
/* Fault injection: when the environment variable P3H_FAIL_ALLOC_AFTER=N is
 * set at p3h_ctx_create, every device allocation of that context after its
 * first N fails with P3H_ENOMEM, as a failed hipMalloc would; tests use it to
 * check that every failure path returns a clean status and frees what it
 * allocated. */
/* release cached device blocks back to HIP */
int p3h_ctx_trim(p3h_ctx* ctx);
/* per-phase HIP-event timing on the context stream (off by default) */
int p3h_ctx_set_profiling(p3h_ctx* ctx, int on);
/* phase: "lde", "leaf_hash", "compress", "fold"; totals since profiling was switched on */
int p3h_phase_stats(const p3h_ctx* ctx, const char* phase, uint64_t* launches, double* total_ms);
/* BabyBear Poseidon2 evaluation engine: 0 = 32-bit integer Montgomery,
 * 1 = exact binary64 on the FP64 VALU (default; fewer instructions per
 * permutation).  Both give identical digests. */
int p3h_ctx_set_hash_engine(p3h_ctx* ctx, int engine);
/* NTT pass shape: log2 columns per LDS tile, max radix-2 stages per pass, and
 * for coset LDEs log2 of the column slab whose passes run back to back
 * (0 = default / whole width) */
int p3h_ctx_set_ntt_tuning(p3h_ctx* ctx, uint32_t log_cols, uint32_t max_stages, uint32_t log_slab);

/* ------------------------------------------------------------- Poseidon2 */
/* Replaces the POD Poseidon2Cuda<F,W,RF,RP,D> (poseidon2_cuda.hpp:37-46) and the
 * generic Poseidon2 constructor create_poseidon2 (poseidon2.hpp:195-214).
 * consts: initial[rounds_f/2][width], terminal[rounds_f/2][width], internal[rounds_p],
 * canonical uint32 (BabyBear, width 16) or uint64 (Goldilocks, width 8).
 * S-box x^7, linear layers of the reference's generic Poseidon2.
 * Limits: rounds_f <= 8 (even), rounds_p <= 32. */
int p3h_poseidon2_set(p3h_ctx* ctx, int field, uint32_t width, uint32_t rounds_f, uint32_t rounds_p,
                      const void* consts);
/* Batch permutation of n states in place (host buffers); the reference's
 * poseidon2_permute_kernel (poseidon2_cuda.hpp:155-166). */
int p3h_poseidon2_permute(p3h_ctx* ctx, int field, void* states, uint64_t n);

/* -------------------------------------------------------------- matrices */
int p3h_mat_upload(p3h_ctx* ctx, int field, const void* host, uint64_t height, uint32_t width,
                   p3h_mat** out);
int p3h_mat_alloc(p3h_ctx* ctx, int field, uint64_t height, uint32_t width, p3h_mat** out);
/* copy out in natural row order (RowMajorMatrix layout) */
int p3h_mat_download(p3h_ctx* ctx, const p3h_mat* m, void* host);
int p3h_mat_info(const p3h_mat* m, int* field, uint64_t* height, uint32_t* width);
/* download rows [row0, row0 + nrows) in natural order */
int p3h_mat_download_rows(p3h_ctx* ctx, const p3h_mat* m, uint64_t row0, uint64_t nrows, void* host);
int p3h_mat_free(p3h_mat* m);
/* Device address of the matrix's first element, for callers that launch
 * their own work on it or pass it to the device-pointer entry points.  Valid
 * while the handle lives.  Rows are in natural order for uploaded and
 * allocated matrices; LDE outputs are stored coset by coset (stored row
 * r * n + i is LDE row i * 2^k + r for 2^k cosets of height n). */
int p3h_mat_device_ptr(const p3h_mat* m, void** out);

/* overwrite a natural-order matrix with new host contents (same shape) */
int p3h_mat_write(p3h_ctx* ctx, p3h_mat* m, const void* host);

/* ------------------------------------------------------------------- DFT */
/* Replaces NttCuda<F> / Radix2Dit<F> (ntt_cuda.hpp:310-378, radix2_dit.hpp:82-233):
 * kind 0 dft_batch, 1 idft_batch, 2 coset_dft_batch(shift), 3 coset_idft_batch(shift).
 * Height must be a power of two (log2_strict_usize, p3_util/util.hpp:17-23). */
int p3h_dft(p3h_ctx* ctx, int kind, const p3h_mat* in, uint64_t shift, p3h_mat** out);
/* Replaces coset_lde_batch(mat, added_bits, shift) (radix2_dit.hpp:221-233,
 * ntt_cuda.hpp:369-378).  The result stays on the device; p3h_mat_download
 * returns it in the reference's natural row order. */
int p3h_coset_lde(p3h_ctx* ctx, const p3h_mat* in, uint32_t added_bits, uint64_t shift, p3h_mat** out);

/* ------------------------------------------------------------------ MMCS */
/* Multi-height MMCS commit: p3_merkle_tree::MerkleTreeMmcs::commit
 * (merkle_tree_mmcs.hpp:100-107, build_merkle_tree merkle_tree.hpp:82-185) and
 * its GPU twin build_merkle_tree_cuda (merkle_tree_cuda.hpp:358-621).
 * All heights must be powers of two; the tree keeps references to the
 * matrices for openings.  cap: 2^cap_height digests. */
int p3h_mmcs_commit(p3h_ctx* ctx, const p3h_mat* const* mats, uint32_t n_mats, uint32_t cap_height,
                    p3h_tree** out);
/* Single-matrix MMCS: p3_merkle::MerkleTreeMmcs::commit_matrix
 * (fri_merkle_tree_mmcs.hpp:55-102): odd layers padded with a zero digest,
 * commitment = root. */
int p3h_flat_commit(p3h_ctx* ctx, const p3h_mat* m, p3h_tree** out);
/* cap (or root) digests; cap_len receives the number of digests */
int p3h_tree_cap(p3h_ctx* ctx, const p3h_tree* t, void* cap_out, uint32_t* cap_len);
int p3h_tree_info(const p3h_tree* t, uint32_t* n_layers, uint64_t* n_leaves, uint32_t* total_width);
/* The committed matrices of a tree in commit order: their count, and, when
 * the arrays are not null, each one's height and width (what verify_batch
 * takes as dimensions, merkle_tree_mmcs.hpp:168-254).  A multi-height tree
 * lists every matrix; a single-matrix tree lists its one matrix (for a PCS
 * commit, the zero-padded row concatenation). */
int p3h_tree_matrices(const p3h_tree* t, uint32_t* n_mats, uint64_t* heights, uint32_t* widths);
/* copy digest layer `layer` (0 = leaves) to the host */
int p3h_tree_layer(p3h_ctx* ctx, const p3h_tree* t, uint32_t layer, void* out);
/* Openings for n_idx row indices at once: open_batch (merkle_tree_mmcs.hpp:120-152)
 * for multi-height trees, open_row (fri_merkle_tree_mmcs.hpp:129-141) for flat
 * trees.  values_out: per index, the opened rows of every matrix in commit
 * order, concatenated (total_width elements); proof_out: per index,
 * (n_layers - 1) sibling digests, leaf level first. */
int p3h_tree_open(p3h_ctx* ctx, const p3h_tree* t, const uint64_t* indices, uint32_t n_idx,
                  void* values_out, void* proof_out);
int p3h_tree_free(p3h_tree* t);
/* Verifiers for an opening, each permutation run on the device:
 * p3_merkle_tree::MerkleTreeMmcs::verify_batch (merkle_tree_mmcs.hpp:168-254):
 * heights/widths: the committed matrices' dimensions in commit order; values:
 * their opened rows concatenated in that order (as p3h_tree_open writes them);
 * proof: proof_len sibling digests, leaf level first; cap: cap_len digests.
 * *ok = 1 when the path reproduces the cap entry, 0 when it does not (a wrong
 * proof length, a sibling or value that does not hash to the cap); an error
 * status only for null pointers, an unknown field, non-canonical elements or
 * an unset Poseidon2. */
int p3h_mmcs_verify(p3h_ctx* ctx, int field, const void* cap, uint32_t cap_len, const uint64_t* heights,
                    const uint32_t* widths, uint32_t n_mats, uint64_t index, const void* values, const void* proof,
                    uint32_t proof_len, int* ok);
/* p3_merkle::MerkleTreeMmcs::verify_row (fri_merkle_tree_mmcs.hpp:146-168):
 * leaf = sponge(row), then one compression per sibling; root: n_root digests
 * (the commitment vector; *ok = 0 when it is empty). */
int p3h_flat_verify(p3h_ctx* ctx, int field, const void* root, uint32_t n_root, uint64_t index, const void* row,
                    uint32_t width, const void* proof, uint32_t proof_len, int* ok);

/* ------------------------------------------------------------------- FRI */
/* fold_matrix_cuda (fri_fold_cuda.hpp:306-348) / TwoAdicFriFolding::fold_matrix
 * (fri_folding.hpp:200-222): `cur` is a (2^log_height x 4) BabyBear matrix
 * holding BabyBear^4 values; beta = 4 canonical coefficients; log_arity 1..3. */
int p3h_fri_fold(p3h_ctx* ctx, const p3h_mat* cur, uint32_t log_arity, const uint32_t* beta,
                 p3h_mat** out);
/* fold_matrix_cuda_device (fri_fold_cuda.hpp:250-277) on device buffers:
 * d_in holds n * 2^log_arity BabyBear^4 values (4 canonical words each),
 * d_out receives n; row i folds d_in[i * arity .. (i + 1) * arity) over the
 * domain of size 2^log_height, as p3h_fri_fold.  n <= 2^(log_height -
 * log_arity).  Like the reference it copies nothing to or from the host and
 * returns when the fold is done; the arity-2 table is cached in the context
 * (allocated on the first call of each height). */
int p3h_fri_fold_device(p3h_ctx* ctx, const uint32_t* d_in, uint32_t* d_out, const uint32_t* beta,
                        uint32_t log_arity, uint32_t log_height, uint64_t n);
/* No counterpart in the reference: p3h_fri_fold_device for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments and result; it
 * calls no hipMalloc/hipFree, keeps the arity-2 table in a static 16 MiB
 * device array per GPU (set up at p3h_ctx_create), and returns once the work
 * is enqueued on the context's stream (p3h_ctx_sync waits for it).  Limit: at
 * arity 2, n <= 2^22 (P3H_EINVAL beyond).  Calls that overlap on one GPU
 * share that array and race. */
int p3h_fri_fold_device_noalloc(p3h_ctx* ctx, const uint32_t* d_in, uint32_t* d_out, const uint32_t* beta,
                                uint32_t log_arity, uint32_t log_height, uint64_t n);
/* prove_fri commit loop (fri_prover.hpp:83-178) for one input codeword of
 * BabyBear^4 (2^log_len x 4 matrix), with the Fiat-Shamir challenges supplied
 * by the caller (betas: 4 words per round).  Each round commits the current
 * vector as a (len/arity) x (4*arity) matrix (ExtensionMmcs,
 * fri_extension_mmcs.hpp:49-75) and folds.  roots_out: 8 words per round;
 * final_poly_out: 4 * 2^log_final_poly_len words; keep (optional) receives the
 * layer trees for later queries. */
int p3h_fri_commit_phase(p3h_ctx* ctx, const p3h_mat* codeword, uint32_t log_blowup,
                         uint32_t log_final_poly_len, uint32_t max_log_arity, const uint32_t* betas,
                         uint32_t* roots_out, uint32_t* final_poly_out, uint32_t* n_rounds,
                         p3h_fri** keep);
/* The same loop with the transcript in the loop, as prove_fri runs it
 * (fri_prover.hpp:83-190) with the FriChallenger of the reference's tests
 * (tests/fri_test.cpp:99-124): per round observe_commitment (the layer root),
 * grind(pow_bits), sample_challenge (beta = 4 samples), fold; afterwards the
 * final polynomial's coefficients and each round's log arity are observed.
 * The transcript lives on the device; with pow_bits = 0 no round waits for
 * the host.  betas_out: 4 words per round; witnesses_out: one per round. */
int p3h_fri_commit_phase_chal(p3h_ctx* ctx, const p3h_mat* codeword, uint32_t log_blowup,
                              uint32_t log_final_poly_len, uint32_t max_log_arity, p3h_chal* challenger,
                              uint32_t pow_bits, uint32_t* roots_out, uint32_t* betas_out, uint64_t* witnesses_out,
                              uint32_t* final_poly_out, uint32_t* n_rounds, p3h_fri** keep);
/* query the kept FRI layers: layer r tree (borrowed pointer, valid until p3h_fri_free) */
int p3h_fri_layer_tree(const p3h_fri* f, uint32_t round, const p3h_tree** out);
int p3h_fri_free(p3h_fri* f);

/* ------------------------------------------------------------ challenger */
/* DuplexChallenger<BabyBear, Poseidon2 W16, 16, 8> (duplex_challenger.hpp:42-268)
 * with the context's BabyBear Poseidon2 constants, its state in device memory.
 * Calls run on the context's stream, in order. */
int p3h_challenger_create(p3h_ctx* ctx, p3h_chal** out);
/* copy constructor (duplex_challenger.hpp:80-82) */
int p3h_challenger_clone(const p3h_chal* ch, p3h_chal** out);
int p3h_challenger_free(p3h_chal* ch);
/* observe_slice (duplex_challenger.hpp:100-113): n words, each reduced mod p
 * as BabyBear(uint32_t) does; n = 0 changes nothing */
int p3h_challenger_observe(p3h_chal* ch, const uint32_t* values, uint32_t n);
/* observe_merkle_cap (duplex_challenger.hpp:206-220) of a BabyBear tree's cap
 * (its root for a flat tree), read from device memory */
int p3h_challenger_observe_cap(p3h_chal* ch, const p3h_tree* tree);
/* n x sample() (duplex_challenger.hpp:120-129); sample_algebra_element<BabyBear4>
 * is n = 4 (coefficients in order) */
int p3h_challenger_sample(p3h_chal* ch, uint32_t n, uint32_t* out);
/* sample_bits (duplex_challenger.hpp:137-148), bits <= 64 */
int p3h_challenger_sample_bits(p3h_chal* ch, uint32_t bits, uint64_t* out);
/* grind (duplex_challenger.hpp:236-254): the least witness whose following
 * sample has `bits` zero low bits (bits < 31), searched on the GPU; then the
 * witness is observed and one sample consumed */
int p3h_challenger_grind(p3h_chal* ch, uint32_t bits, uint64_t* witness);
/* check_witness (duplex_challenger.hpp:262-268) */
int p3h_challenger_check_witness(p3h_chal* ch, uint32_t bits, uint64_t witness, int* ok);
/* state for inspection: sponge[16], input[8], output[8], input length, output length */
int p3h_challenger_state(p3h_chal* ch, uint32_t* out34);

/* ------------------------------------------------------------------- PCS */
/* TwoAdicFriPcs::commit (two_adic_fri_pcs.hpp:207-266): LDE every matrix with
 * log_blowup and its shift, concatenate rows cyclically (row r of matrix m is
 * LDE row r mod height_m), zero-pad the width to a power of two, and commit
 * the result with the single-matrix MMCS.  BabyBear only. */
int p3h_pcs_commit(p3h_ctx* ctx, const p3h_mat* const* mats, const uint64_t* shifts, uint32_t n_mats,
                   uint32_t log_blowup, p3h_tree** out);

/* TwoAdicFriPcs::open (two_adic_fri_pcs.hpp:271-455) for one batch: `tree` is
 * the result of p3h_pcs_commit (its LDEs), shifts the matrices' domain
 * shifts, n_points[m] BabyBear^4 points for matrix m (4 words each,
 * concatenated).  On the device: alpha is sampled from the transcript, every
 * matrix is evaluated at its points by barycentric interpolation on its LDE
 * coset (interpolation.hpp:152-185), the transcript observes the values, one
 * quotient vector per LDE height is accumulated (two_adic_fri_pcs.hpp:335-381)
 * and prove_fri runs over them (fri_prover.hpp:47-280): commit rounds with the
 * smaller inputs mixed in, final polynomial, query proof of work, query
 * indices and the input rows each query opens (PcsQueryInputMmcs::open,
 * two_adic_fri_pcs.hpp:103-121).
 * opened_out: [matrix][point][column] x 4 words; roots_out: 8 per round;
 * query_rows_out: per query, each height group's rows (descending heights,
 * matrices in commit order), total width words; keep (optional): the FRI
 * layers, for the commit-phase openings of the queries. */
int p3h_pcs_open(p3h_ctx* ctx, const p3h_tree* tree, const uint64_t* shifts, const uint32_t* n_points,
                 const uint32_t* points, p3h_chal* challenger, uint32_t log_final_poly_len, uint32_t max_log_arity,
                 uint32_t num_queries, uint32_t commit_pow_bits, uint32_t query_pow_bits, uint32_t* opened_out,
                 uint32_t* roots_out, uint64_t* commit_witnesses_out, uint32_t* final_poly_out,
                 uint64_t* query_witness_out, uint64_t* query_indices_out, uint32_t* query_rows_out,
                 uint32_t* n_rounds, p3h_fri** keep);

/* --------------------------------------------------------------- batched */
/* Commit many BabyBear traces as one multi-height MMCS: every trace gets its
 * own LDE (added bits per trace, common shift) and the LDEs are committed
 * together (multi-height tree, cap_height).  No counterpart in the reference,
 * which commits one matrix per call; the result equals
 * p3h_mmcs_commit over the individual p3h_coset_lde outputs. */
int p3h_batch_commit(p3h_ctx* ctx, const p3h_mat* const* traces, const uint32_t* added_bits,
                     uint32_t n_traces, uint64_t shift, uint32_t cap_height, p3h_tree** out);

/* Planned batched commit: every buffer is allocated and every table uploaded
 * once.  The LDE passes of all traces run as work-list launches grouped by
 * pass shape; with P3H_BATCH_QUEUE_LDE they run instead as one launch whose
 * workgroups take (trace, pass, column tile, row-block chunk) items from a
 * device work list in dependency order (when every pass shape is carried by
 * that kernel; else the grouped launches run).  With P3H_BATCH_GRAPH the whole
 * commit (LDE passes, leaf hashing, tree levels with injection) is captured
 * as one HIP graph and replayed by p3h_batch_run on the traces' current
 * contents.  No counterpart in the reference.  flags: a bitwise or of the
 * P3H_BATCH_* values (the earlier use_graph argument: 1 = graph).  cap_out
 * (nullable): 2^cap_height digests. */
#define P3H_BATCH_GRAPH 1
#define P3H_BATCH_QUEUE_LDE 2
int p3h_batch_plan(p3h_ctx* ctx, const p3h_mat* const* traces, const uint32_t* added_bits, uint32_t n_traces,
                   uint64_t shift, uint32_t cap_height, int flags, p3h_batch** out);
int p3h_batch_run(p3h_batch* b, void* cap_out);
/* kernel launches of the LDE phase per run: 1 for the queue launch, else the
 * number of pass-shape groups */
int p3h_batch_lde_launches(const p3h_batch* b, uint32_t* launches);
/* tree of the last run (borrowed; overwritten by the next run) */
int p3h_batch_tree(const p3h_batch* b, const p3h_tree** out);
int p3h_batch_free(p3h_batch* b);

/* --------------------------------------------------------------- serving */
/* A commit server for BabyBear traces: n_slots worker threads, each with its
 * own context and HIP stream, take requests from one queue.  A request is
 * admitted when its estimated footprint (trace + LDE + Merkle layers) fits the
 * HBM budget next to the results still held; a worker takes every admissible
 * queued request (up to max_batch) and runs their LDEs together, then each
 * tree.  Results stay resident in HBM for openings until released.  Requests
 * on different slots overlap on the device; a request's result does not
 * depend on what else is in flight.  No counterpart in the reference.
 * consts: BabyBear Poseidon2 constants as for p3h_poseidon2_set. */
int p3h_serve_create(int device, uint64_t hbm_budget, uint32_t n_slots, uint32_t max_batch, const uint32_t* consts,
                     uint32_t rounds_f, uint32_t rounds_p, p3h_serve** out);
/* borrows the trace: the host rows must stay valid and unchanged until the
 * request is done (p3h_serve_wait / p3h_serve_poll), when the upload has
 * happened.  Returns P3H_EBUDGET if the request could never fit the budget. */
int p3h_serve_submit(p3h_serve* sv, const uint32_t* trace, uint64_t height, uint32_t width, uint32_t added_bits,
                     uint64_t shift, uint64_t* request_id);
/* blocks until the request is committed; root_out: 8 words; latency: submit -> committed */
int p3h_serve_wait(p3h_serve* sv, uint64_t request_id, uint32_t* root_out, double* latency_ms);
int p3h_serve_poll(p3h_serve* sv, uint64_t request_id, int* done);
/* openings from the resident tree (as p3h_tree_open) */
int p3h_serve_open(p3h_serve* sv, uint64_t request_id, const uint64_t* indices, uint32_t n_idx, uint32_t* values_out,
                   uint32_t* proof_out);
/* drop the request's resident LDE and tree; frees its share of the budget */
int p3h_serve_release(p3h_serve* sv, uint64_t request_id);
int p3h_serve_stats(p3h_serve* sv, uint64_t* completed, uint64_t* batches, uint64_t* reserved_bytes,
                    uint64_t* committed_elems);
const char* p3h_serve_last_error(p3h_serve* sv);
/* Requests still queued fail with P3H_ESTATE (their waiters wake), batches
 * being processed complete, calls in progress on other threads return; then
 * every resident result is freed.  No call on the server may start once
 * destroy has begun (submissions racing with it get P3H_ESTATE). */
int p3h_serve_destroy(p3h_serve* sv);

/* ----------------------------------------------------------- measurement */
/* The VALU ceiling of the NTT arithmetic: `blocks` x 256 threads each take 16
 * words through iters radix-16 rounds of the NTT passes' own register-round
 * code (BabyBear: 32 butterflies of a modular product, sum and difference;
 * Goldilocks: the twisted round, 15 products then 32 power-of-two
 * butterflies), twiddles in registers, with no memory traffic in the timed
 * loop.  *ms: the kernel time (HIP events on the context's stream);
 * *butterflies: the radix-2 butterflies those rounds stand for (32 per round).
 * No counterpart in the reference; the benchmark reports the rate as the
 * passes' VALU roofline. */
int p3h_probe_butterflies(p3h_ctx* ctx, int field, uint32_t blocks, uint32_t iters, double* ms,
                          double* butterflies);

/* ------------------------------------------------------------ autotuner */
/* Time the LDE of one shape under candidate pass shapes (columns per LDS tile,
 * stages per pass, column slab) and record the fastest for that (field,
 * log_n, width, added_bits); later LDEs and batch plans of that shape use it
 * (batch plans ignore the slab).  No counterpart in the reference. */
int p3h_autotune_lde(p3h_ctx* ctx, int field, uint32_t log_n, uint32_t width, uint32_t added_bits,
                     uint32_t iters, uint32_t* best_log_cols, uint32_t* best_max_stages, uint32_t* best_log_slab,
                     double* best_ms);
/* the recorded table as text, one "field log_n width added log_cols max_stages log_slab" per line */
int p3h_tuning_save(const p3h_ctx* ctx, const char* path);
int p3h_tuning_load(p3h_ctx* ctx, const char* path);

#ifdef __cplusplus
}
#endif

#endif /* P3HIP_H */
