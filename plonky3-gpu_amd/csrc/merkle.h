// Host-side launchers for Poseidon2 leaf hashing and tree compression (merkle.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "poseidon2.h"
#include "poseidon2_f64.h"

namespace p3h {

// One matrix taking part in a leaf row.  Logical row j is first wrapped to
// j & hmask (cyclic repetition of a shorter matrix; hmask = ~0 for none), then
// mapped to stored row ((j & (2^k - 1)) << L) | (j >> k) of `ptr` (k = 0:
// natural order; k > 0: a coset-major LDE of 2^k cosets of height 2^L).
// ptr == nullptr stands for w zero columns.
template <class T>
struct RowSeg {
    const T* ptr;
    uint64_t hmask;
    uint32_t w;
    uint32_t k;
    uint32_t L;
    uint32_t pad;
};

// out[j] = sponge(concatenation over segs of row j), j < rows.
// segs lives in device memory.
template <class D>
int hash_rows_device(const RowSeg<typename D::T>* segs, uint32_t nsegs, uint32_t total_w,
                     const RowSeg<typename D::T>& seg0_host, uint64_t rows, typename D::T* out,
                     const P2Params<typename D::T, D::W>& p, hipStream_t st);

// next[i] = compress(prev[2i], prev[2i+1] or zero digest if 2i+1 >= prev_count);
// when ninj > 0 additionally next[i] = compress(next[i], sponge(rows i of inj)).
template <class D>
int compress_layer_device(const typename D::T* prev, uint64_t prev_count, typename D::T* next,
                          uint64_t next_count, const RowSeg<typename D::T>* inj, uint32_t ninj,
                          uint32_t inj_w, const P2Params<typename D::T, D::W>& p, hipStream_t st);

// Batch permutation of n states (in place), canonical in and out.
template <class D>
int permute_device(typename D::T* states, uint64_t n, const P2Params<typename D::T, D::W>& p,
                   hipStream_t st);

// out[q * total_w + c] = column c of the concatenation over segs of row rows[q]
// (each segment through its row map; a null segment reads as zero).
template <class T>
int gather_rows_device(const RowSeg<T>* segs, uint32_t nsegs, uint32_t total_w, const uint64_t* rows,
                       uint32_t nq, T* out, hipStream_t st);

// A batch of Merkle openings in one launch.  Rows: vals[q * total_w + c] =
// column c of the concatenated segment rows, segment s read at row
// rows[q * nsegs + s] through its row map (a null segment reads as zero).
// Proofs: proof[(q * steps + l) * dig + d] = word d of digest sib[q * steps + l]
// of layers[l], or zero where that entry is ~0 (past an odd layer's end).
// segs, rows, layers (a device table of layer pointers) and sib are in device
// memory; total_w or steps may be zero.
template <class T>
int open_rows_device(const RowSeg<T>* segs, uint32_t nsegs, uint32_t total_w, const uint64_t* rows,
                     const T* const* layers, uint32_t steps, uint32_t dig, const uint64_t* sib, uint32_t nq, T* vals,
                     T* proof, hipStream_t st);

// BabyBear kernels with the binary64 permutation (poseidon2_f64.h); same
// contracts as the uint32 versions above.
int hash_rows_f64_device(const RowSeg<uint32_t>* segs, uint32_t nsegs, uint32_t total_w, const RowSeg<uint32_t>& seg0,
                         uint64_t rows, uint32_t* out, const P2ParamsF64& p, hipStream_t st);
// Whether hash_rows_f64_device reads the device table `segs` for these
// arguments (otherwise it takes seg0 by value and segs may be null).
bool hash_rows_f64_reads_table(uint32_t nsegs, const RowSeg<uint32_t>& seg0, uint64_t rows);
// Levels in_count -> counts[0] -> ... -> counts[nlev-1] (each ceil(prev / 2),
// zero digest for a missing right child) in one launch; counts[0] <= 64,
// nlev <= 8.  Binary64 engine, no injection.
int compress_tail_f64_device(const uint32_t* in, uint64_t in_count, uint32_t* const* outs, const uint64_t* counts,
                             uint32_t nlev, const P2ParamsF64& p, hipStream_t st);
int compress_layer_f64_device(const uint32_t* prev, uint64_t prev_count, uint32_t* next, uint64_t next_count,
                              const RowSeg<uint32_t>* inj, uint32_t ninj, uint32_t inj_w, const P2ParamsF64& p,
                              hipStream_t st);
int permute_f64_device(uint32_t* states, uint64_t n, const P2ParamsF64& p, hipStream_t st);

// Independent jobs sharing one launch (binary64 engine): the same level of
// several flat trees, or the leaves of several matrices.  Every workgroup
// belongs to one job.
constexpr uint32_t MULTI_MAX = 32;  // jobs per launch; longer lists take several launches
struct MultiJob {
    const uint32_t* in;
    uint32_t* out;
    uint64_t in_count;   // compress: digests of the level below; leaves: unused
    uint64_t out_count;  // digests written
    uint32_t w;          // leaves: row width in elements, a multiple of 8
};
// out[j] = sponge(row j of `in`, a natural-order out_count x w matrix), 16-byte aligned rows.
int hash_rows_multi_f64_device(const MultiJob* jobs, uint32_t n, const P2ParamsF64& p, hipStream_t st);
// out[i] = compress(in[2i], in[2i+1] or the zero digest when 2i+1 >= in_count),
// out_count = ceil(in_count / 2).
int compress_multi_f64_device(const MultiJob* jobs, uint32_t n, const P2ParamsF64& p, hipStream_t st);

}  // namespace p3h
