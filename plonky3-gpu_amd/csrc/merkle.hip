// Poseidon2 Merkle kernels: padding-free sponge leaf hashing over one or more
// row-concatenated matrices, 2-to-1 truncated-permutation compression with
// optional injection of shorter matrices, batch permutation and row gathers
// for openings.
//
// Sponge (reference: PaddingFreeSponge::hash_iter, padding_free_sponge.hpp:63-87):
// zero state; every chunk of up to RATE inputs overwrites the first slots;
// a permutation follows every non-empty chunk; digest = first DIG lanes.
// Compression (TruncatedPermutation::compress, truncated_permutation.hpp:61-76):
// state = [left | right | 0...], permute, keep the first DIG lanes.
// Injection (build_merkle_tree, merkle_tree.hpp:163-177): node =
// compress(node, sponge(rows of the matrices whose height is this level)).

#include "merkle.h"
#include "poseidon2_f64.h"

namespace p3h {

namespace {

constexpr int HASH_THREADS = 256;

template <class T>
__device__ __forceinline__ uint64_t stored_row(const RowSeg<T>& s, uint64_t j) {
    j &= s.hmask;
    if (s.k == 0) return j;
    return ((j & ((1ull << s.k) - 1)) << s.L) | (j >> s.k);
}

// Absorb the concatenated row j of `segs` into state s (domain form).
// RELOAD: see Poseidon2GL::permute.
template <class D, bool RELOAD>
__device__ __forceinline__ void sponge_row(typename D::T* s, const RowSeg<typename D::T>* segs, uint32_t nsegs,
                                           uint32_t total_w, uint64_t j, const P2Params<typename D::T, D::W>& p) {
    using T = typename D::T;
#pragma unroll
    for (int i = 0; i < D::W; ++i) s[i] = 0;
    uint32_t seg = 0, off = 0;
    RowSeg<T> cur = segs[0];
    const T* rowp = cur.ptr + stored_row(cur, j) * cur.w;
    for (uint32_t start = 0; start < total_w; start += D::RATE) {
#pragma unroll
        for (int q = 0; q < D::RATE; ++q) {
            if (start + q < total_w) {
                while (off >= cur.w) {
                    off = 0;
                    ++seg;
                    cur = segs[seg];
                    rowp = cur.ptr + stored_row(cur, j) * cur.w;
                }
                s[q] = cur.ptr ? D::to_dom(rowp[off]) : T(0);
                ++off;
            }
        }
        Poseidon2<D>::template permute<RELOAD>(s, p);
    }
    (void)nsegs;
}

template <class D>
__global__ __launch_bounds__(HASH_THREADS) void hash_rows_kernel(const RowSeg<typename D::T>* segs,
                                                                 uint32_t nsegs, uint32_t total_w,
                                                                 uint64_t rows, typename D::T* out,
                                                                 P2Params<typename D::T, D::W> p) {
    using T = typename D::T;
    uint64_t j = uint64_t(blockIdx.x) * HASH_THREADS + threadIdx.x;
    if (j >= rows) return;
    T s[D::W];
    sponge_row<D, true>(s, segs, nsegs, total_w, j, p);
#pragma unroll
    for (int i = 0; i < D::DIG; ++i) out[j * D::DIG + i] = D::from_dom(s[i]);
}

// Single matrix whose width is a multiple of RATE: each chunk is RATE
// contiguous elements, fetched with 16-byte loads.
template <class D>
__global__ __launch_bounds__(HASH_THREADS) void hash_rows_dense_kernel(RowSeg<typename D::T> seg, uint64_t rows,
                                                                       typename D::T* out,
                                                                       P2Params<typename D::T, D::W> p) {
    using T = typename D::T;
    constexpr int PER16 = 16 / sizeof(T);
    uint64_t j = uint64_t(blockIdx.x) * HASH_THREADS + threadIdx.x;
    if (j >= rows) return;
    T s[D::W];
#pragma unroll
    for (int i = 0; i < D::W; ++i) s[i] = 0;
    const T* rowp = seg.ptr + stored_row(seg, j) * seg.w;
    for (uint32_t start = 0; start < seg.w; start += D::RATE) {
        const uint4* v = reinterpret_cast<const uint4*>(rowp + start);
#pragma unroll
        for (int q = 0; q < D::RATE / PER16; ++q) {
            uint4 x = v[q];
            T tmp[PER16];
            __builtin_memcpy(tmp, &x, 16);
#pragma unroll
            for (int e = 0; e < PER16; ++e) s[q * PER16 + e] = D::to_dom(tmp[e]);
        }
        Poseidon2<D>::template permute<true>(s, p);
    }
#pragma unroll
    for (int i = 0; i < D::DIG; ++i) out[j * D::DIG + i] = D::from_dom(s[i]);
}

template <class D>
__global__ __launch_bounds__(HASH_THREADS) void compress_kernel(const typename D::T* prev, uint64_t prev_count,
                                                                typename D::T* next, uint64_t next_count,
                                                                const RowSeg<typename D::T>* inj, uint32_t ninj,
                                                                uint32_t inj_w, P2Params<typename D::T, D::W> p) {
    using T = typename D::T;
    uint64_t i = uint64_t(blockIdx.x) * HASH_THREADS + threadIdx.x;
    if (i >= next_count) return;
    T s[D::W];
#pragma unroll
    for (int q = 0; q < D::W; ++q) s[q] = 0;
    const T* l = prev + 2 * i * D::DIG;
#pragma unroll
    for (int q = 0; q < D::DIG; ++q) s[q] = D::to_dom(l[q]);
    if (2 * i + 1 < prev_count) {
#pragma unroll
        for (int q = 0; q < D::DIG; ++q) s[D::DIG + q] = D::to_dom(l[D::DIG + q]);
    }
    Poseidon2<D>::permute(s, p);
    if (ninj) {
        T h[D::W];
        sponge_row<D, false>(h, inj, ninj, inj_w, i, p);
#pragma unroll
        for (int q = 0; q < D::DIG; ++q) s[D::DIG + q] = h[q];
#pragma unroll
        for (int q = 2 * D::DIG; q < D::W; ++q) s[q] = 0;
        Poseidon2<D>::permute(s, p);
    }
#pragma unroll
    for (int q = 0; q < D::DIG; ++q) next[i * D::DIG + q] = D::from_dom(s[q]);
}

template <class D>
__global__ __launch_bounds__(HASH_THREADS) void permute_kernel(typename D::T* st, uint64_t n,
                                                               P2Params<typename D::T, D::W> p) {
    using T = typename D::T;
    uint64_t i = uint64_t(blockIdx.x) * HASH_THREADS + threadIdx.x;
    if (i >= n) return;
    T s[D::W];
#pragma unroll
    for (int q = 0; q < D::W; ++q) s[q] = D::to_dom(st[i * D::W + q]);
    Poseidon2<D>::permute(s, p);
#pragma unroll
    for (int q = 0; q < D::W; ++q) st[i * D::W + q] = D::from_dom(s[q]);
}

// out[q * total_w + c] = concatenated row rows[q], column c
template <class T>
__global__ void gather_rows_kernel(const RowSeg<T>* segs, uint32_t nsegs, uint32_t total_w, const uint64_t* rows,
                                   uint32_t nq, T* out) {
    uint64_t idx = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= uint64_t(nq) * total_w) return;
    uint32_t q = uint32_t(idx / total_w), c = uint32_t(idx % total_w);
    uint32_t seg = 0;
    while (seg + 1 < nsegs && c >= segs[seg].w) {
        c -= segs[seg].w;
        ++seg;
    }
    const RowSeg<T>& s = segs[seg];
    out[idx] = s.ptr ? s.ptr[stored_row(s, rows[q]) * s.w + c] : T(0);
}

// One thread per output word of a batch of openings (open_rows_device).
template <class T>
__global__ void open_rows_kernel(const RowSeg<T>* segs, uint32_t nsegs, uint32_t total_w, const uint64_t* rows,
                                 const T* const* layers, uint32_t steps, uint32_t dig, const uint64_t* sib,
                                 uint32_t nq, T* vals, T* proof) {
    const uint32_t per_q = total_w + steps * dig;
    const uint64_t idx = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= uint64_t(nq) * per_q) return;
    const uint32_t q = uint32_t(idx / per_q), k = uint32_t(idx % per_q);
    if (k < total_w) {
        uint32_t seg = 0, c = k;
        while (seg + 1 < nsegs && c >= segs[seg].w) {
            c -= segs[seg].w;
            ++seg;
        }
        const RowSeg<T>& s = segs[seg];
        vals[uint64_t(q) * total_w + k] =
            s.ptr ? s.ptr[stored_row(s, rows[uint64_t(q) * nsegs + seg]) * s.w + c] : T(0);
    } else {
        const uint32_t r = k - total_w, l = r / dig, d = r % dig;
        const uint64_t si = sib[uint64_t(q) * steps + l];
        proof[(uint64_t(q) * steps + l) * dig + d] = si == ~0ull ? T(0) : layers[l][si * dig + d];
    }
}


// ---------------------------------------------------------------- FP64 ----
// The same sponge / compression with the BabyBear permutation evaluated in
// binary64 (poseidon2_f64.h).  Inputs and outputs are canonical uint32.
// Reads the next n <= 8 words of row j of the concatenated segments into v,
// moving the cursor (segment, offset) forward; a segment without data reads
// as zeros.
struct SegCursor {
    uint32_t seg, off;
    RowSeg<uint32_t> cur;
    const uint32_t* rowp;
};
__device__ __forceinline__ void seg_read8(const RowSeg<uint32_t>* segs, uint64_t j, SegCursor& c, uint32_t n,
                                          uint32_t* v) {
#pragma unroll
    for (uint32_t q = 0; q < 8; ++q) {
        if (q < n) {
            while (c.off >= c.cur.w) {
                c.off = 0;
                ++c.seg;
                c.cur = segs[c.seg];
                c.rowp = c.cur.ptr + stored_row(c.cur, j) * c.cur.w;
            }
            v[q] = c.cur.ptr ? c.rowp[c.off] : 0u;
            ++c.off;
        }
    }
}

// The overwrite sponge over row j of the concatenated segments.  The next
// chunk's words are read before the current chunk's permutation, and the
// rate lanes are left unreduced by a permutation whose successor overwrites
// all of them; on return every lane is reduced (callers use the state as a
// compression input).
template <bool RELOAD>
__device__ __forceinline__ void sponge_row_f64(double* s, const RowSeg<uint32_t>* segs, uint32_t total_w, uint64_t j,
                                               const P2ParamsF64& p) {
#pragma unroll
    for (int i = 0; i < 16; ++i) s[i] = 0.0;
    SegCursor c{0, 0, segs[0], nullptr};
    c.rowp = c.cur.ptr + stored_row(c.cur, j) * c.cur.w;
    uint32_t nx[8];
    seg_read8(segs, j, c, total_w < 8 ? total_w : 8u, nx);
    for (uint32_t start = 0; start < total_w; start += 8) {
        uint32_t v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = nx[q];
        const uint32_t n = total_w - start;  // words left, this chunk takes min(n, 8)
        if (n > 8) seg_read8(segs, j, c, n - 8 < 8 ? n - 8 : 8u, nx);
#pragma unroll
        for (uint32_t q = 0; q < 8; ++q)
            if (q < n) s[q] = double(v[q]);
        Poseidon2F64::permute<false, RELOAD>(s, p);
        if (n < 16) {  // the next chunk is partial or there is none: its untouched rate lanes carry over
#pragma unroll
            for (int q = 0; q < 8; ++q) s[q] = BBF64::reduce(s[q]);
        }
    }
}

__global__ __launch_bounds__(HASH_THREADS) void hash_rows_f64_kernel(const RowSeg<uint32_t>* segs, uint32_t total_w,
                                                                     uint64_t rows, uint32_t* out, P2ParamsF64 p) {
    uint64_t j = uint64_t(blockIdx.x) * HASH_THREADS + threadIdx.x;
    if (j >= rows) return;
    double s[16];
    sponge_row_f64<true>(s, segs, total_w, j, p);
#pragma unroll
    for (int i = 0; i < 8; ++i) out[j * 8 + i] = BBF64::canonical(s[i]);
}

__global__ __launch_bounds__(HASH_THREADS) void hash_rows_dense_f64_kernel(RowSeg<uint32_t> seg, uint64_t rows,
                                                                           uint32_t* out, P2ParamsF64 p) {
    uint64_t j = uint64_t(blockIdx.x) * HASH_THREADS + threadIdx.x;
    if (j >= rows) return;
    double s[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) s[i] = 0.0;
    const uint32_t* rowp = seg.ptr + stored_row(seg, j) * seg.w;
    // the next chunk's loads are issued before the current chunk's permutation
    const uint4* v = reinterpret_cast<const uint4*>(rowp);
    uint4 na = v[0], nb = v[1];
    for (uint32_t start = 0; start < seg.w; start += 8) {
        const uint4 a = na, b = nb;
        if (start + 8 < seg.w) {
            na = v[(start + 8) / 4];
            nb = v[(start + 8) / 4 + 1];
        }
        s[0] = double(a.x);
        s[1] = double(a.y);
        s[2] = double(a.z);
        s[3] = double(a.w);
        s[4] = double(b.x);
        s[5] = double(b.y);
        s[6] = double(b.z);
        s[7] = double(b.w);
        // lanes 0..7 are overwritten by the next chunk or canonicalised below
        Poseidon2F64::permute<false, true>(s, p);
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) out[j * 8 + i] = BBF64::canonical(s[i]);
}

// One matrix of any width: the dense kernel's schedule (the next chunk's loads
// issued before the current chunk's permutation) with 4-byte loads, since rows
// of a width that is not a multiple of 4 are not 16-byte aligned.  A last
// partial chunk overwrites only its own rate slots (PaddingFreeSponge).
__global__ __launch_bounds__(HASH_THREADS) void hash_rows_ragged_f64_kernel(RowSeg<uint32_t> seg, uint64_t rows,
                                                                            uint32_t* out, P2ParamsF64 p) {
    uint64_t j = uint64_t(blockIdx.x) * HASH_THREADS + threadIdx.x;
    if (j >= rows) return;
    double s[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) s[i] = 0.0;
    const uint32_t w = seg.w;
    const uint32_t* rowp = seg.ptr + stored_row(seg, j) * w;
    uint32_t nx[8];
#pragma unroll
    for (uint32_t q = 0; q < 8; ++q) nx[q] = q < w ? rowp[q] : 0u;
    for (uint32_t start = 0; start < w; start += 8) {
        uint32_t cur[8];
#pragma unroll
        for (uint32_t q = 0; q < 8; ++q) cur[q] = nx[q];
        const uint32_t n = w - start;  // words left, this chunk takes min(n, 8)
        if (n > 8) {
#pragma unroll
            for (uint32_t q = 0; q < 8; ++q) nx[q] = start + 8 + q < w ? rowp[start + 8 + q] : 0u;
        }
#pragma unroll
        for (uint32_t q = 0; q < 8; ++q)
            if (q < n) s[q] = double(cur[q]);
        Poseidon2F64::permute<true, true>(s, p);
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) out[j * 8 + i] = BBF64::canonical(s[i]);
}

__global__ __launch_bounds__(HASH_THREADS) void compress_f64_kernel(const uint32_t* prev, uint64_t prev_count,
                                                                    uint32_t* next, uint64_t next_count,
                                                                    const RowSeg<uint32_t>* inj, uint32_t inj_w,
                                                                    P2ParamsF64 p) {
    uint64_t i = uint64_t(blockIdx.x) * HASH_THREADS + threadIdx.x;
    if (i >= next_count) return;
    double s[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) s[q] = 0.0;
    const uint32_t* l = prev + 2 * i * 8;
#pragma unroll
    for (int q = 0; q < 8; ++q) s[q] = double(l[q]);
    if (2 * i + 1 < prev_count) {
#pragma unroll
        for (int q = 0; q < 8; ++q) s[8 + q] = double(l[8 + q]);
    }
    Poseidon2F64::permute(s, p);
    if (inj) {
        double h[16];
        sponge_row_f64<false>(h, inj, inj_w, i, p);
#pragma unroll
        for (int q = 0; q < 8; ++q) s[8 + q] = h[q];
        Poseidon2F64::permute(s, p);
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) next[i * 8 + q] = BBF64::canonical(s[q]);
}

__global__ __launch_bounds__(HASH_THREADS) void permute_f64_kernel(uint32_t* st, uint64_t n, P2ParamsF64 p) {
    uint64_t i = uint64_t(blockIdx.x) * HASH_THREADS + threadIdx.x;
    if (i >= n) return;
    double s[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) s[q] = double(st[i * 16 + q]);
    Poseidon2F64::permute(s, p);
#pragma unroll
    for (int q = 0; q < 16; ++q) st[i * 16 + q] = BBF64::canonical(s[q]);
}

// Lane-parallel kernels: 16 lanes per state (16 states per 256-thread
// workgroup).  Used where a launch has few states, so that the latency of one
// permutation, not the throughput, sets the time.
constexpr int LANE_THREADS = 256;
constexpr uint64_t LANE_MAX = 1ull << 15;  // at most this many states per launch go lane-parallel

__global__ __launch_bounds__(LANE_THREADS) void compress_lanes_f64_kernel(const uint32_t* prev, uint64_t prev_count,
                                                                          uint32_t* next, uint64_t next_count,
                                                                          P2ParamsF64 p) {
    const uint32_t lane = threadIdx.x & 15;
    const auto pc = Poseidon2F64Lanes::consts(p, lane);
    const uint64_t node = (uint64_t(blockIdx.x) * LANE_THREADS + threadIdx.x) >> 4;
    double s = 0.0;
    if (node < next_count) {
        const uint64_t child = 2 * node + (lane >> 3);  // lanes 0-7 left digest, 8-15 right digest
        if (child < prev_count) s = double(prev[child * 8 + (lane & 7)]);
    }
    s = Poseidon2F64Lanes::permute(s, pc, lane);
    if (node < next_count && lane < 8) next[node * 8 + lane] = BBF64::canonical(s);
}

__global__ __launch_bounds__(LANE_THREADS) void hash_rows_lanes_f64_kernel(RowSeg<uint32_t> seg, uint64_t rows,
                                                                           uint32_t* out, P2ParamsF64 p) {
    const uint32_t lane = threadIdx.x & 15;
    const auto pc = Poseidon2F64Lanes::consts(p, lane);
    const uint64_t j = (uint64_t(blockIdx.x) * LANE_THREADS + threadIdx.x) >> 4;
    const bool live = j < rows;
    const uint32_t* rowp = live ? seg.ptr + stored_row(seg, j) * seg.w : nullptr;
    double s = 0.0;
    for (uint32_t start = 0; start < seg.w; start += 8) {
        if (live && lane < 8 && start + lane < seg.w) s = double(rowp[start + lane]);  // overwrite mode
        s = Poseidon2F64Lanes::permute(s, pc, lane);
    }
    if (live && lane < 8) out[j * 8 + lane] = BBF64::canonical(s);
}

// The last levels of a tree in one workgroup of 64 lane-parallel states:
// level l (l < nlev) reads level l-1 (level -1 is `in`, in_count digests),
// writes count[l] = ceil(prev / 2) digests to out[l] and keeps them in LDS for
// level l+1.  A missing right child is the zero digest.  count[0] <= 64.
constexpr uint32_t TAIL_THREADS = 1024;
constexpr uint32_t TAIL_MAX_LEVELS = 8;
struct TailLevels {
    const uint32_t* in;
    uint64_t in_count;
    uint32_t* out[TAIL_MAX_LEVELS];
    uint32_t count[TAIL_MAX_LEVELS];
    uint32_t nlev;
};

__global__ __launch_bounds__(TAIL_THREADS) void compress_tail_lanes_f64_kernel(TailLevels t, P2ParamsF64 p) {
    __shared__ uint32_t buf[2][64 * 8];
    const uint32_t lane = threadIdx.x & 15;
    const auto pc = Poseidon2F64Lanes::consts(p, lane);
    const uint32_t node = threadIdx.x >> 4;  // < 64
    uint32_t prev_count = uint32_t(t.in_count);
    for (uint32_t l = 0; l < t.nlev; ++l) {
        const uint32_t cnt = t.count[l];
        double s = 0.0;
        if (node < cnt) {
            const uint32_t child = 2 * node + (lane >> 3);
            if (child < prev_count)
                s = double(l == 0 ? t.in[uint64_t(child) * 8 + (lane & 7)] : buf[(l - 1) & 1][child * 8 + (lane & 7)]);
        }
        // a wave holds nodes 4w..4w+3; waves past the level's last node skip
        // the permutation, so the live waves do not share their SIMDs' issue
        // slots with idle ones (DPP stays within a 16-lane row, one node)
        if (((threadIdx.x >> 6) << 2) < cnt) s = Poseidon2F64Lanes::permute(s, pc, lane);
        if (node < cnt && lane < 8) {
            const uint32_t v = BBF64::canonical(s);
            buf[l & 1][node * 8 + lane] = v;
            t.out[l][node * 8 + lane] = v;
        }
        __syncthreads();
        prev_count = cnt;
    }
}

__global__ __launch_bounds__(LANE_THREADS) void permute_lanes_f64_kernel(uint32_t* st, uint64_t n, P2ParamsF64 p) {
    const uint32_t lane = threadIdx.x & 15;
    const auto pc = Poseidon2F64Lanes::consts(p, lane);
    const uint64_t i = (uint64_t(blockIdx.x) * LANE_THREADS + threadIdx.x) >> 4;
    double s = i < n ? double(st[i * 16 + lane]) : 0.0;
    s = Poseidon2F64Lanes::permute(s, pc, lane);
    if (i < n) st[i * 16 + lane] = BBF64::canonical(s);
}

// Job table of a multi-job launch, passed by value.  Job t owns workgroups
// [block0[t], block0[t+1]), so the job index is uniform in a workgroup.
struct MultiTable {
    const uint32_t* in[MULTI_MAX];
    uint32_t* out[MULTI_MAX];
    uint64_t in_count[MULTI_MAX];
    uint64_t out_count[MULTI_MAX];
    uint32_t w[MULTI_MAX];
    uint32_t block0[MULTI_MAX + 1];
    uint32_t n;
};

// Kernel arguments are passed by value in the argument block: keep every
// kernel's total within 4 KiB.
static_assert(sizeof(MultiTable) + sizeof(P2ParamsF64) <= 4096, "multi-job kernel arguments past 4 KiB");
static_assert(sizeof(TailLevels) + sizeof(P2ParamsF64) <= 4096, "tail kernel arguments past 4 KiB");

__device__ __forceinline__ uint32_t job_of(const MultiTable& m, uint32_t b) {
    uint32_t t = 0;
    for (uint32_t k = 1; k < m.n; ++k) t += b >= m.block0[k] ? 1u : 0u;
    return t;
}

__global__ __launch_bounds__(HASH_THREADS) void hash_multi_f64_kernel(MultiTable m, P2ParamsF64 p) {
    const uint32_t t = job_of(m, blockIdx.x);
    const uint64_t j = uint64_t(blockIdx.x - m.block0[t]) * HASH_THREADS + threadIdx.x;
    if (j >= m.out_count[t]) return;
    const uint32_t w = m.w[t];
    const uint4* v = reinterpret_cast<const uint4*>(m.in[t] + j * w);
    double s[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) s[i] = 0.0;
    for (uint32_t start = 0; start < w; start += 8) {
        const uint4 a = v[start / 4], b = v[start / 4 + 1];
        s[0] = double(a.x);
        s[1] = double(a.y);
        s[2] = double(a.z);
        s[3] = double(a.w);
        s[4] = double(b.x);
        s[5] = double(b.y);
        s[6] = double(b.z);
        s[7] = double(b.w);
        // lanes 0..7 are overwritten by the next chunk or canonicalised below
        Poseidon2F64::permute<false, true>(s, p);
    }
    uint32_t* o = m.out[t] + j * 8;
#pragma unroll
    for (int q = 0; q < 8; ++q) o[q] = BBF64::canonical(s[q]);
}

__global__ __launch_bounds__(LANE_THREADS) void hash_multi_lanes_f64_kernel(MultiTable m, P2ParamsF64 p) {
    const uint32_t t = job_of(m, blockIdx.x);
    const uint32_t lane = threadIdx.x & 15;
    const auto pc = Poseidon2F64Lanes::consts(p, lane);
    const uint64_t j = uint64_t(blockIdx.x - m.block0[t]) * (LANE_THREADS / 16) + (threadIdx.x >> 4);
    const bool live = j < m.out_count[t];
    const uint32_t w = m.w[t];
    const uint32_t* rowp = m.in[t] + (live ? j : 0) * w;
    double s = 0.0;
    for (uint32_t start = 0; start < w; start += 8) {  // w is uniform: every lane runs every permutation
        if (live && lane < 8) s = double(rowp[start + lane]);
        s = Poseidon2F64Lanes::permute(s, pc, lane);
    }
    if (live && lane < 8) m.out[t][j * 8 + lane] = BBF64::canonical(s);
}

__global__ __launch_bounds__(HASH_THREADS) void compress_multi_f64_kernel(MultiTable m, P2ParamsF64 p) {
    const uint32_t t = job_of(m, blockIdx.x);
    const uint64_t i = uint64_t(blockIdx.x - m.block0[t]) * HASH_THREADS + threadIdx.x;
    if (i >= m.out_count[t]) return;
    const uint32_t* l = m.in[t] + 2 * i * 8;
    double s[16];
#pragma unroll
    for (int q = 0; q < 8; ++q) s[q] = double(l[q]);
#pragma unroll
    for (int q = 8; q < 16; ++q) s[q] = 0.0;
    if (2 * i + 1 < m.in_count[t]) {
#pragma unroll
        for (int q = 0; q < 8; ++q) s[8 + q] = double(l[8 + q]);
    }
    Poseidon2F64::permute(s, p);
    uint32_t* o = m.out[t] + i * 8;
#pragma unroll
    for (int q = 0; q < 8; ++q) o[q] = BBF64::canonical(s[q]);
}

__global__ __launch_bounds__(LANE_THREADS) void compress_multi_lanes_f64_kernel(MultiTable m, P2ParamsF64 p) {
    const uint32_t t = job_of(m, blockIdx.x);
    const uint32_t lane = threadIdx.x & 15;
    const auto pc = Poseidon2F64Lanes::consts(p, lane);
    const uint64_t node = uint64_t(blockIdx.x - m.block0[t]) * (LANE_THREADS / 16) + (threadIdx.x >> 4);
    const bool live = node < m.out_count[t];
    double s = 0.0;
    if (live) {
        const uint64_t child = 2 * node + (lane >> 3);  // lanes 0-7 left digest, 8-15 right digest
        if (child < m.in_count[t]) s = double(m.in[t][child * 8 + (lane & 7)]);
    }
    s = Poseidon2F64Lanes::permute(s, pc, lane);
    if (live && lane < 8) m.out[t][node * 8 + lane] = BBF64::canonical(s);
}

inline uint32_t grid_for(uint64_t n, uint32_t threads) { return uint32_t((n + threads - 1) / threads); }

// Table for jobs[0..n), n <= MULTI_MAX, with `per_block` states per
// workgroup; returns the number of workgroups, or ~0 when that exceeds 2^31 - 1.
uint64_t fill_table(MultiTable& m, const MultiJob* jobs, uint32_t n, uint32_t per_block) {
    uint64_t b = 0;
    m.n = n;
    for (uint32_t k = 0; k < n; ++k) {
        m.in[k] = jobs[k].in;
        m.out[k] = jobs[k].out;
        m.in_count[k] = jobs[k].in_count;
        m.out_count[k] = jobs[k].out_count;
        m.w[k] = jobs[k].w;
        m.block0[k] = uint32_t(b);
        b += (jobs[k].out_count + per_block - 1) / per_block;
        if (b > 0x7FFFFFFFull) return ~0ull;
    }
    m.block0[n] = uint32_t(b);
    return b;
}

}  // namespace

template <class D>
int hash_rows_device(const RowSeg<typename D::T>* segs, uint32_t nsegs, uint32_t total_w,
                     const RowSeg<typename D::T>& seg0, uint64_t rows, typename D::T* out,
                     const P2Params<typename D::T, D::W>& p, hipStream_t st) {
    if (rows == 0) return 0;
    bool dense = nsegs == 1 && seg0.ptr && seg0.w > 0 && seg0.w % D::RATE == 0 &&
                 (reinterpret_cast<uintptr_t>(seg0.ptr) & 15) == 0 && (seg0.w * sizeof(typename D::T)) % 16 == 0;
    if (dense) {
        hipLaunchKernelGGL((hash_rows_dense_kernel<D>), dim3(grid_for(rows, HASH_THREADS)), dim3(HASH_THREADS), 0,
                           st, seg0, rows, out, p);
    } else {
        hipLaunchKernelGGL((hash_rows_kernel<D>), dim3(grid_for(rows, HASH_THREADS)), dim3(HASH_THREADS), 0, st,
                           segs, nsegs, total_w, rows, out, p);
    }
    return int(hipGetLastError());
}

template <class D>
int compress_layer_device(const typename D::T* prev, uint64_t prev_count, typename D::T* next, uint64_t next_count,
                          const RowSeg<typename D::T>* inj, uint32_t ninj, uint32_t inj_w,
                          const P2Params<typename D::T, D::W>& p, hipStream_t st) {
    if (next_count == 0) return 0;
    hipLaunchKernelGGL((compress_kernel<D>), dim3(grid_for(next_count, HASH_THREADS)), dim3(HASH_THREADS), 0, st,
                       prev, prev_count, next, next_count, inj, ninj, inj_w, p);
    return int(hipGetLastError());
}

template <class D>
int permute_device(typename D::T* states, uint64_t n, const P2Params<typename D::T, D::W>& p, hipStream_t st) {
    if (n == 0) return 0;
    hipLaunchKernelGGL((permute_kernel<D>), dim3(grid_for(n, HASH_THREADS)), dim3(HASH_THREADS), 0, st, states, n,
                       p);
    return int(hipGetLastError());
}

template <class T>
int gather_rows_device(const RowSeg<T>* segs, uint32_t nsegs, uint32_t total_w, const uint64_t* rows, uint32_t nq,
                       T* out, hipStream_t st) {
    uint64_t total = uint64_t(nq) * total_w;
    if (total == 0) return 0;
    hipLaunchKernelGGL((gather_rows_kernel<T>), dim3(grid_for(total, 256)), dim3(256), 0, st, segs, nsegs, total_w,
                       rows, nq, out);
    return int(hipGetLastError());
}

template <class T>
int open_rows_device(const RowSeg<T>* segs, uint32_t nsegs, uint32_t total_w, const uint64_t* rows,
                     const T* const* layers, uint32_t steps, uint32_t dig, const uint64_t* sib, uint32_t nq, T* vals,
                     T* proof, hipStream_t st) {
    const uint64_t total = uint64_t(nq) * (total_w + steps * dig);
    if (total == 0) return 0;
    hipLaunchKernelGGL((open_rows_kernel<T>), dim3(grid_for(total, 256)), dim3(256), 0, st, segs, nsegs, total_w,
                       rows, layers, steps, dig, sib, nq, vals, proof);
    return int(hipGetLastError());
}


static bool f64_lanes_form(uint32_t nsegs, const RowSeg<uint32_t>& seg0, uint64_t rows) {
    return nsegs == 1 && seg0.ptr && seg0.w > 0 && rows <= LANE_MAX;
}
static bool f64_dense_form(uint32_t nsegs, const RowSeg<uint32_t>& seg0) {
    return nsegs == 1 && seg0.ptr && seg0.w > 0 && seg0.w % 8 == 0 && (reinterpret_cast<uintptr_t>(seg0.ptr) & 15) == 0;
}

static bool f64_ragged_form(uint32_t nsegs, const RowSeg<uint32_t>& seg0) {
    return nsegs == 1 && seg0.ptr && seg0.w > 0;
}

bool hash_rows_f64_reads_table(uint32_t nsegs, const RowSeg<uint32_t>& seg0, uint64_t rows) {
    return !f64_lanes_form(nsegs, seg0, rows) && !f64_dense_form(nsegs, seg0) && !f64_ragged_form(nsegs, seg0);
}

int hash_rows_f64_device(const RowSeg<uint32_t>* segs, uint32_t nsegs, uint32_t total_w, const RowSeg<uint32_t>& seg0,
                         uint64_t rows, uint32_t* out, const P2ParamsF64& p, hipStream_t st) {
    if (rows == 0) return 0;
    if (f64_lanes_form(nsegs, seg0, rows)) {
        hipLaunchKernelGGL(hash_rows_lanes_f64_kernel, dim3(grid_for(rows * 16, LANE_THREADS)), dim3(LANE_THREADS), 0,
                           st, seg0, rows, out, p);
        return int(hipGetLastError());
    }
    if (f64_dense_form(nsegs, seg0))
        hipLaunchKernelGGL(hash_rows_dense_f64_kernel, dim3(grid_for(rows, HASH_THREADS)), dim3(HASH_THREADS), 0, st,
                           seg0, rows, out, p);
    else if (f64_ragged_form(nsegs, seg0))
        hipLaunchKernelGGL(hash_rows_ragged_f64_kernel, dim3(grid_for(rows, HASH_THREADS)), dim3(HASH_THREADS), 0, st,
                           seg0, rows, out, p);
    else if (segs)
        hipLaunchKernelGGL(hash_rows_f64_kernel, dim3(grid_for(rows, HASH_THREADS)), dim3(HASH_THREADS), 0, st, segs,
                           total_w, rows, out, p);
    else
        return int(hipErrorInvalidValue);
    return int(hipGetLastError());
}

int compress_layer_f64_device(const uint32_t* prev, uint64_t prev_count, uint32_t* next, uint64_t next_count,
                              const RowSeg<uint32_t>* inj, uint32_t ninj, uint32_t inj_w, const P2ParamsF64& p,
                              hipStream_t st) {
    if (next_count == 0) return 0;
    if (ninj == 0 && next_count <= LANE_MAX) {
        hipLaunchKernelGGL(compress_lanes_f64_kernel, dim3(grid_for(next_count * 16, LANE_THREADS)), dim3(LANE_THREADS),
                           0, st, prev, prev_count, next, next_count, p);
        return int(hipGetLastError());
    }
    hipLaunchKernelGGL(compress_f64_kernel, dim3(grid_for(next_count, HASH_THREADS)), dim3(HASH_THREADS), 0, st, prev,
                       prev_count, next, next_count, ninj ? inj : nullptr, inj_w, p);
    return int(hipGetLastError());
}

int compress_tail_f64_device(const uint32_t* in, uint64_t in_count, uint32_t* const* outs, const uint64_t* counts,
                             uint32_t nlev, const P2ParamsF64& p, hipStream_t st) {
    if (nlev == 0) return 0;
    if (nlev > TAIL_MAX_LEVELS || counts[0] > 64 || in_count > 128) return int(hipErrorInvalidValue);
    TailLevels t{};
    t.in = in;
    t.in_count = in_count;
    t.nlev = nlev;
    uint64_t prev = in_count;
    for (uint32_t l = 0; l < nlev; ++l) {
        if (counts[l] != (prev + 1) / 2) return int(hipErrorInvalidValue);
        t.out[l] = outs[l];
        t.count[l] = uint32_t(counts[l]);
        prev = counts[l];
    }
    hipLaunchKernelGGL(compress_tail_lanes_f64_kernel, dim3(1), dim3(TAIL_THREADS), 0, st, t, p);
    return int(hipGetLastError());
}

int permute_f64_device(uint32_t* states, uint64_t n, const P2ParamsF64& p, hipStream_t st) {
    if (n == 0) return 0;
    if (n <= LANE_MAX) {
        hipLaunchKernelGGL(permute_lanes_f64_kernel, dim3(grid_for(n * 16, LANE_THREADS)), dim3(LANE_THREADS), 0, st,
                           states, n, p);
        return int(hipGetLastError());
    }
    hipLaunchKernelGGL(permute_f64_kernel, dim3(grid_for(n, HASH_THREADS)), dim3(HASH_THREADS), 0, st, states, n, p);
    return int(hipGetLastError());
}

// Launches a multi-job kernel over jobs in groups of MULTI_MAX.  With at most
// LANE_MAX states in all, the lane-parallel kernel runs (16 states per
// workgroup), otherwise the thread-per-state kernel.
template <class Dense, class Lanes>
static int launch_multi(const MultiJob* jobs, uint32_t n, Dense dense, Lanes lanes, const P2ParamsF64& p,
                        hipStream_t st) {
    uint64_t total = 0;
    for (uint32_t k = 0; k < n; ++k) total += jobs[k].out_count;
    if (total == 0) return 0;
    const bool lane_form = total <= LANE_MAX;
    for (uint32_t a = 0; a < n; a += MULTI_MAX) {
        const uint32_t c = n - a < MULTI_MAX ? n - a : MULTI_MAX;
        MultiTable m{};
        const uint64_t g = fill_table(m, jobs + a, c, lane_form ? LANE_THREADS / 16 : HASH_THREADS);
        if (g == ~0ull) return int(hipErrorInvalidValue);
        if (g == 0) continue;
        if (lane_form)
            hipLaunchKernelGGL(lanes, dim3(uint32_t(g)), dim3(LANE_THREADS), 0, st, m, p);
        else
            hipLaunchKernelGGL(dense, dim3(uint32_t(g)), dim3(HASH_THREADS), 0, st, m, p);
        const int e = int(hipGetLastError());
        if (e) return e;
    }
    return 0;
}

int hash_rows_multi_f64_device(const MultiJob* jobs, uint32_t n, const P2ParamsF64& p, hipStream_t st) {
    for (uint32_t k = 0; k < n; ++k)
        if (jobs[k].out_count && (!jobs[k].in || !jobs[k].out || jobs[k].w == 0 || jobs[k].w % 8 ||
                                  (reinterpret_cast<uintptr_t>(jobs[k].in) & 15)))
            return int(hipErrorInvalidValue);
    return launch_multi(jobs, n, hash_multi_f64_kernel, hash_multi_lanes_f64_kernel, p, st);
}

int compress_multi_f64_device(const MultiJob* jobs, uint32_t n, const P2ParamsF64& p, hipStream_t st) {
    for (uint32_t k = 0; k < n; ++k)
        if (jobs[k].out_count && (!jobs[k].in || !jobs[k].out || jobs[k].out_count != (jobs[k].in_count + 1) / 2))
            return int(hipErrorInvalidValue);
    return launch_multi(jobs, n, compress_multi_f64_kernel, compress_multi_lanes_f64_kernel, p, st);
}

#define P3H_INST(D)                                                                                            \
    template int hash_rows_device<D>(const RowSeg<D::T>*, uint32_t, uint32_t, const RowSeg<D::T>&, uint64_t,    \
                                     D::T*, const P2Params<D::T, D::W>&, hipStream_t);                          \
    template int compress_layer_device<D>(const D::T*, uint64_t, D::T*, uint64_t, const RowSeg<D::T>*, uint32_t, \
                                          uint32_t, const P2Params<D::T, D::W>&, hipStream_t);                  \
    template int permute_device<D>(D::T*, uint64_t, const P2Params<D::T, D::W>&, hipStream_t);                  \
    template int gather_rows_device<D::T>(const RowSeg<D::T>*, uint32_t, uint32_t, const uint64_t*, uint32_t,   \
                                          D::T*, hipStream_t);                                                     \
    template int open_rows_device<D::T>(const RowSeg<D::T>*, uint32_t, uint32_t, const uint64_t*,               \
                                        const D::T* const*, uint32_t, uint32_t, const uint64_t*, uint32_t, D::T*, \
                                        D::T*, hipStream_t);
P3H_INST(BBDom)
P3H_INST(GLDom)
#undef P3H_INST

}  // namespace p3h
