// Column-batched NTT / coset LDE kernels for gfx950: the kernels and the
// host-side planners and launchers, included by one translation unit per field
// (ntt_bb.hip, ntt_gl.hip), so each field's kernels compile on their own.
//
// The matrix is row-major (height x width, the reference's RowMajorMatrix),
// so every column is an independent transform whose elements sit `width`
// apart.  A pass covers a contiguous range of index bits [a, a+S): each
// workgroup stages a tile of 2^S positions x C columns in LDS, runs the S
// radix-2 stages there and writes back.  Global loads and stores walk the tile
// column-fastest, so each (row, column-tile) segment is one contiguous burst.
//
// Inside a pass the stages run in at most two register rounds of radix 2^R
// (R <= 5): a thread takes a group of 2^R positions of one column, does R
// butterfly stages in VGPRs and writes the group back.  S is a template
// parameter, so every LDS offset and twiddle index inside a round is a
// compile-time immediate; the LDS tile is stored column-major with one word of
// padding per column, which keeps both the global staging and the round
// accesses free of bank conflicts.
//
// Coset LDE (reference: Radix2Dit::coset_lde_batch, radix2_dit.hpp:221-233)
// is computed as
//   1. inverse transform by decimation in frequency (natural order in,
//      bit-reversed coefficients out): outer passes over the high bits, then
//   2. one fused inner pass per block of 2^S rows: the last DIF stages, then for
//      each coset r of the 2^k cosets of the small domain inside the large one
//      the twist a_j * (shift * omega_N^r)^j / n and the first decimation-in-
//      time stages of a forward transform (bit-reversed in, natural out),
//   3. outer DIT passes over the high bits, per coset.
// Coset r's natural index i is LDE row i * 2^k + r.  The device buffer keeps
// the cosets one after another ("coset-major": stored row r * n + i); readers
// (leaf hashing, openings, downloads) map LDE row j to stored row
// (j mod 2^k) * n + j / 2^k.  Zero padding is never materialised.

#pragma once

#include <utility>

#include "field.h"
#include "ntt.h"

namespace p3h {

namespace {

constexpr int NTT_THREADS = 512;
constexpr int TILE_BYTES = 131072;  // largest exchange tile (the LDS is sized per pass)
constexpr uint32_t MAX_STAGES = 10;

template <class T>
constexpr uint32_t tile_log() {
    return 31 - __builtin_clz(uint32_t(TILE_BYTES / sizeof(T)));
}
enum Mode : int {
    OUTER_DIF_INV = 0,     // DIF stages, bits descending
    OUTER_DIT_FWD = 1,     // DIT stages, bits ascending, per region
    INNER_LDE = 2,         // DIF stages + per-coset twist + DIT stages
    INNER_DIT_GATHER = 3,  // load rows bitrev(p) (times scale[row]), DIT stages
    INNER_DIF_SCATTER = 4,  // DIF stages, store position p at row bitrev(p) times scale[row]
    OUTER_DIF_INV_SPLIT = 5,  // OUTER_DIF_INV on a full tile, exchanged half by half
    OUTER_DIT_FWD_SPLIT = 6   // OUTER_DIT_FWD on a full tile, exchanged half by half
};

__device__ __forceinline__ uint32_t bitrev(uint32_t x, uint32_t bits) {
    return bits == 0 ? 0u : (__brev(x) >> (32 - bits));
}

// Twiddles of one pass.  Stage q (bit a+q) at tile position t, low index lo0
// (the whole tile shares one low index), uses
//   omega_{2^(a+q+1)}^(((t mod 2^q) << a) + lo0)
//     = omega_{2^(q+1)}^(t mod 2^q) * omega_{2^(a+q+1)}^lo0,
// stored at tws[(2^q - 1) + (t mod 2^q)].  The first factor comes from the
// contiguous stage table twc (the same 4 or 8 KiB for every pass and size, so
// it stays in cache), the second, one value per stage, from the size-n table
// tw (omega_n^e, e < n/2); inner passes (a = 0) have lo0 = 0 and copy twc.
template <class F>
__device__ __forceinline__ void build_twiddles(typename F::T* tws, const typename F::T* twc,
                                               const typename F::T* tw, uint32_t L, uint32_t a, uint32_t S,
                                               uint32_t lo0) {
    const uint32_t count = (1u << S) - 1;
    for (uint32_t i = threadIdx.x; i < count; i += NTT_THREADS) {
        typename F::T v = twc[i];
        if (lo0) {
            const uint32_t q = 31 - __clz(i + 1);
            v = F::mul_c(v, tw[uint64_t(lo0) << (L - a - q - 1)]);
        }
        tws[i] = v;
    }
}

// Rounds of an S-stage pass: low = index bits [0, R0), high = [R0, S).
__host__ __device__ constexpr uint32_t low_round_bits(uint32_t s) { return s <= 5 ? s : (s + 1) / 2; }

// Words of one direction's twiddle table in a pass's LDS.  Fields without
// power-of-two roots keep the stage table (2^S - 1 words); fields with them
// keep the twisted round table of build_twisted (2^R0 words for the low round,
// 2^S for the high round).
__host__ __device__ constexpr uint32_t tw_table_words(bool twisted, uint32_t s) {
    return s == 0 ? 0u
                  : !twisted ? (1u << s) - 1
                             : (low_round_bits(s) < s ? (1u << s) + (1u << low_round_bits(s)) : (1u << s));
}

// Twisted rounds (fields whose 64th roots of unity are powers of two).  In a
// radix-2^R round at position base u = tb mod 2^Q0, the stage-j twiddle of the
// pair (m, m + 2^j) factors as beta_j * zeta_j^(m mod 2^j) with
// zeta_j = omega_{2^(j+1)} and beta_j = gamma^(2^(R-1-j)),
// gamma = omega_{2^(a+Q0+R)}^(u 2^a + lo0).  The round is then
//   DIT: x[m] *= gamma^bitrev_R(m), then the plain radix-2^R DIT on the zetas;
//   DIF: the plain radix-2^R DIF on the zetas, then x[m] *= gamma^bitrev_R(m),
// (induction on R: the top stage's twiddles beta_j zeta_j^k are exactly what
// the scaled inputs contribute), so a round costs 2^R - 1 general products and
// R 2^(R-1) power-of-two products instead of R 2^(R-1) general products.
// Table: the low round (Q0 = 0, u = 0) at [0, 2^R0), the high round
// (Q0 = R0) at 2^R0 + m 2^R0 + u, so that lanes of neighbouring u read
// neighbouring words.  gamma^k = omega_{2^(Q0+R)}^(k u) times
// omega_{2^(a+Q0+R)}^(k lo0): the first from the stage table twc, the second
// (outer passes only) from the size-n table tw; an exponent in the upper half
// of a table's range is the negated entry of the lower half.
template <class F, int S>
__device__ __forceinline__ void build_twisted(typename F::T* tt, const typename F::T* twc,
                                              const typename F::T* tw, uint32_t L, uint32_t a, uint32_t lo0) {
    using T = typename F::T;
    constexpr uint32_t R0 = low_round_bits(S), R1 = S - R0;
    constexpr uint32_t count = tw_table_words(true, S);
    const uint64_t half = 1ull << (L - 1);  // a pass with S >= 1 has L >= 1
    for (uint32_t i = threadIdx.x; i < count; i += NTT_THREADS) {
        uint32_t q0, r, u, m;
        if (i < (1u << R0)) {
            q0 = 0;
            r = R0;
            u = 0;
            m = i;
        } else {
            const uint32_t j = i - (1u << R0);
            q0 = R0;
            r = R1;
            m = j >> R0;
            u = j & ((1u << R0) - 1);
        }
        const uint32_t k = bitrev(m, r);
        const uint32_t q = q0 + r - 1;  // stage table of the roots of order 2^(q0 + r)
        const uint32_t x = (k * u) & ((2u << q) - 1);
        T v = x < (1u << q) ? twc[(1u << q) - 1 + x] : F::neg_c(twc[(1u << q) - 1 + x - (1u << q)]);
        if (lo0) {
            const uint64_t e = (uint64_t(k) * lo0) << (L - a - q0 - r);
            v = F::mul_c(v, e < half ? tw[e] : F::neg_c(tw[e - half]));
        }
        tt[i] = v;
    }
}

// The twiddles a round's butterflies read, for the group with block index b:
// the stage table at b mod 2^Q0, or the twisted table at that group's u
// (entry m at twb[m << Q0]).
template <class F, int Q0>
__device__ __forceinline__ const typename F::T* round_tw(const typename F::T* tws, uint32_t b) {
    const uint32_t u = b & ((1u << Q0) - 1);
    if constexpr (F::POW2_ROOTS) return tws + (Q0 == 0 ? 0u : (1u << Q0)) + u;
    else return tws + u;
}

// LDS tile: element (t, c) at c * (2^S + 1) + t.
template <int S>
struct Tile {
    static constexpr uint32_t B = 1u << S;
    static constexpr uint32_t STRIDE = B + 1;
};

template <class Fn, int... I>
__device__ __forceinline__ void static_for_impl(Fn& f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}
// f(integral_constant<int, i>) for i = 0 .. N-1, each i a constant expression
template <int N, class Fn>
__device__ __forceinline__ void static_for(Fn f) {
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// The low round of an inner pass (index bits [0, R), a = 0) on a field whose
// small roots of unity are powers of two: the stage-j twiddle of the pair
// (m, m + 2^j) is omega_{2^(j+1)}^(m mod 2^j) = 2^E with
// E = log2(omega_64) (m mod 2^j) 2^(5-j) mod 192, a constant of the unrolled
// code, so every product is a shift and fold (F::dit_p2 / F::dif_p2) and the
// twiddle table is not read.  DIT runs the forward roots, DIF the inverse
// ones: the direction of the fused LDE pass's two halves.
template <class F, int R, bool DIT>
__device__ __forceinline__ void butterflies_p2(typename F::T* x) {
    static_assert(R <= 6, "roots of order at most 64");
    constexpr int M = 1 << R;
    constexpr int LOG2W = DIT ? F::LOG2_ROOT64_FWD : F::LOG2_ROOT64_INV;
    static_for<R>([&](auto jj) {
        constexpr int j = DIT ? int(decltype(jj)::value) : R - 1 - int(decltype(jj)::value);
        static_for<M>([&](auto mm) {
            constexpr int m = decltype(mm)::value;
            if constexpr ((m & (1 << j)) == 0) {
                constexpr int E = LOG2W * (m & ((1 << j) - 1)) * (1 << (5 - j)) % 192;
                if constexpr (DIT) F::template dit_p2<E>(x[m], x[m | (1 << j)]);
                else F::template dif_p2<E>(x[m], x[m | (1 << j)]);
            }
        });
    });
}

// Butterflies of one radix-2^R round on a register group x[0..2^R) holding
// tile positions tb + m * 2^Q0; twb = round_tw(tws, b).
// A0: the pass starts at index bit 0 (a = 0, so lo0 = 0) and this is its low
// round (Q0 = 0).  Then tws[0] = omega^0 and tws[1] = omega^0: the stage of
// bit 0 and the even-m butterflies of bit 1 multiply by 1, and skip it.
// A field with power-of-two roots runs the twisted form (build_twisted);
// with A0 its gamma is 1 and only the power-of-two butterflies remain.  DIT
// passes run the forward roots and DIF passes the inverse ones, the
// directions of butterflies_p2's two forms.
template <class F, int R, int Q0, bool DIT, bool A0 = false>
__device__ __forceinline__ void butterflies(typename F::T* x, const typename F::T* twb) {
    using T = typename F::T;
    constexpr uint32_t M = 1u << R;
    if constexpr (F::POW2_ROOTS) {
        constexpr bool unit = A0 && Q0 == 0;
        if constexpr (DIT && !unit) {
#pragma unroll
            for (uint32_t m = 1; m < M; ++m) x[m] = F::mul_ntt(x[m], twb[m << Q0]);
        }
        butterflies_p2<F, R, DIT>(x);
        if constexpr (!DIT && !unit) {
#pragma unroll
            for (uint32_t m = 1; m < M; ++m) x[m] = F::mul_ntt(x[m], twb[m << Q0]);
        }
    } else if constexpr (DIT) {
#pragma unroll
        for (uint32_t j = 0; j < R; ++j) {
#pragma unroll
            for (uint32_t m = 0; m < M; ++m) {
                if (m & (1u << j)) continue;
                constexpr bool a0 = A0 && Q0 == 0;
                const bool one = a0 && (j == 0 || (j == 1 && (m & 1u) == 0));
                const T w = twb[(1u << (Q0 + j)) - 1 + ((m & ((1u << j) - 1)) << Q0)];
                if (one) F::bf1(x[m], x[m | (1u << j)]);
                else F::dit(x[m], x[m | (1u << j)], w);
            }
        }
    } else {
#pragma unroll
        for (int jj = R - 1; jj >= 0; --jj) {
            const uint32_t j = uint32_t(jj);
#pragma unroll
            for (uint32_t m = 0; m < M; ++m) {
                if (m & (1u << j)) continue;
                constexpr bool a0 = A0 && Q0 == 0;
                const bool one = a0 && (j == 0 || (j == 1 && (m & 1u) == 0));
                const T w = twb[(1u << (Q0 + j)) - 1 + ((m & ((1u << j) - 1)) << Q0)];
                if (one) F::bf1(x[m], x[m | (1u << j)]);
                else F::dif(x[m], x[m | (1u << j)], w);
            }
        }
    }
}

// Group geometry of a round: thread-group gi -> column c and base position tb.
template <int R, int Q0>
struct Group {
    uint32_t c, b, tb;
    __device__ __forceinline__ Group(uint32_t gi, uint32_t logC) {
        c = gi & ((1u << logC) - 1);
        b = gi >> logC;
        tb = ((b >> Q0) << (Q0 + R)) | (b & ((1u << Q0) - 1));
    }
};

// Tile access.  A group's positions tb + (m << Q0), m < 2^R, lie in one column,
// so a group computes its bounds test and base address once and element m is
// the base plus m times a uniform stride.

// The LDS exchange tile, column-major with STRIDE words per column.
template <class T, uint32_t STRIDE>
struct LdsIO {
    T* lds;
    template <int R, int Q0>
    __device__ __forceinline__ void load(uint32_t c, uint32_t tb, T* x) const {
        const T* p = lds + c * STRIDE + tb;
#pragma unroll
        for (uint32_t m = 0; m < (1u << R); ++m) x[m] = p[m << Q0];
    }
    template <int R, int Q0>
    __device__ __forceinline__ void store(uint32_t c, uint32_t tb, const T* x) const {
        T* p = lds + c * STRIDE + tb;
#pragma unroll
        for (uint32_t m = 0; m < (1u << R); ++m) p[m << Q0] = x[m];
    }
};

// Rows in global memory: tile position t, tile column c is at
// base[t * tstride + c]; columns from `cols` on read as zero and are not
// written (the last column tile of a matrix whose width is not a multiple of C).
template <class T>
struct GlobalIO {
    const T* src;
    T* dst;
    uint64_t tstride;
    uint32_t cols;
    template <int R, int Q0>
    __device__ __forceinline__ void load(uint32_t c, uint32_t tb, T* x) const {
        if (c >= cols) {
#pragma unroll
            for (uint32_t m = 0; m < (1u << R); ++m) x[m] = T(0);
            return;
        }
        const T* p = src + uint64_t(tb) * tstride + c;
        const uint64_t step = tstride << Q0;
#pragma unroll
        for (uint32_t m = 0; m < (1u << R); ++m) x[m] = p[m * step];
    }
    template <int R, int Q0>
    __device__ __forceinline__ void store(uint32_t c, uint32_t tb, const T* x) const {
        if (c >= cols) return;
        T* p = dst + uint64_t(tb) * tstride + c;
        const uint64_t step = tstride << Q0;
#pragma unroll
        for (uint32_t m = 0; m < (1u << R); ++m) p[m * step] = ntt_store_form(x[m]);
    }
};

// The same rows through buffer instructions: a group's 2^R elements share one
// 32-bit lane offset (tb * tstride + c) and element m adds the wave-uniform
// m * 2^Q0 * tstride as the scalar offset, so a group needs one address
// register instead of 2^R 64-bit addresses.  Offsets are bytes from the tile's
// origin; the host uses this only when every offset of the pass stays below
// 2^32.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* base) {
    // gfx9 raw buffer: stride 0, the largest range, 32-bit data format word
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, int(0xFFFFFFFFu), 0x00020000);
}
__device__ __forceinline__ void buf_ld(uint32_t& v, __amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so) {
    v = __builtin_amdgcn_raw_buffer_load_b32(r, int(vo), int(so), 0);
}
__device__ __forceinline__ void buf_ld(uint64_t& v, __amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so) {
    v = __builtin_bit_cast(uint64_t, __builtin_amdgcn_raw_buffer_load_b64(r, int(vo), int(so), 0));
}
__device__ __forceinline__ void buf_st(uint32_t v, __amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so) {
    __builtin_amdgcn_raw_buffer_store_b32(v, r, int(vo), int(so), 0);
}
__device__ __forceinline__ void buf_st(uint64_t v, __amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so) {
    using V2 = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V2, v), r, int(vo), int(so), 0);
}

template <class T>
struct BufIO {
    __amdgpu_buffer_rsrc_t src, dst;
    uint32_t tstride_b;  // bytes between tile positions
    uint32_t cols;
    template <int R, int Q0>
    __device__ __forceinline__ void load(uint32_t c, uint32_t tb, T* x) const {
        if (c >= cols) {
#pragma unroll
            for (uint32_t m = 0; m < (1u << R); ++m) x[m] = T(0);
            return;
        }
        const uint32_t vo = tb * tstride_b + c * uint32_t(sizeof(T));
        const uint32_t step = tstride_b << Q0;
#pragma unroll
        for (uint32_t m = 0; m < (1u << R); ++m) buf_ld(x[m], src, vo, m * step);
    }
    template <int R, int Q0>
    __device__ __forceinline__ void store(uint32_t c, uint32_t tb, const T* x) const {
        if (c >= cols) return;
        const uint32_t vo = tb * tstride_b + c * uint32_t(sizeof(T));
        const uint32_t step = tstride_b << Q0;
#pragma unroll
        for (uint32_t m = 0; m < (1u << R); ++m) buf_st(ntt_store_form(x[m]), dst, vo, m * step);
    }
};

// Element-wise access through ld(c, t) / st(c, t, v), for the bit-reversed
// gather and scatter of the plain transforms.
template <class T, class Ld, class St>
struct ElemIO {
    Ld ldf;
    St stf;
    template <int R, int Q0>
    __device__ __forceinline__ void load(uint32_t c, uint32_t tb, T* x) const {
#pragma unroll
        for (uint32_t m = 0; m < (1u << R); ++m) x[m] = ldf(c, tb + (m << Q0));
    }
    template <int R, int Q0>
    __device__ __forceinline__ void store(uint32_t c, uint32_t tb, const T* x) const {
#pragma unroll
        for (uint32_t m = 0; m < (1u << R); ++m) stf(c, tb + (m << Q0), x[m]);
    }
};
template <class T, class Ld, class St>
__device__ __forceinline__ ElemIO<T, Ld, St> elem_io(Ld ld, St st) {
    return ElemIO<T, Ld, St>{ld, st};
}

// A full round: every group is loaded from `in`, transformed, stored to `out`.
template <class F, int S, int R, int Q0, bool DIT, bool A0 = false, class In, class Out>
__device__ __forceinline__ void round_io(const typename F::T* tws, uint32_t logC, const In& in, const Out& out) {
    using T = typename F::T;
    constexpr uint32_t M = 1u << R;
    const uint32_t groups = ((1u << S) >> R) << logC;
    for (uint32_t gi = threadIdx.x; gi < groups; gi += NTT_THREADS) {
        Group<R, Q0> g(gi, logC);
        T x[M];
        in.template load<R, Q0>(g.c, g.tb, x);
        butterflies<F, R, Q0, DIT, A0>(x, round_tw<F, Q0>(tws, g.b));
        out.template store<R, Q0>(g.c, g.tb, x);
    }
}

// Outer pass on a full tile: 2^S positions x 2^LOGC columns = 2^tile_log
// words, LOGC = tile_log - S.  The tile sits in registers, 2^(tile_log - 9)
// words per thread, and goes through an LDS tile of half the columns, one
// half after the other, between the two rounds.  A thread keeps the column
// c = tid mod 2^LOGC in both rounds (its groups are tid + g * NTT_THREADS and
// NTT_THREADS is a multiple of 2^LOGC), so it only ever exchanges its own
// column's data, and 2^LOGC consecutive lanes read and write whole row
// segments of 2^tile_log / 2^S words (128 bytes at S = 10) with the LDS
// footprint of a tile half as wide.  Lanes of the other half idle while one
// half exchanges.
template <class F, bool DIT, int S, class IO>
__device__ __forceinline__ void outer_split(const typename F::T* tws, const IO& gio, typename F::T* lds) {
    using T = typename F::T;
    constexpr int TL = int(tile_log<T>());
    constexpr int LOGC = TL - S;
    static_assert(S > 5 && LOGC >= 1 && LOGC <= 9, "split pass shape");
    constexpr int R0 = (S + 1) / 2, R1 = S - R0;
    constexpr int RI = DIT ? R0 : R1, QI = DIT ? 0 : R0;  // the round that reads HBM
    constexpr int RO = DIT ? R1 : R0, QO = DIT ? R0 : 0;  // the round that writes HBM
    constexpr uint32_t NW = 1u << (TL - 9);
    constexpr uint32_t GI = NW >> RI, GO = NW >> RO, MI = 1u << RI, MO = 1u << RO;
    const LdsIO<T, Tile<S>::STRIDE> half{lds};
    const uint32_t c = threadIdx.x & ((1u << LOGC) - 1);
    const uint32_t phase = c >> (LOGC - 1);
    const uint32_t cl = c & ((1u << (LOGC - 1)) - 1);
    T v[NW];
#pragma unroll
    for (uint32_t g = 0; g < GI; ++g) {
        const Group<RI, QI> gr(threadIdx.x + g * NTT_THREADS, LOGC);
        gio.template load<RI, QI>(gr.c, gr.tb, v + g * MI);
    }
#pragma unroll
    for (uint32_t g = 0; g < GI; ++g) {
        const Group<RI, QI> gr(threadIdx.x + g * NTT_THREADS, LOGC);
        butterflies<F, RI, QI, DIT>(v + g * MI, round_tw<F, QI>(tws, gr.b));
    }
#pragma unroll
    for (uint32_t p = 0; p < 2; ++p) {
        if (p) __syncthreads();  // the first half has read its exchanged data
        if (phase == p) {
#pragma unroll
            for (uint32_t g = 0; g < GI; ++g) {
                const Group<RI, QI> gr(threadIdx.x + g * NTT_THREADS, LOGC);
                half.template store<RI, QI>(cl, gr.tb, v + g * MI);
            }
        }
        __syncthreads();
        if (phase == p) {
#pragma unroll
            for (uint32_t g = 0; g < GO; ++g) {
                const Group<RO, QO> gr(threadIdx.x + g * NTT_THREADS, LOGC);
                half.template load<RO, QO>(cl, gr.tb, v + g * MO);
            }
        }
    }
#pragma unroll
    for (uint32_t g = 0; g < GO; ++g) {
        const Group<RO, QO> gr(threadIdx.x + g * NTT_THREADS, LOGC);
        butterflies<F, RO, QO, DIT>(v + g * MO, round_tw<F, QO>(tws, gr.b));
        gio.template store<RO, QO>(gr.c, gr.tb, v + g * MO);
    }
}

// Dynamic LDS of a pass (pass_lds_bytes on the host): the column-major
// exchange tile with one padding word per column, (2^S + 1) << logC words,
// when the pass has two rounds (half the columns for the split passes); then
// the gathered twiddles, 2^S - 1 words (twice for the fused inner pass).  Sized to the tile actually used, so
// narrow tiles leave room for more resident workgroups per CU.
template <class F, int MODE, int S>
__device__ __forceinline__ void ntt_pass_body(const NttArgs<typename F::T>& A, uint32_t block, typename F::T* smem) {
    using T = typename F::T;
    constexpr uint32_t STRIDE = Tile<S>::STRIDE;
    // the host never plans a pass wider than tile_log - 4 bits for this field
    if constexpr (uint32_t(S) > tile_log<T>() - 4) {
        return;
    } else {
    // rounds: low = bits [0, R0), high = bits [R0, S)
    constexpr int R0 = S <= 5 ? S : (S + 1) / 2;
    constexpr int R1 = S - R0;

    constexpr bool SPLIT = MODE == OUTER_DIF_INV_SPLIT || MODE == OUTER_DIT_FWD_SPLIT;
    const uint32_t a = A.a, logC = A.logC;
    T* lds = smem;
    T* tws = smem + (R1 > 0 ? (STRIDE << (SPLIT ? logC - 1 : logC)) : 0u);
    T* twf = tws + tw_table_words(F::POW2_ROOTS, S);  // forward twiddles of the fused inner pass

    // block -> (column tile, low index, high index, region)
    uint32_t bid = block;
    const uint32_t ct = bid % A.n_col_tiles;
    bid /= A.n_col_tiles;
    const uint32_t lo0 = bid & ((1u << a) - 1);
    bid >>= a;
    const uint32_t hi_bits = A.L - a - S;
    const uint32_t hi = bid & ((1u << hi_bits) - 1);
    const uint32_t region = (MODE == INNER_LDE) ? 0u : (bid >> hi_bits);
    const uint32_t c0 = ct << logC;
    const uint64_t n = 1ull << A.L;
    const uint32_t w = A.w;
    const uint64_t ld = A.ld;

    // tile position t is transform position row0 + (t << a)
    const uint64_t row0 = (uint64_t(hi) << (a + S)) | lo0;
    auto pos = [&](uint32_t t) -> uint64_t { return row0 | (uint64_t(t) << a); };

    if constexpr (S > 0) {
        // the fused inner pass (a = 0, lo0 = 0): forward table after the inverse one
        if constexpr (F::POW2_ROOTS) {
            build_twisted<F, S>(tws, A.twc, A.tw, A.L, a, lo0);
            if (MODE == INNER_LDE) build_twisted<F, S>(twf, A.twc + TWC_LEN, A.tw + (n >> 1), A.L, a, lo0);
        } else {
            build_twiddles<F>(tws, A.twc, A.tw, A.L, a, S, lo0);
            if (MODE == INNER_LDE) build_twiddles<F>(twf, A.twc + TWC_LEN, A.tw + (n >> 1), A.L, a, S, lo0);
        }
        __syncthreads();
    }

    const T* src = A.src + uint64_t(region) * A.src_region_rows * ld;
    T* dst = A.dst + uint64_t(region) * A.dst_region_rows * ld;
    const uint64_t tile_off = row0 * ld + c0;
    const GlobalIO<T> gio{src + tile_off, dst + tile_off, ld << a, w - c0};
    const LdsIO<T, STRIDE> lio{lds};

    if constexpr (MODE == INNER_LDE) {
        // DIF over all S bits, then per coset: twist, DIT over all S bits.
        // The DIF low round and the DIT low round group the same positions, so
        // the coefficients stay in that round's registers across the cosets.
        // The host keeps (2^S >> R0) * C <= NTT_THREADS: one low group per thread.
        // Inner passes have a = 0: tile position t is row row0 + t.
        if constexpr (R1 > 0) {
            round_io<F, S, R1, R0, false>(tws, logC, gio, lio);
            __syncthreads();
        }
        constexpr uint32_t M = 1u << R0;
        const uint32_t groups = ((1u << S) >> R0) << logC;
        const bool active = threadIdx.x < groups;
        Group<R0, 0> g(threadIdx.x, logC);
        T x[M];
        if (active) {
            if constexpr (R1 > 0) lio.template load<R0, 0>(g.c, g.tb, x);
            else gio.template load<R0, 0>(g.c, g.tb, x);
            butterflies<F, R0, 0, false, true>(x, tws);
        }
        for (int r = int(A.n_regions) - 1; r >= 0; --r) {
            const T* twist = A.scale + uint64_t(r) * n + row0 + g.tb;
            T* dst_r = A.dst + uint64_t(r) * A.dst_region_rows * ld + tile_off;
            const GlobalIO<T> gio_r{nullptr, dst_r, ld, w - c0};
            if constexpr (R1 > 0) __syncthreads();  // previous coset's high round has read the tile
            if (active) {
                T y[M];
#pragma unroll
                for (uint32_t m = 0; m < M; ++m) y[m] = F::mul_ntt(x[m], twist[m]);
                butterflies<F, R0, 0, true, true>(y, twf);
                if constexpr (R1 > 0) lio.template store<R0, 0>(g.c, g.tb, y);
                else gio_r.template store<R0, 0>(g.c, g.tb, y);
            }
            if constexpr (R1 > 0) {
                __syncthreads();
                round_io<F, S, R1, R0, true>(twf, logC, lio, gio_r);
            }
        }
        return;
    } else if constexpr (MODE == INNER_DIT_GATHER) {
        // load rows bitrev(p) (times scale[row]) of the natural-order input
        auto ldf = [&](uint32_t c, uint32_t t) -> T {
            const uint32_t col = c0 + c;
            if (col >= w) return T(0);
            const uint64_t row = bitrev(uint32_t(pos(t)), A.L);
            T v = src[row * ld + col];
            return A.scale ? F::mul_ntt(v, A.scale[row]) : v;
        };
        auto stf = [&](uint32_t, uint32_t, T) {};
        const auto gather = elem_io<T>(ldf, stf);
        if constexpr (R1 > 0) {
            round_io<F, S, R0, 0, true, true>(tws, logC, gather, lio);
            __syncthreads();
            round_io<F, S, R1, R0, true>(tws, logC, lio, gio);
        } else {
            round_io<F, S, R0, 0, true, true>(tws, logC, gather, gio);
        }
    } else if constexpr (MODE == INNER_DIF_SCATTER) {
        // store position p at row bitrev(p), times scale[row]
        auto ldf = [&](uint32_t, uint32_t) -> T { return T(0); };
        auto stf = [&](uint32_t c, uint32_t t, T v) {
            const uint32_t col = c0 + c;
            if (col >= w) return;
            const uint64_t row = bitrev(uint32_t(pos(t)), A.L);
            dst[row * ld + col] = ntt_store_form(A.scale ? F::mul_ntt(v, A.scale[row]) : v);
        };
        const auto scatter = elem_io<T>(ldf, stf);
        if constexpr (R1 > 0) {
            round_io<F, S, R1, R0, false>(tws, logC, gio, lio);
            __syncthreads();
            round_io<F, S, R0, 0, false, true>(tws, logC, lio, scatter);
        } else {
            round_io<F, S, R0, 0, false, true>(tws, logC, gio, scatter);
        }
    } else if constexpr (SPLIT) {
        // the host plans the split form only for two-round passes with
        // logC = tile_log - S and a region below 4 GiB
        if constexpr (R1 > 0) {
            const BufIO<T> bio{buf_rsrc(src + tile_off), buf_rsrc(dst + tile_off), uint32_t((ld << a) * sizeof(T)),
                               w - c0};
            outer_split<F, MODE == OUTER_DIT_FWD_SPLIT, S>(tws, bio, lds);
        }
    } else if constexpr (MODE == OUTER_DIT_FWD) {
        if constexpr (R1 > 0) {
            round_io<F, S, R0, 0, true>(tws, logC, gio, lio);
            __syncthreads();
            round_io<F, S, R1, R0, true>(tws, logC, lio, gio);
        } else {
            round_io<F, S, R0, 0, true>(tws, logC, gio, gio);
        }
    } else {  // OUTER_DIF_INV
        if constexpr (R1 > 0) {
            round_io<F, S, R1, R0, false>(tws, logC, gio, lio);
            __syncthreads();
            round_io<F, S, R0, 0, false>(tws, logC, lio, gio);
        } else {
            round_io<F, S, R0, 0, false>(tws, logC, gio, gio);
        }
    }
    }
}

// Workgroups are handed to the 8 XCDs round-robin (workgroup g runs on XCD
// g mod 8).  Renumbering so that each XCD gets one contiguous run of blocks
// puts the column tiles of the same rows -- consecutive blocks -- on one XCD
// at about the same time, where they share that XCD's L2 lines.  A bijection
// of [0, grid).
__device__ __forceinline__ uint32_t xcd_block(uint32_t g, uint32_t grid) {
    constexpr uint32_t XCDS = 8;
    const uint32_t per = (grid + XCDS - 1) / XCDS;  // blocks on the fuller XCDs
    const uint32_t full = grid - (per - 1) * XCDS;  // XCDs that get `per` blocks
    const uint32_t x = g % XCDS, i = g / XCDS;
    return x < full ? x * per + i : full * per + (x - full) * (per - 1) + i;
}

// Waves per SIMD the register allocation must leave room for: the BabyBear
// split passes hold 64 words per lane and would otherwise settle at 3, one
// 8-wave workgroup per CU; 4 lets two workgroups (and their two LDS half
// tiles) share a CU.  The Goldilocks fused LDE pass of 8 stages (the coset
// coefficients of a 16-word low group held across the cosets) takes 140
// registers unbounded, one workgroup per CU; bounded to 128 it spills 52
// bytes per lane and two workgroups share a CU, which took the 2^22 x 64
// blowup-4 LDE from 16.5 to 15.4 ms.  (At 9 and 10 stages the bound would
// spill over 300 bytes per lane; 6 and 7 stages fit without it.)  1 = no bound.
template <class F, int MODE, int S>
constexpr int pass_min_waves() {
    if constexpr (sizeof(typename F::T) == 8) return MODE == INNER_LDE && S >= 6 && S <= 8 ? 4 : 1;
    return (MODE == OUTER_DIF_INV_SPLIT || MODE == OUTER_DIT_FWD_SPLIT) ? 4 : 1;
}

// Workgroups stride over the pass's blocks when there are more blocks than the
// grid holds (a dispatch counts at most 2^32 - 1 work-items).
template <class F, int MODE, int S>
__global__ __launch_bounds__(NTT_THREADS) __attribute__((amdgpu_waves_per_eu(pass_min_waves<F, MODE, S>())))
void ntt_pass(NttArgs<typename F::T> A, uint32_t total) {
    extern __shared__ __align__(16) unsigned char ntt_smem[];
    const uint32_t b0 = xcd_block(blockIdx.x, gridDim.x);
    for (uint32_t b = b0; b < total; b += gridDim.x) {
        if (b != b0) __syncthreads();  // the previous block is done with the LDS tile
        ntt_pass_body<F, MODE, S>(A, b, reinterpret_cast<typename F::T*>(ntt_smem));
    }
}

// Work-list launch: many transforms (one per descriptor) that share MODE and S
// run in one grid; block_start[i] is the first block of descriptor i
// (ascending, block_start[count] = total).
template <class F, int MODE, int S>
__global__ __launch_bounds__(NTT_THREADS) __attribute__((amdgpu_waves_per_eu(pass_min_waves<F, MODE, S>())))
void ntt_pass_list(const NttWork<typename F::T>* work, uint32_t count,
                                                             uint32_t total) {
    extern __shared__ __align__(16) unsigned char ntt_smem[];
    const uint32_t b0 = xcd_block(blockIdx.x, gridDim.x);
    for (uint32_t b = b0; b < total; b += gridDim.x) {
        if (b != b0) __syncthreads();
        // wave-uniform binary search for the descriptor owning this block
        uint32_t lo = 0, hi = count;
        while (hi - lo > 1) {
            uint32_t mid = (lo + hi) >> 1;
            if (work[mid].block_start <= b) lo = mid;
            else hi = mid;
        }
        const NttArgs<typename F::T> A = work[lo].args;
        ntt_pass_body<F, MODE, S>(A, b - work[lo].block_start, reinterpret_cast<typename F::T*>(ntt_smem));
    }
}

// table[i] = base^i * mult for i < count (multiply-constant form of each value)
template <class F>
__global__ void powers_kernel(typename F::T* out, uint64_t count, typename F::T base_c,
                              typename F::T mult_c, typename F::T one_c) {
    uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= count) return;
    using T = typename F::T;
    // square-and-multiply in multiply-constant form: mul_c(x_c, y_c) of two
    // constants gives the constant form of x*y for both fields.
    T acc = one_c, b = base_c;
    uint64_t e = i;
    while (e) {
        if (e & 1) acc = F::mul_c(acc, b);
        b = F::mul_c(b, b);
        e >>= 1;
    }
    out[i] = F::mul_c(acc, mult_c);
}

// LDE twist table for coset r, bit-reversed positions:
// out[r*n + p] = (shift * omega_N^r)^bitrev(p) / n.
template <class F>
__global__ void lde_twist_kernel(typename F::T* out, uint32_t L, uint32_t n_cosets,
                                 const typename F::T* coset_base_c, typename F::T inv_n_c,
                                 typename F::T one_c) {
    using T = typename F::T;
    uint64_t n = 1ull << L;
    uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n * n_cosets) return;
    uint32_t r = uint32_t(i >> L);
    uint32_t p = uint32_t(i & (n - 1));
    uint64_t e = bitrev(p, L);
    T acc = one_c, b = coset_base_c[r];
    while (e) {
        if (e & 1) acc = F::mul_c(acc, b);
        b = F::mul_c(b, b);
        e >>= 1;
    }
    out[i] = F::mul_c(acc, inv_n_c);
}

// coset-major -> natural row order: out[(i << k) | r] = in[r * n + i]
template <class T>
__global__ void coset_to_natural_kernel(const T* in, T* out, uint32_t L, uint32_t k, uint32_t w) {
    uint64_t rows = 1ull << (L + k);
    uint64_t idx = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    uint64_t total = rows * w;
    if (idx >= total) return;
    uint64_t row = idx / w, col = idx - row * w;
    uint64_t r = row & ((1ull << k) - 1), i = row >> k;
    out[idx] = in[((r << L) + i) * w + col];
}

// ------------------------------------------------------------------ host ----

// Pass widths for an L-bit transform: the inner pass takes the low bits, the
// outer passes split the rest as evenly as possible (top pass first).
struct Plan {
    uint32_t s_inner;
    uint32_t outer[32];  // L <= 32 and every pass has at least one stage
    uint32_t n_outer;
};

template <class F>
Plan make_plan(uint32_t L, const NttTuning& tune) {
    using T = typename F::T;
    uint32_t smax = tile_log<T>() - 4;  // keep at least 16 columns per tile
    if (smax > MAX_STAGES) smax = MAX_STAGES;
    if (tune.max_stages && tune.max_stages < smax) smax = tune.max_stages;
    if (smax == 0) smax = 1;
    Plan p{};
    p.s_inner = L < smax ? L : smax;
    uint32_t rest = L - p.s_inner;
    uint32_t np = (rest + smax - 1) / smax;
    p.n_outer = np;
    for (uint32_t i = 0; i < np; ++i) p.outer[i] = rest / np + (i < rest % np ? 1 : 0);
    return p;
}

// columns per tile: as many as the LDS tile holds at this pass width (capped
// by the tuning knob and by the matrix width rounded up to a power of two)
template <class T>
uint32_t cols_log(uint32_t s, uint32_t w, const NttTuning& tune) {
    uint32_t lc = tile_log<T>() - (s < 4 ? 4 : s);
    if (tune.log_cols && tune.log_cols < lc) lc = tune.log_cols;
    uint32_t need = 0;
    while ((1u << need) < w) ++need;
    return lc < need ? lc : need;
}

// Fill the tile geometry of a pass; returns its block count.
template <class T>
uint64_t finalize_args(NttArgs<T>& A, int mode, uint32_t s, const NttTuning& tune) {
    A.logC = cols_log<T>(s, A.w, tune);
    if (s <= 5) {
        // a single-round pass keeps no LDS tile and gives each thread one
        // column's group: 2^logC threads work, so take up to one column per
        // thread (the matrix width rounded up to a power of two, at most 512)
        uint32_t need = 0;
        while ((1u << need) < A.w) ++need;
        A.logC = need < 9 ? need : 9;  // log2(NTT_THREADS) = 9
    }
    if (mode == INNER_LDE) {
        // one low-round group per thread: (2^s >> R0) * C <= NTT_THREADS
        const uint32_t r0 = s <= 5 ? s : (s + 1) / 2;
        const uint32_t max_lc = 9 + r0 - s;  // log2(NTT_THREADS) = 9
        if (A.logC > max_lc) A.logC = max_lc;
    }
    A.n_col_tiles = (A.w + (1u << A.logC) - 1) >> A.logC;
    uint64_t blocks = uint64_t(A.n_col_tiles) << (A.L - s);
    if (mode != INNER_LDE) blocks *= A.n_regions;
    return blocks;
}

// Dynamic LDS bytes of a pass of width s with 2^logC-column tiles (the layout
// of ntt_pass_body).  At most 2^(tile_log) + 2^(tile_log - 4) tile words plus
// 2 (2^10 + 2^5) twiddles: below 153 KiB for both fields, inside the 160 KiB of
// one workgroup; tiles above 80 KiB leave one workgroup per CU.
template <class F>
uint32_t pass_lds_bytes(int mode, uint32_t s, uint32_t logC) {
    const uint32_t r0 = low_round_bits(s);
    const bool split = mode == OUTER_DIF_INV_SPLIT || mode == OUTER_DIT_FWD_SPLIT;
    const uint64_t tile = r0 < s ? (((1ull << s) + 1) << (split ? logC - 1 : logC)) : 0;
    const uint64_t tws = uint64_t(tw_table_words(F::POW2_ROOTS, s)) * (mode == INNER_LDE ? 2 : 1);
    return uint32_t((tile + tws) * sizeof(typename F::T));
}

// The kernel variant of a planned pass: an outer two-round pass whose tile is
// the full 2^tile_log words (one workgroup per CU if held in LDS whole) runs
// in the split form, with half the tile in LDS at a time, when every byte
// offset from a tile's origin fits 32 bits (an n x ld region below 4 GiB).
template <class T>
int pass_mode(int mode, uint32_t s, const NttArgs<T>& A) {
    const uint32_t r0 = s <= 5 ? s : (s + 1) / 2;
    const uint64_t region_bytes = (uint64_t(A.ld) << A.L) * sizeof(T);
    if ((mode == OUTER_DIF_INV || mode == OUTER_DIT_FWD) && r0 < s && s + A.logC == tile_log<T>() &&
        region_bytes <= 0xFFFFFFFFull)
        return mode == OUTER_DIF_INV ? OUTER_DIF_INV_SPLIT : OUTER_DIT_FWD_SPLIT;
    return mode;
}

// grid cap: 2^22 workgroups x 512 work-items = 2^31 work-items per dispatch
constexpr uint64_t MAX_GRID = 1ull << 22;
constexpr int MAX_DYN_LDS = 160 * 1024;  // LDS of one gfx950 workgroup

// Allow every pass kernel the full dynamic LDS of a workgroup (above the
// 64 KiB default).  Done once per device before any launch, so that no first
// launch happens inside a stream capture.
template <class F, int MODE, int S>
hipError_t allow_lds_s() {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&ntt_pass<F, MODE, S>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, MAX_DYN_LDS);
    if (e == hipSuccess)
        e = hipFuncSetAttribute(reinterpret_cast<const void*>(&ntt_pass_list<F, MODE, S>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, MAX_DYN_LDS);
    return e;
}
template <class F, int MODE, int... S>
hipError_t allow_lds_mode(std::integer_sequence<int, S...>) {
    hipError_t e = hipSuccess;
    ((e = e == hipSuccess ? allow_lds_s<F, MODE, S>() : e), ...);
    return e;
}
template <class F>
hipError_t allow_lds_field() {
    using Widths = std::make_integer_sequence<int, int(MAX_STAGES) + 1>;
    hipError_t e = allow_lds_mode<F, OUTER_DIF_INV>(Widths{});
    if (e == hipSuccess) e = allow_lds_mode<F, OUTER_DIT_FWD>(Widths{});
    if (e == hipSuccess) e = allow_lds_mode<F, INNER_LDE>(Widths{});
    if (e == hipSuccess) e = allow_lds_mode<F, INNER_DIT_GATHER>(Widths{});
    if (e == hipSuccess) e = allow_lds_mode<F, INNER_DIF_SCATTER>(Widths{});
    if (e == hipSuccess) e = allow_lds_mode<F, OUTER_DIF_INV_SPLIT>(Widths{});
    if (e == hipSuccess) e = allow_lds_mode<F, OUTER_DIT_FWD_SPLIT>(Widths{});
    return e;
}

template <class F, int MODE, int S>
void launch_s(const NttArgs<typename F::T>& A, uint64_t blocks, uint32_t lds, const NttWork<typename F::T>* work,
              uint32_t count, hipStream_t st) {
    const uint32_t grid = uint32_t(blocks < MAX_GRID ? blocks : MAX_GRID);
    if (work)
        hipLaunchKernelGGL((ntt_pass_list<F, MODE, S>), dim3(grid), dim3(NTT_THREADS), lds, st, work, count,
                           uint32_t(blocks));
    else
        hipLaunchKernelGGL((ntt_pass<F, MODE, S>), dim3(grid), dim3(NTT_THREADS), lds, st, A, uint32_t(blocks));
}

template <class F, int MODE>
void launch_mode(uint32_t s, const NttArgs<typename F::T>& A, uint64_t blocks, uint32_t lds,
                 const NttWork<typename F::T>* work, uint32_t count, hipStream_t st) {
    switch (s) {
        case 0: launch_s<F, MODE, 0>(A, blocks, lds, work, count, st); break;
        case 1: launch_s<F, MODE, 1>(A, blocks, lds, work, count, st); break;
        case 2: launch_s<F, MODE, 2>(A, blocks, lds, work, count, st); break;
        case 3: launch_s<F, MODE, 3>(A, blocks, lds, work, count, st); break;
        case 4: launch_s<F, MODE, 4>(A, blocks, lds, work, count, st); break;
        case 5: launch_s<F, MODE, 5>(A, blocks, lds, work, count, st); break;
        case 6: launch_s<F, MODE, 6>(A, blocks, lds, work, count, st); break;
        case 7: launch_s<F, MODE, 7>(A, blocks, lds, work, count, st); break;
        case 8: launch_s<F, MODE, 8>(A, blocks, lds, work, count, st); break;
        case 9: launch_s<F, MODE, 9>(A, blocks, lds, work, count, st); break;
        default: launch_s<F, MODE, 10>(A, blocks, lds, work, count, st); break;
    }
}

template <class F>
void launch_any(int mode, uint32_t s, const NttArgs<typename F::T>& A, uint64_t blocks, uint32_t lds,
                const NttWork<typename F::T>* work, uint32_t count, hipStream_t st) {
    switch (mode) {
        case OUTER_DIF_INV: launch_mode<F, OUTER_DIF_INV>(s, A, blocks, lds, work, count, st); break;
        case OUTER_DIT_FWD: launch_mode<F, OUTER_DIT_FWD>(s, A, blocks, lds, work, count, st); break;
        case INNER_LDE: launch_mode<F, INNER_LDE>(s, A, blocks, lds, work, count, st); break;
        case INNER_DIT_GATHER: launch_mode<F, INNER_DIT_GATHER>(s, A, blocks, lds, work, count, st); break;
        case OUTER_DIF_INV_SPLIT: launch_mode<F, OUTER_DIF_INV_SPLIT>(s, A, blocks, lds, work, count, st); break;
        case OUTER_DIT_FWD_SPLIT: launch_mode<F, OUTER_DIT_FWD_SPLIT>(s, A, blocks, lds, work, count, st); break;
        default: launch_mode<F, INNER_DIF_SCATTER>(s, A, blocks, lds, work, count, st); break;
    }
}

template <class F, int MODE>
void launch_pass(NttArgs<typename F::T> A, uint32_t s, const NttTuning& tune, hipStream_t st) {
    using T = typename F::T;
    const uint64_t blocks = finalize_args(A, MODE, s, tune);
    const int mode = pass_mode<T>(MODE, s, A);
    launch_any<F>(mode, s, A, blocks, pass_lds_bytes<F>(mode, s, A.logC), nullptr, 0, st);
}

// All passes of one LDE over columns [0, w) of row-stride-ld matrices.
template <class F>
void lde_slab(const typename F::T* in, typename F::T* out, uint32_t L, uint32_t w, uint32_t ld, uint32_t k,
              const typename F::T* tw_inv_fwd, const typename F::T* twist, const NttTuning& tune, hipStream_t st) {
    Plan p = make_plan<F>(L, tune);
    const uint64_t n = 1ull << L;
    NttArgs<typename F::T> A{};
    A.L = L;
    A.w = w;
    A.ld = ld;
    A.n_regions = 1;
    A.tw = tw_inv_fwd;  // inverse half first
    A.twc = twc_of(tw_inv_fwd, L, false);
    // 1. outer DIF passes over the high bits, top down: input -> region 0 of out
    uint32_t top = L;
    const typename F::T* cur = in;
    for (uint32_t i = 0; i < p.n_outer; ++i) {
        A.a = top - p.outer[i];
        A.src = cur;
        A.dst = out;
        launch_pass<F, OUTER_DIF_INV>(A, p.outer[i], tune, st);
        cur = out;
        top = A.a;
    }
    // 2. inner fused pass: rest of the inverse, then per-coset twist + DIT;
    //    the grid covers the blocks, the kernel loops over the cosets
    A.a = 0;
    A.src = cur;
    A.dst = out;
    A.scale = twist;
    A.n_regions = 1u << k;
    A.dst_region_rows = n;
    launch_pass<F, INNER_LDE>(A, p.s_inner, tune, st);
    // 3. outer DIT passes per coset, bottom up, in place
    A.scale = nullptr;
    A.tw = tw_inv_fwd + (n >> 1);
    A.twc = twc_of(tw_inv_fwd, L, true);
    A.src = out;
    A.dst = out;
    A.src_region_rows = A.dst_region_rows = n;
    uint32_t bot = p.s_inner;
    for (int i = int(p.n_outer) - 1; i >= 0; --i) {
        A.a = bot;
        launch_pass<F, OUTER_DIT_FWD>(A, p.outer[i], tune, st);
        bot += p.outer[i];
    }
}

}  // namespace

// With tune.log_slab set, the LDE runs slab by slab (2^log_slab columns each):
// one slab's intermediate rows (2^(L+k) x slab) can stay in the device's
// last-level cache across its passes instead of streaming every pass through
// HBM.
template <class F>
int lde_device(const typename F::T* in, typename F::T* out, uint32_t L, uint32_t w, uint32_t k,
               const typename F::T* tw_inv_fwd, const typename F::T* twist, const NttTuning& tune,
               hipStream_t st) {
    const uint32_t slab = tune.log_slab ? (1u << tune.log_slab) : w;
    for (uint32_t c0 = 0; c0 < w; c0 += slab)
        lde_slab<F>(in + c0, out + c0, L, (w - c0 < slab ? w - c0 : slab), w, k, tw_inv_fwd, twist, tune, st);
    return int(hipGetLastError());
}

template <class F>
int dft_device(int kind, const typename F::T* in, typename F::T* out, typename F::T* tmp, uint32_t L,
               uint32_t w, const typename F::T* tw_inv_fwd, const typename F::T* row_scale,
               const NttTuning& tune, hipStream_t st) {
    Plan p = make_plan<F>(L, tune);
    const uint64_t n = 1ull << L;
    NttArgs<typename F::T> A{};
    A.L = L;
    A.w = w;
    A.ld = w;
    A.n_regions = 1;
    if (kind == 0 || kind == 2) {
        // forward (optionally coset): gather rows in bit-reversed order with
        // the row scale, DIT inner, DIT outer in place.
        A.tw = tw_inv_fwd + (n >> 1);
        A.twc = twc_of(tw_inv_fwd, L, true);
        A.a = 0;
        A.src = in;
        A.dst = out;
        A.scale = row_scale;
        launch_pass<F, INNER_DIT_GATHER>(A, p.s_inner, tune, st);
        A.scale = nullptr;
        A.src = out;
        uint32_t bot = p.s_inner;
        for (int i = int(p.n_outer) - 1; i >= 0; --i) {
            A.a = bot;
            launch_pass<F, OUTER_DIT_FWD>(A, p.outer[i], tune, st);
            bot += p.outer[i];
        }
    } else {
        // inverse (optionally coset): DIF outer (in -> tmp), DIF inner with
        // bit-reversed scatter and the row scale (tmp -> out).
        A.tw = tw_inv_fwd;
        A.twc = twc_of(tw_inv_fwd, L, false);
        uint32_t top = L;
        const typename F::T* cur = in;
        for (uint32_t i = 0; i < p.n_outer; ++i) {
            A.a = top - p.outer[i];
            A.src = cur;
            A.dst = tmp;
            launch_pass<F, OUTER_DIF_INV>(A, p.outer[i], tune, st);
            cur = tmp;
            top = A.a;
        }
        A.a = 0;
        A.src = cur;
        A.dst = out;
        A.scale = row_scale;
        launch_pass<F, INNER_DIF_SCATTER>(A, p.s_inner, tune, st);
    }
    return int(hipGetLastError());
}

template <class F>
int ntt_prepare_field() {
    return int(allow_lds_field<F>());
}

template <class F>
int powers_device(typename F::T* out, uint64_t count, typename F::T base_c, typename F::T mult_c,
                  typename F::T one_c, hipStream_t st) {
    if (count == 0) return 0;
    uint32_t blocks = uint32_t((count + 255) / 256);
    hipLaunchKernelGGL((powers_kernel<F>), dim3(blocks), dim3(256), 0, st, out, count, base_c, mult_c, one_c);
    return int(hipGetLastError());
}

template <class F>
int lde_twist_device(typename F::T* out, uint32_t L, uint32_t n_cosets, const typename F::T* coset_base_c,
                     typename F::T inv_n_c, typename F::T one_c, hipStream_t st) {
    uint64_t total = (1ull << L) * n_cosets;
    uint32_t blocks = uint32_t((total + 255) / 256);
    hipLaunchKernelGGL((lde_twist_kernel<F>), dim3(blocks), dim3(256), 0, st, out, L, n_cosets, coset_base_c,
                       inv_n_c, one_c);
    return int(hipGetLastError());
}

template <class T>
int coset_to_natural_device(const T* in, T* out, uint32_t L, uint32_t k, uint32_t w, hipStream_t st) {
    uint64_t total = (1ull << (L + k)) * w;
    uint32_t blocks = uint32_t((total + 255) / 256);
    hipLaunchKernelGGL((coset_to_natural_kernel<T>), dim3(blocks), dim3(256), 0, st, in, out, L, k, w);
    return int(hipGetLastError());
}


// Every pass of every LDE of a batch, with its slot: the inverse outer passes
// at negative slots (farthest first), the fused inner pass at 0, the forward
// outer passes at positive slots.  Passes of one LDE run in slot order.
template <class T>
struct BatchPass {
    int mode;
    uint32_t s;
    int slot;
    uint32_t job;
    NttArgs<T> A;
    NttTuning tune;
};

template <class F>
std::vector<BatchPass<typename F::T>> batch_passes(const std::vector<LdeJob<typename F::T>>& jobs) {
    using T = typename F::T;
    std::vector<BatchPass<T>> passes;
    for (uint32_t ji = 0; ji < jobs.size(); ++ji) {
        const auto& j = jobs[ji];
        Plan p = make_plan<F>(j.L, j.tune);
        const uint64_t n = 1ull << j.L;
        NttArgs<T> A{};
        A.L = j.L;
        A.w = j.w;
        A.ld = j.w;
        A.n_regions = 1;
        A.tw = j.tw;
        A.twc = twc_of(j.tw, j.L, false);
        uint32_t top = j.L;
        const T* cur = j.in;
        for (uint32_t i = 0; i < p.n_outer; ++i) {
            A.a = top - p.outer[i];
            A.src = cur;
            A.dst = j.out;
            passes.push_back({OUTER_DIF_INV, p.outer[i], -int(p.n_outer - i), ji, A, j.tune});
            cur = j.out;
            top = A.a;
        }
        A.a = 0;
        A.src = cur;
        A.dst = j.out;
        A.scale = j.twist;
        A.n_regions = 1u << j.k;
        A.dst_region_rows = n;
        passes.push_back({INNER_LDE, p.s_inner, 0, ji, A, j.tune});
        A.scale = nullptr;
        A.tw = j.tw + (n >> 1);
        A.twc = twc_of(j.tw, j.L, true);
        A.src = j.out;
        A.dst = j.out;
        A.src_region_rows = A.dst_region_rows = n;
        uint32_t bot = p.s_inner;
        for (int i = int(p.n_outer) - 1; i >= 0; --i) {
            A.a = bot;
            passes.push_back({OUTER_DIT_FWD, p.outer[i], int(p.n_outer) - i, ji, A, j.tune});
            bot += p.outer[i];
        }
    }
    return passes;
}

template <class F>
std::vector<LdePassGroup<typename F::T>> plan_lde_batch(const std::vector<LdeJob<typename F::T>>& jobs) {
    using T = typename F::T;
    const std::vector<BatchPass<T>> passes = batch_passes<F>(jobs);
    std::vector<LdePassGroup<T>> groups;
    int lo = 0, hi = 0;
    for (const auto& ps : passes) {
        lo = ps.slot < lo ? ps.slot : lo;
        hi = ps.slot > hi ? ps.slot : hi;
    }
    // tile geometry and kernel variant of every pass
    std::vector<NttWork<T>> geo(passes.size());
    std::vector<uint64_t> nblk(passes.size());
    std::vector<int> kmode(passes.size());
    for (size_t i = 0; i < passes.size(); ++i) {
        geo[i] = NttWork<T>{};
        geo[i].args = passes[i].A;
        nblk[i] = finalize_args(geo[i].args, passes[i].mode, passes[i].s, passes[i].tune);
        kmode[i] = pass_mode<T>(passes[i].mode, passes[i].s, geo[i].args);
    }
    for (int slot = lo; slot <= hi; ++slot)
        for (uint32_t s = 0; s <= MAX_STAGES; ++s)
            for (int split = 0; split < 2; ++split) {
                const int base = slot < 0 ? OUTER_DIF_INV : (slot == 0 ? INNER_LDE : OUTER_DIT_FWD);
                const int mode = !split ? base
                                        : (base == OUTER_DIF_INV ? OUTER_DIF_INV_SPLIT
                                                                 : (base == OUTER_DIT_FWD ? OUTER_DIT_FWD_SPLIT : -1));
                if (mode < 0) continue;
                LdePassGroup<T> g;
                g.mode = mode;
                g.s = s;
                g.blocks = 0;
                g.lds_bytes = 0;
                for (size_t i = 0; i < passes.size(); ++i) {
                    if (passes[i].slot != slot || passes[i].s != s || kmode[i] != mode) continue;
                    NttWork<T> wk = geo[i];
                    const uint32_t lds = pass_lds_bytes<F>(mode, s, wk.args.logC);
                    g.lds_bytes = lds > g.lds_bytes ? lds : g.lds_bytes;
                    wk.block_start = g.blocks;
                    g.blocks += uint32_t(nblk[i]);
                    g.work.push_back(wk);
                }
                if (!g.work.empty()) groups.push_back(std::move(g));
            }
    return groups;
}

template <class F>
int launch_lde_group(const LdePassGroup<typename F::T>& g, const NttWork<typename F::T>* d_work, hipStream_t st) {
    if (g.blocks == 0) return 0;
    launch_any<F>(g.mode, g.s, g.work[0].args, g.blocks, g.lds_bytes, d_work, uint32_t(g.work.size()), st);
    return int(hipGetLastError());
}

// ---- the queue launch of a batch's LDE passes ----
//
// One launch runs every pass of every LDE of a batch.  A pass's blocks are
// cut into chunks of consecutive blocks (one block for passes of at most 512,
// up to 8 for large ones, so the per-ticket costs below are shared), and the grid has
// one workgroup per chunk; each workgroup takes a ticket from a counter, and
// ticket t is chunk t - first of the pass (item) whose range holds it.  Before
// its chunk the workgroup waits until every block of the pass that item reads
// (the same LDE's previous pass) is counted done, and after it counts its own
// blocks done.  Tickets are handed out in increasing order to
// running workgroups and an item's dependency has lower tickets than the
// item, so every awaited block was taken by a workgroup that is already
// running and waits, if at all, only for lower tickets: the launch completes
// whatever the number of resident workgroups, and the hardware dispatcher
// refills a CU as soon as one of its workgroups ends.  Results pass between
// workgroups on different XCDs (each with its own L2), so a block's stores are
// released at agent scope (written back past the XCD's L2, once per chunk)
// before its completion count, and a waiting workgroup acquires at agent
// scope (once, after its wait) before reading.
//
// Every branch at the level of the queue is uniform across a wave (the ticket
// and the item are scalar, the wave index is read through readfirstlane):
// a region taken by one lane of a wave inside code that holds barriers lets
// the compiler split the wave's lanes between loop nests, which livelocks
// them.  A workgroup takes one ticket, so the only loop around the pass
// bodies runs over the chunk's blocks, with scalar bounds.
//
// The kernel carries the pass variants of batches above 2^5 rows: outer
// passes of 1..10 stages, fused inner passes of 1..3 and 6..10, split outer
// passes of 6..10 (the split form exists only above 5 stages).  It is held to
// 128 VGPRs (two 512-thread workgroups per CU); the single-round fused passes
// of 4 and 5 stages need more (they keep 16 or 32 coefficients and a coset's
// values per lane) and are left to the grouped launches.

constexpr uint32_t QUEUE_SPIN_LIMIT = 1u << 25;  // sleeps of 256 cycles: seconds, far above any pass

__host__ __device__ constexpr bool queue_carries(int mode, uint32_t s) {
    return (mode == OUTER_DIF_INV || mode == OUTER_DIT_FWD) ? (s >= 1 && s <= 10)
           : mode == INNER_LDE ? ((s >= 1 && s <= 3) || (s >= 6 && s <= 10))
           : (mode == OUTER_DIF_INV_SPLIT || mode == OUTER_DIT_FWD_SPLIT) ? (s >= 6 && s <= 10)
                                                                          : false;
}

namespace {

template <class F, int MODE, int... S>
__device__ __forceinline__ void queue_pass_s(uint32_t s, const NttArgs<typename F::T>& A, uint32_t block,
                                             typename F::T* smem, std::integer_sequence<int, S...>) {
    ((s == uint32_t(S) ? (ntt_pass_body<F, MODE, S>(A, block, smem), 0) : 0), ...);
}

template <class F>
__device__ __forceinline__ void queue_pass(uint32_t mode, uint32_t s, const NttArgs<typename F::T>& A,
                                           uint32_t block, typename F::T* smem) {
    using Outer = std::integer_sequence<int, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10>;
    using Inner = std::integer_sequence<int, 1, 2, 3, 6, 7, 8, 9, 10>;
    using Wide = std::integer_sequence<int, 6, 7, 8, 9, 10>;
    switch (mode) {
        case OUTER_DIF_INV: queue_pass_s<F, OUTER_DIF_INV>(s, A, block, smem, Outer{}); break;
        case OUTER_DIT_FWD: queue_pass_s<F, OUTER_DIT_FWD>(s, A, block, smem, Outer{}); break;
        case INNER_LDE: queue_pass_s<F, INNER_LDE>(s, A, block, smem, Inner{}); break;
        case OUTER_DIF_INV_SPLIT: queue_pass_s<F, OUTER_DIF_INV_SPLIT>(s, A, block, smem, Wide{}); break;
        case OUTER_DIT_FWD_SPLIT: queue_pass_s<F, OUTER_DIT_FWD_SPLIT>(s, A, block, smem, Wide{}); break;
        default: break;
    }
}

// ctr: [0] ticket counter, [1] error word, [2 + i] blocks of item i done.
template <class F>
__global__ __launch_bounds__(NTT_THREADS) __attribute__((amdgpu_waves_per_eu(4)))
void ntt_lde_queue(const NttQueueItem<typename F::T>* __restrict__ items, uint32_t n_items, uint32_t total,
                   uint32_t* ctr) {
    using T = typename F::T;
    extern __shared__ __align__(16) unsigned char ntt_smem[];
    __shared__ uint32_t s_ticket;
    uint32_t* const done = ctr + 2;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63u;
    if (wave == 0) {
        // one increment per workgroup: lane 0 adds 1, the other lanes 0, and
        // lane 0's result is the value before its increment
        const uint32_t v = __hip_atomic_fetch_add(ctr, lane == 0 ? 1u : 0u, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT);
        s_ticket = __builtin_amdgcn_readfirstlane(v);
    }
    __syncthreads();
    const uint32_t t = __builtin_amdgcn_readfirstlane(s_ticket);
    if (t >= total) return;  // the grid has exactly `total` workgroups
    uint32_t lo = 0, hi = n_items;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (items[mid].first <= t) lo = mid;
        else hi = mid;
    }
    const int32_t dep = items[lo].dep;
    if (dep >= 0) {
        // wave 0 polls (relaxed: no cache invalidation per poll), then one
        // acquire invalidates this CU's L1 and the XCD's L2 for every wave of
        // the workgroup, which read only after the barrier
        if (wave == 0) {
            const uint32_t need = items[dep].blocks;
            uint32_t spins = 0;
            while (__builtin_amdgcn_readfirstlane(
                       __hip_atomic_load(done + dep, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) < need) {
                __builtin_amdgcn_s_sleep(4);
                if (++spins == QUEUE_SPIN_LIMIT) {
                    __hip_atomic_store(ctr + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        __syncthreads();
    }
    const NttArgs<T> A = items[lo].args;
    const uint32_t mode = items[lo].mode, s = items[lo].s, chunk = items[lo].chunk, blocks = items[lo].blocks;
    const uint32_t b0 = (t - items[lo].first) * chunk;
    const uint32_t b1 = b0 + chunk < blocks ? b0 + chunk : blocks;
    for (uint32_t b = b0; b < b1; ++b) {
        if (b != b0) __syncthreads();  // the previous block is done with the LDS tile
        queue_pass<F>(mode, s, A, b, reinterpret_cast<T*>(ntt_smem));
    }
    // every wave's stores have reached the XCD's L2 (vmcnt 0) before the
    // barrier; then one release by wave 0 writes the L2 back for all of them
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (wave == 0)
        __hip_atomic_fetch_add(done + lo, lane == 0 ? b1 - b0 : 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

template <class U>
__global__ void fill_kernel(U* p, uint32_t n, U v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace

template <class F>
bool plan_lde_queue(const std::vector<LdeJob<typename F::T>>& jobs, LdeQueuePlan<typename F::T>& plan) {
    using T = typename F::T;
    const std::vector<BatchPass<T>> passes = batch_passes<F>(jobs);
    int lo = 0, hi = 0;
    for (const auto& ps : passes) {
        lo = ps.slot < lo ? ps.slot : lo;
        hi = ps.slot > hi ? ps.slot : hi;
    }
    plan = LdeQueuePlan<T>{};
    std::vector<int32_t> last(jobs.size(), -1);  // each LDE's latest planned pass
    uint64_t total = 0;
    for (int slot = lo; slot <= hi; ++slot)
        for (const auto& ps : passes) {
            if (ps.slot != slot) continue;
            NttQueueItem<T> it{};
            it.args = ps.A;
            const uint64_t blocks = finalize_args(it.args, ps.mode, ps.s, ps.tune);
            const int mode = pass_mode<T>(ps.mode, ps.s, it.args);
            if (!queue_carries(mode, ps.s)) return false;
            it.first = uint32_t(total);
            it.blocks = uint32_t(blocks);
            // chunks of consecutive blocks: at least 512 chunks per pass (two
            // per resident workgroup) before any chunk holds two blocks
            it.chunk = uint32_t(blocks >= 512 * 8 ? 8 : (blocks + 511) / 512);
            it.dep = last[ps.job];
            it.mode = uint16_t(mode);
            it.s = uint16_t(ps.s);
            total += (blocks + it.chunk - 1) / it.chunk;
            if (total > MAX_GRID) return false;  // one workgroup per chunk in one dispatch
            const uint32_t lds = pass_lds_bytes<F>(mode, ps.s, it.args.logC);
            // one launch takes the largest pass's LDS for every workgroup: past half
            // a CU's LDS that would hold every pass to one workgroup per CU
            if (lds > uint32_t(MAX_DYN_LDS) / 2 - 1024) return false;
            plan.lds_bytes = lds > plan.lds_bytes ? lds : plan.lds_bytes;
            last[ps.job] = int32_t(plan.items.size());
            plan.items.push_back(it);
        }
    plan.total = uint32_t(total);
    return true;
}

template <class F>
int launch_lde_queue(const LdeQueuePlan<typename F::T>& plan, const NttQueueItem<typename F::T>* d_items,
                     uint32_t* d_counters, hipStream_t st) {
    const uint32_t words = uint32_t(lde_queue_counter_words(plan.items.size()));
    hipLaunchKernelGGL(fill_kernel<uint32_t>, dim3((words + 255) / 256), dim3(256), 0, st, d_counters, words, 0u);
    if (plan.total == 0) return int(hipGetLastError());
    hipLaunchKernelGGL((ntt_lde_queue<F>), dim3(plan.total), dim3(NTT_THREADS), plan.lds_bytes, st, d_items,
                       uint32_t(plan.items.size()), plan.total, d_counters);
    return int(hipGetLastError());
}

// The queue kernel's dynamic LDS: the planner keeps it below half a CU's (the
// kernel also has a few bytes of static LDS, the ticket it shares).
template <class F>
int ntt_prepare_queue() {
    return int(hipFuncSetAttribute(reinterpret_cast<const void*>(&ntt_lde_queue<F>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, MAX_DYN_LDS / 2));
}

namespace {

// VALU ceiling of the NTT arithmetic: each thread takes 16 words through the
// passes' own register-round code -- butterflies<F, 4, 1, DIT>, the round of a
// pass's high bits: 4 stages of 8 butterflies on the stage twiddles for
// BabyBear, the twisted round (15 scalings, then 32 power-of-two butterflies)
// for Goldilocks -- `iters` times, twiddles held in registers (32 per lane,
// lane-dependent like the rows of the LDS table a pass reads), with no memory
// traffic inside the loop; the words are xor-folded into one store so the work
// cannot be dropped.  Its rate is counted in radix-2 butterflies (32 per round).
template <class F>
__global__ __launch_bounds__(256) void butterfly_probe_kernel(typename F::T* out, const typename F::T* tw,
                                                              uint32_t iters) {
    using T = typename F::T;
    constexpr int R = 4, M = 1 << R;
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    T x[M], w[2 * M];
#pragma unroll
    for (int m = 0; m < M; ++m) x[m] = T(tid * 31u + uint32_t(m)) % T(F::P);
#pragma unroll
    for (int i = 0; i < 2 * M; ++i) w[i] = tw[(uint32_t(i) + threadIdx.x) & (2 * M - 1)];
    for (uint32_t it = 0; it < iters; ++it) butterflies<F, R, 1, true>(x, w);
    T acc = 0;
#pragma unroll
    for (int m = 0; m < M; ++m) acc ^= ntt_store_form(x[m]);
    out[tid] = acc;
}

}  // namespace

template <class F>
int butterfly_probe_device(typename F::T* out, const typename F::T* tw, uint32_t blocks, uint32_t iters,
                           hipStream_t st) {
    hipLaunchKernelGGL((butterfly_probe_kernel<F>), dim3(blocks), dim3(256), 0, st, out, tw, iters);
    return int(hipGetLastError());
}

}  // namespace p3h
