// Column-batched NTT / coset LDE kernels for gfx950.
//
// The matrix is row-major (height x width, the reference's RowMajorMatrix),
// so every column is an independent transform whose elements sit `width`
// apart.  A pass covers a contiguous range of index bits [a, a+s): each
// workgroup stages 2^s x G elements in LDS (G = sub-transforms = low indices
// x columns), runs the s radix-2 stages there, and writes back.  Global loads
// and stores walk the tile column-fastest, so each (row, column-tile) segment is
// one contiguous burst.
//
// Coset LDE (reference: Radix2Dit::coset_lde_batch, radix2_dit.hpp:221-233)
// is computed as
//   1. inverse transform by decimation in frequency (natural order in,
//      bit-reversed coefficients out): outer passes over the high bits, then
//   2. one fused inner pass per block of 2^s rows: the last DIF stages, then for
//      each coset r of the 2^k cosets of the small domain inside the large one
//      the twist a_j * (shift * omega_N^r)^j / n and the first decimation-in-
//      time stages of a forward transform (bit-reversed in, natural out),
//   3. outer DIT passes over the high bits, per coset.
// Coset r's natural index i is LDE row i * 2^k + r.  The device buffer keeps
// the cosets one after another ("coset-major": stored row r * n + i); readers
// (leaf hashing, openings, downloads) map LDE row j to stored row
// (j mod 2^k) * n + j / 2^k.  Zero padding is never materialised.

#include "field.h"
#include "ntt.h"

namespace p3h {

namespace {

constexpr int NTT_THREADS = 512;
constexpr int TILE_BYTES = 65536;

template <class T>
struct NttArgs {
    const T* src;
    T* dst;
    const T* tw;     // twiddles omega^(+-e), e < n/2, in multiply-constant form
    const T* scale;  // mode-dependent table (see kernel), or nullptr
    uint64_t src_region_rows;  // row offset between regions in src
    uint64_t dst_region_rows;  // row offset between regions in dst
    uint32_t L;                // log2 transform size n
    uint32_t a, s;             // this pass covers index bits [a, a+s)
    uint32_t w;                // matrix width (row stride in elements)
    uint32_t logC;             // columns per tile = 2^logC
    uint32_t logG;             // sub-transforms per tile = 2^logG (>= logC)
    uint32_t n_col_tiles;
    uint32_t n_regions;        // cosets (LDE) or 1
};

enum Mode : int {
    OUTER_DIF_INV = 0,   // DIF stages, bits descending, in place or src->dst
    OUTER_DIT_FWD = 1,   // DIT stages, bits ascending, per region
    INNER_LDE = 2,       // DIF stages + per-coset twist + DIT stages
    INNER_DIT_GATHER = 3,  // load rows bitrev(p) (times scale[row]), DIT stages
    INNER_DIF_SCATTER = 4  // DIF stages, store position p at row bitrev(p) times scale[row]
};

__device__ __forceinline__ uint32_t bitrev(uint32_t x, uint32_t bits) {
    return bits == 0 ? 0u : (__brev(x) >> (32 - bits));
}

template <class F, bool DIT>
__device__ __forceinline__ void stage(typename F::T* lds, const typename F::T* tw, uint32_t L, uint32_t a,
                                      uint32_t q, uint32_t s, uint32_t logG, uint32_t logC, uint32_t lo0) {
    using T = typename F::T;
    const uint32_t G = 1u << logG;
    const uint32_t nb = (1u << (s - 1)) << logG;  // butterflies in the tile
    const uint32_t qmask = (1u << q) - 1;
    const uint32_t sh = L - (a + q) - 1;           // omega_{2^(a+q+1)} = omega_n^(2^sh)
    for (uint32_t u = threadIdx.x; u < nb; u += NTT_THREADS) {
        uint32_t g = u & (G - 1);
        uint32_t v = u >> logG;
        uint32_t t0 = ((v >> q) << (q + 1)) | (v & qmask);
        uint32_t t1 = t0 | (1u << q);
        uint32_t lo = lo0 + (g >> logC);
        uint32_t e = (((t0 & qmask) << a) | lo) << sh;
        T wv = tw[e];
        T x = lds[(t0 << logG) | g];
        T y = lds[(t1 << logG) | g];
        if (DIT) {
            T yw = F::mul_c(y, wv);
            lds[(t0 << logG) | g] = F::add(x, yw);
            lds[(t1 << logG) | g] = F::sub(x, yw);
        } else {
            lds[(t0 << logG) | g] = F::add(x, y);
            lds[(t1 << logG) | g] = F::mul_c(F::sub(x, y), wv);
        }
    }
}

template <class F, int MODE>
__global__ __launch_bounds__(NTT_THREADS) void ntt_pass(NttArgs<typename F::T> A) {
    using T = typename F::T;
    constexpr uint32_t TILE = TILE_BYTES / sizeof(T);
    __shared__ T lds[TILE];

    const uint32_t s = A.s, a = A.a, logG = A.logG, logC = A.logC;
    const uint32_t G = 1u << logG, C = 1u << logC;
    const uint32_t logNLow = logG - logC;
    const uint32_t B = 1u << s;
    const uint32_t elems = B << logG;

    // block -> (column tile, low tile, high index, region)
    uint32_t bid = blockIdx.x;
    const uint32_t ct = bid % A.n_col_tiles;
    bid /= A.n_col_tiles;
    const uint32_t low_tiles_log = a - logNLow;  // 2^a / NLow
    const uint32_t lt = bid & ((1u << low_tiles_log) - 1);
    bid >>= low_tiles_log;
    const uint32_t hi_bits = A.L - a - s;
    const uint32_t hi = bid & ((1u << hi_bits) - 1);
    const uint32_t region = bid >> hi_bits;
    const uint32_t lo0 = lt << logNLow;
    const uint32_t c0 = ct << logC;
    const uint64_t n = 1ull << A.L;

    // position (within the length-n transform) of tile element (t, g)
    auto pos = [&](uint32_t t, uint32_t g) -> uint64_t {
        return (uint64_t(hi) << (a + s)) | (uint64_t(t) << a) | (lo0 + (g >> logC));
    };

    // ---- load
    const T* src = A.src + uint64_t(region) * A.src_region_rows * A.w;
    for (uint32_t idx = threadIdx.x; idx < elems; idx += NTT_THREADS) {
        uint32_t t = idx >> logG, g = idx & (G - 1);
        uint32_t col = c0 + (g & (C - 1));
        T v = 0;
        if (col < A.w) {
            uint64_t p = pos(t, g);
            if (MODE == INNER_DIT_GATHER) {
                uint64_t row = bitrev(uint32_t(p), A.L);
                v = src[row * A.w + col];
                if (A.scale) v = F::mul_c(v, A.scale[row]);
            } else {
                v = src[p * A.w + col];
            }
        }
        lds[idx] = v;
    }

    if (MODE == OUTER_DIF_INV || MODE == INNER_LDE || MODE == INNER_DIF_SCATTER) {
        for (int q = int(s) - 1; q >= 0; --q) {
            __syncthreads();
            stage<F, false>(lds, A.tw, A.L, a, uint32_t(q), s, logG, logC, lo0);
        }
    }
    if (MODE == OUTER_DIT_FWD || MODE == INNER_DIT_GATHER) {
        for (uint32_t q = 0; q < s; ++q) {
            __syncthreads();
            stage<F, true>(lds, A.tw, A.L, a, q, s, logG, logC, lo0);
        }
    }

    if (MODE == INNER_LDE) {
        // coefficients (bit-reversed order) -> registers, then one forward
        // transform per coset.  A.scale holds, per coset r and position p,
        // (shift * omega_N^r)^bitrev(p) / n in multiply-constant form.
        // A.tw is the inverse table; the forward table follows it at + n/2.
        constexpr uint32_t PER = TILE / NTT_THREADS;
        T coef[PER];
        __syncthreads();
#pragma unroll
        for (uint32_t k = 0; k < PER; ++k) {
            uint32_t idx = threadIdx.x + k * NTT_THREADS;
            coef[k] = idx < elems ? lds[idx] : T(0);
        }
        const T* fwd = A.tw + (n >> 1);
        for (int r = int(A.n_regions) - 1; r >= 0; --r) {
            __syncthreads();
            const T* tws = A.scale + uint64_t(r) * n;
#pragma unroll
            for (uint32_t k = 0; k < PER; ++k) {
                uint32_t idx = threadIdx.x + k * NTT_THREADS;
                if (idx < elems) {
                    uint32_t t = idx >> logG;
                    lds[idx] = F::mul_c(coef[k], tws[pos(t, idx & (G - 1))]);
                }
            }
            for (uint32_t q = 0; q < s; ++q) {
                __syncthreads();
                stage<F, true>(lds, fwd, A.L, 0, q, s, logG, logC, lo0);
            }
            __syncthreads();
            T* dst = A.dst + uint64_t(r) * A.dst_region_rows * A.w;
            for (uint32_t idx = threadIdx.x; idx < elems; idx += NTT_THREADS) {
                uint32_t t = idx >> logG, g = idx & (G - 1);
                uint32_t col = c0 + (g & (C - 1));
                if (col < A.w) dst[pos(t, g) * A.w + col] = lds[idx];
            }
        }
        return;
    }

    // ---- store
    __syncthreads();
    T* dst = A.dst + uint64_t(region) * A.dst_region_rows * A.w;
    for (uint32_t idx = threadIdx.x; idx < elems; idx += NTT_THREADS) {
        uint32_t t = idx >> logG, g = idx & (G - 1);
        uint32_t col = c0 + (g & (C - 1));
        if (col >= A.w) continue;
        uint64_t p = pos(t, g);
        T v = lds[idx];
        if (MODE == INNER_DIF_SCATTER) {
            uint64_t row = bitrev(uint32_t(p), A.L);
            if (A.scale) v = F::mul_c(v, A.scale[row]);
            dst[row * A.w + col] = v;
        } else {
            dst[p * A.w + col] = v;
        }
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

struct Plan {
    uint32_t s_inner;
    uint32_t outer[4];  // outer pass widths, top (for DIF order) first
    uint32_t n_outer;
    uint32_t logC;
};

template <class F>
Plan make_plan(uint32_t L, uint32_t w, const NttTuning& tune) {
    using T = typename F::T;
    const uint32_t tile_log = 31 - __builtin_clz(uint32_t(TILE_BYTES / sizeof(T)));
    uint32_t logC = 0;
    uint32_t cmax = tune.log_cols ? tune.log_cols : 4;
    while (logC < cmax && (1u << logC) < w) ++logC;
    Plan p{};
    p.logC = logC;
    uint32_t smax = tile_log - logC;  // inner pass: G = C
    if (tune.max_stages && tune.max_stages < smax) smax = tune.max_stages;
    p.s_inner = L < smax ? L : smax;
    uint32_t rest = L - p.s_inner;
    uint32_t outer_max = tile_log - 4;  // outer passes keep at least 16 sub-transforms
    if (tune.max_stages && tune.max_stages < outer_max) outer_max = tune.max_stages;
    uint32_t np = (rest + outer_max - 1) / outer_max;
    p.n_outer = np;
    for (uint32_t i = 0; i < np; ++i) p.outer[i] = rest / np + (i < rest % np ? 1 : 0);
    return p;
}

template <class F, int MODE>
void launch_pass(NttArgs<typename F::T> A, hipStream_t st) {
    using T = typename F::T;
    const uint32_t tile_log = 31 - __builtin_clz(uint32_t(TILE_BYTES / sizeof(T)));
    uint32_t logG = tile_log - A.s;
    uint32_t max_logG = A.logC + A.a;  // NLow <= 2^a
    if (logG > max_logG) logG = max_logG;
    if (logG < A.logC) logG = A.logC;
    A.logG = logG;
    A.n_col_tiles = (A.w + (1u << A.logC) - 1) >> A.logC;
    uint64_t blocks = uint64_t(A.n_col_tiles) << (A.L - A.s - (logG - A.logC));
    blocks *= A.n_regions;
    hipLaunchKernelGGL((ntt_pass<F, MODE>), dim3(uint32_t(blocks)), dim3(NTT_THREADS), 0, st, A);
}

}  // namespace

template <class F>
int lde_device(const typename F::T* in, typename F::T* out, uint32_t L, uint32_t w, uint32_t k,
               const typename F::T* tw_inv_fwd, const typename F::T* twist, const NttTuning& tune,
               hipStream_t st) {
    Plan p = make_plan<F>(L, w, tune);
    const uint64_t n = 1ull << L;
    NttArgs<typename F::T> A{};
    A.L = L;
    A.w = w;
    A.logC = p.logC;
    A.n_regions = 1;
    A.tw = tw_inv_fwd;  // inverse half first
    // 1. outer DIF passes over the high bits, top down: input -> region 0 of out
    uint32_t top = L;
    const typename F::T* cur = in;
    for (uint32_t i = 0; i < p.n_outer; ++i) {
        A.s = p.outer[i];
        A.a = top - A.s;
        A.src = cur;
        A.dst = out;
        A.src_region_rows = A.dst_region_rows = 0;
        launch_pass<F, OUTER_DIF_INV>(A, st);
        cur = out;
        top = A.a;
    }
    // 2. inner fused pass: rest of the inverse, then per-coset twist + DIT
    A.a = 0;
    A.s = p.s_inner;
    A.src = cur;
    A.dst = out;
    A.scale = twist;
    A.n_regions = 1u << k;
    A.dst_region_rows = n;
    {
        NttArgs<typename F::T> B = A;
        B.n_regions = 1;  // grid covers blocks only; the kernel loops over cosets
        using T = typename F::T;
        const uint32_t tile_log = 31 - __builtin_clz(uint32_t(TILE_BYTES / sizeof(T)));
        B.logG = B.logC;
        (void)tile_log;
        B.n_col_tiles = (B.w + (1u << B.logC) - 1) >> B.logC;
        uint64_t blocks = uint64_t(B.n_col_tiles) << (L - B.s);
        B.n_regions = 1u << k;
        hipLaunchKernelGGL((ntt_pass<F, INNER_LDE>), dim3(uint32_t(blocks)), dim3(NTT_THREADS), 0, st, B);
    }
    // 3. outer DIT passes per coset, bottom up, in place
    A.scale = nullptr;
    A.tw = tw_inv_fwd + (n >> 1);
    A.src = out;
    A.dst = out;
    A.src_region_rows = A.dst_region_rows = n;
    A.n_regions = 1u << k;
    uint32_t bot = p.s_inner;
    for (int i = int(p.n_outer) - 1; i >= 0; --i) {
        A.a = bot;
        A.s = p.outer[i];
        launch_pass<F, OUTER_DIT_FWD>(A, st);
        bot += A.s;
    }
    return int(hipGetLastError());
}

template <class F>
int dft_device(int kind, const typename F::T* in, typename F::T* out, typename F::T* tmp, uint32_t L,
               uint32_t w, const typename F::T* tw_inv_fwd, const typename F::T* row_scale,
               const NttTuning& tune, hipStream_t st) {
    Plan p = make_plan<F>(L, w, tune);
    const uint64_t n = 1ull << L;
    NttArgs<typename F::T> A{};
    A.L = L;
    A.w = w;
    A.logC = p.logC;
    A.n_regions = 1;
    if (kind == 0 || kind == 2) {
        // forward (optionally coset): gather rows in bit-reversed order with
        // the row scale, DIT inner, DIT outer in place.
        A.tw = tw_inv_fwd + (n >> 1);
        A.a = 0;
        A.s = p.s_inner;
        A.src = in;
        A.dst = out;
        A.scale = row_scale;
        launch_pass<F, INNER_DIT_GATHER>(A, st);
        A.scale = nullptr;
        A.src = out;
        uint32_t bot = p.s_inner;
        for (int i = int(p.n_outer) - 1; i >= 0; --i) {
            A.a = bot;
            A.s = p.outer[i];
            launch_pass<F, OUTER_DIT_FWD>(A, st);
            bot += A.s;
        }
    } else {
        // inverse (optionally coset): DIF outer (in -> tmp), DIF inner with
        // bit-reversed scatter and the row scale (tmp -> out).
        A.tw = tw_inv_fwd;
        uint32_t top = L;
        const typename F::T* cur = in;
        for (uint32_t i = 0; i < p.n_outer; ++i) {
            A.s = p.outer[i];
            A.a = top - A.s;
            A.src = cur;
            A.dst = tmp;
            launch_pass<F, OUTER_DIF_INV>(A, st);
            cur = tmp;
            top = A.a;
        }
        A.a = 0;
        A.s = p.s_inner;
        A.src = cur;
        A.dst = out;
        A.scale = row_scale;
        launch_pass<F, INNER_DIF_SCATTER>(A, st);
    }
    return int(hipGetLastError());
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

template int lde_device<BabyBear>(const uint32_t*, uint32_t*, uint32_t, uint32_t, uint32_t, const uint32_t*,
                                  const uint32_t*, const NttTuning&, hipStream_t);
template int lde_device<Goldilocks>(const uint64_t*, uint64_t*, uint32_t, uint32_t, uint32_t, const uint64_t*,
                                    const uint64_t*, const NttTuning&, hipStream_t);
template int dft_device<BabyBear>(int, const uint32_t*, uint32_t*, uint32_t*, uint32_t, uint32_t,
                                  const uint32_t*, const uint32_t*, const NttTuning&, hipStream_t);
template int dft_device<Goldilocks>(int, const uint64_t*, uint64_t*, uint64_t*, uint32_t, uint32_t,
                                    const uint64_t*, const uint64_t*, const NttTuning&, hipStream_t);
template int powers_device<BabyBear>(uint32_t*, uint64_t, uint32_t, uint32_t, uint32_t, hipStream_t);
template int powers_device<Goldilocks>(uint64_t*, uint64_t, uint64_t, uint64_t, uint64_t, hipStream_t);
template int lde_twist_device<BabyBear>(uint32_t*, uint32_t, uint32_t, const uint32_t*, uint32_t, uint32_t,
                                        hipStream_t);
template int lde_twist_device<Goldilocks>(uint64_t*, uint32_t, uint32_t, const uint64_t*, uint64_t, uint64_t,
                                          hipStream_t);
template int coset_to_natural_device<uint32_t>(const uint32_t*, uint32_t*, uint32_t, uint32_t, uint32_t,
                                               hipStream_t);
template int coset_to_natural_device<uint64_t>(const uint64_t*, uint64_t*, uint32_t, uint32_t, uint32_t,
                                               hipStream_t);

}  // namespace p3h
