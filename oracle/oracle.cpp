// CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.h).  A restatement of the
// reference's CPU algorithms in plain C++ with canonical-integer arithmetic.
// Each function cites the reference file:line it follows (paths relative to
// the reference's src/include/).  Written for clarity, single-threaded, exact.
//
// Parity pins (which check catches which mistake) are listed in the header of
// tests/test_oracle.py; every function here is pinned by at least one of:
// the compiled reference (oracle/_ref), committed golden fixtures, the
// reference's own known answers, or closed forms / brute force.

#include "oracle.h"

#include <algorithm>
#include <map>
#include <array>
#include <cstring>
#include <numeric>
#include <random>
#include <stdexcept>
#include <thread>
#include <vector>

namespace {

// ---------------------------------------------------------------- fields ----

// BabyBear, canonical u32 (baby_bear.hpp:23-88): add/sub/mul via u64 and % p.
struct BB {
    using Raw = uint32_t;
    static constexpr uint32_t P = 0x78000001u;
    static uint32_t add(uint32_t a, uint32_t b) { return uint32_t((uint64_t(a) + b) % P); }
    static uint32_t sub(uint32_t a, uint32_t b) { return uint32_t((uint64_t(a) + P - b) % P); }
    static uint32_t mul(uint32_t a, uint32_t b) { return uint32_t((uint64_t(a) * b) % P); }
    static uint32_t from_u64(uint64_t v) { return uint32_t(v % P); }
    // two_adic_generator (baby_bear.hpp:131-145): 440564289^(2^(27-bits)).
    static uint32_t root(size_t bits);
};

// Goldilocks (goldilocks.hpp:21-189); the reference keeps non-canonical
// representatives internally but its add/sub/reduce128 are exact mod p, so a
// canonical u128 implementation yields the same canonical outputs.
struct GL {
    using Raw = uint64_t;
    static constexpr uint64_t P = 0xFFFFFFFF00000001ull;
    static uint64_t add(uint64_t a, uint64_t b) {
        return uint64_t((static_cast<unsigned __int128>(a) + b) % P);
    }
    static uint64_t sub(uint64_t a, uint64_t b) {
        return uint64_t((static_cast<unsigned __int128>(a) + P - b) % P);
    }
    // reduce128 (goldilocks.hpp:135-162): with x = lo + 2^64 hl + 2^96 hh and
    // 2^64 = 2^32 - 1, 2^96 = -1 (mod p), x = lo - hh + hl (2^32 - 1); each
    // wrap of the 64-bit word is worth 2^32 - 1; the result is made canonical.
    static uint64_t mul(uint64_t a, uint64_t b) {
        constexpr uint64_t E = 0xFFFFFFFFull;  // 2^64 mod p
        const unsigned __int128 x = static_cast<unsigned __int128>(a) * b;
        const uint64_t lo = uint64_t(x), hi = uint64_t(x >> 64);
        const uint64_t hh = hi >> 32, hl = hi & E;
        uint64_t t0 = lo - hh;
        if (lo < hh) t0 -= E;
        const uint64_t t1 = hl * E;
        uint64_t t2 = t0 + t1;
        if (t2 < t1) t2 += E;
        return t2 >= P ? t2 - P : t2;
    }
    static uint64_t from_u64(uint64_t v) { return v % P; }
    // Not defined by the reference: generator 7, omega_{2^32} = 7^((p-1)/2^32).
    static uint64_t root(size_t bits);
};

template <class F>
typename F::Raw powf(typename F::Raw a, uint64_t e) {
    typename F::Raw r = 1;
    while (e) {
        if (e & 1) r = F::mul(r, a);
        a = F::mul(a, a);
        e >>= 1;
    }
    return r;
}
template <class F>
typename F::Raw invf(typename F::Raw a) {
    if (a == 0) throw std::runtime_error("inverse of zero");
    return powf<F>(a, F::P - 2);
}

uint32_t BB::root(size_t bits) {
    if (bits > 27) throw std::invalid_argument("bits > 27");
    return powf<BB>(440564289u, uint64_t(1) << (27 - bits));
}
uint64_t GL::root(size_t bits) {
    if (bits > 32) throw std::invalid_argument("bits > 32");
    uint64_t r32 = powf<GL>(7, (GL::P - 1) >> 32);
    return powf<GL>(r32, uint64_t(1) << (32 - bits));
}

// BabyBear^4 = F[a]/(a^4 - 11), schoolbook (extension_field.hpp:106-122).
using E4 = std::array<uint32_t, 4>;
E4 e4_add(const E4& a, const E4& b) {
    E4 r;
    for (int i = 0; i < 4; ++i) r[i] = BB::add(a[i], b[i]);
    return r;
}
E4 e4_sub(const E4& a, const E4& b) {
    E4 r;
    for (int i = 0; i < 4; ++i) r[i] = BB::sub(a[i], b[i]);
    return r;
}
E4 e4_mul(const E4& a, const E4& b) {
    E4 r{0, 0, 0, 0};
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            uint32_t p = BB::mul(a[i], b[j]);
            if (i + j < 4) r[i + j] = BB::add(r[i + j], p);
            else r[i + j - 4] = BB::add(r[i + j - 4], BB::mul(p, 11));
        }
    return r;
}
E4 e4_scale(const E4& a, uint32_t s) {
    E4 r;
    for (int i = 0; i < 4; ++i) r[i] = BB::mul(a[i], s);
    return r;
}
E4 e4_base(uint32_t s) { return E4{s, 0, 0, 0}; }
// Inverse by exponentiation a^(p^4 - 2) would be slow; the Lagrange fold only
// inverts (beta - t_i) * base, done here through the norm map:
// a^-1 = a^(p + p^2 + p^3) / N(a), N(a) = a^(1 + p + p^2 + p^3) in F_p.
E4 e4_pow(E4 a, unsigned __int128 e) {
    E4 r{1, 0, 0, 0};
    while (e) {
        if (e & 1) r = e4_mul(r, a);
        a = e4_mul(a, a);
        e >>= 1;
    }
    return r;
}
E4 e4_inv(const E4& a) {
    if (a[0] == 0 && a[1] == 0 && a[2] == 0 && a[3] == 0) throw std::runtime_error("inverse of zero");
    unsigned __int128 p = BB::P;
    E4 t = e4_pow(a, p + p * p + p * p * p);  // product of the three conjugates
    E4 n = e4_mul(t, a);                       // the norm, lies in F_p
    return e4_scale(t, invf<BB>(n[0]));
}

size_t log2_strict(size_t n) {
    if (n == 0 || (n & (n - 1))) throw std::invalid_argument("not a power of two");
    return size_t(__builtin_ctzll(n));
}
size_t rev_bits(size_t x, size_t bits) {
    size_t r = 0;
    for (size_t i = 0; i < bits; ++i) r = (r << 1) | ((x >> i) & 1);
    return r;
}

// -------------------------------------------------------------- Poseidon2 ---
// Poseidon2::permute_mut (poseidon2.hpp:108-112) with GenericPoseidon2External
// /Internal (poseidon2.hpp:136-183): MDS-light first (external.hpp:258-275),
// then rf/2 rounds of [add rc to all lanes, x^7 on all lanes, MDS-light],
// rp rounds of [add rc to lane 0, x^7 on lane 0, internal layer], then rf/2
// full rounds again.  MDS-light (external.hpp:121-162): the 4x4 block
// [[2,3,1,1],[1,2,3,1],[1,1,2,3],[3,1,1,2]] on each group of 4 lanes, then add to
// lane i the sum of all lanes congruent to i mod 4.  Internal layer
// (generic.hpp:58-142): s_i <- 2^i * s_i + sum(s).  S-box: x^7 (field.hpp:119-123).
template <class F, size_t W>
struct Perm {
    using R = typename F::Raw;
    std::vector<std::array<R, W>> init, term;
    std::vector<R> internal;

    Perm(const R* c, int rf, int rp) {
        if (rf < 0 || rf % 2 || rp < 0) throw std::invalid_argument("round counts");
        size_t k = 0;
        init.resize(rf / 2);
        term.resize(rf / 2);
        for (auto& r : init)
            for (auto& x : r) x = F::from_u64(c[k++]);
        for (auto& r : term)
            for (auto& x : r) x = F::from_u64(c[k++]);
        internal.resize(rp);
        for (auto& x : internal) x = F::from_u64(c[k++]);
    }
    static R sbox(R x) {
        R x2 = F::mul(x, x);
        R x3 = F::mul(x2, x);
        R x4 = F::mul(x2, x2);
        return F::mul(x3, x4);
    }
    static void mat4(R* x) {
        static const int M[4][4] = {{2, 3, 1, 1}, {1, 2, 3, 1}, {1, 1, 2, 3}, {3, 1, 1, 2}};
        R y[4];
        for (int i = 0; i < 4; ++i) {
            R acc = 0;
            for (int j = 0; j < 4; ++j) acc = F::add(acc, F::mul(R(M[i][j]), x[j]));
            y[i] = acc;
        }
        for (int i = 0; i < 4; ++i) x[i] = y[i];
    }
    static void mds_light(std::array<R, W>& s) {
        for (size_t i = 0; i < W; i += 4) mat4(&s[i]);
        R sums[4] = {0, 0, 0, 0};
        for (size_t i = 0; i < W; ++i) sums[i % 4] = F::add(sums[i % 4], s[i]);
        for (size_t i = 0; i < W; ++i) s[i] = F::add(s[i], sums[i % 4]);
    }
    static void internal_layer(std::array<R, W>& s) {
        R sum = 0;
        for (size_t i = 0; i < W; ++i) sum = F::add(sum, s[i]);
        for (size_t i = 0; i < W; ++i) s[i] = F::add(F::mul(s[i], powf<F>(2, i)), sum);
    }
    void full_round(std::array<R, W>& s, const std::array<R, W>& rc) const {
        for (size_t i = 0; i < W; ++i) s[i] = sbox(F::add(s[i], rc[i]));
        mds_light(s);
    }
    void permute(std::array<R, W>& s) const {
        mds_light(s);
        for (const auto& rc : init) full_round(s, rc);
        for (R rc : internal) {
            s[0] = sbox(F::add(s[0], rc));
            internal_layer(s);
        }
        for (const auto& rc : term) full_round(s, rc);
    }
};

template <class F>
struct Hcfg;
template <>
struct Hcfg<BB> {
    static constexpr size_t W = 16, RATE = 8, DIG = 8;
};
template <>
struct Hcfg<GL> {
    static constexpr size_t W = 8, RATE = 4, DIG = 4;
};

template <class F>
using Digest = std::array<typename F::Raw, Hcfg<F>::DIG>;

// PaddingFreeSponge::hash_iter (padding_free_sponge.hpp:63-87): zero state;
// each chunk of up to RATE inputs OVERWRITES the first slots; a full chunk is
// followed by a permutation; a partial last chunk keeps the older values in
// the slots it does not reach and is permuted iff it is non-empty; the digest
// is the first DIG lanes.
template <class F>
Digest<F> sponge(const Perm<F, Hcfg<F>::W>& perm, const typename F::Raw* d, size_t len) {
    constexpr size_t W = Hcfg<F>::W, RATE = Hcfg<F>::RATE, DIG = Hcfg<F>::DIG;
    std::array<typename F::Raw, W> s{};
    size_t pos = 0;
    while (pos < len) {
        size_t take = std::min(RATE, len - pos);
        for (size_t i = 0; i < take; ++i) s[i] = d[pos + i];
        pos += take;
        perm.permute(s);
    }
    Digest<F> out;
    for (size_t i = 0; i < DIG; ++i) out[i] = s[i];
    return out;
}

// TruncatedPermutation::compress (truncated_permutation.hpp:61-76).
template <class F>
Digest<F> compress(const Perm<F, Hcfg<F>::W>& perm, const Digest<F>& l, const Digest<F>& r) {
    constexpr size_t W = Hcfg<F>::W, DIG = Hcfg<F>::DIG;
    std::array<typename F::Raw, W> s{};
    for (size_t i = 0; i < DIG; ++i) {
        s[i] = l[i];
        s[DIG + i] = r[i];
    }
    perm.permute(s);
    Digest<F> out;
    for (size_t i = 0; i < DIG; ++i) out[i] = s[i];
    return out;
}

// -------------------------------------------------------------------- DFT ---
// Radix2Dit (radix2_dit.hpp:82-233): dft_batch bit-reverses the rows then runs
// DIT layers l = 0..log_h-1 with twiddle omega_h^(j * (h/2 >> l)); idft_batch
// = dft_batch, reverse rows 1..h-1, scale by 2^-log_h; coset_dft_batch scales
// row i by shift^i then dft_batch; coset_idft_batch = idft then row i by
// shift^-i; coset_lde_batch = idft, zero-pad to h << added, coset_dft.
template <class F>
struct Mat {
    using R = typename F::Raw;
    std::vector<R> v;
    size_t h, w;
    R* row(size_t i) { return v.data() + i * w; }
};

template <class F>
void dft(Mat<F>& m) {
    using R = typename F::Raw;
    size_t h = m.h, w = m.w;
    if (h <= 1) return;
    size_t lh = log2_strict(h);
    for (size_t i = 0; i < h; ++i) {
        size_t j = rev_bits(i, lh);
        if (i < j)
            for (size_t c = 0; c < w; ++c) std::swap(m.v[i * w + c], m.v[j * w + c]);
    }
    R root = F::root(lh);
    std::vector<R> tw(h / 2);
    R cur = 1;
    for (size_t i = 0; i < h / 2; ++i) {
        tw[i] = cur;
        cur = F::mul(cur, root);
    }
    for (size_t l = 0; l < lh; ++l) {
        size_t half = size_t(1) << l, stride = (h / 2) >> l;
        for (size_t blk = 0; blk < h; blk += 2 * half)
            for (size_t j = 0; j < half; ++j) {
                R t = tw[j * stride];
                R* a = m.row(blk + j);
                R* b = m.row(blk + half + j);
                for (size_t c = 0; c < w; ++c) {
                    R bt = F::mul(b[c], t);
                    R ac = a[c];
                    a[c] = F::add(ac, bt);
                    b[c] = F::sub(ac, bt);
                }
            }
    }
}

template <class F>
void idft(Mat<F>& m) {
    size_t h = m.h, w = m.w;
    if (h <= 1) return;
    dft(m);
    for (size_t lo = 1, hi = h - 1; lo < hi; ++lo, --hi)
        for (size_t c = 0; c < w; ++c) std::swap(m.v[lo * w + c], m.v[hi * w + c]);
    auto inv_n = powf<F>(invf<F>(2), log2_strict(h));
    for (auto& x : m.v) x = F::mul(x, inv_n);
}

template <class F>
void scale_rows_by_powers(Mat<F>& m, typename F::Raw s) {
    typename F::Raw p = 1;
    for (size_t i = 0; i < m.h; ++i) {
        for (size_t c = 0; c < m.w; ++c) m.v[i * m.w + c] = F::mul(m.v[i * m.w + c], p);
        p = F::mul(p, s);
    }
}

template <class F>
Mat<F> load(const typename F::Raw* in, size_t h, size_t w) {
    Mat<F> m{std::vector<typename F::Raw>(in, in + h * w), h, w};
    for (auto& x : m.v) x = F::from_u64(x);
    return m;
}

template <class F>
int dft_family(int kind, const typename F::Raw* in, size_t log_h, size_t w, typename F::Raw shift,
               typename F::Raw* out) {
    if (w == 0 || kind < 0 || kind > 3) return -1;
    Mat<F> m = load<F>(in, size_t(1) << log_h, w);
    if (kind == 2) scale_rows_by_powers(m, shift);
    if (kind == 0 || kind == 2) dft(m);
    if (kind == 1 || kind == 3) idft(m);
    if (kind == 3) scale_rows_by_powers(m, invf<F>(shift));
    std::copy(m.v.begin(), m.v.end(), out);
    return 0;
}

template <class F>
Mat<F> lde(const typename F::Raw* in, size_t log_h, size_t w, size_t added, typename F::Raw shift) {
    Mat<F> m = load<F>(in, size_t(1) << log_h, w);
    idft(m);
    m.h <<= added;
    m.v.resize(m.h * w, 0);
    scale_rows_by_powers(m, shift);
    dft(m);
    return m;
}

// Direct evaluation: coefficients a_j = (1/n) sum_i x_i omega_n^(-ij), then
// lde[k] = sum_j a_j (shift * omega_N^k)^j.  Independent of the FFT structure.
template <class F>
void lde_brute(const typename F::Raw* in, size_t log_h, size_t w, size_t added,
               typename F::Raw shift, typename F::Raw* out) {
    using R = typename F::Raw;
    size_t n = size_t(1) << log_h, N = n << added;
    R wn_inv = invf<F>(F::root(log_h)), wN = F::root(log_h + added);
    R n_inv = invf<F>(F::from_u64(n));
    std::vector<R> a(n);
    for (size_t c = 0; c < w; ++c) {
        for (size_t j = 0; j < n; ++j) {
            R acc = 0, step = powf<F>(wn_inv, j), pw = 1;
            for (size_t i = 0; i < n; ++i) {
                acc = F::add(acc, F::mul(in[i * w + c] % F::P, pw));
                pw = F::mul(pw, step);
            }
            a[j] = F::mul(acc, n_inv);
        }
        for (size_t k = 0; k < N; ++k) {
            R x = F::mul(shift, powf<F>(wN, k)), acc = 0, pw = 1;
            for (size_t j = 0; j < n; ++j) {
                acc = F::add(acc, F::mul(a[j], pw));
                pw = F::mul(pw, x);
            }
            out[k * w + c] = acc;
        }
    }
}

// Runs body(lo, hi) over [0, n) split into `threads` contiguous ranges.  The
// threaded entry points below use it only for work whose parts are
// independent (columns of a transform, leaves, nodes of one tree level, rows of
// a fold), so their results equal the single-threaded ones.
template <class Body>
void parallel_for(size_t n, int threads, Body body) {
    const size_t t = std::max<size_t>(1, std::min<size_t>(size_t(threads), n));
    if (t == 1) {
        body(size_t(0), n);
        return;
    }
    std::vector<std::thread> pool;
    for (size_t i = 0; i < t; ++i) pool.emplace_back(body, n * i / t, n * (i + 1) / t);
    for (auto& th : pool) th.join();
}

// ------------------------------------------------------------------- MMCS ---
// p3_merkle::MerkleTreeMmcs::commit_matrix (fri_merkle_tree_mmcs.hpp:55-102):
// leaf i = sponge(row i); each layer pairs (2j, 2j+1), a missing right child of
// an odd-sized layer is the all-zero digest; stop at one node.
template <class F>
std::vector<std::vector<Digest<F>>> flat_tree(const Perm<F, Hcfg<F>::W>& perm,
                                              const typename F::Raw* d, size_t h, size_t w, int threads = 1) {
    if (w == 0 || h == 0) throw std::invalid_argument("empty matrix");
    std::vector<std::vector<Digest<F>>> tree(1);
    tree[0].resize(h);
    parallel_for(h, threads, [&](size_t i0, size_t i1) {
        for (size_t i = i0; i < i1; ++i) tree[0][i] = sponge<F>(perm, d + i * w, w);
    });
    while (tree.back().size() > 1) {
        const auto& prev = tree.back();
        size_t nxt = (prev.size() + 1) / 2;
        std::vector<Digest<F>> layer(nxt);
        parallel_for(nxt, nxt >= 4096 ? threads : 1, [&](size_t j0, size_t j1) {
            for (size_t j = j0; j < j1; ++j) {
                Digest<F> r{};
                if (2 * j + 1 < prev.size()) r = prev[2 * j + 1];
                layer[j] = compress<F>(perm, prev[2 * j], r);
            }
        });
        tree.push_back(std::move(layer));
    }
    return tree;
}

// p3_merkle_tree::build_merkle_tree (merkle_tree.hpp:82-185): stable sort by
// height (descending); leaf i = sponge over the concatenation of row i of every
// tallest matrix (original order); per level, compress pairs, then for the
// matrices whose height equals the new level size, node = compress(node,
// sponge(concat rows)); stop at 2^cap_height nodes.
template <class F>
struct MultiTree {
    std::vector<std::vector<Digest<F>>> layers;
};

template <class F>
MultiTree<F> multi_tree(const Perm<F, Hcfg<F>::W>& perm, size_t n, const size_t* hs, const size_t* ws,
                        const typename F::Raw* const* mats, size_t cap_height, int threads = 1) {
    using R = typename F::Raw;
    if (n == 0) throw std::invalid_argument("no matrices");
    std::vector<size_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return hs[a] > hs[b]; });
    size_t maxh = hs[order[0]];
    log2_strict(maxh);
    size_t next = 0;
    auto take = [&](size_t h) {
        std::vector<size_t> idx;
        while (next < n && hs[order[next]] == h) idx.push_back(order[next++]);
        return idx;
    };
    auto hash_rows = [&](size_t row, const std::vector<size_t>& idx) {
        std::vector<R> buf;
        for (size_t m : idx) buf.insert(buf.end(), mats[m] + row * ws[m], mats[m] + (row + 1) * ws[m]);
        return sponge<F>(perm, buf.data(), buf.size());
    };
    MultiTree<F> t;
    auto top = take(maxh);
    t.layers.emplace_back(maxh);
    parallel_for(maxh, threads, [&](size_t i0, size_t i1) {
        for (size_t i = i0; i < i1; ++i) t.layers[0][i] = hash_rows(i, top);
    });
    size_t level = maxh, cap = size_t(1) << cap_height;
    while (level > cap) {
        level /= 2;
        std::vector<Digest<F>> nl(level);
        const auto& prev = t.layers.back();
        auto inj = take(level);
        parallel_for(level, level >= 1024 ? threads : 1, [&](size_t i0, size_t i1) {
            for (size_t i = i0; i < i1; ++i) {
                nl[i] = compress<F>(perm, prev[2 * i], prev[2 * i + 1]);
                if (!inj.empty()) nl[i] = compress<F>(perm, nl[i], hash_rows(i, inj));
            }
        });
        t.layers.push_back(std::move(nl));
    }
    return t;
}

// MerkleTreeMmcs::open_batch (merkle_tree_mmcs.hpp:120-152): row index /
// (max_height / height) of every matrix; siblings from the leaf layer up to
// (not including) the cap.
template <class F>
int mmcs_entry(const typename F::Raw* consts, int rf, int rp, size_t n, const size_t* hs,
               const size_t* ws, const typename F::Raw* const* mats, size_t cap_height,
               typename F::Raw* cap_out, long long open_index, typename F::Raw* values_out,
               typename F::Raw* proof_out, size_t* proof_len) {
    constexpr size_t DIG = Hcfg<F>::DIG;
    Perm<F, Hcfg<F>::W> perm(consts, rf, rp);
    auto t = multi_tree<F>(perm, n, hs, ws, mats, cap_height);
    size_t k = 0;
    for (const auto& d : t.layers.back())
        for (size_t i = 0; i < DIG; ++i) cap_out[k++] = d[i];
    if (open_index >= 0) {
        size_t maxh = t.layers[0].size(), idx = size_t(open_index);
        if (idx >= maxh) return -2;
        size_t v = 0;
        for (size_t m = 0; m < n; ++m) {
            size_t row = idx / (maxh / hs[m]);
            for (size_t c = 0; c < ws[m]; ++c) values_out[v++] = mats[m][row * ws[m] + c];
        }
        size_t steps = t.layers.size() - 1, p = 0, cur = idx;
        for (size_t s = 0; s < steps; ++s) {
            const auto& d = t.layers[s][cur ^ 1];
            for (size_t i = 0; i < DIG; ++i) proof_out[p++] = d[i];
            cur >>= 1;
        }
        *proof_len = steps;
    }
    return 0;
}

// MerkleTreeMmcs::verify_batch (merkle_tree_mmcs.hpp:168-254).
template <class F>
int mmcs_verify(const typename F::Raw* consts, int rf, int rp, size_t n, const size_t* hs,
                const size_t* ws, size_t cap_height, const typename F::Raw* cap, size_t index,
                const typename F::Raw* values, const typename F::Raw* proof, size_t proof_len) {
    using R = typename F::Raw;
    constexpr size_t DIG = Hcfg<F>::DIG;
    if (n == 0) return 0;
    Perm<F, Hcfg<F>::W> perm(consts, rf, rp);
    std::vector<size_t> off(n + 1, 0);
    for (size_t m = 0; m < n; ++m) off[m + 1] = off[m] + ws[m];
    std::vector<size_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return hs[a] > hs[b]; });
    size_t maxh = hs[order[0]], cap_size = size_t(1) << cap_height, levels = 0;
    if (maxh == 0) return 0;
    for (size_t h = maxh; h > cap_size; h /= 2) ++levels;
    if (proof_len != levels) return 0;
    size_t next = 0;
    auto take = [&](size_t h) {
        std::vector<size_t> idx;
        while (next < n && hs[order[next]] == h) idx.push_back(order[next++]);
        return idx;
    };
    auto hash_vals = [&](const std::vector<size_t>& idx) {
        std::vector<R> buf;
        for (size_t m : idx) buf.insert(buf.end(), values + off[m], values + off[m + 1]);
        return sponge<F>(perm, buf.data(), buf.size());
    };
    Digest<F> cur = hash_vals(take(maxh));
    size_t ci = index;
    for (size_t s = 0; s < levels; ++s) {
        Digest<F> sib;
        for (size_t i = 0; i < DIG; ++i) sib[i] = proof[s * DIG + i];
        cur = (ci & 1) == 0 ? compress<F>(perm, cur, sib) : compress<F>(perm, sib, cur);
        ci >>= 1;
        auto inj = take(maxh >> (s + 1));
        if (!inj.empty()) cur = compress<F>(perm, cur, hash_vals(inj));
    }
    if (ci >= cap_size) return 0;
    for (size_t i = 0; i < DIG; ++i)
        if (cur[i] != cap[ci * DIG + i]) return 0;
    return 1;
}

// -------------------------------------------------------------------- FRI ---
// TwoAdicFriFolding::fold_row (fri_folding.hpp:74-189): row i of the folded
// domain holds evaluations at t_j = x * w^bitrev(j) with
// x = omega_{2^log_height}^bitrev(i, log_height - log_arity) and w the primitive
// arity-th root.  The reference combines them with the weights
//   L_j = prod_k (beta - t_k) / ((beta - t_j) * x^(arity-1) * w^(-j) * arity)
// (fri_folding.hpp:140-186), with the fast path beta == t_j returning evals[j].
// For arity 2 these are exactly the Lagrange weights of the interpolant
// through (x, e0), (-x, e1); for arity 4 and 8 the reference pairs w^(-j) with
// the bit-reversed node order, and the oracle keeps that formula as written.
E4 fold_row(size_t i, size_t log_height, size_t log_arity, const E4& beta, const E4* ev) {
    size_t arity = size_t(1) << log_arity;
    uint32_t x = powf<BB>(BB::root(log_height), rev_bits(i, log_height - log_arity));
    uint32_t w = BB::root(log_arity);
    std::vector<uint32_t> t(arity);
    for (size_t j = 0; j < arity; ++j) t[j] = BB::mul(x, powf<BB>(w, rev_bits(j, log_arity)));
    for (size_t j = 0; j < arity; ++j)
        if (beta == e4_base(t[j])) return ev[j];
    E4 full{1, 0, 0, 0};
    for (size_t k = 0; k < arity; ++k) full = e4_mul(full, e4_sub(beta, e4_base(t[k])));
    uint32_t arity_inv = powf<BB>(invf<BB>(2), log_arity);
    uint32_t x_am1 = powf<BB>(x, arity - 1), w_inv = powf<BB>(w, arity - 1);
    E4 res{0, 0, 0, 0};
    for (size_t j = 0; j < arity; ++j) {
        uint32_t base = BB::mul(x_am1, powf<BB>(w_inv, j));
        E4 den = e4_scale(e4_sub(beta, e4_base(t[j])), base);
        E4 lj = e4_mul(e4_scale(full, arity_inv), e4_inv(den));
        res = e4_add(res, e4_mul(ev[j], lj));
    }
    return res;
}

std::vector<E4> fold(size_t log_height, size_t log_arity, const E4& beta, const std::vector<E4>& cur,
                     int threads = 1) {
    size_t arity = size_t(1) << log_arity;
    if (cur.size() % arity) throw std::invalid_argument("fold: size");
    std::vector<E4> out(cur.size() / arity);
    parallel_for(out.size(), out.size() >= 1024 ? threads : 1, [&](size_t i0, size_t i1) {
        for (size_t i = i0; i < i1; ++i) out[i] = fold_row(i, log_height, log_arity, beta, &cur[i * arity]);
    });
    return out;
}

std::vector<E4> load_e4(const uint32_t* p, size_t n) {
    std::vector<E4> v(n);
    for (size_t i = 0; i < n; ++i)
        for (int k = 0; k < 4; ++k) v[i][k] = p[4 * i + k] % BB::P;
    return v;
}

// DuplexChallenger<BabyBear, Poseidon2 width 16, 16, 8> (duplex_challenger.hpp:42-265):
// observe pushes onto the input buffer (a full buffer of 8 triggers a
// duplexing); sample duplexes when inputs are pending or no output is left and
// pops the last output; a duplexing overwrites state[0..len) with the inputs,
// permutes and exposes state[0..8) as outputs.
struct Challenger {
    const Perm<BB, 16>* perm;
    std::array<uint32_t, 16> st{};
    std::vector<uint32_t> in, out;

    void duplex() {
        std::copy(in.begin(), in.end(), st.begin());
        in.clear();
        perm->permute(st);
        out.assign(st.begin(), st.begin() + 8);
    }
    void absorb(uint32_t v) {  // absorb_without_invalidation (:51-56)
        in.push_back(v);
        if (in.size() == 8) duplex();
    }
    void observe(uint32_t v) {  // :93-97
        out.clear();
        absorb(v % BB::P);
    }
    uint32_t sample() {  // :120-129
        if (!in.empty() || out.empty()) duplex();
        uint32_t v = out.back();
        out.pop_back();
        return v;
    }
    uint64_t sample_bits(size_t bits) {  // :137-148, 31 bits per sample
        uint64_t r = 0;
        size_t rem = bits;
        while (rem > 0) {
            uint32_t v = sample();
            size_t chunk = std::min<size_t>(rem, 31);
            r |= uint64_t(v & ((1u << chunk) - 1u)) << (bits - rem);
            rem -= chunk;
        }
        return r;
    }
    void observe_witness(uint64_t w) {  // :58-62
        out.clear();
        absorb(uint32_t(w & 0xFFFFFFFFu) % BB::P);
        absorb(uint32_t(w >> 32) % BB::P);
    }
    uint64_t grind(size_t bits) {  // :236-254: the smallest witness whose sample has `bits` zero low bits
        const uint64_t mask = (uint64_t(1) << bits) - 1;
        for (uint64_t w = 0;; ++w) {
            Challenger a = *this;
            a.observe_witness(w);
            if ((a.sample() & mask) == 0) {
                observe_witness(w);
                sample();
                return w;
            }
        }
    }
    bool check_witness(size_t bits, uint64_t w) {  // :262-268
        observe_witness(w);
        return (sample() & ((uint64_t(1) << bits) - 1)) == 0;
    }
};

// Challenger script: op = code | arg << 8.  0 observe(arg), 1 sample,
// 2 sample_bits(arg), 3 grind(arg), 4 check_witness(arg & 63, arg >> 6),
// 5 sample_algebra_element<BabyBear4> (4 results).
size_t run_script(Challenger& ch, const uint64_t* ops, size_t n_ops, uint64_t* out) {
    size_t k = 0;
    for (size_t i = 0; i < n_ops; ++i) {
        const uint64_t code = ops[i] & 0xFF, arg = ops[i] >> 8;
        switch (code) {
            case 0: ch.observe(uint32_t(arg)); break;
            case 1: out[k++] = ch.sample(); break;
            case 2: out[k++] = ch.sample_bits(arg); break;
            case 3: out[k++] = ch.grind(arg); break;
            case 4: out[k++] = ch.check_witness(arg & 63, arg >> 6) ? 1 : 0; break;
            case 5:
                for (int q = 0; q < 4; ++q) out[k++] = ch.sample();
                break;
            default: throw std::invalid_argument("challenger op");
        }
    }
    return k;
}

// Final polynomial of the FRI commit phase (fri_prover.hpp:147-178): truncate
// to n = 2^lfpl, bit-reverse, naive inverse DFT over <omega_n>.
std::vector<E4> final_poly(std::vector<E4> cur, size_t lfpl) {
    size_t n = size_t(1) << lfpl;
    cur.resize(n);
    std::vector<E4> br(n), res(n);
    for (size_t i = 0; i < n; ++i) br[rev_bits(i, lfpl)] = cur[i];
    uint32_t winv = invf<BB>(BB::root(lfpl)), ninv = invf<BB>(uint32_t(n));
    for (size_t k = 0; k < n; ++k) {
        E4 s{0, 0, 0, 0};
        for (size_t j = 0; j < n; ++j) s = e4_add(s, e4_scale(br[j], powf<BB>(winv, k * j)));
        res[k] = e4_scale(s, ninv);
    }
    return res;
}

// LDE on `threads` host threads: every column's transform is independent, so
// the columns are split into slabs, one per thread, each run through lde()
// above.  Natural row order, N x w.
template <class F>
std::vector<typename F::Raw> lde_mt(const typename F::Raw* in, size_t log_h, size_t w, size_t added,
                                    typename F::Raw shift, int threads) {
    using R = typename F::Raw;
    if (w == 0) throw std::invalid_argument("empty matrix");
    const size_t h = size_t(1) << log_h, N = h << added;
    std::vector<R> out(N * w);
    parallel_for(w, threads, [&](size_t c0, size_t c1) {
        if (c0 == c1) return;
        const size_t wc = c1 - c0;
        std::vector<R> slab(h * wc);
        for (size_t i = 0; i < h; ++i) std::copy(in + i * w + c0, in + i * w + c1, slab.begin() + i * wc);
        Mat<F> m = lde<F>(slab.data(), log_h, wc, added, shift);
        slab = std::vector<R>();
        for (size_t i = 0; i < N; ++i) std::copy(m.row(i), m.row(i) + wc, out.begin() + i * w + c0);
    });
    return out;
}

// LDE + flat commit on `threads` host threads: the CPU baseline's multi-core
// figure and the source of the full-size fixtures.  The arithmetic is the
// single-threaded code's, so the root is the same (tests/test_oracle.py checks
// it).
template <class F>
Digest<F> lde_commit_mt(const typename F::Raw* c, int rf, int rp, const typename F::Raw* in, size_t log_h,
                        size_t w, size_t added, typename F::Raw shift, int threads) {
    auto lde_m = lde_mt<F>(in, log_h, w, added, shift, threads);
    const Perm<F, Hcfg<F>::W> perm(c, rf, rp);
    return flat_tree<F>(perm, lde_m.data(), (size_t(1) << log_h) << added, w, threads).back()[0];
}

// Opening of a flat tree (open_row, fri_merkle_tree_mmcs.hpp:129-141): the row
// and one sibling per level below the root, leaf level first.
template <class F>
void flat_open_rows(const std::vector<std::vector<Digest<F>>>& t, const typename F::Raw* m, size_t w,
                    const uint64_t* idx, size_t n_idx, typename F::Raw* values, typename F::Raw* proof) {
    constexpr size_t DIG = Hcfg<F>::DIG;
    const size_t steps = t.size() - 1, h = t[0].size();
    for (size_t q = 0; q < n_idx; ++q) {
        if (idx[q] >= h) throw std::invalid_argument("index out of range");
        std::copy(m + idx[q] * w, m + (idx[q] + 1) * w, values + q * w);
        size_t cur = idx[q];
        for (size_t s = 0; s < steps; ++s) {
            Digest<F> sib{};
            if ((cur ^ 1) < t[s].size()) sib = t[s][cur ^ 1];
            std::copy(sib.begin(), sib.end(), proof + (q * steps + s) * DIG);
            cur >>= 1;
        }
    }
}

// LDE + flat commit on `threads` host threads with openings at n_idx LDE rows:
// values_out n_idx x w (natural row order), proof_out n_idx x (log_h + added)
// digests, leaf level first.
template <class F>
int lde_commit_open_entry(const typename F::Raw* c, int rf, int rp, const typename F::Raw* in, size_t log_h,
                          size_t w, size_t added, typename F::Raw shift, int threads, const uint64_t* idx,
                          size_t n_idx, typename F::Raw* root, typename F::Raw* values, typename F::Raw* proof) {
    auto lde_m = lde_mt<F>(in, log_h, w, added, shift % F::P, threads);
    const Perm<F, Hcfg<F>::W> perm(c, rf, rp);
    auto t = flat_tree<F>(perm, lde_m.data(), (size_t(1) << log_h) << added, w, threads);
    std::copy(t.back()[0].begin(), t.back()[0].end(), root);
    flat_open_rows<F>(t, lde_m.data(), w, idx, n_idx, values, proof);
    return 0;
}

template <class Fn>
int guarded(Fn&& f) {
    try {
        return f();
    } catch (const std::exception&) {
        return -1;
    }
}

}  // namespace

// ======================================================================= C ===
extern "C" {

uint32_t orc_bb_mul(uint32_t a, uint32_t b) { return BB::mul(a, b); }
uint32_t orc_bb_pow(uint32_t a, uint64_t e) { return powf<BB>(a % BB::P, e); }
uint32_t orc_bb_two_adic_generator(uint32_t bits) { return BB::root(bits); }
uint64_t orc_gl_mul(uint64_t a, uint64_t b) { return GL::mul(a, b); }
uint64_t orc_gl_pow(uint64_t a, uint64_t e) { return powf<GL>(a % GL::P, e); }
uint64_t orc_gl_two_adic_generator(uint32_t bits) { return GL::root(bits); }
void orc_bb4_mul(const uint32_t* a, const uint32_t* b, uint32_t* out) {
    E4 r = e4_mul(E4{a[0], a[1], a[2], a[3]}, E4{b[0], b[1], b[2], b[3]});
    std::copy(r.begin(), r.end(), out);
}

void orc_bb_fill(uint64_t seed, uint32_t* out, size_t n) {
    std::mt19937_64 rng(seed);
    for (size_t i = 0; i < n; ++i) out[i] = uint32_t(rng() % BB::P);
}
void orc_gl_fill(uint64_t seed, uint64_t* out, size_t n) {
    std::mt19937_64 rng(seed);
    for (size_t i = 0; i < n; ++i) out[i] = rng() % GL::P;
}

int orc_bb_permute(const uint32_t* c, int rf, int rp, uint32_t* st, size_t n) {
    return guarded([&] {
        Perm<BB, 16> perm(c, rf, rp);
        for (size_t s = 0; s < n; ++s) {
            std::array<uint32_t, 16> x;
            for (int i = 0; i < 16; ++i) x[i] = st[s * 16 + i] % BB::P;
            perm.permute(x);
            std::copy(x.begin(), x.end(), st + s * 16);
        }
        return 0;
    });
}
int orc_gl_permute(const uint64_t* c, int rf, int rp, uint64_t* st, size_t n) {
    return guarded([&] {
        Perm<GL, 8> perm(c, rf, rp);
        for (size_t s = 0; s < n; ++s) {
            std::array<uint64_t, 8> x;
            for (int i = 0; i < 8; ++i) x[i] = st[s * 8 + i] % GL::P;
            perm.permute(x);
            std::copy(x.begin(), x.end(), st + s * 8);
        }
        return 0;
    });
}

int orc_bb_sponge(const uint32_t* c, int rf, int rp, const uint32_t* d, size_t len, uint32_t* out) {
    return guarded([&] {
        Perm<BB, 16> perm(c, rf, rp);
        std::vector<uint32_t> v(d, d + len);
        for (auto& x : v) x %= BB::P;
        auto r = sponge<BB>(perm, v.data(), len);
        std::copy(r.begin(), r.end(), out);
        return 0;
    });
}
int orc_gl_sponge(const uint64_t* c, int rf, int rp, const uint64_t* d, size_t len, uint64_t* out) {
    return guarded([&] {
        Perm<GL, 8> perm(c, rf, rp);
        std::vector<uint64_t> v(d, d + len);
        for (auto& x : v) x %= GL::P;
        auto r = sponge<GL>(perm, v.data(), len);
        std::copy(r.begin(), r.end(), out);
        return 0;
    });
}
int orc_bb_compress(const uint32_t* c, int rf, int rp, const uint32_t* in, uint32_t* out) {
    return guarded([&] {
        Perm<BB, 16> perm(c, rf, rp);
        Digest<BB> l, r;
        for (int i = 0; i < 8; ++i) {
            l[i] = in[i] % BB::P;
            r[i] = in[8 + i] % BB::P;
        }
        auto d = compress<BB>(perm, l, r);
        std::copy(d.begin(), d.end(), out);
        return 0;
    });
}
int orc_gl_compress(const uint64_t* c, int rf, int rp, const uint64_t* in, uint64_t* out) {
    return guarded([&] {
        Perm<GL, 8> perm(c, rf, rp);
        Digest<GL> l, r;
        for (int i = 0; i < 4; ++i) {
            l[i] = in[i] % GL::P;
            r[i] = in[4 + i] % GL::P;
        }
        auto d = compress<GL>(perm, l, r);
        std::copy(d.begin(), d.end(), out);
        return 0;
    });
}

int orc_bb_dft(int kind, const uint32_t* in, size_t log_h, size_t w, uint32_t shift, uint32_t* out) {
    return guarded([&] { return dft_family<BB>(kind, in, log_h, w, shift % BB::P, out); });
}
int orc_gl_dft(int kind, const uint64_t* in, size_t log_h, size_t w, uint64_t shift, uint64_t* out) {
    return guarded([&] { return dft_family<GL>(kind, in, log_h, w, shift % GL::P, out); });
}
int orc_bb_lde(const uint32_t* in, size_t log_h, size_t w, size_t added, uint32_t shift, uint32_t* out) {
    return guarded([&] {
        if (w == 0) return -1;
        auto m = lde<BB>(in, log_h, w, added, shift % BB::P);
        std::copy(m.v.begin(), m.v.end(), out);
        return 0;
    });
}
int orc_gl_lde(const uint64_t* in, size_t log_h, size_t w, size_t added, uint64_t shift, uint64_t* out) {
    return guarded([&] {
        if (w == 0) return -1;
        auto m = lde<GL>(in, log_h, w, added, shift % GL::P);
        std::copy(m.v.begin(), m.v.end(), out);
        return 0;
    });
}
int orc_bb_lde_brute(const uint32_t* in, size_t log_h, size_t w, size_t added, uint32_t shift, uint32_t* out) {
    return guarded([&] {
        lde_brute<BB>(in, log_h, w, added, shift % BB::P, out);
        return 0;
    });
}
int orc_gl_lde_brute(const uint64_t* in, size_t log_h, size_t w, size_t added, uint64_t shift, uint64_t* out) {
    return guarded([&] {
        lde_brute<GL>(in, log_h, w, added, shift % GL::P, out);
        return 0;
    });
}

int orc_bb_flat_commit(const uint32_t* c, int rf, int rp, const uint32_t* d, size_t h, size_t w,
                       uint32_t* root, uint32_t* tree_out) {
    return guarded([&] {
        Perm<BB, 16> perm(c, rf, rp);
        auto t = flat_tree<BB>(perm, d, h, w);
        std::copy(t.back()[0].begin(), t.back()[0].end(), root);
        if (tree_out) {
            size_t k = 0;
            for (const auto& layer : t)
                for (const auto& dg : layer)
                    for (uint32_t x : dg) tree_out[k++] = x;
        }
        return 0;
    });
}
int orc_gl_flat_commit(const uint64_t* c, int rf, int rp, const uint64_t* d, size_t h, size_t w,
                       uint64_t* root, uint64_t* tree_out) {
    return guarded([&] {
        Perm<GL, 8> perm(c, rf, rp);
        auto t = flat_tree<GL>(perm, d, h, w);
        std::copy(t.back()[0].begin(), t.back()[0].end(), root);
        if (tree_out) {
            size_t k = 0;
            for (const auto& layer : t)
                for (const auto& dg : layer)
                    for (uint64_t x : dg) tree_out[k++] = x;
        }
        return 0;
    });
}
int orc_bb_lde_commit_mt(const uint32_t* c, int rf, int rp, const uint32_t* in, size_t log_h, size_t w,
                         size_t added, uint32_t shift, int threads, uint32_t* root) {
    return guarded([&] {
        auto d = lde_commit_mt<BB>(c, rf, rp, in, log_h, w, added, shift % BB::P, threads);
        std::copy(d.begin(), d.end(), root);
        return 0;
    });
}
int orc_gl_lde_commit_mt(const uint64_t* c, int rf, int rp, const uint64_t* in, size_t log_h, size_t w,
                         size_t added, uint64_t shift, int threads, uint64_t* root) {
    return guarded([&] {
        auto d = lde_commit_mt<GL>(c, rf, rp, in, log_h, w, added, shift % GL::P, threads);
        std::copy(d.begin(), d.end(), root);
        return 0;
    });
}

int orc_bb_lde_commit_open_mt(const uint32_t* c, int rf, int rp, const uint32_t* in, size_t log_h, size_t w,
                              size_t added, uint32_t shift, int threads, const uint64_t* idx, size_t n_idx,
                              uint32_t* root, uint32_t* values, uint32_t* proof) {
    return guarded([&] {
        return lde_commit_open_entry<BB>(c, rf, rp, in, log_h, w, added, shift, threads, idx, n_idx, root, values,
                                         proof);
    });
}

int orc_gl_lde_commit_open_mt(const uint64_t* c, int rf, int rp, const uint64_t* in, size_t log_h, size_t w,
                              size_t added, uint64_t shift, int threads, const uint64_t* idx, size_t n_idx,
                              uint64_t* root, uint64_t* values, uint64_t* proof) {
    return guarded([&] {
        return lde_commit_open_entry<GL>(c, rf, rp, in, log_h, w, added, shift, threads, idx, n_idx, root, values,
                                         proof);
    });
}

// Batched commit of BabyBear traces: each trace's coset LDE (coset_lde_batch,
// radix2_dit.hpp:221-233, with its own added bits and the common shift), then
// one multi-height MMCS over the LDEs in trace order (build_merkle_tree,
// merkle_tree.hpp:82-185) with openings (open_batch,
// merkle_tree_mmcs.hpp:120-152) at n_idx indices of the tallest LDE.
// values_out: per index, the opened row of every LDE in trace order;
// proof_out: per index, *proof_len digests.
int orc_bb_batch_commit_mt(const uint32_t* c, int rf, int rp, size_t n, const size_t* log_hs, const size_t* ws,
                           const size_t* adds, const uint32_t* const* traces, uint32_t shift, size_t cap_height,
                           int threads, const uint64_t* idx, size_t n_idx, uint32_t* cap_out, uint32_t* values_out,
                           uint32_t* proof_out, size_t* proof_len) {
    return guarded([&] {
        std::vector<std::vector<uint32_t>> ldes(n);
        std::vector<size_t> hs(n);
        std::vector<const uint32_t*> ptrs(n);
        for (size_t m = 0; m < n; ++m) {
            ldes[m] = lde_mt<BB>(traces[m], log_hs[m], ws[m], adds[m], shift % BB::P, threads);
            hs[m] = (size_t(1) << log_hs[m]) << adds[m];
            ptrs[m] = ldes[m].data();
        }
        const Perm<BB, 16> perm(c, rf, rp);
        auto t = multi_tree<BB>(perm, n, hs.data(), ws, ptrs.data(), cap_height, threads);
        size_t k = 0;
        for (const auto& d : t.layers.back())
            for (size_t i = 0; i < 8; ++i) cap_out[k++] = d[i];
        const size_t maxh = t.layers[0].size(), steps = t.layers.size() - 1;
        size_t tw = 0;
        for (size_t m = 0; m < n; ++m) tw += ws[m];
        for (size_t q = 0; q < n_idx; ++q) {
            if (idx[q] >= maxh) return -2;
            size_t v = q * tw;
            for (size_t m = 0; m < n; ++m) {
                const size_t row = idx[q] / (maxh / hs[m]);
                std::copy(ptrs[m] + row * ws[m], ptrs[m] + (row + 1) * ws[m], values_out + v);
                v += ws[m];
            }
            size_t cur = idx[q];
            for (size_t s = 0; s < steps; ++s) {
                std::copy(t.layers[s][cur ^ 1].begin(), t.layers[s][cur ^ 1].end(), proof_out + (q * steps + s) * 8);
                cur >>= 1;
            }
        }
        *proof_len = steps;
        return 0;
    });
}

// MerkleTreeMmcs::open_row (fri_merkle_tree_mmcs.hpp:129-141): siblings, with a
// zero digest where the sibling falls past an odd layer's end.
int orc_bb_flat_open(const uint32_t* c, int rf, int rp, const uint32_t* d, size_t h, size_t w,
                     size_t index, uint32_t* row_out, uint32_t* proof_out) {
    return guarded([&] {
        if (index >= h) return -2;
        Perm<BB, 16> perm(c, rf, rp);
        auto t = flat_tree<BB>(perm, d, h, w);
        std::copy(d + index * w, d + (index + 1) * w, row_out);
        size_t cur = index, p = 0;
        for (size_t l = 0; l + 1 < t.size(); ++l) {
            size_t sib = cur ^ 1;
            Digest<BB> z{};
            const Digest<BB>& s = sib < t[l].size() ? t[l][sib] : z;
            for (uint32_t x : s) proof_out[p++] = x;
            cur >>= 1;
        }
        return 0;
    });
}

int orc_bb_mmcs(const uint32_t* c, int rf, int rp, size_t n, const size_t* hs, const size_t* ws,
                const uint32_t* const* mats, size_t cap_h, uint32_t* cap, long long idx, uint32_t* vals,
                uint32_t* proof, size_t* plen) {
    return guarded([&] { return mmcs_entry<BB>(c, rf, rp, n, hs, ws, mats, cap_h, cap, idx, vals, proof, plen); });
}
int orc_gl_mmcs(const uint64_t* c, int rf, int rp, size_t n, const size_t* hs, const size_t* ws,
                const uint64_t* const* mats, size_t cap_h, uint64_t* cap, long long idx, uint64_t* vals,
                uint64_t* proof, size_t* plen) {
    return guarded([&] { return mmcs_entry<GL>(c, rf, rp, n, hs, ws, mats, cap_h, cap, idx, vals, proof, plen); });
}
int orc_bb_mmcs_verify(const uint32_t* c, int rf, int rp, size_t n, const size_t* hs, const size_t* ws,
                       size_t cap_h, const uint32_t* cap, size_t index, const uint32_t* vals,
                       const uint32_t* proof, size_t plen) {
    return guarded([&] { return mmcs_verify<BB>(c, rf, rp, n, hs, ws, cap_h, cap, index, vals, proof, plen); });
}

int orc_bb4_fold(size_t log_height, size_t log_arity, const uint32_t* beta, const uint32_t* cur, size_t len,
                 uint32_t* out) {
    return guarded([&] {
        if (log_arity < 1 || log_arity > log_height) return -1;
        auto v = load_e4(cur, len);
        E4 b{beta[0] % BB::P, beta[1] % BB::P, beta[2] % BB::P, beta[3] % BB::P};
        auto f = fold(log_height, log_arity, b, v);
        for (size_t i = 0; i < f.size(); ++i)
            for (int k = 0; k < 4; ++k) out[4 * i + k] = f[i][k];
        return 0;
    });
}

// prove_fri commit loop (fri_prover.hpp:83-139) for one input vector, with the
// arity schedule of compute_log_arity_for_round (fri_params.hpp:36-57); each
// layer is committed by ExtensionMmcs (fri_extension_mmcs.hpp:49-75): a
// (len/arity) x arity matrix of BabyBear^4 flattened to 4*arity base elements
// per row, then the flat MMCS.  Final polynomial (fri_prover.hpp:147-178):
// truncate to 2^log_final_poly_len, bit-reverse, inverse DFT (naive).
int orc_bb4_fri_commit_mt(const uint32_t* c, int rf, int rp, const uint32_t* codeword, size_t log_len,
                          size_t log_blowup, size_t log_final_poly_len, size_t max_log_arity,
                          const uint32_t* betas, int threads, uint32_t* roots_out, uint32_t* final_out,
                          size_t* n_rounds) {
    return guarded([&] {
        Perm<BB, 16> perm(c, rf, rp);
        auto cur = load_e4(codeword, size_t(1) << log_len);
        size_t lfh = log_blowup + log_final_poly_len, round = 0;
        while (cur.size() > (size_t(1) << lfh)) {
            size_t lh = log2_strict(cur.size());
            size_t la = std::min(lh - lfh, max_log_arity);
            if (la == 0) break;
            size_t arity = size_t(1) << la, rows = cur.size() / arity;
            std::vector<uint32_t> flat(cur.size() * 4);
            for (size_t i = 0; i < cur.size(); ++i)
                for (int k = 0; k < 4; ++k) flat[4 * i + k] = cur[i][k];
            auto t = flat_tree<BB>(perm, flat.data(), rows, 4 * arity, threads);
            std::copy(t.back()[0].begin(), t.back()[0].end(), roots_out + 8 * round);
            const uint32_t* bp = betas + 4 * round;
            cur = fold(lh, la, E4{bp[0] % BB::P, bp[1] % BB::P, bp[2] % BB::P, bp[3] % BB::P}, cur, threads);
            ++round;
        }
        size_t n = size_t(1) << log_final_poly_len;
        cur.resize(n);
        size_t ln = log2_strict(n);
        std::vector<E4> br(n);
        for (size_t i = 0; i < n; ++i) br[rev_bits(i, ln)] = cur[i];
        uint32_t winv = invf<BB>(BB::root(ln)), ninv = invf<BB>(uint32_t(n));
        for (size_t k = 0; k < n; ++k) {
            E4 s{0, 0, 0, 0};
            for (size_t j = 0; j < n; ++j) s = e4_add(s, e4_scale(br[j], powf<BB>(winv, k * j)));
            s = e4_scale(s, ninv);
            for (int q = 0; q < 4; ++q) final_out[4 * k + q] = s[q];
        }
        *n_rounds = round;
        return 0;
    });
}

int orc_bb4_fri_commit(const uint32_t* c, int rf, int rp, const uint32_t* codeword, size_t log_len,
                       size_t log_blowup, size_t log_final_poly_len, size_t max_log_arity,
                       const uint32_t* betas, uint32_t* roots_out, uint32_t* final_out, size_t* n_rounds) {
    return orc_bb4_fri_commit_mt(c, rf, rp, codeword, log_len, log_blowup, log_final_poly_len, max_log_arity, betas,
                                 1, roots_out, final_out, n_rounds);
}

// TwoAdicFriPcs::commit (two_adic_fri_pcs.hpp:207-266).
int orc_bb_pcs_commit(const uint32_t* c, int rf, int rp, size_t n, const size_t* log_hs, const size_t* ws,
                      const uint32_t* shifts, const uint32_t* const* mats, size_t log_blowup, uint32_t* root) {
    return guarded([&] {
        Perm<BB, 16> perm(c, rf, rp);
        std::vector<Mat<BB>> ldes;
        size_t maxh = 0, tw = 0;
        for (size_t m = 0; m < n; ++m) {
            ldes.push_back(lde<BB>(mats[m], log_hs[m], ws[m], log_blowup, shifts[m] % BB::P));
            maxh = std::max(maxh, ldes.back().h);
            tw += ws[m];
        }
        if (tw == 0) return -1;
        size_t pw = 1;
        while (pw < tw) pw <<= 1;
        std::vector<uint32_t> flat;
        flat.reserve(maxh * pw);
        for (size_t r = 0; r < maxh; ++r) {
            for (auto& l : ldes) {
                const uint32_t* row = l.row(r % l.h);
                flat.insert(flat.end(), row, row + l.w);
            }
            flat.insert(flat.end(), pw - tw, 0u);
        }
        auto t = flat_tree<BB>(perm, flat.data(), maxh, pw);
        std::copy(t.back()[0].begin(), t.back()[0].end(), root);
        return 0;
    });
}

// DuplexChallenger driven by a script (encoding in run_script); results in
// order into out, their count into *n_out.
int orc_bb_challenger_run(const uint32_t* c, int rf, int rp, const uint64_t* ops, size_t n_ops, uint64_t* out,
                          size_t* n_out) {
    return guarded([&] {
        Perm<BB, 16> perm(c, rf, rp);
        Challenger ch{&perm, {}, {}, {}};
        *n_out = run_script(ch, ops, n_ops, out);
        return 0;
    });
}

// prove_fri commit phase (fri_prover.hpp:83-183) with the FriChallenger of the
// reference's tests (tests/fri_test.cpp:99-124) over DuplexChallenger: after
// `pre` observations, each round commits the layer, observes its root
// (observe_merkle_cap), grinds pow_bits, samples beta (4 samples) and folds;
// then the final polynomial's coefficients and every round's log arity are
// observed.  post_out receives 8 further samples (the transcript state).
int orc_bb4_fri_commit_chal(const uint32_t* c, int rf, int rp, const uint32_t* codeword, size_t log_len,
                            size_t log_blowup, size_t log_final_poly_len, size_t max_log_arity, size_t pow_bits,
                            const uint32_t* pre, size_t n_pre, uint32_t* roots_out, uint32_t* betas_out,
                            uint64_t* witnesses_out, uint32_t* final_out, uint32_t* post_out, size_t* n_rounds) {
    return guarded([&] {
        Perm<BB, 16> perm(c, rf, rp);
        Challenger ch{&perm, {}, {}, {}};
        for (size_t i = 0; i < n_pre; ++i) ch.observe(pre[i]);
        auto cur = load_e4(codeword, size_t(1) << log_len);
        size_t lfh = log_blowup + log_final_poly_len, round = 0;
        std::vector<size_t> arities;
        while (cur.size() > (size_t(1) << lfh)) {
            size_t lh = log2_strict(cur.size());
            size_t la = std::min(lh - lfh, max_log_arity);
            if (la == 0) break;
            size_t arity = size_t(1) << la, rows = cur.size() / arity;
            std::vector<uint32_t> flat(cur.size() * 4);
            for (size_t i = 0; i < cur.size(); ++i)
                for (int k = 0; k < 4; ++k) flat[4 * i + k] = cur[i][k];
            auto t = flat_tree<BB>(perm, flat.data(), rows, 4 * arity);
            for (int k = 0; k < 8; ++k) {
                roots_out[8 * round + k] = t.back()[0][k];
                ch.observe(t.back()[0][k]);
            }
            witnesses_out[round] = ch.grind(pow_bits);
            E4 beta;
            for (int q = 0; q < 4; ++q) beta[q] = ch.sample();
            std::copy(beta.begin(), beta.end(), betas_out + 4 * round);
            cur = fold(lh, la, beta, cur);
            arities.push_back(la);
            ++round;
        }
        auto fp = final_poly(cur, log_final_poly_len);
        for (size_t k = 0; k < fp.size(); ++k)
            for (int q = 0; q < 4; ++q) {
                final_out[4 * k + q] = fp[k][q];
                ch.observe(fp[k][q]);
            }
        for (size_t la : arities) ch.observe(uint32_t(la));
        for (int k = 0; k < 8; ++k) post_out[k] = ch.sample();
        *n_rounds = round;
        return 0;
    });
}

// TwoAdicFriPcs::commit then ::open (two_adic_fri_pcs.hpp:207-455) with the
// FriChallenger transcript of the reference's tests (tests/fri_test.cpp:99-124),
// one batch of n matrices: the transcript observes `pre` and the commitment
// root, then open samples alpha (two_adic_fri_pcs.hpp:277), evaluates every
// matrix at its points by barycentric interpolation on its LDE coset
// (interpolation.hpp:43-185), observes the opened values, builds one quotient
// vector per LDE height (:335-395; bit-reversed, descending heights) and runs
// prove_fri (fri_prover.hpp:47-280) over them: commit rounds with inputs mixed
// in by beta^arity, final polynomial, query grinding, query indices
// (sample_bits of the largest log height) and the input rows each query opens
// (PcsQueryInputMmcs::open, two_adic_fri_pcs.hpp:103-121).
// Outputs: root (8), opened [matrix][point][col] x 4, roots 8/round, commit
// witnesses 1/round, final poly 4 x 2^lfpl, query witness, query indices,
// query rows (per query: each height group's concatenated row, descending),
// post 8 samples after the proof.
int orc_bb_pcs_open(const uint32_t* c, int rf, int rp, size_t n, const size_t* log_hs, const size_t* ws,
                    const uint32_t* shifts, const uint32_t* const* mats, size_t log_blowup, const size_t* n_points,
                    const uint32_t* points, const uint32_t* pre, size_t n_pre, size_t lfpl, size_t mla,
                    size_t num_queries, size_t commit_pow_bits, size_t query_pow_bits, uint32_t* root_out,
                    uint32_t* opened_out, uint32_t* roots_out, uint64_t* cwit_out, uint32_t* final_out,
                    uint64_t* qwit_out, uint64_t* qidx_out, uint32_t* qrows_out, uint32_t* post_out,
                    size_t* n_rounds) {
    return guarded([&] {
        Perm<BB, 16> perm(c, rf, rp);
        Challenger ch{&perm, {}, {}, {}};
        for (size_t i = 0; i < n_pre; ++i) ch.observe(pre[i]);
        // commit (the flat MMCS over the cyclically repeated, width-padded LDE rows)
        std::vector<Mat<BB>> ldes;
        size_t maxh = 0, tw = 0;
        for (size_t m = 0; m < n; ++m) {
            ldes.push_back(lde<BB>(mats[m], log_hs[m], ws[m], log_blowup, shifts[m] % BB::P));
            maxh = std::max(maxh, ldes.back().h);
            tw += ws[m];
        }
        size_t pw = 1;
        while (pw < tw) pw <<= 1;
        std::vector<uint32_t> flat;
        for (size_t r = 0; r < maxh; ++r) {
            for (auto& l : ldes) flat.insert(flat.end(), l.row(r % l.h), l.row(r % l.h) + l.w);
            flat.insert(flat.end(), pw - tw, 0u);
        }
        auto t = flat_tree<BB>(perm, flat.data(), maxh, pw);
        for (int k = 0; k < 8; ++k) {
            root_out[k] = t.back()[0][k];
            ch.observe(t.back()[0][k]);  // observe_commitment
        }
        // open
        E4 alpha;
        for (int q = 0; q < 4; ++q) alpha[q] = ch.sample();
        std::vector<std::vector<std::vector<E4>>> opened(n);
        size_t pt = 0;
        for (size_t m = 0; m < n; ++m) {
            const Mat<BB>& L = ldes[m];
            const size_t N = L.h, lN = log2_strict(N);
            const uint32_t g = shifts[m] % BB::P, om = BB::root(lN);
            // diff_invs[i] = h_i / (N g^(N-1)) (interpolation.hpp:81-99)
            const uint32_t common = invf<BB>(BB::mul(uint32_t(N % BB::P), powf<BB>(g, N - 1)));
            for (size_t pi = 0; pi < n_points[m]; ++pi, ++pt) {
                const E4 z{points[4 * pt] % BB::P, points[4 * pt + 1] % BB::P, points[4 * pt + 2] % BB::P,
                           points[4 * pt + 3] % BB::P};
                std::vector<E4> acc(L.w, E4{0, 0, 0, 0});
                uint32_t h = 1;
                for (size_t i = 0; i < N; ++i, h = BB::mul(h, om)) {
                    const uint32_t x = BB::mul(g, h);
                    const E4 wgt = e4_scale(e4_inv(e4_sub(z, e4_base(x))), BB::mul(h, common));
                    for (size_t col = 0; col < L.w; ++col) acc[col] = e4_add(acc[col], e4_scale(wgt, L.v[i * L.w + col]));
                }
                const E4 van = e4_sub(e4_pow(z, N), e4_base(powf<BB>(g, N)));
                std::vector<E4> vals(L.w);
                for (size_t col = 0; col < L.w; ++col) vals[col] = e4_mul(van, acc[col]);
                opened[m].push_back(vals);
            }
        }
        size_t o = 0;
        for (auto& mv : opened)
            for (auto& pv : mv)
                for (auto& v : pv)
                    for (int q = 0; q < 4; ++q) {
                        opened_out[o++] = v[q];
                        ch.observe(v[q]);  // observe_challenge
                    }
        // quotients, one vector per LDE height, descending
        std::map<size_t, std::vector<E4>, std::greater<size_t>> groups;
        std::map<size_t, std::vector<size_t>, std::greater<size_t>> group_mats;
        size_t apow = 0;
        pt = 0;
        for (size_t m = 0; m < n; ++m) {
            const Mat<BB>& L = ldes[m];
            const size_t N = L.h, lN = log2_strict(N);
            auto& qv = groups[lN];
            if (qv.empty()) qv.assign(N, E4{0, 0, 0, 0});
            group_mats[lN].push_back(m);
            const uint32_t g = shifts[m] % BB::P, om = BB::root(lN);
            std::vector<E4> ap(n_points[m] * L.w);
            E4 cur = e4_pow(alpha, apow);
            for (auto& a : ap) {
                a = cur;
                cur = e4_mul(cur, alpha);
            }
            apow += ap.size();
            for (size_t k = 0; k < N; ++k) {
                const E4 x = e4_base(BB::mul(g, powf<BB>(om, k)));
                for (size_t pi = 0; pi < n_points[m]; ++pi) {
                    const size_t pp = pt + pi;
                    const E4 z{points[4 * pp] % BB::P, points[4 * pp + 1] % BB::P, points[4 * pp + 2] % BB::P,
                               points[4 * pp + 3] % BB::P};
                    const E4 dinv = e4_inv(e4_sub(z, x));
                    for (size_t col = 0; col < L.w; ++col) {
                        const E4 num = e4_sub(opened[m][pi][col], e4_base(L.v[k * L.w + col]));
                        qv[k] = e4_add(qv[k], e4_mul(e4_mul(ap[pi * L.w + col], num), dinv));
                    }
                }
            }
            pt += n_points[m];
        }
        std::vector<std::vector<E4>> inputs;
        for (auto& [lh, qv] : groups) {
            std::vector<E4> br(qv.size());
            for (size_t i = 0; i < qv.size(); ++i) br[i] = qv[rev_bits(i, lh)];
            inputs.push_back(br);
        }
        // prove_fri
        const size_t lfh = log_blowup + lfpl;
        if (log2_strict(inputs.back().size()) <= lfh) throw std::invalid_argument("input too small for FRI");
        std::vector<E4> cur = inputs[0];
        size_t next_in = 1, round = 0;
        std::vector<size_t> arities;
        std::vector<std::vector<E4>> layers;
        while (cur.size() > (size_t(1) << lfh)) {
            const size_t lh = log2_strict(cur.size());
            const bool has_next = next_in < inputs.size();
            const size_t target = has_next ? log2_strict(inputs[next_in].size()) : lfh;
            if (lh <= target) break;
            const size_t la = std::min(lh - target, mla);
            const size_t arity = size_t(1) << la, rows = cur.size() / arity;
            std::vector<uint32_t> fl(cur.size() * 4);
            for (size_t i = 0; i < cur.size(); ++i)
                for (int q = 0; q < 4; ++q) fl[4 * i + q] = cur[i][q];
            auto tr = flat_tree<BB>(perm, fl.data(), rows, 4 * arity);
            for (int k = 0; k < 8; ++k) {
                roots_out[8 * round + k] = tr.back()[0][k];
                ch.observe(tr.back()[0][k]);
            }
            cwit_out[round] = ch.grind(commit_pow_bits);
            E4 beta;
            for (int q = 0; q < 4; ++q) beta[q] = ch.sample();
            cur = fold(lh, la, beta, cur);
            if (has_next && inputs[next_in].size() == cur.size()) {
                E4 bp = beta;
                for (size_t s2 = 0; s2 < la; ++s2) bp = e4_mul(bp, bp);
                for (size_t k = 0; k < cur.size(); ++k) cur[k] = e4_add(cur[k], e4_mul(bp, inputs[next_in][k]));
                ++next_in;
            }
            arities.push_back(la);
            ++round;
        }
        auto fp = final_poly(cur, lfpl);
        for (size_t k = 0; k < fp.size(); ++k)
            for (int q = 0; q < 4; ++q) {
                final_out[4 * k + q] = fp[k][q];
                ch.observe(fp[k][q]);
            }
        for (size_t la : arities) ch.observe(uint32_t(la));
        *qwit_out = ch.grind(query_pow_bits);
        size_t log_max = lfh;
        for (size_t la : arities) log_max += la;
        size_t qr = 0;
        for (size_t q = 0; q < num_queries; ++q) {
            const uint64_t idx = ch.sample_bits(log_max);
            qidx_out[q] = idx;
            for (auto& [lh, ms] : group_mats) {
                const size_t k = rev_bits(idx >> (log_max - lh), lh);
                for (size_t m : ms)
                    for (size_t col = 0; col < ldes[m].w; ++col) qrows_out[qr++] = ldes[m].v[k * ldes[m].w + col];
            }
        }
        for (int k = 0; k < 8; ++k) post_out[k] = ch.sample();
        *n_rounds = round;
        return 0;
    });
}

}  // extern "C"
