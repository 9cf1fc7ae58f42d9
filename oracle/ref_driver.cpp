// Test infrastructure only: a C-ABI driver over the reference's own header-only
// C++ library (dloghin/plonky3-gpu, src/include).  It is compiled from the
// reference sources where they lie (see oracle/Makefile, target `ref`) into
// oracle/_ref/libp3ref.so, which is never committed.  Its only uses are
// (1) writing the golden fixtures under tests/golden/ and (2) checking the C++
// restatement in oracle/oracle.cpp, and (3) timing the reference CPU path as a
// baseline.  Nothing in the product links or loads it.
//
// Every entry point returns 0 on success and -1 if the reference threw; the
// message is available from ref_last_error().
//
// Poseidon2 constant layout (canonical values): initial[rf/2][W], then
// terminal[rf/2][W], then internal[rp] -- the field order of the reference's
// POD Poseidon2Cuda (src/include/poseidon2_cuda.hpp:37-46).

#include "baby_bear.hpp"
#include "goldilocks.hpp"
#include "extension_field.hpp"
#include "poseidon2.hpp"
#include "padding_free_sponge.hpp"
#include "truncated_permutation.hpp"
#include "traits.hpp"
#include "radix2_dit.hpp"
#include "dense_matrix.hpp"
#include "merkle_tree_mmcs.hpp"
#include "fri_merkle_tree_mmcs.hpp"
#include "fri_extension_mmcs.hpp"
#include "fri_folding.hpp"
#include "fri_params.hpp"
#include "domain.hpp"
#include "two_adic_fri_pcs.hpp"
#include "duplex_challenger.hpp"
#include "p3_util/util.hpp"

#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

using BB = p3_field::BabyBear;
using GL = p3_field::Goldilocks;
using BB4 = p3_field::BabyBear4;

// The reference defines no two-adic structure for Goldilocks
// (src/include/goldilocks.hpp has no two_adic_generator).  This build fixes
// generator 7 and omega_{2^32} = 7^((p-1)/2^32); this specialisation lets the
// reference's Radix2Dit run unchanged with those constants.
namespace p3_dft {
template <>
struct TwoAdicFieldTraits<GL> {
    static GL two_adic_generator(size_t bits) {
        if (bits > 32) throw std::invalid_argument("bits > 32 for Goldilocks");
        GL g(static_cast<uint64_t>(7));
        GL root = g.exp_u64((GL::PRIME - 1) >> 32);
        return root.exp_u64(uint64_t(1) << (32 - bits));
    }
};
}  // namespace p3_dft

namespace {

thread_local std::string g_err;

template <class F, size_t W>
struct RefPerm {
    std::shared_ptr<poseidon2::Poseidon2<F, F, W, 7>> perm;
    void permute_mut(std::array<F, W>& s) const { perm->permute_mut(s); }
};

template <class F, class Raw, size_t W>
RefPerm<F, W> make_perm(const Raw* consts, int rf, int rp) {
    std::vector<std::array<F, W>> init(rf / 2), term(rf / 2);
    size_t k = 0;
    for (int r = 0; r < rf / 2; ++r)
        for (size_t i = 0; i < W; ++i) init[r][i] = F(consts[k++]);
    for (int r = 0; r < rf / 2; ++r)
        for (size_t i = 0; i < W; ++i) term[r][i] = F(consts[k++]);
    std::vector<F> internal(rp);
    for (int r = 0; r < rp; ++r) internal[r] = F(consts[k++]);
    poseidon2::ExternalLayerConstants<F, W> ext(init, term);
    return RefPerm<F, W>{poseidon2::create_poseidon2<F, F, W, 7>(ext, internal)};
}

inline uint32_t raw(const BB& x) { return x.value(); }
inline uint64_t raw(const GL& x) { return x.as_canonical_u64(); }

// Field-specific type bundle: BabyBear uses width 16 / rate 8 / digest 8,
// Goldilocks width 8 / rate 4 / digest 4 (the configurations of the
// reference's tests and examples).
template <class F>
struct Cfg;
template <>
struct Cfg<BB> {
    using Raw = uint32_t;
    static constexpr size_t W = 16, RATE = 8, DIG = 8;
};
template <>
struct Cfg<GL> {
    using Raw = uint64_t;
    static constexpr size_t W = 8, RATE = 4, DIG = 4;
};

template <class F>
struct Stack {
    using C = Cfg<F>;
    using Perm = RefPerm<F, C::W>;
    using Hasher = p3_symmetric::PaddingFreeSponge<Perm, F, C::W, C::RATE, C::DIG>;
    using Compress = p3_symmetric::TruncatedPermutation<Perm, F, 2, C::DIG, C::W>;
    using FlatMmcs = p3_merkle::MerkleTreeMmcs<F, F, Hasher, Compress, 2, C::DIG>;
    using MultiMmcs = p3_merkle_tree::MerkleTreeMmcs<F, F, Hasher, Compress, C::DIG>;
};

template <class Fn>
int guard(Fn&& fn) {
    try {
        fn();
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

template <class F>
p3_matrix::RowMajorMatrix<F> load_mat(const typename Cfg<F>::Raw* in, size_t h, size_t w) {
    std::vector<F> v(h * w);
    for (size_t i = 0; i < h * w; ++i) v[i] = F(in[i]);
    return p3_matrix::RowMajorMatrix<F>(std::move(v), w);
}

template <class F>
void store_mat(const p3_matrix::RowMajorMatrix<F>& m, typename Cfg<F>::Raw* out) {
    for (size_t i = 0; i < m.values.size(); ++i) out[i] = raw(m.values[i]);
}

template <class F>
int permute_impl(const typename Cfg<F>::Raw* consts, int rf, int rp, typename Cfg<F>::Raw* states,
                 size_t n) {
    return guard([&] {
        constexpr size_t W = Cfg<F>::W;
        auto perm = make_perm<F, typename Cfg<F>::Raw, W>(consts, rf, rp);
        for (size_t s = 0; s < n; ++s) {
            std::array<F, W> st;
            for (size_t i = 0; i < W; ++i) st[i] = F(states[s * W + i]);
            perm.permute_mut(st);
            for (size_t i = 0; i < W; ++i) states[s * W + i] = raw(st[i]);
        }
    });
}

template <class F>
int sponge_impl(const typename Cfg<F>::Raw* consts, int rf, int rp, const typename Cfg<F>::Raw* data,
                size_t len, typename Cfg<F>::Raw* out) {
    return guard([&] {
        using S = Stack<F>;
        auto perm = make_perm<F, typename Cfg<F>::Raw, Cfg<F>::W>(consts, rf, rp);
        typename S::Hasher h(perm);
        std::vector<F> v(len);
        for (size_t i = 0; i < len; ++i) v[i] = F(data[i]);
        auto d = h.hash_iter(v);
        for (size_t i = 0; i < Cfg<F>::DIG; ++i) out[i] = raw(d[i]);
    });
}

template <class F>
int compress_impl(const typename Cfg<F>::Raw* consts, int rf, int rp, const typename Cfg<F>::Raw* in,
                  typename Cfg<F>::Raw* out) {
    return guard([&] {
        using S = Stack<F>;
        constexpr size_t D = Cfg<F>::DIG;
        auto perm = make_perm<F, typename Cfg<F>::Raw, Cfg<F>::W>(consts, rf, rp);
        typename S::Compress c(perm);
        std::array<std::array<F, D>, 2> pr;
        for (size_t i = 0; i < D; ++i) {
            pr[0][i] = F(in[i]);
            pr[1][i] = F(in[D + i]);
        }
        auto d = c.compress(pr);
        for (size_t i = 0; i < D; ++i) out[i] = raw(d[i]);
    });
}

// kind: 0 dft, 1 idft, 2 coset_dft, 3 coset_idft
template <class F>
int dft_impl(int kind, const typename Cfg<F>::Raw* in, size_t log_h, size_t w,
             typename Cfg<F>::Raw shift, typename Cfg<F>::Raw* out) {
    return guard([&] {
        p3_dft::Radix2Dit<F> dft;
        auto m = load_mat<F>(in, size_t(1) << log_h, w);
        F s(shift);
        if (kind == 0) m = dft.dft_batch(std::move(m));
        else if (kind == 1) m = dft.idft_batch(std::move(m));
        else if (kind == 2) m = dft.coset_dft_batch(std::move(m), s);
        else m = dft.coset_idft_batch(std::move(m), s);
        store_mat<F>(m, out);
    });
}

template <class F>
int lde_impl(const typename Cfg<F>::Raw* in, size_t log_h, size_t w, size_t added_bits,
             typename Cfg<F>::Raw shift, typename Cfg<F>::Raw* out) {
    return guard([&] {
        p3_dft::Radix2Dit<F> dft;
        auto m = load_mat<F>(in, size_t(1) << log_h, w);
        m = dft.coset_lde_batch(std::move(m), added_bits, F(shift));
        store_mat<F>(m, out);
    });
}

// tree_out receives every layer, leaves first: sum over layers of size*DIG words.
template <class F>
int flat_commit_impl(const typename Cfg<F>::Raw* consts, int rf, int rp,
                     const typename Cfg<F>::Raw* data, size_t height, size_t width,
                     typename Cfg<F>::Raw* root_out, typename Cfg<F>::Raw* tree_out) {
    return guard([&] {
        using S = Stack<F>;
        constexpr size_t D = Cfg<F>::DIG;
        auto perm = make_perm<F, typename Cfg<F>::Raw, Cfg<F>::W>(consts, rf, rp);
        typename S::FlatMmcs mmcs{typename S::Hasher(perm), typename S::Compress(perm)};
        std::vector<F> v(height * width);
        for (size_t i = 0; i < v.size(); ++i) v[i] = F(data[i]);
        auto [commit, pd] = mmcs.commit_matrix(v, width);
        for (size_t i = 0; i < D; ++i) root_out[i] = raw(commit[0][i]);
        if (tree_out) {
            size_t k = 0;
            for (const auto& layer : pd.tree)
                for (const auto& dg : layer)
                    for (size_t i = 0; i < D; ++i) tree_out[k++] = raw(dg[i]);
        }
    });
}

// Multi-height MMCS commit; optionally opens row `open_index`:
// values_out gets the opened rows (original matrix order, concatenated),
// proof_out gets num_levels digests.
template <class F>
int mmcs_impl(const typename Cfg<F>::Raw* consts, int rf, int rp, size_t n_mats,
              const size_t* heights, const size_t* widths,
              const typename Cfg<F>::Raw* const* mats, size_t cap_height,
              typename Cfg<F>::Raw* cap_out, long long open_index,
              typename Cfg<F>::Raw* values_out, typename Cfg<F>::Raw* proof_out,
              size_t* proof_len) {
    return guard([&] {
        using S = Stack<F>;
        constexpr size_t D = Cfg<F>::DIG;
        auto perm = make_perm<F, typename Cfg<F>::Raw, Cfg<F>::W>(consts, rf, rp);
        typename S::MultiMmcs mmcs{typename S::Hasher(perm), typename S::Compress(perm), cap_height};
        std::vector<p3_matrix::RowMajorMatrix<F>> ms;
        std::vector<p3_matrix::Dimensions> dims;
        for (size_t m = 0; m < n_mats; ++m) {
            ms.push_back(load_mat<F>(mats[m], heights[m], widths[m]));
            dims.push_back(p3_matrix::Dimensions{widths[m], heights[m]});
        }
        auto [cap, tree] = mmcs.commit(std::move(ms));
        size_t k = 0;
        for (const auto& dg : cap.cap)
            for (size_t i = 0; i < D; ++i) cap_out[k++] = raw(dg[i]);
        if (open_index >= 0) {
            auto op = mmcs.open_batch(static_cast<size_t>(open_index), tree);
            size_t v = 0;
            for (const auto& row : op.opened_values)
                for (const auto& x : row) values_out[v++] = raw(x);
            size_t p = 0;
            for (const auto& dg : op.proof)
                for (size_t i = 0; i < D; ++i) proof_out[p++] = raw(dg[i]);
            *proof_len = op.proof.size();
            if (!mmcs.verify_batch(cap, dims, static_cast<size_t>(open_index), op))
                throw std::runtime_error("reference verify_batch rejected its own opening");
        }
    });
}

}  // namespace

extern "C" {

const char* ref_last_error() { return g_err.c_str(); }

int ref_bb_permute(const uint32_t* c, int rf, int rp, uint32_t* st, size_t n) {
    return permute_impl<BB>(c, rf, rp, st, n);
}
int ref_gl_permute(const uint64_t* c, int rf, int rp, uint64_t* st, size_t n) {
    return permute_impl<GL>(c, rf, rp, st, n);
}
int ref_bb_sponge(const uint32_t* c, int rf, int rp, const uint32_t* d, size_t len, uint32_t* out) {
    return sponge_impl<BB>(c, rf, rp, d, len, out);
}
int ref_gl_sponge(const uint64_t* c, int rf, int rp, const uint64_t* d, size_t len, uint64_t* out) {
    return sponge_impl<GL>(c, rf, rp, d, len, out);
}
int ref_bb_compress(const uint32_t* c, int rf, int rp, const uint32_t* in, uint32_t* out) {
    return compress_impl<BB>(c, rf, rp, in, out);
}
int ref_gl_compress(const uint64_t* c, int rf, int rp, const uint64_t* in, uint64_t* out) {
    return compress_impl<GL>(c, rf, rp, in, out);
}
int ref_bb_dft(int kind, const uint32_t* in, size_t log_h, size_t w, uint32_t shift, uint32_t* out) {
    return dft_impl<BB>(kind, in, log_h, w, shift, out);
}
int ref_gl_dft(int kind, const uint64_t* in, size_t log_h, size_t w, uint64_t shift, uint64_t* out) {
    return dft_impl<GL>(kind, in, log_h, w, shift, out);
}
int ref_bb_lde(const uint32_t* in, size_t log_h, size_t w, size_t added, uint32_t shift, uint32_t* out) {
    return lde_impl<BB>(in, log_h, w, added, shift, out);
}
int ref_gl_lde(const uint64_t* in, size_t log_h, size_t w, size_t added, uint64_t shift, uint64_t* out) {
    return lde_impl<GL>(in, log_h, w, added, shift, out);
}
int ref_bb_flat_commit(const uint32_t* c, int rf, int rp, const uint32_t* d, size_t h, size_t w,
                       uint32_t* root, uint32_t* tree) {
    return flat_commit_impl<BB>(c, rf, rp, d, h, w, root, tree);
}
int ref_gl_flat_commit(const uint64_t* c, int rf, int rp, const uint64_t* d, size_t h, size_t w,
                       uint64_t* root, uint64_t* tree) {
    return flat_commit_impl<GL>(c, rf, rp, d, h, w, root, tree);
}
int ref_bb_mmcs(const uint32_t* c, int rf, int rp, size_t n, const size_t* hs, const size_t* ws,
                const uint32_t* const* mats, size_t cap_h, uint32_t* cap, long long idx,
                uint32_t* vals, uint32_t* proof, size_t* plen) {
    return mmcs_impl<BB>(c, rf, rp, n, hs, ws, mats, cap_h, cap, idx, vals, proof, plen);
}
int ref_gl_mmcs(const uint64_t* c, int rf, int rp, size_t n, const size_t* hs, const size_t* ws,
                const uint64_t* const* mats, size_t cap_h, uint64_t* cap, long long idx,
                uint64_t* vals, uint64_t* proof, size_t* plen) {
    return mmcs_impl<GL>(c, rf, rp, n, hs, ws, mats, cap_h, cap, idx, vals, proof, plen);
}

// BabyBear^4 multiplication (x^4 - 11), for the extension-field known answers.
int ref_bb4_mul(const uint32_t* a, const uint32_t* b, uint32_t* out) {
    return guard([&] {
        BB4 x(std::array<BB, 4>{BB(a[0]), BB(a[1]), BB(a[2]), BB(a[3])});
        BB4 y(std::array<BB, 4>{BB(b[0]), BB(b[1]), BB(b[2]), BB(b[3])});
        BB4 z = x * y;
        for (int i = 0; i < 4; ++i) out[i] = z.coeffs[i].value();
    });
}

// TwoAdicFriFolding::fold_matrix over BabyBear^4 (flat coefficient arrays).
int ref_bb4_fold(size_t log_height, size_t log_arity, const uint32_t* beta, const uint32_t* cur,
                 size_t len, uint32_t* out) {
    return guard([&] {
        std::vector<BB4> v(len);
        for (size_t i = 0; i < len; ++i)
            v[i] = BB4(std::array<BB, 4>{BB(cur[4 * i]), BB(cur[4 * i + 1]), BB(cur[4 * i + 2]),
                                         BB(cur[4 * i + 3])});
        BB4 b(std::array<BB, 4>{BB(beta[0]), BB(beta[1]), BB(beta[2]), BB(beta[3])});
        auto f = p3_fri::TwoAdicFriFolding<BB, BB4>::fold_matrix(log_height, log_arity, b, v);
        for (size_t i = 0; i < f.size(); ++i)
            for (int k = 0; k < 4; ++k) out[4 * i + k] = f[i].coeffs[k].value();
    });
}

// FRI commit phase as in prove_fri (src/include/fri_prover.hpp:83-178) with the
// challenger replaced by caller-supplied betas (one BabyBear^4 per round) and a
// single input vector.  roots_out: 8 words per round; final_out: BB4 coeffs of
// the final polynomial.  *n_rounds receives the number of commit rounds.
int ref_bb4_fri_commit(const uint32_t* c, int rf, int rp, const uint32_t* codeword, size_t log_len,
                       size_t log_blowup, size_t log_final_poly_len, size_t max_log_arity,
                       const uint32_t* betas, uint32_t* roots_out, uint32_t* final_out,
                       size_t* n_rounds) {
    return guard([&] {
        using S = Stack<BB>;
        auto perm = make_perm<BB, uint32_t, 16>(c, rf, rp);
        typename S::FlatMmcs inner{typename S::Hasher(perm), typename S::Compress(perm)};
        p3_merkle::ExtensionMmcs<BB, BB4, typename S::FlatMmcs> mmcs(inner);
        size_t len = size_t(1) << log_len;
        std::vector<BB4> cur(len);
        for (size_t i = 0; i < len; ++i)
            cur[i] = BB4(std::array<BB, 4>{BB(codeword[4 * i]), BB(codeword[4 * i + 1]),
                                           BB(codeword[4 * i + 2]), BB(codeword[4 * i + 3])});
        size_t log_final_height = log_blowup + log_final_poly_len;
        size_t round = 0;
        while (cur.size() > (size_t(1) << log_final_height)) {
            size_t log_h = p3_util::log2_strict_usize(cur.size());
            size_t la = p3_fri::compute_log_arity_for_round(log_h, false, 0, log_final_height,
                                                            max_log_arity);
            if (la == 0) break;
            auto [commit, pd] = mmcs.commit_matrix(cur, size_t(1) << la);
            for (int i = 0; i < 8; ++i) roots_out[8 * round + i] = commit[0][i].value();
            const uint32_t* bp = betas + 4 * round;
            BB4 beta(std::array<BB, 4>{BB(bp[0]), BB(bp[1]), BB(bp[2]), BB(bp[3])});
            cur = p3_fri::TwoAdicFriFolding<BB, BB4>::fold_matrix(log_h, la, beta, cur);
            ++round;
        }
        // Final polynomial: truncate, bit-reverse, naive inverse DFT.
        size_t fpl = size_t(1) << log_final_poly_len;
        cur.resize(fpl);
        p3_util::reverse_slice_index_bits(cur);
        size_t n = cur.size();
        BB omega = BB::two_adic_generator(p3_util::log2_strict_usize(n));
        BB omega_inv = omega.exp_u64(n - 1);
        BB n_inv = BB(static_cast<uint32_t>(n)).inv();
        BB wk = BB::one_val();
        for (size_t k = 0; k < n; ++k) {
            BB4 sum;
            BB wkj = BB::one_val();
            for (size_t j = 0; j < n; ++j) {
                sum = sum + cur[j] * BB4(wkj);
                wkj = wkj * wk;
            }
            BB4 ck = sum * BB4(n_inv);
            for (int i = 0; i < 4; ++i) final_out[4 * k + i] = ck.coeffs[i].value();
            wk = wk * omega_inv;
        }
        *n_rounds = round;
    });
}


}  // extern "C"

using RefChallenger = p3_challenger::DuplexChallenger<BB, RefPerm<BB, 16>, 16, 8>;

// Transcript adapter with the methods prove_fri and TwoAdicFriPcs::open call,
// of the same shape as the FriChallenger in src/tests/fri_test.cpp:99-124.
struct PcsTranscript {
    RefChallenger inner;
    template <class Commit>
    void observe_commitment(const Commit& cm) { inner.observe_merkle_cap(cm); }
    template <class EF = BB4>
    EF sample_challenge() { return inner.template sample_algebra_element<EF>(); }
    template <class EF>
    void observe_challenge(const EF& v) { inner.template observe_algebra_element<EF>(v); }
    void observe_arity(size_t la) { inner.observe(BB(static_cast<uint32_t>(la))); }
    uint64_t grind(size_t bits) { return inner.grind(bits); }
    size_t sample_bits(size_t bits) { return inner.sample_bits(bits); }
};

extern "C" {

// The reference's DuplexChallenger driven by a script: op = code | arg << 8;
// 0 observe(arg), 1 sample, 2 sample_bits(arg), 3 grind(arg),
// 4 check_witness(arg & 63, arg >> 6), 5 sample_algebra_element<BabyBear4>.
int ref_bb_challenger_run(const uint32_t* c, int rf, int rp, const uint64_t* ops, size_t n_ops, uint64_t* out,
                          size_t* n_out) {
    return guard([&] {
        RefChallenger ch(make_perm<BB, uint32_t, 16>(c, rf, rp));
        size_t k = 0;
        for (size_t i = 0; i < n_ops; ++i) {
            const uint64_t code = ops[i] & 0xFF, arg = ops[i] >> 8;
            switch (code) {
                case 0: ch.observe(BB(static_cast<uint32_t>(arg))); break;
                case 1: out[k++] = ch.sample().value(); break;
                case 2: out[k++] = ch.sample_bits(arg); break;
                case 3: out[k++] = ch.grind(arg); break;
                case 4: out[k++] = ch.check_witness(arg & 63, arg >> 6) ? 1 : 0; break;
                case 5: {
                    BB4 e = ch.sample_algebra_element<BB4>();
                    for (int q = 0; q < 4; ++q) out[k++] = e.coeffs[q].value();
                    break;
                }
                default: throw std::invalid_argument("challenger op");
            }
        }
        *n_out = k;
    });
}

// prove_fri's commit phase (src/include/fri_prover.hpp:83-190) with the
// FriChallenger of src/tests/fri_test.cpp:99-124: per round observe_commitment
// (observe_merkle_cap), grind(pow_bits), sample_challenge (4 samples), fold;
// then observe_challenge for each final-polynomial coefficient and
// observe_arity for each round.  pre: values observed before the loop;
// post_out: 8 samples taken after the loop.
int ref_bb4_fri_commit_chal(const uint32_t* c, int rf, int rp, const uint32_t* codeword, size_t log_len,
                            size_t log_blowup, size_t log_final_poly_len, size_t max_log_arity, size_t pow_bits,
                            const uint32_t* pre, size_t n_pre, uint32_t* roots_out, uint32_t* betas_out,
                            uint64_t* witnesses_out, uint32_t* final_out, uint32_t* post_out, size_t* n_rounds) {
    return guard([&] {
        using S = Stack<BB>;
        auto perm = make_perm<BB, uint32_t, 16>(c, rf, rp);
        typename S::FlatMmcs inner{typename S::Hasher(perm), typename S::Compress(perm)};
        p3_merkle::ExtensionMmcs<BB, BB4, typename S::FlatMmcs> mmcs(inner);
        RefChallenger ch(perm);
        for (size_t i = 0; i < n_pre; ++i) ch.observe(BB(pre[i]));
        size_t len = size_t(1) << log_len;
        std::vector<BB4> cur(len);
        for (size_t i = 0; i < len; ++i)
            cur[i] = BB4(std::array<BB, 4>{BB(codeword[4 * i]), BB(codeword[4 * i + 1]),
                                           BB(codeword[4 * i + 2]), BB(codeword[4 * i + 3])});
        size_t log_final_height = log_blowup + log_final_poly_len;
        size_t round = 0;
        std::vector<size_t> arities;
        while (cur.size() > (size_t(1) << log_final_height)) {
            size_t log_h = p3_util::log2_strict_usize(cur.size());
            size_t la = p3_fri::compute_log_arity_for_round(log_h, false, 0, log_final_height, max_log_arity);
            if (la == 0) break;
            auto [commit, pd] = mmcs.commit_matrix(cur, size_t(1) << la);
            for (int i = 0; i < 8; ++i) roots_out[8 * round + i] = commit[0][i].value();
            ch.observe_merkle_cap(commit);
            witnesses_out[round] = ch.grind(pow_bits);
            BB4 beta = ch.sample_algebra_element<BB4>();
            for (int q = 0; q < 4; ++q) betas_out[4 * round + q] = beta.coeffs[q].value();
            cur = p3_fri::TwoAdicFriFolding<BB, BB4>::fold_matrix(log_h, la, beta, cur);
            arities.push_back(la);
            ++round;
        }
        size_t fpl = size_t(1) << log_final_poly_len;
        cur.resize(fpl);
        p3_util::reverse_slice_index_bits(cur);
        size_t n = cur.size();
        BB omega = BB::two_adic_generator(p3_util::log2_strict_usize(n));
        BB omega_inv = omega.exp_u64(n - 1);
        BB n_inv = BB(static_cast<uint32_t>(n)).inv();
        BB wk = BB::one_val();
        for (size_t k = 0; k < n; ++k) {
            BB4 sum;
            BB wkj = BB::one_val();
            for (size_t j = 0; j < n; ++j) {
                sum = sum + cur[j] * BB4(wkj);
                wkj = wkj * wk;
            }
            BB4 ck = sum * BB4(n_inv);
            for (int i = 0; i < 4; ++i) final_out[4 * k + i] = ck.coeffs[i].value();
            ch.observe_algebra_element<BB4>(ck);
            wk = wk * omega_inv;
        }
        for (size_t la : arities) ch.observe(BB(static_cast<uint32_t>(la)));
        for (int k = 0; k < 8; ++k) post_out[k] = ch.sample().value();
        *n_rounds = round;
    });
}

// TwoAdicFriPcs::commit (src/include/two_adic_fri_pcs.hpp:207-266) with
// Radix2Dit and the flat MMCS, as instantiated by src/tests/fri_test.cpp.
int ref_bb_pcs_commit(const uint32_t* c, int rf, int rp, size_t n, const size_t* log_hs, const size_t* ws,
                      const uint32_t* shifts, const uint32_t* const* mats, size_t log_blowup, uint32_t* root) {
    return guard([&] {
        using S = Stack<BB>;
        using EFMmcs = p3_merkle::ExtensionMmcs<BB, BB4, typename S::FlatMmcs>;
        using Pcs = p3_fri::TwoAdicFriPcs<BB, BB4, p3_dft::Radix2Dit<BB>, typename S::FlatMmcs, EFMmcs>;
        auto perm = make_perm<BB, uint32_t, 16>(c, rf, rp);
        typename S::FlatMmcs val_mmcs{typename S::Hasher(perm), typename S::Compress(perm)};
        EFMmcs ef_mmcs(val_mmcs);
        p3_fri::FriParameters<EFMmcs> params{log_blowup, 0, 1, 1, 0, 0, ef_mmcs};
        Pcs pcs(p3_dft::Radix2Dit<BB>(), val_mmcs, params);
        std::vector<std::pair<typename Pcs::Domain, p3_matrix::RowMajorMatrix<BB>>> evals;
        for (size_t m = 0; m < n; ++m) {
            typename Pcs::Domain d;
            d.log_n = log_hs[m];
            d.shift = BB(shifts[m]);
            evals.emplace_back(d, load_mat<BB>(mats[m], size_t(1) << log_hs[m], ws[m]));
        }
        auto [commit, pd] = pcs.commit(std::move(evals));
        for (int i = 0; i < 8; ++i) root[i] = commit[0][i].value();
    });
}

// TwoAdicFriPcs::commit + ::open (src/include/two_adic_fri_pcs.hpp:207-455)
// as src/tests/fri_test.cpp instantiates it, with a transcript adapter of the
// same shape as that test's FriChallenger; outputs as orc_bb_pcs_open.
int ref_bb_pcs_open(const uint32_t* c, int rf, int rp, size_t n, const size_t* log_hs, const size_t* ws,
                    const uint32_t* shifts, const uint32_t* const* mats, size_t log_blowup, const size_t* n_points,
                    const uint32_t* points, const uint32_t* pre, size_t n_pre, size_t lfpl, size_t mla,
                    size_t num_queries, size_t commit_pow_bits, size_t query_pow_bits, uint32_t* root_out,
                    uint32_t* opened_out, uint32_t* roots_out, uint64_t* cwit_out, uint32_t* final_out,
                    uint64_t* qwit_out, uint64_t* qidx_out, uint32_t* qrows_out, uint32_t* post_out,
                    size_t* n_rounds) {
    return guard([&] {
        using S = Stack<BB>;
        using EFMmcs = p3_merkle::ExtensionMmcs<BB, BB4, typename S::FlatMmcs>;
        using Pcs = p3_fri::TwoAdicFriPcs<BB, BB4, p3_dft::Radix2Dit<BB>, typename S::FlatMmcs, EFMmcs>;
        auto perm = make_perm<BB, uint32_t, 16>(c, rf, rp);
        typename S::FlatMmcs inner{typename S::Hasher(perm), typename S::Compress(perm)};
        EFMmcs ef_mmcs(inner);
        p3_fri::FriParameters<EFMmcs> params{log_blowup, lfpl, mla, num_queries, commit_pow_bits, query_pow_bits,
                                             ef_mmcs};
        Pcs pcs(p3_dft::Radix2Dit<BB>{}, inner, params);
        std::vector<std::pair<typename Pcs::Domain, p3_matrix::RowMajorMatrix<BB>>> evals;
        for (size_t m = 0; m < n; ++m) {
            typename Pcs::Domain d;
            d.log_n = log_hs[m];
            d.shift = BB(shifts[m]);
            evals.emplace_back(d, load_mat<BB>(mats[m], size_t(1) << log_hs[m], ws[m]));
        }
        auto [commit, pd] = pcs.commit(std::move(evals));
        for (int i = 0; i < 8; ++i) root_out[i] = commit[0][i].value();
        PcsTranscript tr{RefChallenger(perm)};
        for (size_t i = 0; i < n_pre; ++i) tr.inner.observe(BB(pre[i]));
        tr.observe_commitment(commit);
        std::vector<std::vector<BB4>> pts(n);
        size_t k = 0;
        for (size_t m = 0; m < n; ++m)
            for (size_t i = 0; i < n_points[m]; ++i, ++k)
                pts[m].push_back(BB4(std::array<BB, 4>{BB(points[4 * k]), BB(points[4 * k + 1]), BB(points[4 * k + 2]),
                                                       BB(points[4 * k + 3])}));
        std::vector<std::pair<const typename Pcs::PcsProverData*, std::vector<std::vector<BB4>>>> od;
        od.emplace_back(&pd, pts);
        auto [opened, proof] = pcs.open(std::move(od), tr);
        size_t o = 0;
        for (auto& mv : opened[0])
            for (auto& pv : mv)
                for (auto& v : pv)
                    for (int q = 0; q < 4; ++q) opened_out[o++] = v.coeffs[q].value();
        for (size_t r = 0; r < proof.commit_phase_commits.size(); ++r)
            for (int i = 0; i < 8; ++i) roots_out[8 * r + i] = proof.commit_phase_commits[r][0][i].value();
        for (size_t r = 0; r < proof.commit_pow_witnesses.size(); ++r) cwit_out[r] = proof.commit_pow_witnesses[r];
        for (size_t i = 0; i < proof.final_poly.size(); ++i)
            for (int q = 0; q < 4; ++q) final_out[4 * i + q] = proof.final_poly[i].coeffs[q].value();
        *qwit_out = proof.query_pow_witness;
        size_t qr = 0;
        for (size_t q = 0; q < proof.query_proofs.size(); ++q) {
            qidx_out[q] = proof.query_proofs[q].input_proof.row_index;
            for (const BB& v : proof.query_proofs[q].input_proof.row_values) qrows_out[qr++] = v.value();
        }
        for (int i = 0; i < 8; ++i) post_out[i] = tr.inner.sample().value();
        *n_rounds = proof.commit_phase_commits.size();
    });
}

}  // extern "C"
