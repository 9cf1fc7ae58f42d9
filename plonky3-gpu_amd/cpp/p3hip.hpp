// C++ host adapters over libp3hip (include/p3hip.h) with the signatures of the
// reference's DFT / MMCS / FRI-fold API (dloghin/plonky3-gpu, src/include/), so
// the reference's callers (TwoAdicFriPcs, prove_fri, the examples) can take
// them as their template parameters unchanged.
//
//   HipDft<F>                 ~ Radix2Dit<F> / NttCuda<F>      (radix2_dit.hpp:82-233, ntt_cuda.hpp:310-378)
//   HipMerkleTreeMmcs<F, D>   ~ p3_merkle_tree::MerkleTreeMmcs (merkle_tree_mmcs.hpp:75-254)
//   HipFlatMmcs<F, D>         ~ p3_merkle::MerkleTreeMmcs      (fri_merkle_tree_mmcs.hpp:29-168)
//   hip_fold_matrix<EF>       ~ fold_matrix_cuda                (fri_fold_cuda.hpp:306-348)
//   hip_fold_matrix_device    ~ fold_matrix_cuda_device         (fri_fold_cuda.hpp:250-277)
//   HipDuplexChallenger<F>    ~ p3_challenger::DuplexChallenger (duplex_challenger.hpp:42-268)
//
// Matrix types: any type with `values` (std::vector<F>), `width()` and
// `height()` and a (std::vector<F>, size_t width) constructor -- the shape of
// p3_matrix::RowMajorMatrix (dense_matrix.hpp:20-60).  Field types: anything
// with a FieldRaw<F> specialisation converting to and from the canonical
// integer (uint32_t for BabyBear, uint64_t for Goldilocks); the default works
// for types exposing as_canonical_u64() and a uint64_t constructor.
//
// Errors: std::invalid_argument for bad shapes or arguments, std::runtime_error
// for device and budget failures -- the reference's exception types.
#pragma once

#include <array>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "p3hip.h"

namespace p3hip {

inline void check(p3h_ctx* ctx, int rc) {
    if (rc == P3H_OK) return;
    std::string msg = p3h_last_error(ctx);
    if (rc == P3H_EINVAL) throw std::invalid_argument(msg);
    throw std::runtime_error("libp3hip status " + std::to_string(rc) + ": " + msg);
}

// Shared device context (one HIP stream, caching allocator, HBM budget).
class Context {
   public:
    explicit Context(int device = 0, uint64_t hbm_budget = 0) {
        p3h_ctx* c = nullptr;
        int rc = p3h_ctx_create(device, hbm_budget, &c);
        if (rc != P3H_OK) throw std::runtime_error("p3h_ctx_create failed: status " + std::to_string(rc));
        ctx_.reset(c, [](p3h_ctx* p) { p3h_ctx_destroy(p); });
    }
    p3h_ctx* get() const { return ctx_.get(); }

   private:
    std::shared_ptr<p3h_ctx> ctx_;
};

// Canonical integer form of a field element.
template <class F, class Raw>
struct FieldRaw {
    static Raw to(const F& x) { return static_cast<Raw>(x.as_canonical_u64()); }
    static F from(Raw r) { return F(static_cast<uint64_t>(r)); }
};

template <class Raw>
constexpr int field_id() {
    static_assert(sizeof(Raw) == 4 || sizeof(Raw) == 8, "raw type must be uint32_t or uint64_t");
    return sizeof(Raw) == 4 ? P3H_BABYBEAR : P3H_GOLDILOCKS;
}

// RAII device matrix
class DeviceMat {
   public:
    DeviceMat() = default;
    DeviceMat(p3h_ctx* ctx, p3h_mat* m) : ctx_(ctx), m_(m, [](p3h_mat* p) { p3h_mat_free(p); }) {}
    p3h_mat* get() const { return m_.get(); }
    p3h_ctx* ctx() const { return ctx_; }

   private:
    p3h_ctx* ctx_ = nullptr;
    std::shared_ptr<p3h_mat> m_;
};

template <class Raw, class F>
DeviceMat upload(p3h_ctx* ctx, const std::vector<F>& values, size_t width) {
    std::vector<Raw> raw(values.size());
    for (size_t i = 0; i < values.size(); ++i) raw[i] = FieldRaw<F, Raw>::to(values[i]);
    size_t height = width ? values.size() / width : 0;
    p3h_mat* m = nullptr;
    check(ctx, p3h_mat_upload(ctx, field_id<Raw>(), raw.data(), height, uint32_t(width), &m));
    return DeviceMat(ctx, m);
}

template <class Raw, class F>
std::vector<F> download(const DeviceMat& m) {
    int field = 0;
    uint64_t h = 0;
    uint32_t w = 0;
    check(m.ctx(), p3h_mat_info(m.get(), &field, &h, &w));
    std::vector<Raw> raw(h * w);
    check(m.ctx(), p3h_mat_download(m.ctx(), m.get(), raw.data()));
    std::vector<F> out(raw.size());
    for (size_t i = 0; i < raw.size(); ++i) out[i] = FieldRaw<F, Raw>::from(raw[i]);
    return out;
}

// ------------------------------------------------------------------- DFT ---
template <class F, class Raw = uint32_t>
class HipDft {
   public:
    explicit HipDft(Context ctx = Context()) : ctx_(std::move(ctx)) {}

    template <class Mat>
    Mat dft_batch(Mat mat) { return run(0, std::move(mat), F()); }
    template <class Mat>
    Mat idft_batch(Mat mat) { return run(1, std::move(mat), F()); }
    template <class Mat>
    Mat coset_dft_batch(Mat mat, const F& shift) { return run(2, std::move(mat), shift); }
    template <class Mat>
    Mat coset_idft_batch(Mat mat, const F& shift) { return run(3, std::move(mat), shift); }

    template <class Mat>
    Mat coset_lde_batch(Mat mat, size_t added_bits, const F& shift) {
        p3h_ctx* c = ctx_.get();
        size_t w = mat.width();
        DeviceMat in = upload<Raw>(c, mat.values, w);
        p3h_mat* out = nullptr;
        check(c, p3h_coset_lde(c, in.get(), uint32_t(added_bits), FieldRaw<F, Raw>::to(shift), &out));
        DeviceMat o(c, out);
        return Mat(download<Raw, F>(o), w);
    }

    const Context& context() const { return ctx_; }

   private:
    template <class Mat>
    Mat run(int kind, Mat mat, const F& shift) {
        if (mat.height() <= 1) return mat;  // the reference returns such inputs unchanged
        p3h_ctx* c = ctx_.get();
        size_t w = mat.width();
        DeviceMat in = upload<Raw>(c, mat.values, w);
        p3h_mat* out = nullptr;
        uint64_t s = kind >= 2 ? uint64_t(FieldRaw<F, Raw>::to(shift)) : 1;
        check(c, p3h_dft(c, kind, in.get(), s, &out));
        DeviceMat o(c, out);
        return Mat(download<Raw, F>(o), w);
    }
    Context ctx_;
};

// -------------------------------------------------------------- Poseidon2 --
// Constants in the Poseidon2Cuda field order: initial[rf/2][W], terminal[rf/2][W], internal[rp].
template <class Raw>
void set_poseidon2(const Context& ctx, const std::vector<Raw>& consts, uint32_t rounds_f, uint32_t rounds_p) {
    constexpr int field = field_id<Raw>();
    uint32_t width = field == P3H_BABYBEAR ? 16 : 8;
    if (consts.size() != size_t(rounds_f) * width + rounds_p)
        throw std::invalid_argument("Poseidon2 constant count does not match rounds");
    check(ctx.get(), p3h_poseidon2_set(ctx.get(), field, width, rounds_f, rounds_p, consts.data()));
}

// ------------------------------------------------------------------ MMCS ---
template <class F, size_t DIGEST>
struct HipBatchOpening {
    std::vector<std::vector<F>> opened_values;
    std::vector<std::array<F, DIGEST>> proof;
};

template <class F, size_t DIGEST>
struct HipCap {
    std::vector<std::array<F, DIGEST>> cap;
    bool operator==(const HipCap& o) const { return cap == o.cap; }
};

// Device-resident prover data: the tree keeps the committed matrices.
class HipTree {
   public:
    HipTree() = default;
    HipTree(p3h_ctx* ctx, p3h_tree* t) : ctx_(ctx), t_(t, [](p3h_tree* p) { p3h_tree_free(p); }) {}
    p3h_tree* get() const { return t_.get(); }
    p3h_ctx* ctx() const { return ctx_; }
    // widths of the committed matrices in commit order, as the native tree records them
    std::vector<size_t> widths() const {
        uint32_t n = 0;
        check(ctx_, p3h_tree_matrices(t_.get(), &n, nullptr, nullptr));
        std::vector<uint32_t> w(n);
        check(ctx_, p3h_tree_matrices(t_.get(), &n, nullptr, w.data()));
        return std::vector<size_t>(w.begin(), w.end());
    }

   private:
    p3h_ctx* ctx_ = nullptr;
    std::shared_ptr<p3h_tree> t_;
};

template <class F, class Raw, size_t DIGEST>
std::vector<std::array<F, DIGEST>> digests_of(const std::vector<Raw>& raw) {
    std::vector<std::array<F, DIGEST>> out(raw.size() / DIGEST);
    for (size_t i = 0; i < out.size(); ++i)
        for (size_t k = 0; k < DIGEST; ++k) out[i][k] = FieldRaw<F, Raw>::from(raw[i * DIGEST + k]);
    return out;
}

template <class F, class Raw, size_t DIGEST>
std::vector<std::array<F, DIGEST>> tree_cap(const HipTree& t) {
    uint32_t n = 0;
    check(t.ctx(), p3h_tree_cap(t.ctx(), t.get(), nullptr, &n));
    std::vector<Raw> raw(size_t(n) * DIGEST);
    check(t.ctx(), p3h_tree_cap(t.ctx(), t.get(), raw.data(), &n));
    return digests_of<F, Raw, DIGEST>(raw);
}

// Multi-height MMCS with cap (p3_merkle_tree::MerkleTreeMmcs).
template <class F, class Raw = uint32_t, size_t DIGEST = (sizeof(Raw) == 4 ? 8 : 4)>
class HipMerkleTreeMmcs {
   public:
    using Commitment = HipCap<F, DIGEST>;
    using ProverData = HipTree;
    using Opening = HipBatchOpening<F, DIGEST>;

    explicit HipMerkleTreeMmcs(Context ctx, size_t cap_height = 0) : ctx_(std::move(ctx)), cap_height_(cap_height) {}

    template <class Mat>
    std::pair<Commitment, ProverData> commit(std::vector<Mat> matrices) {
        p3h_ctx* c = ctx_.get();
        std::vector<DeviceMat> dev;
        std::vector<const p3h_mat*> ptrs;
        for (auto& m : matrices) {
            dev.push_back(upload<Raw>(c, m.values, m.width()));
            ptrs.push_back(dev.back().get());
        }
        p3h_tree* t = nullptr;
        check(c, p3h_mmcs_commit(c, ptrs.data(), uint32_t(ptrs.size()), uint32_t(cap_height_), &t));
        HipTree tree(c, t);
        return {Commitment{tree_cap<F, Raw, DIGEST>(tree)}, tree};
    }

    Opening open_batch(size_t index, const ProverData& tree) const {
        uint32_t n_layers = 0, total_w = 0;
        uint64_t leaves = 0;
        check(tree.ctx(), p3h_tree_info(tree.get(), &n_layers, &leaves, &total_w));
        std::vector<Raw> vals(total_w), proof(size_t(n_layers - 1) * DIGEST);
        uint64_t idx = index;
        check(tree.ctx(), p3h_tree_open(tree.ctx(), tree.get(), &idx, 1, vals.data(), proof.data()));
        Opening op;
        op.proof = digests_of<F, Raw, DIGEST>(proof);
        size_t at = 0;  // one row per committed matrix, in commit order
        for (size_t w : tree.widths()) {
            op.opened_values.emplace_back();
            for (size_t k = 0; k < w; ++k) op.opened_values.back().push_back(FieldRaw<F, Raw>::from(vals[at + k]));
            at += w;
        }
        return op;
    }

    // verify_batch (merkle_tree_mmcs.hpp:168-254); Dims has .width and .height
    // (p3_matrix::Dimensions), in commit order.
    template <class Dims>
    bool verify_batch(const Commitment& commitment, const std::vector<Dims>& dims, size_t index,
                      const Opening& opening) const {
        if (dims.size() != opening.opened_values.size()) return false;
        std::vector<uint64_t> heights;
        std::vector<uint32_t> widths;
        std::vector<Raw> vals, cap, proof;
        for (size_t i = 0; i < dims.size(); ++i) {
            if (opening.opened_values[i].size() != dims[i].width) return false;
            heights.push_back(dims[i].height);
            widths.push_back(uint32_t(dims[i].width));
            for (const F& x : opening.opened_values[i]) vals.push_back(FieldRaw<F, Raw>::to(x));
        }
        for (const auto& d : commitment.cap)
            for (const F& x : d) cap.push_back(FieldRaw<F, Raw>::to(x));
        for (const auto& d : opening.proof)
            for (const F& x : d) proof.push_back(FieldRaw<F, Raw>::to(x));
        int ok = 0;
        check(ctx_.get(), p3h_mmcs_verify(ctx_.get(), field_id<Raw>(), cap.data(), uint32_t(commitment.cap.size()),
                                          heights.data(), widths.data(), uint32_t(dims.size()), index, vals.data(),
                                          proof.data(), uint32_t(opening.proof.size()), &ok));
        return ok != 0;
    }

   private:
    Context ctx_;
    size_t cap_height_;
};

// Single-matrix MMCS (p3_merkle::MerkleTreeMmcs): commitment = {root}.
template <class F, class Raw = uint32_t, size_t DIGEST = (sizeof(Raw) == 4 ? 8 : 4)>
class HipFlatMmcs {
   public:
    using Digest = std::array<F, DIGEST>;
    using Commitment = std::vector<Digest>;
    using OpeningProof = std::vector<Digest>;
    struct ProverData {
        HipTree tree;
        size_t height = 0, width = 0;
    };

    explicit HipFlatMmcs(Context ctx) : ctx_(std::move(ctx)) {}

    std::pair<Commitment, ProverData> commit_matrix(const std::vector<F>& data, size_t width) const {
        if (width == 0 || data.size() % width != 0) throw std::invalid_argument("commit_matrix: bad dimensions");
        if (data.empty()) throw std::invalid_argument("commit_matrix: empty matrix");
        p3h_ctx* c = ctx_.get();
        DeviceMat m = upload<Raw>(c, data, width);
        p3h_tree* t = nullptr;
        check(c, p3h_flat_commit(c, m.get(), &t));
        ProverData pd{HipTree(c, t), data.size() / width, width};
        return {tree_cap<F, Raw, DIGEST>(pd.tree), pd};
    }

    // get_row_as_vals (fri_merkle_tree_mmcs.hpp:115-118): the committed row,
    // read from the device-resident matrix; ExtensionMmcs::get_row calls it.
    std::vector<F> get_row_as_vals(const ProverData& d, size_t row) const {
        if (row >= d.height) throw std::invalid_argument("get_row_as_vals: row out of range");
        std::vector<Raw> vals(d.width), proof;
        uint32_t n_layers = 0;
        check(d.tree.ctx(), p3h_tree_info(d.tree.get(), &n_layers, nullptr, nullptr));
        proof.resize(size_t(n_layers - 1) * DIGEST + 1);
        uint64_t idx = row;
        check(d.tree.ctx(), p3h_tree_open(d.tree.ctx(), d.tree.get(), &idx, 1, vals.data(), proof.data()));
        std::vector<F> out;
        for (Raw r : vals) out.push_back(FieldRaw<F, Raw>::from(r));
        return out;
    }
    // get_row: the reference's alias of get_row_as_vals (fri_merkle_tree_mmcs.hpp:120-124)
    std::vector<F> get_row(const ProverData& d, size_t row) const { return get_row_as_vals(d, row); }

    void open_row(const ProverData& d, size_t row_index, OpeningProof& proof) const {
        uint32_t n_layers = 0;
        check(d.tree.ctx(), p3h_tree_info(d.tree.get(), &n_layers, nullptr, nullptr));
        std::vector<Raw> raw(size_t(n_layers - 1) * DIGEST + 1);
        uint64_t idx = row_index;
        check(d.tree.ctx(), p3h_tree_open(d.tree.ctx(), d.tree.get(), &idx, 1, nullptr, raw.data()));
        raw.resize(size_t(n_layers - 1) * DIGEST);
        proof = digests_of<F, Raw, DIGEST>(raw);
    }

    // verify_row (fri_merkle_tree_mmcs.hpp:146-168)
    bool verify_row(const Commitment& commit, size_t row_index, const std::vector<F>& row_vals,
                    const OpeningProof& proof) const {
        std::vector<Raw> root, row, pr;
        for (const auto& d : commit)
            for (const F& x : d) root.push_back(FieldRaw<F, Raw>::to(x));
        for (const F& x : row_vals) row.push_back(FieldRaw<F, Raw>::to(x));
        for (const auto& d : proof)
            for (const F& x : d) pr.push_back(FieldRaw<F, Raw>::to(x));
        int ok = 0;
        check(ctx_.get(), p3h_flat_verify(ctx_.get(), field_id<Raw>(), root.data(), uint32_t(commit.size()),
                                          row_index, row.data(), uint32_t(row.size()), pr.data(),
                                          uint32_t(proof.size()), &ok));
        return ok != 0;
    }

    size_t log_height(const ProverData& d) const { return log2_strict(d.height); }
    size_t log_width(const ProverData& d) const { return log2_strict(d.width); }

   private:
    static size_t log2_strict(size_t n) {
        if (n == 0 || (n & (n - 1))) throw std::invalid_argument("log2_strict_usize: value must be a non-zero power of 2");
        return size_t(__builtin_ctzll(n));
    }
    Context ctx_;
};

// ------------------------------------------------------------------- FRI ---
// EF: BabyBear^4 with `coeffs` (std::array of 4 base elements), as
// BinomialExtensionField (extension_field.hpp:32-60).
template <class EF>
std::vector<EF> hip_fold_matrix(const Context& ctx, const EF& beta, size_t log_arity, size_t log_height,
                                const std::vector<EF>& current) {
    using Base = typename std::decay<decltype(current[0].coeffs[0])>::type;
    p3h_ctx* c = ctx.get();
    if (current.size() != (size_t(1) << log_height)) throw std::invalid_argument("fold: size is not 2^log_height");
    std::vector<Base> flat;
    for (const EF& e : current)
        for (const Base& b : e.coeffs) flat.push_back(b);
    DeviceMat m = upload<uint32_t>(c, flat, 4);
    uint32_t b[4];
    for (int i = 0; i < 4; ++i) b[i] = FieldRaw<Base, uint32_t>::to(beta.coeffs[i]);
    p3h_mat* out = nullptr;
    check(c, p3h_fri_fold(c, m.get(), uint32_t(log_arity), b, &out));
    DeviceMat o(c, out);
    std::vector<Base> vals = download<uint32_t, Base>(o);
    std::vector<EF> res(vals.size() / 4);
    for (size_t i = 0; i < res.size(); ++i)
        for (int k = 0; k < 4; ++k) res[i].coeffs[k] = vals[4 * i + k];
    return res;
}

// fold_matrix_cuda_device (fri_fold_cuda.hpp:250-277): fold n rows of
// 2^log_arity values from device buffer d_input into d_output, both already
// on the device; nothing is copied to or from the host.  EF must be laid out
// as 4 canonical uint32 coefficients (BinomialExtensionField<BabyBear, 4>).
template <class Val, class EF>
void hip_fold_matrix_device(const Context& ctx, const EF* d_input, EF* d_output, const EF& beta, size_t log_arity,
                            size_t log_height, size_t n) {
    static_assert(sizeof(EF) == 4 * sizeof(uint32_t), "EF must be four 32-bit coefficients");
    uint32_t b[4];
    for (int i = 0; i < 4; ++i) b[i] = FieldRaw<Val, uint32_t>::to(beta.coeffs[i]);
    check(ctx.get(), p3h_fri_fold_device(ctx.get(), reinterpret_cast<const uint32_t*>(d_input),
                                         reinterpret_cast<uint32_t*>(d_output), b, uint32_t(log_arity),
                                         uint32_t(log_height), uint64_t(n)));
}

// ------------------------------------------------------------ challenger ---
// DuplexChallenger<F, Perm, 16, 8> (duplex_challenger.hpp:42-268) for BabyBear
// with the context's Poseidon2 constants, its state on the device.  The method
// names are the reference's, so FriChallenger-style wrappers
// (tests/fri_test.cpp:99-124) compile unchanged.
template <class F>
class HipDuplexChallenger {
   public:
    explicit HipDuplexChallenger(Context ctx) : ctx_(std::move(ctx)) {
        p3h_chal* h = nullptr;
        check(ctx_.get(), p3h_challenger_create(ctx_.get(), &h));
        h_.reset(h, [](p3h_chal* p) { p3h_challenger_free(p); });
    }
    HipDuplexChallenger(const HipDuplexChallenger& o) : ctx_(o.ctx_) {  // deep copy, as the reference's
        p3h_chal* h = nullptr;
        check(ctx_.get(), p3h_challenger_clone(o.h_.get(), &h));
        h_.reset(h, [](p3h_chal* p) { p3h_challenger_free(p); });
    }
    HipDuplexChallenger& operator=(const HipDuplexChallenger& o) {
        if (this != &o) *this = HipDuplexChallenger(o);
        return *this;
    }
    HipDuplexChallenger(HipDuplexChallenger&&) = default;
    HipDuplexChallenger& operator=(HipDuplexChallenger&&) = default;

    void observe(const F& v) {
        uint32_t r = FieldRaw<F, uint32_t>::to(v);
        check(ctx_.get(), p3h_challenger_observe(h_.get(), &r, 1));
    }
    void observe_slice(const std::vector<F>& vs) {
        std::vector<uint32_t> r;
        for (const F& v : vs) r.push_back(FieldRaw<F, uint32_t>::to(v));
        check(ctx_.get(), p3h_challenger_observe(h_.get(), r.data(), uint32_t(r.size())));
    }
    template <class Hash>
    void observe_merkle_cap(const std::vector<Hash>& cap) {
        std::vector<uint32_t> r;
        for (const Hash& h : cap)
            for (const F& v : h) r.push_back(FieldRaw<F, uint32_t>::to(v));
        check(ctx_.get(), p3h_challenger_observe(h_.get(), r.data(), uint32_t(r.size())));
    }
    template <class EF>
    void observe_algebra_element(const EF& e) {
        std::vector<F> v(e.coeffs.begin(), e.coeffs.end());
        observe_slice(v);
    }
    F sample() {
        uint32_t r = 0;
        check(ctx_.get(), p3h_challenger_sample(h_.get(), 1, &r));
        return FieldRaw<F, uint32_t>::from(r);
    }
    size_t sample_bits(size_t bits) {
        uint64_t r = 0;
        check(ctx_.get(), p3h_challenger_sample_bits(h_.get(), uint32_t(bits), &r));
        return size_t(r);
    }
    template <class EF>
    EF sample_algebra_element() {
        uint32_t r[4];
        check(ctx_.get(), p3h_challenger_sample(h_.get(), 4, r));
        EF e{};
        for (int i = 0; i < 4; ++i) e.coeffs[i] = FieldRaw<F, uint32_t>::from(r[i]);
        return e;
    }
    uint64_t grind(size_t bits) {
        uint64_t w = 0;
        check(ctx_.get(), p3h_challenger_grind(h_.get(), uint32_t(bits), &w));
        return w;
    }
    bool check_witness(size_t bits, uint64_t w) {
        int ok = 0;
        check(ctx_.get(), p3h_challenger_check_witness(h_.get(), uint32_t(bits), w, &ok));
        return ok != 0;
    }
    p3h_chal* handle() const { return h_.get(); }

   private:
    Context ctx_;
    std::shared_ptr<p3h_chal> h_;
};

}  // namespace p3hip
