// Compile check: the reference's own PCS and FRI code, unchanged, over this
// library's adapters (plonky3-gpu_amd/cpp/p3hip.hpp).  It instantiates
//   TwoAdicFriPcs<BabyBear, BabyBear4, HipDft<BabyBear>, HipFlatMmcs<BabyBear>,
//                 ExtensionMmcs<BabyBear, BabyBear4, HipFlatMmcs<BabyBear>>>
// with commit, open and verify (two_adic_fri_pcs.hpp:207-560; open drives
// prove_fri, fri_prover.hpp:47-280, and verify the reference's FRI verifier),
// a FriChallenger over HipDuplexChallenger shaped as the reference's test
// wrapper (tests/fri_test.cpp:99-124), the multi-height MMCS adapter beside
// the reference's MerkleTreeMmcs call pattern, and the device-pointer fold.
// The flow is the reference's FRI round trip (tests/fri_test.cpp:186-270).
//
// tests/test_reference_compile.py builds it with -fsyntax-only against the
// reference headers where they lie (skipped when they are absent); nothing is
// linked or run.
#include "baby_bear.hpp"
#include "dense_matrix.hpp"
#include "extension_field.hpp"
#include "fri_extension_mmcs.hpp"
#include "fri_params.hpp"
#include "fri_proof.hpp"
#include "fri_prover.hpp"
#include "fri_verifier.hpp"
#include "two_adic_fri_pcs.hpp"

#include "p3hip.hpp"

#include <cstdint>
#include <utility>
#include <vector>

using BB = p3_field::BabyBear;
using BB4 = p3_field::BabyBear4;

using Dft = p3hip::HipDft<BB>;
using ValMmcs = p3hip::HipFlatMmcs<BB>;
using EFMmcs = p3_merkle::ExtensionMmcs<BB, BB4, ValMmcs>;
using FriParams = p3_fri::FriParameters<EFMmcs>;
using Pcs = p3_fri::TwoAdicFriPcs<BB, BB4, Dft, ValMmcs, EFMmcs>;

// The reference test's FriChallenger, with the device transcript inside.
struct FriChallenger {
    p3hip::HipDuplexChallenger<BB> inner;

    explicit FriChallenger(p3hip::Context ctx) : inner(std::move(ctx)) {}

    template <typename Commit>
    void observe_commitment(const Commit& c) { inner.observe_merkle_cap(c); }
    template <typename EF = BB4>
    EF sample_challenge() { return inner.template sample_algebra_element<EF>(); }
    template <typename EF>
    void observe_challenge(const EF& c) { inner.template observe_algebra_element<EF>(c); }
    void observe_arity(size_t la) { inner.observe(BB(static_cast<uint32_t>(la))); }
    uint64_t grind(size_t bits) { return inner.grind(bits); }
    bool check_witness(size_t bits, uint64_t w) { return inner.check_witness(bits, w); }
    size_t sample_bits(size_t n) { return inner.sample_bits(n); }
    void observe(BB v) { inner.observe(v); }
    void observe_slice(const std::vector<BB>& vals) { inner.observe_slice(vals); }
    BB sample() { return inner.sample(); }
};

bool pcs_round_trip(p3hip::Context ctx, std::vector<p3_matrix::RowMajorMatrix<BB>> mats) {
    using Domain = p3_fri::TwoAdicMultiplicativeCoset<BB>;
    FriParams fri{1, 2, 1, 10, 0, 8, EFMmcs(ValMmcs(ctx))};
    Pcs pcs{Dft{ctx}, ValMmcs{ctx}, std::move(fri)};

    std::vector<std::pair<Domain, p3_matrix::RowMajorMatrix<BB>>> evals;
    std::vector<Domain> domains;
    for (auto& m : mats) {
        Domain d;
        d.log_n = p3_util::log2_strict_usize(m.height());
        d.shift = BB::one_val();
        domains.push_back(d);
        evals.push_back({d, std::move(m)});
    }
    FriChallenger prover(ctx);
    auto [commitment, prover_data] = pcs.commit(std::move(evals));
    prover.observe_commitment(commitment);
    BB4 zeta = prover.template sample_challenge<BB4>();
    std::vector<std::vector<BB4>> points(domains.size(), std::vector<BB4>{zeta});
    std::vector<std::pair<const Pcs::PcsProverData*, std::vector<std::vector<BB4>>>> open_data;
    open_data.push_back({&prover_data, points});
    auto [opened, proof] = pcs.open(open_data, prover);

    FriChallenger verifier(ctx);
    verifier.observe_commitment(commitment);
    BB4 vz = verifier.template sample_challenge<BB4>();
    Pcs::VerifyCommitment vc;
    vc.commitment = commitment;
    for (size_t i = 0; i < domains.size(); ++i) {
        vc.domains.push_back(domains[i]);
        vc.points.push_back({vz});
        vc.opened_values.push_back(opened[0][i]);
    }
    return pcs.verify({vc}, proof, verifier);
}

// The multi-height MMCS adapter in the reference's MerkleTreeMmcs call
// pattern (merkle_tree_mmcs.hpp:100-254), and the device-pointer fold.
bool mmcs_and_fold(p3hip::Context ctx, std::vector<p3_matrix::RowMajorMatrix<BB>> mats, const BB4* d_in, BB4* d_out,
                   const BB4& beta) {
    p3hip::HipMerkleTreeMmcs<BB> mmcs(ctx, 1);
    std::vector<p3_matrix::Dimensions> dims;
    for (const auto& m : mats) dims.push_back(p3_matrix::Dimensions{m.width(), m.height()});
    auto [cap, tree] = mmcs.commit(std::move(mats));
    auto opening = mmcs.open_batch(3, tree);
    p3hip::hip_fold_matrix_device<BB, BB4>(ctx, d_in, d_out, beta, 1, 10, 512);
    return mmcs.verify_batch(cap, dims, 3, opening);
}
