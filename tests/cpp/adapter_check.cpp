// Drives the C++ adapters (plonky3-gpu_amd/cpp/p3hip.hpp) the way the
// reference's callers drive Radix2Dit / MerkleTreeMmcs / fold_matrix_cuda, with
// minimal stand-ins for the reference's BabyBear and RowMajorMatrix types.
// Prints results as text; tests/test_gpu_cpp_adapter.py checks them against
// the oracle on the same seeded inputs.
#include <cstdio>
#include <random>

#include "p3hip.hpp"

struct BB {  // canonical BabyBear value, like p3_field::BabyBear
    uint32_t v = 0;
    BB() = default;
    explicit BB(uint64_t x) : v(uint32_t(x % 0x78000001u)) {}
    uint64_t as_canonical_u64() const { return v; }
};
struct BB4 {
    std::array<BB, 4> coeffs;
};
template <class F>
struct RowMajorMatrix {  // shape of p3_matrix::RowMajorMatrix
    std::vector<F> values;
    size_t width_;
    RowMajorMatrix(std::vector<F> v, size_t w) : values(std::move(v)), width_(w) {}
    size_t width() const { return width_; }
    size_t height() const { return width_ ? values.size() / width_ : 0; }
};

static std::vector<BB> fill(uint64_t seed, size_t n) {
    std::mt19937_64 rng(seed);
    std::vector<BB> out(n);
    for (auto& x : out) x = BB(rng() % 0x78000001u);
    return out;
}

int main() {
    p3hip::Context ctx(0);
    std::vector<uint32_t> consts;
    for (const BB& x : fill(1 ^ 0x5EEDC0DE00000002ull, 8 * 16 + 13)) consts.push_back(x.v);
    p3hip::set_poseidon2<uint32_t>(ctx, consts, 8, 13);

    RowMajorMatrix<BB> trace(fill(5, 1024 * 8), 8);
    p3hip::HipDft<BB> dft(ctx);
    auto lde = dft.coset_lde_batch(trace, 1, BB(31));
    std::printf("lde_rows %zu\n", lde.height());
    std::printf("lde_row3");
    for (size_t c = 0; c < 8; ++c) std::printf(" %u", lde.values[3 * 8 + c].v);
    std::printf("\n");
    auto d = dft.dft_batch(RowMajorMatrix<BB>(fill(6, 64 * 2), 2));
    std::printf("dft_row1 %u %u\n", d.values[2].v, d.values[3].v);

    p3hip::HipFlatMmcs<BB> flat(ctx);
    auto [root, pd] = flat.commit_matrix(lde.values, 8);
    std::printf("flat_root");
    for (const BB& x : root[0]) std::printf(" %u", x.v);
    std::printf("\n");
    std::vector<std::array<BB, 8>> proof;
    flat.open_row(pd, 77, proof);
    std::printf("flat_proof_len %zu first %u\n", proof.size(), proof[0][0].v);
    auto row77 = flat.get_row(pd, 77);
    bool flat_ok = flat.verify_row(root, 77, row77, proof);
    row77[5] = BB(row77[5].v + 1);
    std::printf("flat_verify %d %d\n", int(flat_ok), int(flat.verify_row(root, 77, row77, proof)));

    p3hip::HipMerkleTreeMmcs<BB> mmcs(ctx, 1);
    std::vector<RowMajorMatrix<BB>> mats{lde, trace};
    auto [cap, tree] = mmcs.commit(mats);
    std::printf("mmcs_cap");
    for (const auto& dg : cap.cap)
        for (const BB& x : dg) std::printf(" %u", x.v);
    std::printf("\n");
    auto op = mmcs.open_batch(9, tree);
    std::printf("mmcs_open_len %zu %zu %zu proof_len %zu\n", op.opened_values.size(), op.opened_values[0].size(),
                op.opened_values[1].size(), op.proof.size());
    struct Dim {
        size_t width, height;
    };
    std::vector<Dim> dims{{8, 2048}, {8, 1024}};
    const bool honest = mmcs.verify_batch(cap, dims, 9, op);
    auto bad_value = op;
    bad_value.opened_values[1][2] = BB(bad_value.opened_values[1][2].v + 1);
    auto bad_sibling = op;
    bad_sibling.proof[4][0] = BB(bad_sibling.proof[4][0].v + 1);
    std::printf("mmcs_verify %d %d %d %d\n", int(honest), int(mmcs.verify_batch(cap, dims, 9, bad_value)),
                int(mmcs.verify_batch(cap, dims, 9, bad_sibling)), int(mmcs.verify_batch(cap, dims, 8, op)));

    std::vector<BB4> cw(64);
    auto vals = fill(7, 64 * 4);
    for (size_t i = 0; i < 64; ++i)
        for (int k = 0; k < 4; ++k) cw[i].coeffs[k] = vals[4 * i + k];
    BB4 beta{{BB(3), BB(1), BB(4), BB(1)}};
    auto folded = p3hip::hip_fold_matrix(ctx, beta, 1, 6, cw);
    std::printf("fold0 %u %u %u %u\n", folded[0].coeffs[0].v, folded[0].coeffs[1].v, folded[0].coeffs[2].v,
                folded[0].coeffs[3].v);

    p3hip::HipDuplexChallenger<BB> chal(ctx);
    chal.observe_slice(fill(9, 11));
    chal.observe_merkle_cap(root);
    BB4 z = chal.sample_algebra_element<BB4>();
    auto copy = chal;  // deep copy: both continue from the same transcript
    std::printf("chal_ext %u %u %u %u\n", z.coeffs[0].v, z.coeffs[1].v, z.coeffs[2].v, z.coeffs[3].v);
    uint64_t w = chal.grind(6);
    std::printf("chal_grind %llu %u %u\n", (unsigned long long)w, chal.sample().v, copy.sample().v);

    try {
        dft.coset_lde_batch(RowMajorMatrix<BB>(fill(8, 3 * 4), 4), 1, BB(31));
        std::printf("error_not_raised\n");
    } catch (const std::invalid_argument&) {
        std::printf("invalid_argument_raised\n");
    }
    return 0;
}
