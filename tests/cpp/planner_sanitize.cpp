// Sanitizer driver for the host-only planners of libp3hip: plan_lde_batch
// (csrc/ntt.hip) over many batches of LDE jobs of both fields, and the serving
// queue's footprint and admission rule (csrc/host_plan.h).  Built by
// tests/test_sanitize.py with hipcc, the host code under
// -fsanitize=address,undefined; no device call is made (the job pointers are
// never dereferenced on the host).  Besides the sanitizers it checks the
// plans' structure: work descriptors in ascending block order covering each
// group's grid, tiles within the matrix, dynamic LDS within a workgroup's
// 160 KiB; and that admission is FIFO, within max_batch and within the budget.
#include "host_plan.h"
#include "ntt.h"

#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

using namespace p3h;

namespace {

int fails = 0;
void expect(bool ok, const char* what) {
    if (!ok && fails++ < 20) std::printf("check failed: %s\n", what);
}

template <class F>
void plan_many(std::mt19937_64& rng, int rounds) {
    using T = typename F::T;
    const uint32_t max_l = sizeof(T) == 4 ? 22 : 20;
    for (int r = 0; r < rounds; ++r) {
        std::vector<LdeJob<T>> jobs;
        const int n = 1 + int(rng() % 12);
        for (int i = 0; i < n; ++i) {
            LdeJob<T> j{};
            j.L = uint32_t(rng() % (max_l + 1));
            j.w = 1 + uint32_t(rng() % 300);
            j.k = uint32_t(rng() % 4);
            // distinct fake device addresses: only copied into descriptors
            j.in = reinterpret_cast<const T*>(uintptr_t(0x100000) * (2 * i + 1));
            j.out = reinterpret_cast<T*>(uintptr_t(0x100000) * (2 * i + 2));
            j.tw = reinterpret_cast<const T*>(uintptr_t(0x7000000));
            j.twist = reinterpret_cast<const T*>(uintptr_t(0x8000000));
            j.tune.log_cols = uint32_t(rng() % 8);
            j.tune.max_stages = uint32_t(rng() % 11);
            j.tune.log_slab = 0;
            jobs.push_back(j);
        }
        const auto groups = plan_lde_batch<F>(jobs);
        expect(!groups.empty(), "a plan has at least one group");
        for (const auto& g : groups) {
            expect(!g.work.empty(), "a group has work");
            expect(g.lds_bytes <= 160u * 1024u, "dynamic LDS within one workgroup");
            uint32_t prev = 0;
            for (size_t i = 0; i < g.work.size(); ++i) {
                const auto& wk = g.work[i];
                expect(wk.block_start >= prev, "descriptors in ascending block order");
                prev = wk.block_start;
                expect(wk.args.w >= 1 && wk.args.w <= wk.args.ld, "columns within the row stride");
                expect(wk.args.a + g.s <= wk.args.L || wk.args.L == 0, "pass bits within the transform");
                expect((wk.args.n_col_tiles << wk.args.logC) >= wk.args.w, "column tiles cover the width");
            }
            expect(g.work.front().block_start == 0, "the first descriptor starts at block 0");
            expect(g.work.back().block_start < g.blocks, "every descriptor has blocks");
        }
    }
}

void admission(std::mt19937_64& rng, int rounds) {
    for (int r = 0; r < rounds; ++r) {
        std::vector<uint64_t> fp;
        const size_t n = rng() % 20;
        for (size_t i = 0; i < n; ++i) {
            const uint64_t h = uint64_t(1) << (rng() % 21);
            fp.push_back(serve_footprint(h, 1 + uint32_t(rng() % 256), uint32_t(rng() % 3)));
        }
        const uint64_t budget = (rng() % 4) == 0 ? ~uint64_t(0) : (uint64_t(1) << (20 + rng() % 16));
        const uint64_t reserved = budget == ~uint64_t(0) ? rng() : rng() % (budget + 1);
        const size_t max_batch = 1 + rng() % 8;
        const size_t take = serve_admit(reserved, budget, max_batch, fp.data(), fp.size());
        expect(take <= max_batch && take <= n, "admission within max_batch and the queue");
        uint64_t sum = reserved;
        for (size_t i = 0; i < take; ++i) sum += fp[i];
        expect(sum >= reserved && sum <= budget, "admitted requests fit the budget");
        if (take < n && take < max_batch) expect(fp[take] > budget - sum, "admission stops at the first misfit");
    }
    for (uint64_t b = 0; b < (uint64_t(1) << 24); b = b * 3 + 1) {
        const size_t c = size_class(b);
        expect(c >= b && c >= 256, "a size class holds the request");
    }
}

}  // namespace

int main() {
    std::mt19937_64 rng(20240601);
    plan_many<BabyBear>(rng, 400);
    plan_many<Goldilocks>(rng, 200);
    admission(rng, 5000);
    if (fails) {
        std::printf("%d checks failed\n", fails);
        return 1;
    }
    std::printf("ok\n");
    return 0;
}
