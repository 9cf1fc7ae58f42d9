// Sanitizer driver for the host-only planners of libp3hip: plan_lde_batch and
// plan_lde_queue (csrc/ntt_impl.h) over many batches of LDE jobs, and the serving
// queue's footprint and admission rule (csrc/host_plan.h).  Built by
// tests/test_sanitize.py with hipcc, the host code under
// -fsanitize=address,undefined; no device call is made (the job pointers are
// never dereferenced on the host).  Besides the sanitizers it checks the
// plans' structure: work descriptors in ascending block order covering each
// group's grid, tiles within the matrix, dynamic LDS within a workgroup's
// 160 KiB; the queue plan's tickets in dependency order, covering the same
// blocks as the grouped plan, and the queue kernel's protocol (take a ticket,
// wait for the pass it reads, count the block done) run to completion by a
// random interleaving with 1 to 64 resident workgroups; and that admission is FIFO,
// within max_batch and within the budget.
#include "host_plan.h"
#include "ntt.h"

#include <cstdint>
#include <cstdio>
#include <random>
#include <type_traits>
#include <vector>

using namespace p3h;

namespace {

int fails = 0;
void expect(bool ok, const char* what) {
    if (!ok && fails++ < 20) std::printf("check failed: %s\n", what);
}

int protocol_plans = 0;    // queue plans whose protocol was run
int protocol_chunked = 0;  // of them, plans with chunks of more than one block

// The queue kernel's protocol on the host: g resident workgroup slots, one
// chosen at random takes a step at a time -- an empty slot receives the next
// workgroup, which takes the next ticket (the slot retires when none is
// left); a workgroup whose ticket's pass dependency is complete runs its
// block, counts it done and ends, freeing its slot.  Returns false on a state
// where no slot can step before all have retired (a deadlock), or when the
// counts do not end at every block.
template <class T>
bool run_queue_protocol(const LdeQueuePlan<T>& q, uint32_t g, std::mt19937_64& rng) {
    std::vector<uint32_t> done(q.items.size(), 0);
    std::vector<int64_t> held(g, -1);
    std::vector<bool> left(g, false);
    uint32_t next = 0, gone = 0;
    auto item_of = [&](uint32_t t) {
        size_t lo = 0, hi = q.items.size();
        while (hi - lo > 1) {
            const size_t mid = (lo + hi) / 2;
            if (q.items[mid].first <= t) lo = mid;
            else hi = mid;
        }
        return lo;
    };
    std::vector<uint32_t> can;
    while (gone < g) {
        can.clear();
        for (uint32_t w = 0; w < g; ++w) {
            if (left[w]) continue;
            if (held[w] < 0) {
                can.push_back(w);
            } else {
                const auto& it = q.items[item_of(uint32_t(held[w]))];
                if (it.dep < 0 || done[size_t(it.dep)] == q.items[size_t(it.dep)].blocks) can.push_back(w);
            }
        }
        if (can.empty()) return false;
        const uint32_t w = can[rng() % can.size()];
        if (held[w] < 0) {
            if (next >= q.total) {
                left[w] = true;
                ++gone;
            } else {
                held[w] = next++;
            }
        } else {
            const auto& it = q.items[item_of(uint32_t(held[w]))];
            const uint32_t b0 = (uint32_t(held[w]) - it.first) * it.chunk;
            done[item_of(uint32_t(held[w]))] += (b0 + it.chunk < it.blocks ? it.chunk : it.blocks - b0);
            held[w] = -1;
        }
    }
    for (size_t i = 0; i < q.items.size(); ++i)
        if (done[i] != q.items[i].blocks) return false;
    return next == q.total;
}

template <class T>
void check_queue_plan(const std::vector<LdeJob<T>>& jobs, const LdeQueuePlan<T>& q, uint64_t grouped_blocks) {
    expect(q.lds_bytes <= 160u * 1024u, "queue: dynamic LDS within one workgroup");
    uint64_t sum = 0, tickets = 0;
    for (size_t i = 0; i < q.items.size(); ++i) {
        const auto& it = q.items[i];
        expect(it.first == tickets, "queue: items hold consecutive ticket ranges");
        expect(it.blocks > 0, "queue: every item has blocks");
        expect(it.chunk >= 1 && it.chunk <= 8, "queue: chunks of 1 to 8 blocks");
        tickets += (it.blocks + it.chunk - 1) / it.chunk;
        expect(it.dep < int32_t(i), "queue: a dependency comes before its dependent");
        if (it.dep >= 0) {
            const auto& d = q.items[size_t(it.dep)];
            expect(d.first + (d.blocks + d.chunk - 1) / d.chunk <= it.first, "queue: a dependency's tickets are all lower");
            expect(d.args.dst == it.args.src || it.args.src == it.args.dst, "queue: a pass reads its LDE's buffer");
        }
        sum += it.blocks;
    }
    expect(tickets == q.total, "queue: the tickets are the items' chunks");
    expect(sum == grouped_blocks, "queue: the same blocks as the grouped launches");
    size_t heads = 0;
    for (const auto& it : q.items) heads += it.dep < 0;
    expect(heads == jobs.size(), "queue: one dependency chain per LDE");
}

template <class F>
void queue_protocol(std::mt19937_64& rng, int rounds) {
    using T = typename F::T;
    for (int r = 0; r < rounds; ++r) {
        std::vector<LdeJob<T>> jobs;
        const int n = 1 + int(rng() % 6);
        for (int i = 0; i < n; ++i) {
            LdeJob<T> j{};
            j.L = 6 + uint32_t(rng() % 15);
            j.w = 1 + uint32_t(rng() % 48);
            j.k = uint32_t(rng() % 3);
            j.in = reinterpret_cast<const T*>(uintptr_t(0x100000) * (2 * i + 1));
            j.out = reinterpret_cast<T*>(uintptr_t(0x100000) * (2 * i + 2));
            j.tw = reinterpret_cast<const T*>(uintptr_t(0x7000000));
            j.twist = reinterpret_cast<const T*>(uintptr_t(0x8000000));
            j.tune.max_stages = 6 + uint32_t(rng() % 5);
            jobs.push_back(j);
        }
        LdeQueuePlan<T> q;
        if (!plan_lde_queue<F>(jobs, q)) {
            expect(false, "queue: a batch of transforms of 2^6 rows and up is carried");
            continue;
        }
        if (q.total > 40000) continue;
        protocol_chunked += q.total < [&] { uint64_t b = 0; for (const auto& it : q.items) b += it.blocks; return b; }();
        ++protocol_plans;
        for (uint32_t g : {1u, 2u, 7u, 64u}) expect(run_queue_protocol(q, g, rng), "queue: the protocol completes");
    }
}

int queue_tried = 0, queue_ok = 0;  // batches with carried pass widths; of them, planned as a queue

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
        if constexpr (std::is_same<F, BabyBear>::value) {
            uint64_t grouped = 0;
            for (const auto& g : groups) grouped += g.blocks;
            bool uncarried = false;  // a fused inner pass of 0, 4 or 5 stages
            for (const auto& j : jobs) {
                const uint32_t cap = j.tune.max_stages && j.tune.max_stages < 10 ? j.tune.max_stages : 10;
                const uint32_t s_inner = j.L < cap ? j.L : cap;
                uncarried = uncarried || s_inner == 0 || s_inner == 4 || s_inner == 5;
            }
            LdeQueuePlan<T> q;
            const bool ok = plan_lde_queue<F>(jobs, q);
            expect(!(ok && uncarried), "queue: no fused pass of 0, 4 or 5 stages is carried");
            if (!uncarried) ++queue_tried, queue_ok += ok;
            if (ok) check_queue_plan(jobs, q, grouped);
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
    expect(queue_ok * 4 >= queue_tried * 3, "queue: most batches of carried widths fit the queue's LDS rule");
    queue_protocol<BabyBear>(rng, 300);
    admission(rng, 5000);
    if (fails) {
        std::printf("%d checks failed\n", fails);
        return 1;
    }
    expect(protocol_plans >= 100, "queue: the protocol ran on enough plans");
    expect(protocol_chunked >= 10, "queue: the protocol ran on plans with chunks");
    std::printf("%d of %d batches planned as a queue; protocol run on %d plans, %d with chunks\nok\n", queue_ok,
                queue_tried, protocol_plans, protocol_chunked);
    return 0;
}
