// Sanitizer driver for the oracle (test infrastructure): every entry point of
// oracle/oracle.h on small and edge-case inputs, with buffers of exactly the
// documented sizes, built with -fsanitize=address,undefined by
// tests/test_sanitize.py.  Any out-of-bounds access, use after free, integer
// overflow of a signed type or other undefined behaviour aborts the run; the
// program prints "ok" and exits 0 otherwise.  Values are not checked here
// (tests/test_oracle.py pins them).
#include "oracle.h"

#include <cstdint>
#include <cstdio>
#include <vector>

namespace {

template <class T>
std::vector<T> fill(uint64_t seed, size_t n, bool bb) {
    std::vector<T> v(n);
    if (bb) orc_bb_fill(seed, reinterpret_cast<uint32_t*>(v.data()), n);
    else orc_gl_fill(seed, reinterpret_cast<uint64_t*>(v.data()), n);
    return v;
}

int fails = 0;
void expect(bool ok, const char* what) {
    if (!ok) {
        std::printf("unexpected result: %s\n", what);
        ++fails;
    }
}

size_t flat_proof_len(size_t h) {
    size_t steps = 0;
    while (h > 1) {
        h = (h + 1) / 2;
        ++steps;
    }
    return steps;
}

}  // namespace

int main() {
    const int RF = 8, RPB = 13, RPG = 22;
    auto cb = fill<uint32_t>(1, 16 * RF + RPB, true);
    auto cg = fill<uint64_t>(2, 8 * RF + RPG, false);

    // permutations and sponges of every length 0..33 (17 for Goldilocks)
    auto sb = fill<uint32_t>(3, 16 * 5, true);
    expect(orc_bb_permute(cb.data(), RF, RPB, sb.data(), 5) == 0, "bb permute");
    auto sg = fill<uint64_t>(4, 8 * 5, false);
    expect(orc_gl_permute(cg.data(), RF, RPG, sg.data(), 5) == 0, "gl permute");
    for (size_t len = 0; len <= 33; ++len) {
        auto d = fill<uint32_t>(10 + len, len, true);
        std::vector<uint32_t> out(8);
        expect(orc_bb_sponge(cb.data(), RF, RPB, d.data(), len, out.data()) == 0, "bb sponge");
    }
    for (size_t len = 0; len <= 17; ++len) {
        auto d = fill<uint64_t>(50 + len, len, false);
        std::vector<uint64_t> out(4);
        expect(orc_gl_sponge(cg.data(), RF, RPG, d.data(), len, out.data()) == 0, "gl sponge");
    }
    {
        auto in = fill<uint32_t>(5, 16, true);
        std::vector<uint32_t> out(8);
        expect(orc_bb_compress(cb.data(), RF, RPB, in.data(), out.data()) == 0, "bb compress");
        auto ing = fill<uint64_t>(6, 8, false);
        std::vector<uint64_t> outg(4);
        expect(orc_gl_compress(cg.data(), RF, RPG, ing.data(), outg.data()) == 0, "gl compress");
    }
    // every DFT kind, LDEs (brute force too), heights 2^0..2^6, ragged widths
    for (size_t lh = 0; lh <= 6; ++lh)
        for (size_t w : {size_t(1), size_t(3), size_t(8)}) {
            const size_t h = size_t(1) << lh;
            auto m = fill<uint32_t>(100 + lh * 7 + w, h * w, true);
            auto mg = fill<uint64_t>(200 + lh * 7 + w, h * w, false);
            for (int kind = 0; kind < 4; ++kind) {
                std::vector<uint32_t> o(h * w);
                expect(orc_bb_dft(kind, m.data(), lh, w, 31, o.data()) == 0, "bb dft");
                std::vector<uint64_t> og(h * w);
                expect(orc_gl_dft(kind, mg.data(), lh, w, 7, og.data()) == 0, "gl dft");
            }
            for (size_t k = 0; k <= 2; ++k) {
                std::vector<uint32_t> o((h << k) * w), ob((h << k) * w);
                expect(orc_bb_lde(m.data(), lh, w, k, 31, o.data()) == 0, "bb lde");
                expect(orc_bb_lde_brute(m.data(), lh, w, k, 31, ob.data()) == 0, "bb lde brute");
                std::vector<uint64_t> og((h << k) * w), obg((h << k) * w);
                expect(orc_gl_lde(mg.data(), lh, w, k, 7, og.data()) == 0, "gl lde");
                expect(orc_gl_lde_brute(mg.data(), lh, w, k, 7, obg.data()) == 0, "gl lde brute");
            }
            std::vector<uint32_t> bad(8);
            expect(orc_bb_dft(4, m.data(), lh, w, 31, bad.data()) != 0, "bad dft kind rejected");
        }
    // flat trees of every height 1..19 (odd layers padded), openings of every row
    for (size_t h = 1; h <= 19; ++h) {
        const size_t w = 5;
        auto m = fill<uint32_t>(300 + h, h * w, true);
        size_t total = 0, cnt = h;
        while (true) {
            total += cnt;
            if (cnt == 1) break;
            cnt = (cnt + 1) / 2;
        }
        std::vector<uint32_t> root(8), tree(total * 8);
        expect(orc_bb_flat_commit(cb.data(), RF, RPB, m.data(), h, w, root.data(), tree.data()) == 0, "flat commit");
        auto mg = fill<uint64_t>(400 + h, h * w, false);
        std::vector<uint64_t> rootg(4);
        expect(orc_gl_flat_commit(cg.data(), RF, RPG, mg.data(), h, w, rootg.data(), nullptr) == 0, "gl flat");
        for (size_t i = 0; i < h; ++i) {
            std::vector<uint32_t> row(w), proof(flat_proof_len(h) * 8 + (h == 1 ? 8 : 0));
            expect(orc_bb_flat_open(cb.data(), RF, RPB, m.data(), h, w, i, row.data(), proof.data()) == 0,
                   "flat open");
        }
    }
    // threaded LDE + commit, with openings
    for (int threads : {1, 3, 9}) {
        const size_t lh = 6, w = 7, k = 1, N = size_t(1) << (lh + k);
        auto m = fill<uint32_t>(500 + threads, (size_t(1) << lh) * w, true);
        std::vector<uint32_t> root(8);
        expect(orc_bb_lde_commit_mt(cb.data(), RF, RPB, m.data(), lh, w, k, 31, threads, root.data()) == 0, "mt");
        std::vector<uint64_t> idx = {0, 1, N - 1, 77};
        std::vector<uint32_t> vals(idx.size() * w), proof(idx.size() * (lh + k) * 8);
        expect(orc_bb_lde_commit_open_mt(cb.data(), RF, RPB, m.data(), lh, w, k, 31, threads, idx.data(), idx.size(),
                                         root.data(), vals.data(), proof.data()) == 0,
               "mt open");
        auto mg = fill<uint64_t>(600 + threads, (size_t(1) << lh) * w, false);
        std::vector<uint64_t> rootg(4), valsg(idx.size() * w), proofg(idx.size() * (lh + k) * 4);
        expect(orc_gl_lde_commit_open_mt(cg.data(), RF, RPG, mg.data(), lh, w, k, 7, threads, idx.data(),
                                         idx.size(), rootg.data(), valsg.data(), proofg.data()) == 0,
               "gl mt open");
    }
    // multi-height MMCS: height ties, injection, caps 0..2, every index, verify + tamper
    {
        const std::vector<size_t> hs = {16, 4, 16, 8, 2, 4}, ws = {3, 5, 1, 8, 2, 9};
        std::vector<std::vector<uint32_t>> mats;
        std::vector<const uint32_t*> ptrs;
        size_t tw = 0;
        for (size_t i = 0; i < hs.size(); ++i) {
            mats.push_back(fill<uint32_t>(700 + i, hs[i] * ws[i], true));
            ptrs.push_back(mats.back().data());
            tw += ws[i];
        }
        for (size_t cap = 0; cap <= 2; ++cap)
            for (long long idx = 0; idx < 16; ++idx) {
                std::vector<uint32_t> capv((size_t(1) << cap) * 8), vals(tw), proof((4 - cap) * 8);
                size_t plen = 0;
                expect(orc_bb_mmcs(cb.data(), RF, RPB, hs.size(), hs.data(), ws.data(), ptrs.data(), cap,
                                   capv.data(), idx, vals.data(), proof.data(), &plen) == 0,
                       "mmcs");
                expect(plen == 4 - cap, "proof length");
                expect(orc_bb_mmcs_verify(cb.data(), RF, RPB, hs.size(), hs.data(), ws.data(), cap, capv.data(),
                                          size_t(idx), vals.data(), proof.data(), plen) == 1,
                       "verify honest");
                vals[0] ^= 1;
                expect(orc_bb_mmcs_verify(cb.data(), RF, RPB, hs.size(), hs.data(), ws.data(), cap, capv.data(),
                                          size_t(idx), vals.data(), proof.data(), plen) == 0,
                       "verify tampered");
            }
        std::vector<size_t> adds = {1, 2, 1, 0, 2, 1}, lhs;
        for (size_t h : hs) lhs.push_back(__builtin_ctzll(h));
        std::vector<uint64_t> idx = {0, 5, 31};
        size_t maxn = 0, t2 = 0;
        for (size_t i = 0; i < hs.size(); ++i) {
            maxn = std::max(maxn, hs[i] << adds[i]);
            t2 += ws[i];
        }
        size_t levels = __builtin_ctzll(maxn) - 1;
        std::vector<uint32_t> capv(16), vals(idx.size() * t2), proof(idx.size() * levels * 8);
        size_t plen = 0;
        expect(orc_bb_batch_commit_mt(cb.data(), RF, RPB, hs.size(), lhs.data(), ws.data(), adds.data(), ptrs.data(),
                                      31, 1, 4, idx.data(), idx.size(), capv.data(), vals.data(), proof.data(),
                                      &plen) == 0,
               "batch mt");
        expect(plen == levels, "batch proof length");
        std::vector<uint64_t> capg(4), valsg(tw), proofg(4 * 4);
        std::vector<std::vector<uint64_t>> mg;
        std::vector<const uint64_t*> pg;
        for (size_t i = 0; i < hs.size(); ++i) {
            mg.push_back(fill<uint64_t>(800 + i, hs[i] * ws[i], false));
            pg.push_back(mg.back().data());
        }
        expect(orc_gl_mmcs(cg.data(), RF, RPG, hs.size(), hs.data(), ws.data(), pg.data(), 0, capg.data(), 9,
                           valsg.data(), proofg.data(), &plen) == 0,
               "gl mmcs");
    }
    // folds of every arity and the FRI commit phase (plain, threaded, with the transcript)
    for (size_t la = 1; la <= 3; ++la) {
        const size_t lh = 7, n = size_t(1) << lh;
        auto cur = fill<uint32_t>(900 + la, 4 * n, true);
        auto beta = fill<uint32_t>(910 + la, 4, true);
        std::vector<uint32_t> out(4 * (n >> la));
        expect(orc_bb4_fold(lh, la, beta.data(), cur.data(), n, out.data()) == 0, "fold");
    }
    for (size_t mla = 1; mla <= 3; ++mla) {
        const size_t log_len = 9, lb = 1, lfpl = 2;
        auto cw = fill<uint32_t>(950 + mla, 4 * (size_t(1) << log_len), true);
        auto betas = fill<uint32_t>(960 + mla, 4 * 16, true);
        std::vector<uint32_t> roots(64 * 8), final4(4 << lfpl), roots2(64 * 8), final2(4 << lfpl);
        size_t nr = 0, nr2 = 0;
        expect(orc_bb4_fri_commit(cb.data(), RF, RPB, cw.data(), log_len, lb, lfpl, mla, betas.data(), roots.data(),
                                  final4.data(), &nr) == 0,
               "fri");
        expect(orc_bb4_fri_commit_mt(cb.data(), RF, RPB, cw.data(), log_len, lb, lfpl, mla, betas.data(), 4,
                                     roots2.data(), final2.data(), &nr2) == 0 && nr == nr2 && roots == roots2,
               "fri mt");
        std::vector<uint32_t> pre = {1, 2, 3}, betas_out(64 * 4), post(8);
        std::vector<uint64_t> wit(64);
        expect(orc_bb4_fri_commit_chal(cb.data(), RF, RPB, cw.data(), log_len, lb, lfpl, mla, 3, pre.data(),
                                       pre.size(), roots.data(), betas_out.data(), wit.data(), final4.data(),
                                       post.data(), &nr) == 0,
               "fri chal");
    }
    // challenger scripts: observe runs of every length, samples, grinds, checks
    {
        std::vector<uint64_t> ops;
        for (uint64_t i = 0; i < 20; ++i) ops.push_back(0 | ((1000 + i) << 8));
        ops.push_back(1);
        ops.push_back(2 | (20ull << 8));
        ops.push_back(3 | (6ull << 8));
        ops.push_back(4 | ((6ull | (5ull << 6)) << 8));
        ops.push_back(5);
        std::vector<uint64_t> out(64);
        size_t n_out = 0;
        expect(orc_bb_challenger_run(cb.data(), RF, RPB, ops.data(), ops.size(), out.data(), &n_out) == 0, "chal");
    }
    // PCS commit and open: three heights, two points on one matrix
    {
        const std::vector<size_t> lhs = {6, 4, 6}, ws = {5, 3, 4}, npts = {1, 2, 1};
        std::vector<std::vector<uint32_t>> mats;
        std::vector<const uint32_t*> ptrs;
        for (size_t i = 0; i < lhs.size(); ++i) {
            mats.push_back(fill<uint32_t>(1100 + i, (size_t(1) << lhs[i]) * ws[i], true));
            ptrs.push_back(mats.back().data());
        }
        std::vector<uint32_t> shifts = {31, 31, 31}, root(8);
        expect(orc_bb_pcs_commit(cb.data(), RF, RPB, lhs.size(), lhs.data(), ws.data(), shifts.data(), ptrs.data(), 1,
                                 root.data()) == 0,
               "pcs commit");
        auto points = fill<uint32_t>(1200, 4 * 4 + 4, true);
        std::vector<uint32_t> pre = {7}, opened(4 * (5 * 1 + 3 * 2 + 4 * 1)), roots(64 * 8), final4(4 << 1),
            qrows(6 * 12), post(8);
        std::vector<uint64_t> cwit(64), qwit(1), qidx(6);
        size_t nr = 0;
        expect(orc_bb_pcs_open(cb.data(), RF, RPB, lhs.size(), lhs.data(), ws.data(), shifts.data(), ptrs.data(), 1,
                               npts.data(), points.data(), pre.data(), pre.size(), 1, 1, 6, 2, 3, root.data(),
                               opened.data(), roots.data(), cwit.data(), final4.data(), qwit.data(), qidx.data(),
                               qrows.data(), post.data(), &nr) == 0,
               "pcs open");
    }
    if (fails) return 1;
    std::printf("ok\n");
    return 0;
}
