"""Write tests/golden/golden.npz from the reference build (oracle/_ref/libp3ref.so).

The reference (dloghin/plonky3-gpu) is header-only C++; `make -C oracle ref`
compiles a driver over its headers where they are present.  This script runs
that build on small seeded inputs and stores inputs and outputs as plain
arrays, so the oracle and the GPU path can be checked against the reference on
machines that do not have it.  Nothing from the reference is copied: only data.

    make -C oracle ref && python tests/golden/gen_golden.py
"""

from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from native import U32, U64, p2_consts, reference, seeded  # noqa: E402


def challenger_scripts():
    """Op lists (code, arg) in the encoding of oracle/oracle.cpp run_script."""
    from native import CHECK_WITNESS, GRIND, OBSERVE, SAMPLE, SAMPLE_BITS, SAMPLE_EXT

    rng = np.random.default_rng(77)
    # 0: every input-buffer length, then samples across the output-buffer refill
    s0 = []
    for n_obs in range(18):
        s0 += [(OBSERVE, int(v)) for v in rng.integers(0, 1 << 32, size=n_obs)]
        s0 += [(SAMPLE, 0)] * int(n_obs % 10 + 1)
    # 1: grinding from every input-buffer length (one or two permutations per witness)
    s1 = []
    for ell in range(8):
        s1 += [(OBSERVE, int(v)) for v in rng.integers(0, 0x78000001, size=8 + ell)]
        s1 += [(GRIND, int(ell % 7 + 1)), (SAMPLE, 0)]
    # 2: a random mix, including check_witness on accepted and rejected witnesses
    s2 = []
    for _ in range(120):
        r = int(rng.integers(0, 12))
        if r < 5:
            s2.append((OBSERVE, int(rng.integers(0, 1 << 32))))
        elif r < 7:
            s2.append((SAMPLE, 0))
        elif r == 7:
            s2.append((SAMPLE_BITS, int(rng.integers(1, 64))))
        elif r == 8:
            s2.append((SAMPLE_EXT, 0))
        elif r == 9:
            s2.append((GRIND, int(rng.integers(0, 11))))
        else:
            s2.append((CHECK_WITNESS, int(rng.integers(0, 4)) | (int(rng.integers(0, 40)) << 6)))
    return [s0, s1, s2]


def main():
    ref = reference()
    if ref is None:
        raise SystemExit("oracle/_ref/libp3ref.so missing: run `make -C oracle ref` first")
    g = {}
    # Poseidon2 constants: seeded BabyBear W16 RF8 RP13, seeded Goldilocks W8 RF8
    # RP22, and the all-zero BabyBear set with 22 partial rounds used by the
    # reference's own FRI test (src/tests/fri_test.cpp:57-71).
    cb = p2_consts(1, "bb", 8, 13)
    cg = p2_consts(2, "gl", 8, 22)
    cz = np.zeros(8 * 16 + 22, dtype=U32)
    g["consts_bb"], g["consts_gl"], g["consts_bb_zero22"] = cb, cg, cz

    st = seeded(10, 16 * 6, "bb")
    g["perm_bb_in"], g["perm_bb_out"] = st, ref.permute("bb", cb, 8, 13, st)
    g["perm_bb_zero22_out"] = ref.permute("bb", cz, 8, 22, st)
    stg = seeded(11, 8 * 6, "gl")
    g["perm_gl_in"], g["perm_gl_out"] = stg, ref.permute("gl", cg, 8, 22, stg)

    stream = seeded(12, 40, "bb")
    g["sponge_bb_stream"] = stream
    g["sponge_bb_out"] = np.stack([ref.sponge("bb", cb, 8, 13, stream[:n]) for n in range(34)])
    sg = seeded(13, 20, "gl")
    g["sponge_gl_stream"] = sg
    g["sponge_gl_out"] = np.stack([ref.sponge("gl", cg, 8, 22, sg[:n]) for n in range(18)])

    pairs = seeded(14, 16 * 4, "bb").reshape(4, 16)
    g["compress_bb_in"] = pairs
    g["compress_bb_out"] = np.stack([ref.compress("bb", cb, 8, 13, p) for p in pairs])
    pg = seeded(15, 8 * 4, "gl").reshape(4, 8)
    g["compress_gl_in"] = pg
    g["compress_gl_out"] = np.stack([ref.compress("gl", cg, 8, 22, p) for p in pg])

    m = seeded(16, 32 * 3, "bb").reshape(32, 3)
    g["dft_bb_in"] = m
    for kind in range(4):
        g[f"dft_bb_k{kind}"] = ref.dft("bb", kind, m, 31)
    mg = seeded(17, 16 * 3, "gl").reshape(16, 3)
    g["dft_gl_in"] = mg
    for kind in range(4):
        g[f"dft_gl_k{kind}"] = ref.dft("gl", kind, mg, 7)

    lm = seeded(18, 64 * 5, "bb").reshape(64, 5)
    g["lde_bb_in"] = lm
    g["lde_bb_a1"] = ref.lde("bb", lm, 1, 31)
    g["lde_bb_a2"] = ref.lde("bb", lm, 2, 31)
    lg = seeded(19, 32 * 3, "gl").reshape(32, 3)
    g["lde_gl_in"] = lg
    g["lde_gl_a2"] = ref.lde("gl", lg, 2, 7)

    flat_shapes = [(1, 8), (3, 8), (5, 16), (16, 32), (7, 5), (6, 13)]
    for i, (h, w) in enumerate(flat_shapes):
        d = seeded(100 + i, h * w, "bb").reshape(h, w)
        g[f"flat_bb_{i}_in"] = d
        g[f"flat_bb_{i}_root"] = ref.flat_commit("bb", cb, 8, 13, d)
    for i, (h, w) in enumerate([(6, 4), (8, 8), (5, 3)]):
        d = seeded(120 + i, h * w, "gl").reshape(h, w)
        g[f"flat_gl_{i}_in"] = d
        g[f"flat_gl_{i}_root"] = ref.flat_commit("gl", cg, 8, 22, d)

    # multi-height trees: matrices in commit order, several heights incl. ties
    mm_sets = [
        [(8, 4), (4, 3), (8, 2), (2, 5)],
        [(16, 8), (16, 3), (4, 7), (1, 2), (8, 9)],
        [(32, 1)],
        [(4, 16), (8, 8), (2, 11)],
    ]
    for s, shapes in enumerate(mm_sets):
        mats = [seeded(200 + 10 * s + j, h * w, "bb").reshape(h, w) for j, (h, w) in enumerate(shapes)]
        for j, mat in enumerate(mats):
            g[f"mm_bb_{s}_m{j}"] = mat
        maxh = max(h for h, _ in shapes)
        for cap in (0, 1):
            idx = (5 * s + 3) % maxh
            c, vals, proof = ref.mmcs("bb", cb, 8, 13, mats, cap, idx)
            g[f"mm_bb_{s}_cap{cap}"] = c
            g[f"mm_bb_{s}_cap{cap}_idx"] = np.array([idx], dtype=U64)
            g[f"mm_bb_{s}_cap{cap}_vals"] = vals
            g[f"mm_bb_{s}_cap{cap}_proof"] = proof
    gl_shapes = [(8, 4), (4, 6), (8, 1)]
    mats = [seeded(300 + j, h * w, "gl").reshape(h, w) for j, (h, w) in enumerate(gl_shapes)]
    for j, mat in enumerate(mats):
        g[f"mm_gl_m{j}"] = mat
    c, vals, proof = ref.mmcs("gl", cg, 8, 22, mats, 1, 6)
    g["mm_gl_cap1"], g["mm_gl_vals"], g["mm_gl_proof"] = c, vals, proof

    # BabyBear^4 multiplication known answers (src/tests/test_extension_field.cpp:175-180)
    g["bb4_mul"] = ref.bb4_mul([1, 2, 3, 4], [5, 6, 7, 8])

    # FRI folds (arity 2, 4, 8) and commit phases
    for la, lh in ((1, 6), (2, 6), (3, 7)):
        cw = seeded(400 + la, (1 << lh) * 4, "bb").reshape(-1, 4)
        beta = seeded(410 + la, 4, "bb")
        g[f"fold_a{la}_in"], g[f"fold_a{la}_beta"] = cw, beta
        g[f"fold_a{la}_out"] = ref.fold(lh, la, beta, cw)
    for case, (log_len, lb, lfpl, mla) in enumerate(((8, 1, 2, 1), (9, 1, 1, 2), (8, 2, 0, 3))):
        cw = seeded(500 + case, (1 << log_len) * 4, "bb").reshape(-1, 4)
        betas = seeded(510 + case, 4 * 16, "bb").reshape(-1, 4)
        roots, final = ref.fri_commit(cb, 8, 13, cw, lb, lfpl, mla, betas)
        g[f"fri_{case}_cw"], g[f"fri_{case}_betas"] = cw, betas
        g[f"fri_{case}_params"] = np.array([log_len, lb, lfpl, mla], dtype=U64)
        g[f"fri_{case}_roots"], g[f"fri_{case}_final"] = roots, final

    # PCS commit (TwoAdicFriPcs::commit): uneven heights, width padded to 2^k
    for case, (shapes, shifts, lb) in enumerate((
        ([(16, 3), (32, 4), (16, 1)], [31, 1, 31], 1),
        ([(8, 8)], [31], 2),
        ([(4, 5), (16, 2)], [1, 31], 1),
    )):
        mats = [seeded(600 + 10 * case + j, h * w, "bb").reshape(h, w) for j, (h, w) in enumerate(shapes)]
        for j, mat in enumerate(mats):
            g[f"pcs_{case}_m{j}"] = mat
        g[f"pcs_{case}_shifts"] = np.array(shifts, dtype=U32)
        g[f"pcs_{case}_lb"] = np.array([lb], dtype=U64)
        g[f"pcs_{case}_root"] = ref.pcs_commit(cb, 8, 13, mats, shifts, lb)

    # DuplexChallenger (duplex_challenger.hpp:42-268) on scripted transcripts.
    for i, ops in enumerate(challenger_scripts()):
        g[f"chal_{i}_ops"] = np.array(ops, dtype=U64).reshape(-1, 2)
        g[f"chal_{i}_out"] = ref.challenger_run(cb, 8, 13, ops)
    # prove_fri's commit phase with the challenger in the loop (grinding per round)
    for case, (log_len, lb, lfpl, mla, pow_bits, npre) in enumerate(
            ((9, 1, 2, 1, 0, 3), (10, 1, 1, 2, 6, 0), (9, 2, 0, 3, 4, 11), (11, 1, 3, 1, 9, 8))):
        cw = seeded(700 + case, (1 << log_len) * 4, "bb").reshape(-1, 4)
        pre = seeded(710 + case, npre, "bb") if npre else np.zeros(0, dtype=U32)
        r = ref.fri_commit_chal(cb, 8, 13, cw, lb, lfpl, mla, pow_bits, pre)
        g[f"fchal_{case}_cw"], g[f"fchal_{case}_pre"] = cw, pre
        g[f"fchal_{case}_params"] = np.array([log_len, lb, lfpl, mla, pow_bits], dtype=U64)
        for k, v in r.items():
            g[f"fchal_{case}_{k}"] = v

    # TwoAdicFriPcs::commit + open with the transcript (opened values, FRI proof parts)
    for case, (shapes, shifts, lb, npts, lfpl, mla, nq, cpow, qpow) in enumerate((
        ([(16, 3), (32, 4), (16, 1)], [31, 1, 31], 1, [1, 2, 1], 2, 1, 4, 0, 0),
        ([(64, 5)], [31], 1, [2], 1, 2, 6, 3, 4),
        ([(32, 2), (8, 3), (64, 1)], [7, 31, 1], 2, [1, 1, 3], 0, 3, 3, 2, 1),
    )):
        mats = [seeded(800 + 10 * case + j, h * w, "bb").reshape(h, w) for j, (h, w) in enumerate(shapes)]
        pts = [seeded(850 + 10 * case + j, 4 * k, "bb").reshape(k, 4) for j, k in enumerate(npts)]
        pre = seeded(890 + case, 3, "bb")
        r = ref.pcs_open(cb, 8, 13, mats, shifts, lb, pts, pre, lfpl, mla, nq, cpow, qpow)
        for j, mat in enumerate(mats):
            g[f"popen_{case}_m{j}"], g[f"popen_{case}_p{j}"] = mat, pts[j]
        g[f"popen_{case}_shifts"], g[f"popen_{case}_pre"] = np.array(shifts, dtype=U32), pre
        g[f"popen_{case}_params"] = np.array([lb, lfpl, mla, nq, cpow, qpow], dtype=U64)
        for k, v in r.items():
            g[f"popen_{case}_{k}"] = v

    out = os.path.join(HERE, "golden.npz")
    np.savez_compressed(out, **g)
    print(f"wrote {out}: {len(g)} arrays, {os.path.getsize(out)} bytes")


if __name__ == "__main__":
    main()
