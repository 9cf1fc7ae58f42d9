"""GPU parity: libp3hip (through the C ABI) against the golden fixtures of the
reference build and against the CPU oracle on the same seeded inputs.
Bit-exact everywhere: this is integer field arithmetic."""

import numpy as np
import pytest

import native
import p3hip
from native import BB_P, GL_P, U32, U64

pytestmark = pytest.mark.gpu
RF = 8
BB, GL = p3hip.BABYBEAR, p3hip.GOLDILOCKS


def rnd(seed, shape, f="bb"):
    rng = np.random.default_rng(seed)
    p = BB_P if f == "bb" else GL_P
    return rng.integers(0, p, size=shape, dtype=native.DT[f])


def mats_of(golden, prefix):
    out, j = [], 0
    while f"{prefix}_m{j}" in golden:
        out.append(golden[f"{prefix}_m{j}"])
        j += 1
    return out


# -------------------------------------------------------------- Poseidon2 --
def test_permutation(ctx, golden):
    assert (ctx.permute(BB, golden["perm_bb_in"]) == golden["perm_bb_out"]).all()
    assert (ctx.permute(GL, golden["perm_gl_in"]) == golden["perm_gl_out"]).all()


def test_permutation_many_states(ctx, golden, orc):
    st = rnd(1, 16 * 3000)
    assert (ctx.permute(BB, st) == orc.permute("bb", golden["consts_bb"], RF, 13, st)).all()
    stg = rnd(2, 8 * 1000, "gl")
    assert (ctx.permute(GL, stg) == orc.permute("gl", golden["consts_gl"], RF, 22, stg)).all()


def test_permutation_both_binary64_layouts(ctx, golden, orc):
    """Up to 2^15 states run one state per 16-lane row; more run one state per
    thread.  Both against the oracle, including all-(p-1) and all-zero states."""
    for n in (7, 1 << 15, (1 << 15) + 5):
        st = rnd(n, 16 * n)
        st[:16] = BB_P - 1
        st[16:32] = 0
        assert (ctx.permute(BB, st) == orc.permute("bb", golden["consts_bb"], RF, 13, st)).all(), n


@pytest.mark.parametrize("kind", ["max", "zero", "random"])
def test_permutation_under_extreme_round_constants(kind, orc):
    """Both engines' constant handling against the oracle: the binary64
    engine's constants folded through MDS-light^-1 into the previous round
    (balanced residues computed on the host) and the Goldilocks partial rounds
    on split accumulators, with every round constant p - 1, every one zero,
    and random; states include all-(p-1) and all-zero rows, and the sponge
    and the dense leaf kernel (whose rate lanes stay unreduced between
    chunks) run on the same constants."""
    def consts(f, n, p):
        if kind == "max":
            return np.full(n, p - 1, dtype=np.uint32 if f == "bb" else np.uint64)
        if kind == "zero":
            return np.zeros(n, dtype=np.uint32 if f == "bb" else np.uint64)
        return rnd(77, n, f)
    cb, cg = consts("bb", RF * 16 + 13, BB_P), consts("gl", RF * 8 + 22, GL_P)
    with p3hip.Context(0) as c:
        c.set_poseidon2(BB, cb, RF, 13)
        c.set_poseidon2(GL, cg, RF, 22)
        st = rnd(5, 16 * 4100)
        st[:16] = BB_P - 1
        st[16:32] = 0
        assert (c.permute(BB, st) == orc.permute("bb", cb, RF, 13, st)).all()
        stg = rnd(6, 8 * 1100, "gl")
        stg[:8] = GL_P - 1
        stg[8:16] = 0
        assert (c.permute(GL, stg) == orc.permute("gl", cg, RF, 22, stg)).all()
        rows = rnd(8, (1 << 16) * 24).reshape(1 << 16, 24)  # past 2^15 leaves: the dense leaf kernel
        rows[0] = BB_P - 1
        tree = c.mmcs_commit([c.upload(rows)], 0)
        assert (tree.root() == orc.flat_commit("bb", cb, RF, 13, rows)).all()
        rows_gl = rnd(9, 2048 * 12, "gl").reshape(2048, 12)
        rows_gl[0] = GL_P - 1
        tree_gl = c.mmcs_commit([c.upload(rows_gl)], 0)
        assert (tree_gl.root() == orc.flat_commit("gl", cg, RF, 22, rows_gl)).all()


# ------------------------------------------------------------------- DFT --
def test_dft_golden(ctx, golden):
    dft = p3hip.Radix2DitHip(ctx)
    for kind, fn in enumerate((dft.dft_batch, dft.idft_batch)):
        assert (fn(golden["dft_bb_in"]) == golden[f"dft_bb_k{kind}"]).all(), kind
        assert (fn(golden["dft_gl_in"]) == golden[f"dft_gl_k{kind}"]).all(), kind
    assert (dft.coset_dft_batch(golden["dft_bb_in"], 31) == golden["dft_bb_k2"]).all()
    assert (dft.coset_idft_batch(golden["dft_bb_in"], 31) == golden["dft_bb_k3"]).all()
    assert (dft.coset_dft_batch(golden["dft_gl_in"], 7) == golden["dft_gl_k2"]).all()
    assert (dft.coset_idft_batch(golden["dft_gl_in"], 7) == golden["dft_gl_k3"]).all()


@pytest.mark.parametrize("f,lh,w", [("bb", 1, 3), ("bb", 7, 5), ("bb", 10, 16), ("bb", 11, 3), ("bb", 14, 20),
                                    ("bb", 17, 2), ("gl", 3, 2), ("gl", 9, 4), ("gl", 12, 5)])
def test_dft_kinds_vs_oracle(ctx, orc, f, lh, w):
    m = rnd(lh * 7 + w, (1 << lh, w), f)
    shift = 31 if f == "bb" else 7
    for kind in range(4):
        got = ctx.dft(kind, ctx.upload(m), shift).download()
        assert (got == orc.dft(f, kind, m, shift)).all(), kind


# ------------------------------------------------------------------- LDE --
def test_lde_golden(ctx, golden):
    dft = p3hip.Radix2DitHip(ctx)
    assert (dft.coset_lde_batch(golden["lde_bb_in"], 1, 31) == golden["lde_bb_a1"]).all()
    assert (dft.coset_lde_batch(golden["lde_bb_in"], 2, 31) == golden["lde_bb_a2"]).all()
    assert (dft.coset_lde_batch(golden["lde_gl_in"], 2, 7) == golden["lde_gl_a2"]).all()


@pytest.mark.parametrize("f,lh,w,k", [("bb", 0, 5, 1), ("bb", 1, 3, 1), ("bb", 4, 8, 1), ("bb", 6, 5, 2),
                                      ("bb", 10, 16, 1), ("bb", 11, 3, 1), ("bb", 12, 33, 2), ("bb", 13, 7, 3),
                                      ("bb", 16, 32, 1), ("bb", 15, 64, 0), ("gl", 0, 3, 2), ("gl", 5, 3, 2),
                                      ("gl", 9, 4, 2), ("gl", 10, 5, 1), ("gl", 13, 3, 2)])
def test_lde_vs_oracle(ctx, orc, f, lh, w, k):
    m = rnd(100 + lh * 3 + w + k, (1 << lh, w), f)
    shift = 31 if f == "bb" else 7
    got = ctx.coset_lde(ctx.upload(m), k, shift).download()
    assert (got == orc.lde(f, m, k, shift)).all()


@pytest.fixture
def shape_ctx(ctx):
    yield ctx
    ctx.set_ntt_tuning(0, 0, 0)


@pytest.mark.parametrize("max_stages", [1, 2, 3, 4, 5, 6, 7, 8, 9, 10])
def test_bb_lde_every_pass_width(shape_ctx, orc, max_stages):
    """Every pass width the planner can choose (1..10 stages), one and two
    outer pass levels, extreme values: an all p-1 and an all-zero column
    beside random ones."""
    ctx = shape_ctx
    ctx.set_ntt_tuning(0, max_stages, 0)
    for L, w, k in ((12, 6, 1), (14, 3, 2)):
        m = rnd(L * 31 + max_stages, (1 << L, w))
        m[:, 0] = BB_P - 1
        m[:, 1] = 0
        assert (ctx.coset_lde(ctx.upload(m), k, 31).download() == orc.lde("bb", m, k, 31)).all(), (L, w, k)


@pytest.mark.parametrize("log_cols", [1, 3, 5, 7])
def test_bb_lde_column_tiles(shape_ctx, orc, log_cols):
    """Column tiles narrower and wider than the matrix, a ragged last tile."""
    ctx = shape_ctx
    ctx.set_ntt_tuning(log_cols, 0, 0)
    m = rnd(log_cols, (1 << 13, 45))
    assert (ctx.coset_lde(ctx.upload(m), 1, 31).download() == orc.lde("bb", m, 1, 31)).all()


@pytest.mark.parametrize("f,L,ms,w,k", [("bb", 12, 6, 520, 2), ("bb", 14, 7, 300, 1), ("bb", 16, 8, 130, 1),
                                         ("gl", 14, 7, 70, 2), ("gl", 16, 8, 40, 1)])
def test_lde_split_exchange_passes(shape_ctx, orc, f, L, ms, w, k):
    """Outer passes whose tile is the full 128 KiB (2^S positions x 2^(15-S)
    BabyBear or 2^(14-S) Goldilocks columns) keep the tile in registers and
    exchange it through LDS one column half at a time.  Every such pass width
    from 6 to 8 stages, both outer directions (the LDE's inverse and forward
    halves, and all four DFT kinds), ragged last column tiles, and tiles whose
    second half lies partly or wholly past the matrix width."""
    ctx = shape_ctx
    ctx.set_ntt_tuning(0, ms, 0)
    shift = 31 if f == "bb" else 7
    m = rnd(L * 13 + w, (1 << L, w), f)
    m[:, 0] = (BB_P if f == "bb" else GL_P) - 1
    assert (ctx.coset_lde(ctx.upload(m), k, shift).download() == orc.lde(f, m, k, shift)).all()
    for kind in range(4):
        assert (ctx.dft(kind, ctx.upload(m), shift).download() == orc.dft(f, kind, m, shift)).all(), kind


@pytest.mark.parametrize("f,w,tunes", [("bb", 45, ((0, 0, 0), (4, 10, 0), (3, 7, 0))),
                                       ("gl", 12, ((0, 0, 0), (3, 10, 0), (2, 7, 0)))])
def test_lde_full_tile_passes_agree_at_scale(shape_ctx, f, w, tunes):
    """2^20 rows, one added bit: the default plan runs both outer passes at
    10 stages on full tiles (32 BabyBear / 16 Goldilocks columns, exchanged
    half by half; the BabyBear width leaves a 13-column last tile whose second
    half is empty); the other two plans use narrower tiles held in LDS whole.
    All three LDEs must be identical."""
    ctx = shape_ctx
    shift = 31 if f == "bb" else 7
    m = ctx.upload(rnd(4545 + w, (1 << 20, w), f))
    outs = []
    for tune in tunes:
        ctx.set_ntt_tuning(*tune)
        lde = ctx.coset_lde(m, 1, shift)
        outs.append(lde.download())
        lde.free()
    assert (outs[0] == outs[1]).all() and (outs[0] == outs[2]).all()


def test_lde_region_of_4_gib(shape_ctx):
    """2^20 x 1024 BabyBear: each 4 GiB coset region is past the 32-bit byte
    offsets of the split outer passes, so the default plan runs its 10-stage
    outer passes with the full 32-column tile in LDS; the first 512 columns
    alone (2 GiB regions) run them split.  An LDE is column-wise, so the two
    must agree on those columns (rows sampled at the start, middle and end)."""
    ctx = shape_ctx
    ctx.set_ntt_tuning(0, 0, 0)
    rng = np.random.default_rng(4096)
    t = rng.integers(0, BB_P, size=(1 << 20, 1024), dtype=np.uint32)
    full = ctx.coset_lde(ctx.upload(t), 1, 31)
    half = ctx.coset_lde(ctx.upload(np.ascontiguousarray(t[:, :512])), 1, 31)
    del t
    for r0 in (0, (1 << 20) - 1024, (1 << 21) - 1024):
        assert (full.download_rows(r0, 1024)[:, :512] == half.download_rows(r0, 1024)).all(), r0
    full.free()
    half.free()


@pytest.mark.parametrize("f,L,w,k", [("bb", 22, 64, 1), ("gl", 22, 8, 2)])
def test_lde_round_trip_at_full_height(ctx, f, L, w, k):
    """Full-height property check (no oracle at this size): LDE row i * 2^k is
    P(shift * omega_n^i), so the coset inverse DFT of those rows gives P's
    coefficients and the plain DFT of the coefficients gives the trace back,
    bit for bit.  2^22 rows: two outer passes plus the fused inner pass for
    the LDE, and the inverse and forward DFT kinds at their full depth."""
    shift = 31 if f == "bb" else 7
    trace = rnd(9000 + L + k, (1 << L, w), f)
    lde = ctx.coset_lde(ctx.upload(trace), k, shift)
    coset0 = lde.download()[:: 1 << k].copy()
    lde.free()
    coefs = ctx.dft(3, ctx.upload(coset0), shift)
    back = ctx.dft(0, coefs, 1).download()
    assert (back == trace).all()


def test_bb_lde_pass_shapes_agree_at_scale(shape_ctx):
    """2^20 x 64 with 2 added bits under three pass shapes (5-, 3- and
    2-pass plans): identical Merkle roots of the 2^22 x 64 LDEs."""
    ctx = shape_ctx
    m = ctx.upload(rnd(4242, (1 << 20, 64)))
    roots = []
    for tune in ((0, 0, 0), (4, 7, 0), (4, 10, 0)):
        ctx.set_ntt_tuning(*tune)
        lde = ctx.coset_lde(m, 2, 31)
        t = ctx.mmcs_commit([lde], 0)
        roots.append(t.root())
        t.free()
        lde.free()
    assert (roots[0] == roots[1]).all() and (roots[0] == roots[2]).all()


# ------------------------------------------------------------------ MMCS --
def test_flat_commit_golden(ctx, golden):
    mmcs = p3hip.FlatMerkleTreeMmcsHip(ctx)
    for i in range(6):
        d = golden[f"flat_bb_{i}_in"]
        root, _ = mmcs.commit_matrix(d.reshape(-1), d.shape[1])
        assert (root == golden[f"flat_bb_{i}_root"]).all(), i
    for i in range(3):
        d = golden[f"flat_gl_{i}_in"]
        root, _ = mmcs.commit_matrix(d.reshape(-1), d.shape[1])
        assert (root == golden[f"flat_gl_{i}_root"]).all(), i


@pytest.mark.parametrize("h,w", [(1, 1), (2, 8), (3, 9), (999, 16), (1024, 24), (513, 7), (4096, 256)])
def test_flat_commit_and_open_vs_oracle(ctx, golden, orc, h, w):
    d = rnd(h + w, (h, w))
    t = ctx.flat_commit(ctx.upload(d))
    root, layers = orc.flat_commit("bb", golden["consts_bb"], RF, 13, d, with_tree=True)
    assert (t.root() == root).all()
    assert t.n_layers == len(layers)
    idx = [0, h - 1, h // 2]
    vals, proof = t.open(idx)
    for q, i in enumerate(idx):
        row, pr = orc.flat_open(golden["consts_bb"], RF, 13, d, i)
        assert (vals[q] == row).all()
        assert (proof[q] == pr).all()


@pytest.mark.parametrize("w", [5, 12, 13, 61])
def test_single_matrix_leaves_of_any_width(ctx, golden, orc, w):
    """Leaves of one matrix taller than the lane-parallel cut-off (2^15 rows)
    at widths that are not a multiple of 8 (12 rows are 16-byte aligned, the
    others are not), straight and through an LDE's coset row map."""
    d = rnd(700 + w, (1 << 16, w))
    want = orc.flat_commit("bb", golden["consts_bb"], RF, 13, d)
    assert (ctx.flat_commit(ctx.upload(d)).root() == want).all()
    assert (ctx.mmcs_commit([ctx.upload(d)], 0).root() == want).all()
    t = rnd(800 + w, (1 << 15, w))
    lde = ctx.coset_lde(ctx.upload(t), 1, 31)
    want = orc.flat_commit("bb", golden["consts_bb"], RF, 13, orc.lde("bb", t, 1, 31))
    assert (ctx.mmcs_commit([lde], 0).root() == want).all()


def test_mmcs_golden(ctx, golden):
    for s in range(4):
        mats = mats_of(golden, f"mm_bb_{s}")
        for cap in (0, 1):
            t = ctx.mmcs_commit([ctx.upload(m) for m in mats], cap)
            assert (t.cap() == golden[f"mm_bb_{s}_cap{cap}"]).all(), (s, cap)
            idx = int(golden[f"mm_bb_{s}_cap{cap}_idx"][0])
            vals, proof = t.open([idx])
            assert (vals[0] == golden[f"mm_bb_{s}_cap{cap}_vals"]).all()
            assert (proof[0] == golden[f"mm_bb_{s}_cap{cap}_proof"]).all()
    mats = mats_of(golden, "mm_gl")
    t = ctx.mmcs_commit([ctx.upload(m) for m in mats], 1)
    assert (t.cap() == golden["mm_gl_cap1"]).all()
    vals, proof = t.open([6])
    assert (vals[0] == golden["mm_gl_vals"]).all() and (proof[0] == golden["mm_gl_proof"]).all()


def test_mmcs_random_heights_vs_oracle(ctx, golden, orc):
    rng = np.random.default_rng(5)
    shapes = [(1 << int(rng.integers(0, 11)), int(rng.integers(1, 70))) for _ in range(9)]
    shapes.append((1024, 8))
    mats = [rnd(300 + i, s) for i, s in enumerate(shapes)]
    for cap in (0, 3):
        t = ctx.mmcs_commit([ctx.upload(m) for m in mats], cap)
        for idx in (0, 513, 1023):
            c, vals, proof = orc.mmcs("bb", golden["consts_bb"], RF, 13, mats, cap, idx)
            assert (t.cap() == c).all()
            v, p = t.open([idx])
            assert (v[0] == vals).all() and (p[0] == proof).all()
            assert orc.mmcs_verify(golden["consts_bb"], RF, 13, shapes, cap, c, idx, v[0], p[0]) == 1


def test_mmcs_verify_accepts_openings_and_rejects_tampering(ctx, golden, orc):
    """verify_batch on the device against the oracle's verifier: honest
    openings pass at every cap height; a changed value, sibling or index, a
    short proof and a wrong cap entry fail in both."""
    rng = np.random.default_rng(11)
    shapes = [(1 << int(rng.integers(0, 10)), int(rng.integers(1, 40))) for _ in range(6)] + [(512, 8)]
    mats = [rnd(700 + i, sh) for i, sh in enumerate(shapes)]
    mmcs = p3hip.MerkleTreeMmcsHip(ctx, 0)
    for cap_h in (0, 2):
        mmcs.cap_height = cap_h
        cap, tree = mmcs.commit(mats)
        for idx in (0, 301, 511):
            opened, proof = mmcs.open_batch(idx, tree)
            assert [len(v) for v in opened] == [w for _, w in shapes]
            assert mmcs.verify_batch(cap, shapes, idx, opened, proof)
            flat_vals = np.concatenate(opened)
            cases = []
            bad = [v.copy() for v in opened]
            bad[-1][0] = (int(bad[-1][0]) + 1) % BB_P
            cases.append((idx, bad, proof))
            if len(proof):
                p2 = proof.copy()
                p2[len(proof) // 2][3] = (int(p2[len(proof) // 2][3]) + 1) % BB_P
                cases.append((idx, opened, p2))
                cases.append((idx, opened, proof[:-1]))
            cases.append((idx ^ 1, opened, proof))
            for i, vals, pr in cases:
                got = mmcs.verify_batch(cap, shapes, i, vals, pr)
                want = orc.mmcs_verify(golden["consts_bb"], RF, 13, shapes, cap_h, cap, i, np.concatenate(vals), pr)
                assert got == bool(want) and not got, (cap_h, idx, i)
            assert orc.mmcs_verify(golden["consts_bb"], RF, 13, shapes, cap_h, cap, idx, flat_vals, proof) == 1
        wrong_cap = cap.copy()
        wrong_cap[0][0] = (int(wrong_cap[0][0]) + 1) % BB_P
        opened, proof = mmcs.open_batch(0, tree)
        assert not mmcs.verify_batch(wrong_cap, shapes, 0, opened, proof)
    # Goldilocks: the golden multi-height tree
    gm = mats_of(golden, "mm_gl")
    mmcs = p3hip.MerkleTreeMmcsHip(ctx, 1)
    cap, tree = mmcs.commit(gm)
    dims = [m.shape for m in gm]
    opened, proof = mmcs.open_batch(6, tree)
    assert (np.concatenate(opened) == golden["mm_gl_vals"]).all()
    assert mmcs.verify_batch(cap, dims, 6, opened, proof)
    opened[0][0] = (int(opened[0][0]) + 1) % GL_P
    assert not mmcs.verify_batch(cap, dims, 6, opened, proof)


def test_flat_verify_row(ctx):
    """verify_row: every opened row of a flat tree (odd layers padded with the
    zero digest) verifies; a changed element or sibling does not."""
    for f, dt, P in (("bb", np.uint32, BB_P), ("gl", np.uint64, GL_P)):
        d = rnd(900, (37, 5), f)
        mmcs = p3hip.FlatMerkleTreeMmcsHip(ctx)
        root, tree = mmcs.commit_matrix(d.reshape(-1), 5)
        for r in (0, 17, 36):
            row, proof = mmcs.open_row(tree, r)
            assert (row == d[r]).all()
            assert mmcs.verify_row(root, r, row, proof), (f, r)
            bad = row.copy()
            bad[4] = (int(bad[4]) + 1) % P
            assert not mmcs.verify_row(root, r, bad, proof)
            p2 = proof.copy()
            p2[0][0] = (int(p2[0][0]) + 1) % P
            assert not mmcs.verify_row(root, r, row, p2)


def test_config1_commit_root(ctx, golden, orc):
    """BabyBear 2^16 x 32, log_blowup 1, shift 31, flat and multi-height roots."""
    trace = native.seeded(2024, (1 << 16) * 32).reshape(1 << 16, 32)
    lde = ctx.coset_lde(ctx.upload(trace), 1, 31)
    want_lde = orc.lde("bb", trace, 1, 31)
    assert (lde.download() == want_lde).all()
    want_root = orc.flat_commit("bb", golden["consts_bb"], RF, 13, want_lde)
    assert (ctx.flat_commit(lde).root() == want_root).all()
    assert (ctx.mmcs_commit([lde], 0).root() == want_root).all()


def test_goldilocks_lde_commit(ctx, golden, orc):
    trace = rnd(77, (1 << 10, 8), "gl")
    lde = ctx.coset_lde(ctx.upload(trace), 2, 7)
    want = orc.lde("gl", trace, 2, 7)
    assert (lde.download() == want).all()
    assert (ctx.mmcs_commit([lde]).root() == orc.flat_commit("gl", golden["consts_gl"], RF, 22, want)).all()


# ------------------------------------------------------------------- FRI --
def test_fold_golden(ctx, golden):
    for la in (1, 2, 3):
        cw = golden[f"fold_a{la}_in"]
        out = ctx.fri_fold(ctx.upload(cw), la, golden[f"fold_a{la}_beta"]).download()
        assert (out == golden[f"fold_a{la}_out"]).all(), la


def test_fold_device_pointers(ctx, golden, orc):
    """p3h_fri_fold_device (fold_matrix_cuda_device) on device buffers: the
    golden folds at every arity, a partial fold (n below the folded height:
    rows 0..n-1 only, the rest of the output untouched), and a 2^16 fold
    against the oracle; argument errors raise."""
    for la in (1, 2, 3):
        cw = golden[f"fold_a{la}_in"]
        src = ctx.upload(cw)
        lh = cw.shape[0].bit_length() - 1
        want = golden[f"fold_a{la}_out"]
        dst = ctx.upload(np.full_like(want, 7))
        ctx.fri_fold_device(src.device_ptr(), dst.device_ptr(), golden[f"fold_a{la}_beta"], la, lh, want.shape[0])
        assert (dst.download() == want).all(), la
        part = max(1, want.shape[0] // 2 - 1)
        dst2 = ctx.upload(np.full_like(want, 7))
        ctx.fri_fold_device(src.device_ptr(), dst2.device_ptr(), golden[f"fold_a{la}_beta"], la, lh, part)
        got = dst2.download()
        assert (got[:part] == want[:part]).all() and (got[part:] == 7).all(), la
    cw = rnd(21, (1 << 16, 4))
    beta = rnd(22, (4,))
    src, dst = ctx.upload(cw), ctx.alloc(p3hip.BABYBEAR, 1 << 15, 4)
    ctx.fri_fold_device(src.device_ptr(), dst.device_ptr(), beta, 1, 16, 1 << 15)
    assert (dst.download() == orc.fold(16, 1, beta, cw)).all()
    with pytest.raises(ValueError):
        ctx.fri_fold_device(src.device_ptr(), dst.device_ptr(), beta, 4, 16, 1 << 12)
    with pytest.raises(ValueError):
        ctx.fri_fold_device(src.device_ptr(), dst.device_ptr(), beta, 1, 16, (1 << 15) + 1)
    with pytest.raises(ValueError):
        ctx.fri_fold_device(src.device_ptr(), dst.device_ptr(), [BB_P, 0, 0, 0], 1, 16, 8)


def test_tree_matrices_and_open_batch(ctx, golden, orc):
    """p3h_tree_matrices reports the committed dimensions in commit order for
    trees from any commit call, and open_batch splits the opened row by them."""
    mats = [rnd(40 + i, (h, w)) for i, (h, w) in enumerate([(64, 5), (256, 3), (64, 9), (128, 2)])]
    tree = ctx.mmcs_commit([ctx.upload(m) for m in mats], 1)
    assert tree.dims() == [(m.shape[0], m.shape[1]) for m in mats]
    mm = p3hip.MerkleTreeMmcsHip(ctx, 1)
    vals, proof = mm.open_batch(200, tree)
    for m, v in zip(mats, vals):
        assert (v == m[200 // (256 // m.shape[0])]).all()
    assert mm.verify_batch(tree.cap(), tree.dims(), 200, vals, proof)
    flat = ctx.flat_commit(ctx.upload(mats[1]))
    assert flat.dims() == [(256, 3)]


def test_fri_commit_golden(ctx, golden):
    for case in range(3):
        log_len, lb, lfpl, mla = (int(v) for v in golden[f"fri_{case}_params"])
        roots, final, _ = ctx.fri_commit_phase(ctx.upload(golden[f"fri_{case}_cw"]), lb, lfpl, mla,
                                               golden[f"fri_{case}_betas"])
        assert (roots == golden[f"fri_{case}_roots"]).all(), case
        assert (final == golden[f"fri_{case}_final"]).all(), case


def test_fri_commit_vs_oracle(ctx, golden, orc):
    cw = rnd(9, ((1 << 14), 4))
    betas = rnd(10, (20, 4))
    roots, final, fri = ctx.fri_commit_phase(ctx.upload(cw), 1, 3, 1, betas, keep=True)
    want_roots, want_final = orc.fri_commit(golden["consts_bb"], RF, 13, cw, 1, 3, 1, betas)
    assert len(roots) == 10
    assert (roots == want_roots).all() and (final == want_final).all()
    t0 = fri.layer_tree(0)
    assert (t0.root() == want_roots[0]).all()
    fri.free()


def test_fri_commit_layers_built_together(ctx, golden, orc):
    """With the challenges given, every layer is folded first and all layer
    trees are built together (one leaf-hash launch, then one launch per level).
    At 2^17 the leaf launch and the wide levels take the thread-per-state
    kernels, the narrow levels the lane-parallel ones; arity 8 gives 32-wide
    leaf rows.  Roots against the oracle, and every layer's openings against
    the per-round trees of the integer engine."""
    cw = rnd(21, ((1 << 17), 4))
    betas = rnd(22, (20, 4))
    idx = np.array([0, 1, 2, 3, 5, 8, 13, 21], dtype=np.uint64)
    for mla in (1, 3):
        roots, final, fri = ctx.fri_commit_phase(ctx.upload(cw), 1, 3, mla, betas, keep=True)
        want_roots, want_final = orc.fri_commit(golden["consts_bb"], RF, 13, cw, 1, 3, mla, betas)
        assert len(roots) == len(want_roots)
        assert (roots == want_roots).all() and (final == want_final).all(), mla
        opened = []
        for r in range(len(roots)):
            t = fri.layer_tree(r)
            opened.append(t.open(idx % np.uint64(t.n_leaves)))
        fri.free()
        ctx.set_hash_engine(0)
        try:
            roots0, _, fri0 = ctx.fri_commit_phase(ctx.upload(cw), 1, 3, mla, betas, keep=True)
            assert (roots0 == roots).all()
            for r in range(len(roots)):
                t = fri0.layer_tree(r)
                vals, proof = t.open(idx % np.uint64(t.n_leaves))
                assert (vals == opened[r][0]).all() and (proof == opened[r][1]).all(), (mla, r)
            fri0.free()
        finally:
            ctx.set_hash_engine(1)


# ------------------------------------------------------------------- PCS --
def test_pcs_commit_golden(ctx, golden):
    for case in range(3):
        mats = [ctx.upload(m) for m in mats_of(golden, f"pcs_{case}")]
        t = ctx.pcs_commit(mats, golden[f"pcs_{case}_shifts"], int(golden[f"pcs_{case}_lb"][0]))
        assert (t.root() == golden[f"pcs_{case}_root"]).all(), case


def test_batch_commit_vs_oracle(ctx, golden, orc):
    rng = np.random.default_rng(11)
    traces, ks, ldes = [], [], []
    for i in range(6):
        h, w, k = 1 << int(rng.integers(3, 10)), int(rng.integers(8, 40)), int(rng.integers(1, 3))
        tr = rnd(500 + i, (h, w))
        traces.append(tr)
        ks.append(k)
        ldes.append(orc.lde("bb", tr, k, 31))
    t = ctx.batch_commit([ctx.upload(x) for x in traces], ks, 31, 1)
    c, vals, proof = orc.mmcs("bb", golden["consts_bb"], RF, 13, ldes, 1, 77)
    assert (t.cap() == c).all()
    v, p = t.open([77])
    assert (v[0] == vals).all() and (p[0] == proof).all()


# ---------------------------------------------------------------- errors --
def test_errors_are_raised(ctx):
    with pytest.raises(ValueError):
        ctx.coset_lde(ctx.upload(rnd(1, (3, 4))), 1, 31)  # height not a power of two
    with pytest.raises(ValueError):
        ctx.coset_lde(ctx.upload(rnd(1, (4, 4))), 1, BB_P)  # non-canonical shift
    with pytest.raises(ValueError):
        ctx.fri_fold(ctx.upload(rnd(1, (8, 4))), 4, [1, 2, 3, 4])


def test_budget_rejects_oversized_request(golden):
    with p3hip.Context(0, hbm_budget=1 << 20) as small:
        small.set_poseidon2(BB, golden["consts_bb"], 8, 13)
        with pytest.raises(p3hip.BudgetError):
            small.upload(np.zeros((1 << 20, 4), dtype=U32))
        m = small.upload(np.zeros((16, 8), dtype=U32))  # still usable afterwards
        assert small.flat_commit(m).n_leaves == 16


# ------------------------------------------------------- batch plan / graph --
def _batch_inputs(seed, count, hmin=3, hmax=10):
    rng = np.random.default_rng(seed)
    traces, ks = [], []
    for i in range(count):
        h, w, k = 1 << int(rng.integers(hmin, hmax + 1)), int(rng.integers(8, 40)), int(rng.integers(1, 3))
        traces.append(rnd(seed * 100 + i, (h, w)))
        ks.append(k)
    return traces, ks


def _oracle_batch(orc, golden, traces, ks, cap):
    ldes = [orc.lde("bb", t, k, 31) for t, k in zip(traces, ks)]
    return orc.mmcs("bb", golden["consts_bb"], RF, 13, ldes, cap, 3)


@pytest.mark.parametrize("use_graph,lde_queue,hmin,hmax", [(False, True, 6, 10), (True, True, 6, 10),
                                                        (True, False, 3, 10), (False, True, 11, 14),
                                                        (True, True, 11, 14), (True, False, 11, 14)])
def test_batch_plan_replays_match_oracle(ctx, golden, orc, use_graph, lde_queue, hmin, hmax):
    """The planned batch (LDE as one work-queue launch or grouped launches,
    eager or replayed as a HIP graph) against the oracle's per-trace LDEs and
    multi-height MMCS.  Heights up to 2^10 are one fused inner pass per trace;
    2^11..2^14 add outer passes, so queue items wait on their trace's previous
    pass.  (Transforms of 2^4 and 2^5 rows keep the grouped launches.)"""
    traces, ks = _batch_inputs(21 if hmax <= 10 else 23, 7, hmin, hmax)
    mats = [ctx.upload(t) for t in traces]
    plan = ctx.batch_plan(mats, ks, 31, cap_height=1, use_graph=use_graph, lde_queue=lde_queue)
    assert (plan.lde_launches() == 1) == lde_queue
    c, vals, proof = _oracle_batch(orc, golden, traces, ks, 1)
    for rep in range(3):
        cap = plan.run()
        assert (cap == c).all(), rep
        v, p = plan.tree().open([3])
        assert (v[0] == vals).all() and (p[0] == proof).all()
        # eager work on the same stream between replays
        ctx.coset_lde(ctx.upload(rnd(rep, (64, 8))), 1, 31).download()
    # new contents in the same device buffers: the replay recomputes
    traces2 = [rnd(900 + i, t.shape) for i, t in enumerate(traces)]
    for m, t in zip(mats, traces2):
        m.write(t)
    c2, _, _ = _oracle_batch(orc, golden, traces2, ks, 1)
    for rep in range(2):
        assert (plan.run() == c2).all(), rep
    plan.free()


def test_batch_graph_as_first_gpu_work(tmp_path):
    """A fresh process whose first GPU work is a batch plan captured as a HIP
    graph (so every pass kernel's first launch is inside the capture) gives
    the cap of the eager batched commit."""
    import os
    import subprocess
    import sys

    pkg = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "plonky3-gpu_amd")
    script = tmp_path / "first_work.py"
    script.write_text(f"""
import sys
import numpy as np
sys.path.insert(0, {pkg!r})
import p3hip
g = np.load({os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden.npz")!r})
rng = np.random.default_rng(5)
traces = [rng.integers(0, {BB_P}, size=(1 << h, w), dtype=np.uint32) for h, w in ((12, 20), (10, 33), (12, 7))]
with p3hip.Context(0) as ctx:
    ctx.set_poseidon2(p3hip.BABYBEAR, g["consts_bb"], 8, 13)
    mats = [ctx.upload(t) for t in traces]
    plan = ctx.batch_plan(mats, [1, 2, 1], 31, cap_height=0, use_graph=True)
    cap = plan.run()
    want = ctx.batch_commit(mats, [1, 2, 1], 31, 0).cap()
    assert (cap == want).all(), (cap, want)
    plan.free()
print("ok")
""")
    r = subprocess.run([sys.executable, str(script)], capture_output=True, text=True, timeout=110)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout + r.stderr


def test_autotuned_lde_matches_oracle(ctx, orc):
    best = ctx.autotune_lde(p3hip.BABYBEAR, 12, 24, 1, iters=1)
    assert best["ms"] > 0
    m = rnd(31, (1 << 12, 24))
    assert (ctx.coset_lde(ctx.upload(m), 1, 31).download() == orc.lde("bb", m, 1, 31)).all()


@pytest.fixture
def tuned_ctx(ctx):
    yield ctx
    ctx.set_ntt_tuning(0, 0, 0)


@pytest.mark.parametrize("log_cols,max_stages,log_slab", [(2, 0, 4), (4, 3, 5), (3, 4, 4), (4, 0, 6)])
def test_lde_column_slabs_vs_oracle(tuned_ctx, orc, log_cols, max_stages, log_slab):
    """Slabbed LDEs (every pass of one column slab before the next slab), ragged
    last slab, both fields, 1 and 2 added bits."""
    ctx = tuned_ctx
    ctx.set_ntt_tuning(log_cols, max_stages, log_slab)
    for f, fid, shift, L, w in (("bb", BB, 31, 11, 37), ("bb", BB, 31, 13, 70), ("gl", GL, 7, 10, 21)):
        m = rnd(L * 100 + w, (1 << L, w), f)
        for k in (1, 2):
            got = ctx.coset_lde(ctx.upload(m), k, shift).download()
            assert (got == orc.lde(f, m, k, shift)).all(), (f, L, w, k)


def test_lde_grid_stride_matches_default(tuned_ctx):
    """A pass shape whose block count exceeds one dispatch's grid (4-column
    tiles, 3-stage passes on a 2^20 x 256 trace: 2^24 blocks) gives the same
    LDE as the default shape: equal Merkle roots of the 2^21 x 256 results."""
    ctx = tuned_ctx
    m = ctx.upload(rnd(77, (1 << 20, 256)))
    roots = []
    for tune in ((0, 0, 0), (2, 3, 0)):
        ctx.set_ntt_tuning(*tune)
        lde = ctx.coset_lde(m, 1, 31)
        t = ctx.mmcs_commit([lde], 0)
        roots.append(t.root())
        t.free()
        lde.free()
    assert (roots[0] == roots[1]).all()


# ---------------------------------------------- uint32 Montgomery engine --
@pytest.fixture
def ctx64(ctx):
    """The session context runs the default binary64 engine; this fixture
    switches to the uint32 Montgomery engine for the same checks."""
    ctx.set_hash_engine(0)
    yield ctx
    ctx.set_hash_engine(1)


def test_u32_engine_permutation_and_sponge(ctx64, golden, orc):
    assert (ctx64.permute(BB, golden["perm_bb_in"]) == golden["perm_bb_out"]).all()
    st = rnd(3, 16 * 5000)
    assert (ctx64.permute(BB, st) == orc.permute("bb", golden["consts_bb"], RF, 13, st)).all()
    # extreme inputs: all p-1, all zero
    ext = np.concatenate([np.full(16 * 64, BB_P - 1, dtype=U32), np.zeros(16 * 64, dtype=U32)])
    assert (ctx64.permute(BB, ext) == orc.permute("bb", golden["consts_bb"], RF, 13, ext)).all()
    mmcs = p3hip.FlatMerkleTreeMmcsHip(ctx64)
    for i in range(6):
        d = golden[f"flat_bb_{i}_in"]
        root, _ = mmcs.commit_matrix(d.reshape(-1), d.shape[1])
        assert (root == golden[f"flat_bb_{i}_root"]).all(), i


def test_u32_engine_trees(ctx64, golden, orc):
    for s in range(4):
        mats = mats_of(golden, f"mm_bb_{s}")
        for cap in (0, 1):
            t = ctx64.mmcs_commit([ctx64.upload(m) for m in mats], cap)
            assert (t.cap() == golden[f"mm_bb_{s}_cap{cap}"]).all(), (s, cap)
    trace = native.seeded(2024, (1 << 14) * 40).reshape(1 << 14, 40)
    lde = ctx64.coset_lde(ctx64.upload(trace), 1, 31)
    want = orc.flat_commit("bb", golden["consts_bb"], RF, 13, orc.lde("bb", trace, 1, 31))
    assert (ctx64.mmcs_commit([lde], 0).root() == want).all()
    full = np.full((256, 24), BB_P - 1, dtype=U32)
    assert (ctx64.flat_commit(ctx64.upload(full)).root() == orc.flat_commit("bb", golden["consts_bb"], RF, 13,
                                                                           full)).all()
    for case in range(3):
        log_len, lb, lfpl, mla = (int(v) for v in golden[f"fri_{case}_params"])
        roots, final, _ = ctx64.fri_commit_phase(ctx64.upload(golden[f"fri_{case}_cw"]), lb, lfpl, mla,
                                                 golden[f"fri_{case}_betas"])
        assert (roots == golden[f"fri_{case}_roots"]).all(), case
    traces, ks = _batch_inputs(44, 5)
    plan = ctx64.batch_plan([ctx64.upload(t) for t in traces], ks, 31, cap_height=1, use_graph=True)
    c, _, _ = _oracle_batch(orc, golden, traces, ks, 1)
    assert (plan.run() == c).all() and (plan.run() == c).all()
    plan.free()


@pytest.mark.parametrize("seed", range(16))
def test_random_lde_commit_open_sweep(shape_ctx, golden, orc, seed):
    """Seeded random shapes through the whole path against the oracle: field,
    height 2^0..2^12, width 1..70, added bits 0..3, pass-width cap and column
    tile from the tuner's ranges; LDE, single-matrix commit root and (BabyBear)
    openings of random rows."""
    ctx = shape_ctx
    rng = np.random.default_rng(31337 + seed)
    f = "bb" if seed % 3 else "gl"
    lh, w, k = int(rng.integers(0, 13)), int(rng.integers(1, 71)), int(rng.integers(0, 4))
    if f == "gl":
        lh = min(lh, 10)
    ctx.set_ntt_tuning(int(rng.choice([0, 2, 3, 4, 5, 6, 7])), int(rng.choice([0, 3, 5, 7, 9, 10])), 0)
    m = rnd(seed * 7919 + lh, (1 << lh, w), f)
    shift = 31 if f == "bb" else 7
    want = orc.lde(f, m, k, shift)
    lde = ctx.coset_lde(ctx.upload(m), k, shift)
    assert (lde.download() == want).all(), (f, lh, w, k)
    consts, rp = (golden["consts_bb"], 13) if f == "bb" else (golden["consts_gl"], 22)
    t = ctx.mmcs_commit([lde], 0)
    assert (t.root() == orc.flat_commit(f, consts, RF, rp, want)).all(), (f, lh, w, k)
    if f == "bb":
        idx = rng.integers(0, want.shape[0], size=3)
        vals, proof = t.open(idx)
        for q, i in enumerate(idx):
            row, pr = orc.flat_open(consts, RF, 13, want, int(i))
            assert (vals[q] == row).all() and (proof[q] == pr).all(), (lh, w, k, int(i))
