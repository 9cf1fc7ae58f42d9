"""GPU parity of TwoAdicFriPcs::open (p3h_pcs_open): opened values, the FRI
proof parts over the quotient vectors, query indices and input rows, against
the reference build's golden proofs and against the oracle."""

import numpy as np
import pytest

import native
import p3hip

pytestmark = pytest.mark.gpu
RF = 8


def _case(golden, case):
    mats, j = [], 0
    while f"popen_{case}_m{j}" in golden:
        mats.append(golden[f"popen_{case}_m{j}"])
        j += 1
    pts = [golden[f"popen_{case}_p{j}"] for j in range(len(mats))]
    lb, lfpl, mla, nq, cpow, qpow = (int(v) for v in golden[f"popen_{case}_params"])
    return mats, pts, lb, lfpl, mla, nq, cpow, qpow


def _gpu_open(ctx, mats, shifts, lb, pts, pre, lfpl, mla, nq, cpow, qpow, keep=False):
    tree = ctx.pcs_commit([ctx.upload(m) for m in mats], shifts, lb)
    ch = ctx.challenger()
    ch.observe(pre)
    ch.observe_cap(tree)
    r = ctx.pcs_open(tree, shifts, pts, ch, lfpl, mla, nq, cpow, qpow, keep=keep)
    r["root"] = tree.root()
    r["post"] = ch.sample(8)
    return r


def test_pcs_open_golden(ctx, golden):
    for case in range(3):
        mats, pts, lb, lfpl, mla, nq, cpow, qpow = _case(golden, case)
        r = _gpu_open(ctx, mats, golden[f"popen_{case}_shifts"], lb, pts, golden[f"popen_{case}_pre"], lfpl, mla, nq,
                      cpow, qpow)
        for k in ("root", "opened", "roots", "commit_witnesses", "final", "query_witness", "query_indices",
                  "query_rows", "post"):
            want = golden[f"popen_{case}_{k}"]
            assert r[k].shape == want.shape and (r[k] == want).all(), (case, k)


def test_pcs_open_vs_oracle(ctx, golden, orc):
    """Larger and more varied than the fixtures: three heights, up to 6 points on
    one matrix (two launches of the interpolation), a width above 256, proof of
    work in both phases."""
    rng = np.random.default_rng(4)
    shapes = [(1 << 10, 300), (1 << 9, 5), (1 << 10, 7), (1 << 8, 2)]
    shifts = [31, 7, 1, 31]
    mats = [rng.integers(0, native.BB_P, size=s, dtype=np.uint32) for s in shapes]
    pts = [rng.integers(0, native.BB_P, size=(k, 4), dtype=np.uint32) for k in (2, 6, 1, 3)]
    args = (1, 2, 2, 12, 4, 6)  # log_blowup, lfpl, max_log_arity, queries, commit pow, query pow
    want = orc.pcs_open(golden["consts_bb"], RF, 13, mats, shifts, args[0], pts, [9], *args[1:])
    r = _gpu_open(ctx, mats, shifts, args[0], pts, [9], *args[1:])
    for k in want:
        assert r[k].shape == want[k].shape and (r[k] == want[k]).all(), k


def test_pcs_open_commit_phase_openings(ctx, golden):
    """The kept FRI layers serve the commit-phase openings of the queries: each
    query's row in every layer opens against that layer's root."""
    mats, pts, lb, lfpl, mla, nq, cpow, qpow = _case(golden, 1)
    r = _gpu_open(ctx, mats, golden["popen_1_shifts"], lb, pts, golden["popen_1_pre"], lfpl, mla, nq, cpow, qpow,
                  keep=True)
    fri = r["fri"]
    for rnd in range(len(r["roots"])):
        t = fri.layer_tree(rnd)
        assert (t.root() == r["roots"][rnd]).all()
    fri.free()
