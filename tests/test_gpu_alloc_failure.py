"""Allocation failures (P3H_FAIL_ALLOC_AFTER): for each path, a failure is
injected at every device allocation in turn -- the first, the second, ... until
the path needs fewer allocations than allowed and succeeds.  Every failure must
surface as a clean P3H_ENOMEM status, and once the caller has freed the
handles it holds, nothing but the context's table caches may be left in use.
The final, uninjected run must match the oracle."""

import os

import numpy as np
import pytest

import p3hip
from native import BB_P

pytestmark = pytest.mark.gpu

RF = 8


def _rnd(seed, shape):
    return np.random.default_rng(seed).integers(0, BB_P, size=shape, dtype=np.uint32)


def _sweep(golden, path, max_allocs=120):
    """Runs path(ctx, keep) with the failure injected after 0, 1, 2, ...
    allocations; keep collects the handles the path made.  Returns the number
    of injected failures and the successful run's result."""
    failures = 0
    for n in range(max_allocs):
        os.environ["P3H_FAIL_ALLOC_AFTER"] = str(n)
        try:
            ctx = p3hip.Context(0)
        finally:
            del os.environ["P3H_FAIL_ALLOC_AFTER"]
        ctx.set_poseidon2(p3hip.BABYBEAR, golden["consts_bb"], RF, 13)
        keep, result, failed = [], None, False
        try:
            result = path(ctx, keep)
        except RuntimeError as e:
            assert "[status 2]" in str(e) and "injected allocation failure" in str(e), str(e)
            failed = True
        for h in reversed(keep):
            h.free()
        ctx.sync()
        assert ctx.memory()["in_use"] == ctx.table_bytes(), (n, ctx.memory(), ctx.table_bytes())
        ctx.close()
        if not failed:
            return failures, result
        failures += 1
    raise AssertionError(f"path still failing after {max_allocs} allocations")


def test_lde_commit_open_under_allocation_failures(golden, orc):
    trace = _rnd(1, (1 << 10, 24))

    def path(ctx, keep):
        m = ctx.upload(trace)
        keep.append(m)
        lde = ctx.coset_lde(m, 1, 31)
        keep.append(lde)
        tree = ctx.mmcs_commit([lde], 1)
        keep.append(tree)
        vals, proof = tree.open([0, 5, 2047])
        flat = ctx.flat_commit(lde)
        keep.append(flat)
        return tree.cap(), vals, flat.root()

    failures, (cap, vals, root) = _sweep(golden, path)
    assert failures >= 5
    lde = orc.lde("bb", trace, 1, 31)
    assert (root == orc.flat_commit("bb", golden["consts_bb"], RF, 13, lde)).all()
    assert (cap == orc.mmcs("bb", golden["consts_bb"], RF, 13, [lde], cap_height=1)).all()
    assert (vals == lde[[0, 5, 2047]]).all()


def test_fri_and_fold_under_allocation_failures(golden, orc):
    cw = _rnd(2, (1 << 12, 4))
    betas = _rnd(3, (16, 4))

    def path(ctx, keep):
        m = ctx.upload(cw)
        keep.append(m)
        f = ctx.fri_fold(m, 1, betas[0])
        keep.append(f)
        roots, final, fri = ctx.fri_commit_phase(m, 1, 2, 1, betas, keep=True)
        keep.append(fri)
        return f.download(), roots, final

    failures, (folded, roots, final) = _sweep(golden, path)
    assert failures >= 3
    assert (folded == orc.fold(12, 1, betas[0], cw)).all()
    want_roots, want_final = orc.fri_commit(golden["consts_bb"], RF, 13, cw, 1, 2, 1, betas)
    assert (roots == want_roots).all() and (final == want_final).all()


def test_batch_and_pcs_commit_under_allocation_failures(golden, orc):
    traces = [_rnd(10 + i, (1 << h, w)) for i, (h, w) in enumerate([(8, 8), (6, 16), (8, 24)])]
    ks = [1, 2, 1]

    def path(ctx, keep):
        mats = []
        for t in traces:
            mats.append(ctx.upload(t))
            keep.append(mats[-1])
        tree = ctx.batch_commit(mats, ks, 31, 0)
        keep.append(tree)
        pcs = ctx.pcs_commit(mats[:2], [31, 31], 1)
        keep.append(pcs)
        return tree.root(), pcs.root()

    failures, (root, pcs_root) = _sweep(golden, path)
    assert failures >= 5
    ldes = [orc.lde("bb", t, k, 31) for t, k in zip(traces, ks)]
    assert (root == orc.mmcs("bb", golden["consts_bb"], RF, 13, ldes)[0]).all()
    assert (pcs_root == orc.pcs_commit(golden["consts_bb"], RF, 13, traces[:2], [31, 31], 1)).all()
