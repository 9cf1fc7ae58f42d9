"""Serving queue: every request's commitment and openings are the same alone
(plain per-request API) and in the mix (continuous batching over several
streams), and admission respects the HBM budget."""

import numpy as np
import pytest

import p3hip
from native import BB_P

pytestmark = pytest.mark.gpu


def _requests(seed, count, hmin=5, hmax=12):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        h, w = 1 << int(rng.integers(hmin, hmax + 1)), int(rng.integers(8, 48))
        out.append(rng.integers(0, BB_P, size=(h, w), dtype=np.uint32))
    return out


def test_results_same_alone_and_in_the_mix(ctx, golden, orc):
    reqs = _requests(5, 16)
    with p3hip.Server(golden["consts_bb"], 8, 13, n_slots=3, max_batch=4) as sv:
        ids = [sv.submit(t, 1, 31) for t in reqs]
        rng = np.random.default_rng(9)
        for rid, t in zip(ids, reqs):
            root, lat = sv.wait(rid)
            assert lat >= 0
            lde = ctx.coset_lde(ctx.upload(t), 1, 31)
            tree = ctx.mmcs_commit([lde], 0)
            assert (root == tree.root()).all()
            idx = rng.integers(0, 2 * t.shape[0], size=4)
            v1, p1 = sv.open(rid, idx, t.shape[1], tree.n_layers)
            v2, p2 = tree.open(idx)
            assert (v1 == v2).all() and (p1 == p2).all()
            sv.release(rid)
        st = sv.stats()
        assert st["completed"] == 16 and st["reserved_bytes"] == 0
    # small requests also against the oracle
    t = reqs[0]
    want = orc.flat_commit("bb", golden["consts_bb"], 8, 13, orc.lde("bb", t, 1, 31))
    with p3hip.Server(golden["consts_bb"], 8, 13, n_slots=1, max_batch=1) as sv:
        rid = sv.submit(t, 1, 31)
        assert (sv.wait(rid)[0] == want).all()


def test_admission_under_a_small_budget(golden):
    reqs = _requests(7, 8, 10, 10)  # equal sizes
    one = 4 * (1 << 10) * 48 * 3 + (1 << 11) * 64 + (64 << 10)  # generous single-request footprint
    with p3hip.Server(golden["consts_bb"], 8, 13, hbm_budget=3 * one, n_slots=2, max_batch=8) as sv:
        ids = [sv.submit(t, 1, 31) for t in reqs]
        roots = []
        for rid in ids:
            roots.append(sv.wait(rid)[0])
            assert sv.stats()["reserved_bytes"] <= 3 * one
            sv.release(rid)
        assert sv.stats()["completed"] == 8
    with p3hip.Server(golden["consts_bb"], 8, 13, hbm_budget=1 << 16, n_slots=1) as sv:
        with pytest.raises(p3hip.BudgetError):
            sv.submit(reqs[0], 1, 31)


def test_pinned_host_buffers(ctx, golden, orc):
    """Uploads and serving submissions from page-locked host memory give the
    oracle's commitment, and the memory reads back what was written."""
    rng = np.random.default_rng(21)
    t = rng.integers(0, BB_P, size=(1 << 9, 13), dtype=np.uint32)
    pt = p3hip.pinned_empty(t.shape, np.uint32)
    pt[...] = t
    assert isinstance(pt, p3hip.PinnedArray) and (pt == t).all()
    want = orc.flat_commit("bb", golden["consts_bb"], 8, 13, orc.lde("bb", t, 1, 31))
    m = ctx.upload(pt)
    assert (m.download() == t).all()
    lde = ctx.coset_lde(m, 1, 31)
    assert (ctx.mmcs_commit([lde], 0).root() == want).all()
    with p3hip.Server(golden["consts_bb"], 8, 13, n_slots=2, max_batch=2) as sv:
        ids = [sv.submit(pt, 1, 31), sv.submit(pt[: 1 << 8], 1, 31)]
        assert (sv.wait(ids[0])[0] == want).all()
        want_half = orc.flat_commit("bb", golden["consts_bb"], 8, 13, orc.lde("bb", t[: 1 << 8], 1, 31))
        assert (sv.wait(ids[1])[0] == want_half).all()
    # a view outlives the array it was cut from
    tail = pt[-4:]
    del pt
    assert (tail == t[-4:]).all()


def test_destroy_with_requests_queued_and_a_waiter(golden):
    """Destroying the server while requests are still queued and another thread
    waits on the last one: the waiter returns -- with the root, or with
    P3H_ESTATE when its request never ran -- and nothing hangs."""
    import ctypes
    import threading

    reqs = _requests(13, 24, 11, 12)
    sv = p3hip.Server(golden["consts_bb"], 8, 13, n_slots=1, max_batch=1)
    ids = [sv.submit(t, 1, 31) for t in reqs]
    h, out = sv.h, {}
    root = np.zeros(8, dtype=np.uint32)

    def waiter():
        out["rc"] = sv.lib.p3h_serve_wait(h, ids[-1], root.ctypes.data_as(ctypes.c_void_p), None)

    th = threading.Thread(target=waiter)
    th.start()
    sv.wait(ids[0])
    sv.close()
    th.join(timeout=120)
    assert not th.is_alive(), "a waiter hung across destroy"
    assert out["rc"] in (0, 5), out  # P3H_OK or P3H_ESTATE
