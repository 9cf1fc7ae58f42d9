"""p3h_fri_fold_device_noalloc (the fold for callers that cannot allocate)
against p3h_fri_fold_device and the oracle, through the C ABI: two calls in a
row with the second smaller (the static scratch keeps the first call's table),
every arity, and no growth of the context's allocator."""

import ctypes as C

import numpy as np
import pytest

import p3hip
from native import BB_P

pytestmark = pytest.mark.gpu


def _noalloc(ctx, d_in, d_out, beta, la, lh, n):
    b = np.ascontiguousarray(beta, dtype=np.uint32)
    return ctx.lib.p3h_fri_fold_device_noalloc(ctx.h, d_in, d_out, b.ctypes.data_as(C.c_void_p), la, lh, n)


def _rnd(seed, shape):
    return np.random.default_rng(seed).integers(0, BB_P, size=shape, dtype=np.uint32)


@pytest.mark.parametrize("la,sizes", [(1, (18, 12)), (2, (14, 9)), (3, (13, 6))])
def test_noalloc_fold_matches_fold_device(ctx, orc, la, sizes):
    for call, lh in enumerate(sizes):
        cw = _rnd(1000 * la + lh, (1 << lh, 4))
        beta = _rnd(7 * la + lh, (4,))
        n = 1 << (lh - la)
        src = ctx.upload(cw)
        want_dev = ctx.alloc(p3hip.BABYBEAR, n, 4)
        ctx.fri_fold_device(src.device_ptr(), want_dev.device_ptr(), beta, la, lh, n)
        want = want_dev.download()
        dst = ctx.upload(np.full((n, 4), 5, dtype=np.uint32))
        ctx.sync()
        mem0 = ctx.memory()
        rc = _noalloc(ctx, src.device_ptr(), dst.device_ptr(), beta, la, lh, n)
        assert rc == 0, ctx.lib.p3h_last_error(ctx.h)
        ctx.sync()
        assert ctx.memory() == mem0, "the allocation-free call changed the context's allocator"
        got = dst.download()
        assert (got == want).all(), (la, lh, call)
        assert (got == orc.fold(lh, la, beta, cw)).all(), (la, lh, call)
        for m in (src, want_dev, dst):
            m.free()


def test_noalloc_fold_rejects_beyond_its_scratch(ctx):
    src = ctx.upload(np.zeros((16, 4), dtype=np.uint32))
    p = src.device_ptr()
    assert _noalloc(ctx, p, p, [1, 2, 3, 4], 1, 24, (1 << 22) + 1) == 1  # P3H_EINVAL, before any device work
    assert b"scratch" in ctx.lib.p3h_last_error(ctx.h)
    assert _noalloc(ctx, p, p, [BB_P, 0, 0, 0], 1, 4, 8) == 1
    assert _noalloc(ctx, p, p, [1, 2, 3, 4], 4, 8, 8) == 1
    src.free()
