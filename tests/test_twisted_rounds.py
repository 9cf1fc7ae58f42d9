"""The Goldilocks NTT's twisted register rounds (csrc/ntt_impl.h, build_twisted and
the POW2_ROOTS branch of butterflies), checked on the CPU in exact integers.

A radix-2^R round whose stage-j twiddle for the pair (m, m + 2^j) is
omega_{2^(a+q+1)}^(((t mod 2^q) << a) + lo0) (q = Q0 + j, the per-stage form of
the pass kernels, which the reference's per-layer kernels use at
ntt_cuda.hpp:121-153) must equal: DIT -- scale x[m] by the table entry
gamma^bitrev_R(m), then the plain radix-2^R DIT whose stage-j twiddles are
2^E with E = log2(omega_64) (m mod 2^j) 2^(5-j) mod 192; DIF -- the plain DIF,
then the scaling.  The table is built exactly as build_twisted builds it (stage
table twc, size-n table tw with e < n/2, negation for the upper half).  A wrong
bit reversal, a wrong exponent of the table, a transposed (u, m) layout or a
wrong direction of the roots makes one of these cases fail.
"""

import random

import pytest

P = 2**64 - 2**32 + 1


def root(k, inv=False):
    w = pow(7, (P - 1) >> k, P)
    return pow(w, P - 2, P) if inv else w


def bitrev(x, b):
    return int(format(x, f"0{b}b")[::-1], 2) if b else 0


def twisted_table(L, a, S, lo0, inv):
    """build_twisted: the low round at [0, 2^R0), the high round at 2^R0 + m 2^R0 + u."""
    n = 1 << L
    tw = [pow(root(L, inv), e, P) for e in range(n // 2)]
    twc = {(1 << q) - 1 + i: pow(root(q + 1, inv), i, P) for q in range(10) for i in range(1 << q)}
    R0 = S if S <= 5 else (S + 1) // 2
    R1 = S - R0
    count = (1 << S) + (1 << R0) if R1 else (1 << S)
    tt = []
    for i in range(count):
        if i < (1 << R0):
            q0, r, u, m = 0, R0, 0, i
        else:
            j = i - (1 << R0)
            q0, r, m, u = R0, R1, j >> R0, j & ((1 << R0) - 1)
        k = bitrev(m, r)
        q = q0 + r - 1
        x = (k * u) & ((2 << q) - 1)
        v = twc[(1 << q) - 1 + x] if x < (1 << q) else (P - twc[(1 << q) - 1 + x - (1 << q)]) % P
        if lo0:
            e = (k * lo0) << (L - a - q0 - r)
            v = v * (tw[e] if e < n // 2 else (P - tw[e - n // 2]) % P) % P
        tt.append(v)
    return tt


def stage_round(x, R, Q0, u, a, lo0, dit):
    inv = not dit
    y = list(x)
    for j in (range(R) if dit else range(R - 1, -1, -1)):
        for m in range(1 << R):
            if m & (1 << j):
                continue
            q = Q0 + j
            t = u + ((m & ((1 << j) - 1)) << Q0)
            w = pow(root(a + q + 1, inv), (t << a) + lo0, P)
            A, B = y[m], y[m | (1 << j)]
            if dit:
                y[m], y[m | (1 << j)] = (A + w * B) % P, (A - w * B) % P
            else:
                y[m], y[m | (1 << j)] = (A + B) % P, (A - B) * w % P
    return y


def twisted_round(x, R, Q0, u, tt, dit):
    log2w = 39 if dit else 153
    assert pow(2, log2w, P) == root(6, not dit)  # omega_64 (inverse) is a power of two
    base = (0 if Q0 == 0 else (1 << Q0)) + u
    z = list(x)
    if dit:
        z = [z[m] * (tt[base + (m << Q0)] if m else 1) % P for m in range(1 << R)]
    for j in (range(R) if dit else range(R - 1, -1, -1)):
        for m in range(1 << R):
            if m & (1 << j):
                continue
            w = pow(2, log2w * (m & ((1 << j) - 1)) * (1 << (5 - j)) % 192, P)
            A, B = z[m], z[m | (1 << j)]
            if dit:
                z[m], z[m | (1 << j)] = (A + w * B) % P, (A - w * B) % P
            else:
                z[m], z[m | (1 << j)] = (A + B) % P, (A - B) * w % P
    if not dit:
        z = [z[m] * (tt[base + (m << Q0)] if m else 1) % P for m in range(1 << R)]
    return z


@pytest.mark.parametrize("L,a,S", [(6, 0, 6), (9, 0, 9), (8, 0, 3), (10, 3, 7), (12, 5, 7), (12, 6, 6),
                                   (13, 9, 4), (16, 6, 10), (14, 4, 9), (11, 1, 5)])
@pytest.mark.parametrize("dit", [True, False])
def test_twisted_round_equals_stage_round(L, a, S, dit):
    rng = random.Random(L * 1000 + a * 10 + S + dit)
    R0 = S if S <= 5 else (S + 1) // 2
    rounds = [(R0, 0)] + ([(S - R0, R0)] if S > R0 else [])
    for lo0 in ([0] if a == 0 else [0, 1, (1 << a) - 1, rng.randrange(1 << a)]):
        tt = twisted_table(L, a, S, lo0, not dit)
        for R, Q0 in rounds:
            us = range(1 << Q0) if S < 9 else rng.sample(range(1 << Q0), min(4, 1 << Q0))
            for u in us:
                x = [rng.randrange(P) for _ in range(1 << R)]
                assert twisted_round(x, R, Q0, u, tt, dit) == stage_round(x, R, Q0, u, a, lo0, dit), (lo0, R, Q0, u)
