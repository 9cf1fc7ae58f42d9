"""Round-constant folding of the binary64 Poseidon2 engine.

The one-lane-per-state engine (plonky3-gpu_amd/csrc/poseidon2_f64.h,
Poseidon2F64::permute) adds the constants of init round 0, term round 0 and
partial rounds 1.. itself; the constants of every other full round, and of
partial round 0, arrive folded into the S-boxes of the round before: with M the
MDS-light matrix, M (y + c) = M y + rc (mod p) for c = M^-1 rc.  The host
builds c in runtime.cpp (bb_f64_params).  Here the folded schedule is run in
exact integers (no reduction, so every value is a representative of the
engine's) and its canonical output is compared with the oracle's permutation,
with M^-1 computed independently by elimination over F_p; a wrong round for
a folded constant, a wrong lane for the partial-round constant or a missing
addition fails the comparison.  The mul_q bounds of the same engine are in
test_lazy_schedule.py."""

import numpy as np
import pytest

import native

P = native.BB_P
RF, RP = 8, 13


def mat4(v):
    t01, t23 = v[0] + v[1], v[2] + v[3]
    t0123 = t01 + t23
    return [t0123 + v[1] + t01, t0123 + v[1] + 2 * v[2], t0123 + v[3] + t23, t0123 + v[3] + 2 * v[0]]


def mds_light(s):
    y = mat4(s[0:4]) + mat4(s[4:8]) + mat4(s[8:12]) + mat4(s[12:16])
    sums = [y[k] + y[k + 4] + y[k + 8] + y[k + 12] for k in range(4)]
    return [y[i] + sums[i & 3] for i in range(16)]


def mds_inverse_times(v):
    """x with mds_light(x) = v (mod p), by Gauss-Jordan elimination."""
    cols = [mds_light([1 if k == j else 0 for k in range(16)]) for j in range(16)]
    a = [[cols[j][i] % P for j in range(16)] + [v[i] % P] for i in range(16)]
    for c in range(16):
        piv = next(r for r in range(c, 16) if a[r][c])
        a[c], a[piv] = a[piv], a[c]
        inv = pow(a[c][c], P - 2, P)
        a[c] = [x * inv % P for x in a[c]]
        for r in range(16):
            if r != c and a[r][c]:
                f = a[r][c]
                a[r] = [(x - f * y) % P for x, y in zip(a[r], a[c])]
    return [a[i][16] for i in range(16)]


def balanced(x):
    return x - P if x > P // 2 else x


def folded_permute(state, consts):
    hf = RF // 2
    init = [list(consts[16 * r:16 * r + 16]) for r in range(hf)]
    term = [list(consts[16 * (hf + r):16 * (hf + r) + 16]) for r in range(hf)]
    internal = list(consts[32 * hf:32 * hf + RP])
    fold_init = [mds_inverse_times(init[r + 1]) if r + 1 < hf else mds_inverse_times([internal[0]] + [0] * 15)
                 for r in range(hf)]
    fold_term = [mds_inverse_times(term[r + 1]) if r + 1 < hf else [0] * 16 for r in range(hf)]
    s = mds_light([int(x) for x in state])
    s = [x + c for x, c in zip(s, init[0])]
    for r in range(hf):
        s = mds_light([x ** 7 + balanced(c) for x, c in zip(s, fold_init[r])])
        s = [x % P for x in s]  # any representative: keeps Python integers small
    for r in range(RP):
        x = s[0] if r == 0 else s[0] + internal[r]
        s[0] = x ** 7
        tot = sum(s)
        s = [s[0] + tot] + [(s[i] << i) + tot for i in range(1, 16)]
        s = [x % P for x in s]
    s = [x + c for x, c in zip(s, term[0])]
    for r in range(hf):
        s = mds_light([x ** 7 + balanced(c) for x, c in zip(s, fold_term[r])])
        s = [x % P for x in s]
    return [x % P for x in s]


def test_mds_inverse_solves():
    rng = np.random.default_rng(3)
    for _ in range(4):
        v = [int(x) for x in rng.integers(0, P, 16)]
        x = mds_inverse_times(v)
        assert [y % P for y in mds_light(x)] == v


@pytest.mark.parametrize("seed", [1, 2])
def test_folded_schedule_matches_oracle(seed):
    consts = native.p2_consts(seed, "bb", RF, RP)
    states = native.seeded(seed + 100, 4 * 16).reshape(4, 16)
    want = native.oracle().permute("bb", consts, RF, RP, states)
    for st, w in zip(states, want):
        assert folded_permute(st, [int(c) for c in consts]) == [int(x) for x in w]


def test_fold_placement_is_checked():
    """The comparison above is sensitive to the folded constants: partial
    round 0's constant doubled (as if it were both folded and added) fails."""
    consts = [int(c) for c in native.p2_consts(1, "bb", RF, RP)]
    st = native.seeded(101, 16)
    want = [int(x) for x in native.oracle().permute("bb", consts, RF, RP, st.reshape(1, 16))[0]]
    doubled = list(consts)
    doubled[32 * (RF // 2)] = (2 * doubled[32 * (RF // 2)]) % P  # the fold and an explicit add
    assert folded_permute(st, doubled) != want
