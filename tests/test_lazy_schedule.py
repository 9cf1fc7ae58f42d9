"""The deferred-reduction schedule of the binary64 Poseidon2 engine.

Poseidon2F64 (plonky3-gpu_amd/csrc/poseidon2_f64.h) skips the modular
reduction of some lanes after each of the 13 partial rounds.  The values stay
exact integers as long as every magnitude stays below 2^53, and the S-box
stays exact as long as its input squared stays inside the range of the
magic-constant reduction (|v| < 2^51 p).  Products through mul_q (a shared
multiplier b/p and the quotient correction -q (p - 1), exact while
15 q < 2^53) need |a b| < 2^79 instead: every full-round S-box (x^2, x2 x,
x3^2, x6 x) and the partial-round products after the square.  Random GPU
inputs never reach the worst case, so these bounds are checked here, on the
CPU, by worst-case magnitude recursion over the schedule read from the
header -- and the checker is shown to reject schedules that break each
bound."""

import math
import os
import re

import pytest

HEADER = os.path.join(os.path.dirname(__file__), "..", "plonky3-gpu_amd", "csrc", "poseidon2_f64.h")
P = 2013265921
REDUCED = P // 2 + 1          # |reduce(v)| <= p/2 + 1 for |v| < 2^53
SBOX_OUT = 2 ** 31            # S-box output, with a folded constant (< 0.51 p + p/2)
RC = P - 1                    # canonical round constant
MDS_OUT = 35 * SBOX_OUT       # MDS-light: coefficient sum 7 per quad + 28 from the column sums
REDUCE_RANGE = 2 ** 51 * P    # reduce() needs |v / p| < 2^51
MULQ_RANGE = 2 ** 79          # mul_q needs |a b| < 2^79 (15 q < 2^53)
SQUARE_OUT = P * 0.74 + 2 ** 28  # two-product square of a lane below 2^41: reduce(h) + l
EXACT = 2 ** 53


def mulq_out(ab):
    """|a b - q p| for |a b| <= ab: q = round(a (b/p)) is off from a b / p by at
    most 1/2 plus the relative rounding 2^-52 of b/p."""
    assert ab < MULQ_RANGE, f"mul_q product 2^{math.log2(ab):.2f}"
    return P * (0.5 + ab / P * 2.0 ** -52)


def check_full_rounds():
    """Full rounds: inputs are MDS-light outputs, plus a canonical constant in
    round 0 of each half (init: after MDS-light of a state below p; term: after
    the last partial round, whose lanes are reduced); every product is a
    mul_q, and the last one carries a balanced folded constant (|c| <= p/2)."""
    for x in (MDS_OUT + RC, MDS_OUT, REDUCED + RC):
        x2 = mulq_out(x * x)
        x3 = mulq_out(x2 * x)
        x6 = mulq_out(x3 * x3)
        out = mulq_out(x6 * x) + P // 2
        assert out < SBOX_OUT, f"S-box output 2^{math.log2(out):.4f}"


def header_schedule():
    src = open(HEADER).read()
    m = re.search(r"LAZY13\[13\]\s*=\s*\{([^}]*)\}", src)
    assert m, "schedule not found in the header"
    return [int(x, 16) for x in m.group(1).replace("\n", " ").split(",")]


def check(masks):
    """Worst-case magnitudes through the partial rounds and into the terminal
    full rounds; raises AssertionError naming the first violated bound."""
    b = [MDS_OUT] * 16
    for r, m in enumerate(masks):
        x = b[0] + RC
        assert x * x < REDUCE_RANGE, f"round {r}: S-box square 2^{math.log2(x * x):.2f}"
        x3 = mulq_out(x * SQUARE_OUT)
        x6 = mulq_out(x3 * x3)
        assert mulq_out(x6 * x) <= SBOX_OUT
        s = SBOX_OUT + sum(b[1:])
        nb = [SBOX_OUT + s] + [(1 << i) * b[i] + s for i in range(1, 16)]
        assert max(nb) < EXACT, f"round {r}: lane 2^{math.log2(max(nb)):.2f} not exact"
        b = [REDUCED if (m >> i) & 1 else nb[i] for i in range(16)]
    for i in range(16):
        x = b[i] + RC
        assert x * x < MULQ_RANGE, f"terminal S-box input of lane {i}: 2^{math.log2(x):.2f}"
    return b


def test_full_rounds_within_bounds():
    check_full_rounds()


def test_header_schedule_is_within_bounds():
    masks = header_schedule()
    assert len(masks) == 13
    check(masks)
    assert sum(bin(m).count("1") for m in masks) == 132


def test_all_reduced_schedule_passes():
    check([0xFFFF] * 13)


@pytest.mark.parametrize("bad", [
    [0x7FFF] + [0xFFFF] * 12,        # lane 15 left unreduced after round 0: 2^15 growth overflows 2^53
    [0xFFFF] * 12 + [0xEFFF],        # lane 12 unreduced into the terminal rounds: S-box square too large
    [0x0000] * 13,                   # nothing reduced
    [0xFF80, 0xFC70] + [0xFFFF] * 11,  # round 1 leaves lane 0 unreduced after an unreduced round 0
])
def test_checker_rejects_broken_schedules(bad):
    with pytest.raises(AssertionError):
        check(bad)
