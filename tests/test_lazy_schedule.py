"""The deferred-reduction schedule of the binary64 Poseidon2 engine.

Poseidon2F64 (plonky3-gpu_amd/csrc/poseidon2_f64.h) skips the modular
reduction of some lanes after each of the 13 partial rounds.  The values stay
exact integers as long as every magnitude stays below 2^53, and the S-box
stays exact as long as its input squared stays inside the range of the
magic-constant reduction (|v| < 2^51 p).  Random GPU inputs never reach the
worst case, so these bounds are checked here, on the CPU, by worst-case
magnitude recursion over the schedule read from the header -- and the checker
is shown to reject schedules that break each bound."""

import math
import os
import re

import pytest

HEADER = os.path.join(os.path.dirname(__file__), "..", "plonky3-gpu_amd", "csrc", "poseidon2_f64.h")
P = 2013265921
REDUCED = P // 2 + 1          # |reduce(v)| <= p/2 + 1 for |v| < 2^53
SBOX_OUT = 2 ** 31            # reduce(h) + l with |l| <= 2^28
RC = P - 1                    # canonical round constant
MDS_OUT = 35 * SBOX_OUT       # MDS-light: coefficient sum 7 per quad + 28 from the column sums
REDUCE_RANGE = 2 ** 51 * P    # reduce() needs |v / p| < 2^51
EXACT = 2 ** 53


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
        assert x * SBOX_OUT < REDUCE_RANGE
        s = SBOX_OUT + sum(b[1:])
        nb = [SBOX_OUT + s] + [(1 << i) * b[i] + s for i in range(1, 16)]
        assert max(nb) < EXACT, f"round {r}: lane 2^{math.log2(max(nb)):.2f} not exact"
        b = [REDUCED if (m >> i) & 1 else nb[i] for i in range(16)]
    for i in range(16):
        x = b[i] + RC
        assert x * x < REDUCE_RANGE, f"terminal S-box input of lane {i}: 2^{math.log2(x):.2f}"
    return b


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
