"""The fold schedule of the Goldilocks Poseidon2 partial rounds.

Poseidon2GL (plonky3-gpu_amd/csrc/poseidon2.h) runs the 22 partial rounds of
the width-8 permutation on split accumulators: lane i is a pair of 64-bit
words (L_i, H_i) with value H_i 2^32 + L_i, the internal layer updates
L_i <- 2^i L_i + sum(L) and H_i <- 2^i H_i + sum(H), and a lane is folded
back to one 64-bit word (and split again) only before its S-box (lane 0) or
after the rounds whose bit is set in FOLD22; a round constant joins lane 0
as its 32-bit halves just before that lane's fold, and term round 0's
constants join every lane at the final folds.  The arithmetic is exact while
every word stays below 2^64 and every folded H below 2^56, the range of
fold().  Random GPU inputs never reach the worst case, so the schedule read
from the header is checked here by worst-case magnitude recursion, and the
checker is shown to reject schedules that break a bound."""

import math
import os
import re

import pytest

HEADER = os.path.join(os.path.dirname(__file__), "..", "plonky3-gpu_amd", "csrc", "poseidon2.h")
HALF = 2 ** 32 - 1      # a split word: the low or the high 32 bits of a 64-bit value
FOLD_H = 2 ** 56        # fold() needs H < 2^56 (any L below 2^64)
WORD = 2 ** 64


def header_schedule():
    src = open(HEADER).read()
    m = re.search(r"FOLD22\[22\]\s*=\s*\{([^}]*)\}", src)
    assert m, "schedule not found in the header"
    return [int(x, 16) for x in m.group(1).replace("\n", " ").split(",")]


def check(masks):
    """Worst-case word bounds through the partial rounds; returns the number
    of folds of lanes 1..7 and raises AssertionError at the first broken
    bound.  Every lane enters split from a 64-bit word."""
    bl, bh = [HALF] * 8, [HALF] * 8
    folds = 0
    for r, m in enumerate(masks):
        # the round constant's high half joins lane 0's H just before its fold
        assert bh[0] + HALF < FOLD_H, f"round {r}: lane 0 folded with H 2^{math.log2(bh[0]):.1f}"
        assert bl[0] + HALF < WORD
        bl[0] = bh[0] = HALF  # the S-box output, split
        sl, sh = sum(bl), sum(bh)
        bl = [bl[0] + sl] + [(bl[i] << i) + sl for i in range(1, 8)]
        bh = [bh[0] + sh] + [(bh[i] << i) + sh for i in range(1, 8)]
        assert max(bl + bh) < WORD, f"round {r}: a word passes 2^64"
        assert m & 1 == 0, "lane 0 is folded before its S-box, not by the schedule"
        for i in range(1, 8):
            if (m >> i) & 1:
                assert bh[i] < FOLD_H, f"round {r}: lane {i} folded with H 2^{math.log2(bh[i]):.1f}"
                bl[i] = bh[i] = HALF
                folds += 1
    for i in range(8):  # the final folds take term round 0's constants as halves
        assert bh[i] + HALF < FOLD_H, f"final fold of lane {i}: H 2^{math.log2(bh[i]):.1f}"
        assert bl[i] + HALF < WORD
    return folds


def test_header_schedule_is_within_bounds():
    masks = header_schedule()
    assert len(masks) == 22
    assert check(masks) == 53


def test_fold_every_lane_every_round_passes():
    assert check([0xFE] * 22) == 7 * 22


@pytest.mark.parametrize("bad", [
    [0x00] * 22,                                # nothing folded: lane 7 grows 2^7 a round past 2^56
    [0x00, 0x00, 0x40] + [0xFE] * 19,           # lane 7 left unfolded after round 2 grows past 2^56
    [0x00, 0x00, 0xC0, 0xBE, 0xC0, 0x7E] + [0xFE] * 16,  # lane 7 not folded in round 5
    [0x01] + [0xFE] * 21,                       # a schedule that claims lane 0
])
def test_checker_rejects_broken_schedules(bad):
    with pytest.raises(AssertionError):
        check(bad)
