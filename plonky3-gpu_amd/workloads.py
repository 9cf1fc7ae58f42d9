"""Synthetic workloads of the benchmark configurations: shapes, seeds, and the
std::mt19937_64 draws that define every input.

Every input element is ``std::mt19937_64(seed)() % p`` in row-major fill order
(``p3hip.seeded_fill`` on the host); Poseidon2 constants come from an
independent stream, seed XOR ``CONST_TAG``, in the order initial[RF/2][W],
terminal[RF/2][W], internal[RP].  The small draws that pick shapes and opening
indices use a pure-Python mt19937_64 (``MT19937_64``), so a workload is fully
defined by this file.  The same definitions drive the benchmark, the GPU
parity tests and the generator of the full-size fixtures.

Configurations (BabyBear p = 2^31 - 2^27 + 1, shift 31; Goldilocks
p = 2^64 - 2^32 + 1, shift 7):

* ``bb_2^16x32_b1``  BabyBear 2^16 x 32, log_blowup 1 (CPU-runnable commit)
* ``bb_2^20x256_b1`` BabyBear 2^20 x 256, log_blowup 1
* ``bb_2^22x64_b1``  BabyBear 2^22 x 64, log_blowup 1
* ``gl_2^22x64_b2``  Goldilocks 2^22 x 64, log_blowup 2 (Poseidon2 W8 RP22)
* ``FRI_BB4``        BabyBear^4 codeword of 2^23, arity 2, log_blowup 1,
                     final polynomial 2^3, 19 supplied challenges
* ``BATCH``          64 BabyBear traces, h in [12, 20], w in [8, 128],
                     log_blowup 1 or 2, one multi-height MMCS
* ``SERVE_MIX``      16 BabyBear requests from the same height/width
                     distribution, log_blowup 1, 64 openings each
"""

from __future__ import annotations

BB_P = 0x78000001
GL_P = 0xFFFFFFFF00000001
CONST_TAG = 0x5EED_C0DE_0000_0002
CONST_SEED = 1
RF = 8
RP = {"bb": 13, "gl": 22}
P2_WIDTH = {"bb": 16, "gl": 8}
SHIFT = {"bb": 31, "gl": 7}
N_OPEN = 64

_M64 = (1 << 64) - 1


class MT19937_64:
    """std::mt19937_64 (C++11 [rand.predef]): the 64-bit Mersenne Twister with
    n = 312, m = 156, r = 31, a = 0xB5026F5AA96619E9, tempering (29, 0x5555...,
    17, 0x71D67FFFEDA60000, 37, 0xFFF7EEE000000000, 43) and seeding multiplier
    6364136223846793005.  Pure Python, for the few hundred draws that pick
    shapes and indices."""

    N, M = 312, 156
    UPPER, LOWER = _M64 & ~((1 << 31) - 1), (1 << 31) - 1

    def __init__(self, seed: int):
        mt = [seed & _M64]
        for i in range(1, self.N):
            prev = mt[-1]
            mt.append((6364136223846793005 * (prev ^ (prev >> 62)) + i) & _M64)
        self.mt, self.idx = mt, self.N

    def _twist(self):
        mt = self.mt
        for i in range(self.N):
            y = (mt[i] & self.UPPER) | (mt[(i + 1) % self.N] & self.LOWER)
            v = mt[(i + self.M) % self.N] ^ (y >> 1)
            if y & 1:
                v ^= 0xB5026F5AA96619E9
            mt[i] = v
        self.idx = 0

    def __call__(self) -> int:
        if self.idx >= self.N:
            self._twist()
        y = self.mt[self.idx]
        self.idx += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000
        y ^= (y << 37) & 0xFFF7EEE000000000
        y ^= y >> 43
        return y & _M64


def draws(seed: int, n: int) -> list:
    g = MT19937_64(seed)
    return [g() for _ in range(n)]


def field_elems(seed: int, n: int, field: str) -> list:
    p = BB_P if field == "bb" else GL_P
    return [v % p for v in draws(seed, n)]


def p2_consts(field: str, seed: int = CONST_SEED) -> list:
    """Poseidon2 constants: initial[RF/2][W], terminal[RF/2][W], internal[RP]."""
    return field_elems(seed ^ CONST_TAG, RF * P2_WIDTH[field] + RP[field], field)


def open_indices(seed: int, height: int, n: int = N_OPEN) -> list:
    """Opening indices: n draws of the stream seed XOR 0x0BE4, each mod height."""
    return [d % height for d in draws(seed ^ 0x0BE4, n)]


# Single-trace commits: field, log_n, width, log_blowup, trace seed.
SINGLE = {
    "bb_2^16x32_b1": {"field": "bb", "log_n": 16, "width": 32, "added": 1, "seed": 0x1001},
    "bb_2^20x256_b1": {"field": "bb", "log_n": 20, "width": 256, "added": 1, "seed": 0x2A},
    "bb_2^22x64_b1": {"field": "bb", "log_n": 22, "width": 64, "added": 1, "seed": 0x2B},
    "gl_2^22x64_b2": {"field": "gl", "log_n": 22, "width": 64, "added": 2, "seed": 0x3A},
}

# FRI commit phase over BabyBear^4: codeword stream `seed` (2^log_len x 4
# words), challenges from the stream seed + 1 (4 words per round).
FRI_BB4 = {"log_len": 23, "log_blowup": 1, "log_final_poly_len": 3, "max_log_arity": 1, "seed": 0x3B, "rounds": 19}

# Batched commit: 64 traces, cap of 2^4 digests.
BATCH = {"count": 64, "seed": 0x40, "cap_height": 4, "hmin": 12, "hmax": 20, "wmin": 8, "wmax": 128}

# Serving mix: 16 requests, log_blowup 1.
SERVE_MIX = {"count": 16, "seed": 0x50}


def batch_shapes(seed: int = BATCH["seed"], count: int = BATCH["count"], blowups=(1, 2)) -> list:
    """(log_height, width, log_blowup, trace_seed) per trace: three draws of
    the stream `seed` per trace pick h = hmin + d0 mod 9, w = wmin + d1 mod 121
    and the blowup d2 mod len(blowups); trace i's elements come from the
    stream seed * 1000 + i."""
    d = draws(seed, 3 * count)
    nh = BATCH["hmax"] - BATCH["hmin"] + 1
    nw = BATCH["wmax"] - BATCH["wmin"] + 1
    out = []
    for i in range(count):
        h = BATCH["hmin"] + d[3 * i] % nh
        w = BATCH["wmin"] + d[3 * i + 1] % nw
        k = blowups[d[3 * i + 2] % len(blowups)]
        out.append((h, w, k, seed * 1000 + i))
    return out


def serve_shapes(seed: int = SERVE_MIX["seed"], count: int = SERVE_MIX["count"]) -> list:
    """The serving mix: the batch distribution with log_blowup 1."""
    return batch_shapes(seed, count, blowups=(1,))


def committed_elems(shapes) -> int:
    return sum((1 << (h + k)) * w for h, w, k, _ in shapes)
