"""GPU parity of the device-resident DuplexChallenger and of the FRI commit
phase with the challenger in the loop: against the reference build's golden
transcripts and against the oracle, bit-exact."""

import numpy as np
import pytest

import native
import p3hip
from native import BB_P

pytestmark = pytest.mark.gpu
RF = 8


def _run_script(ch, ops):
    out = []
    for code, arg in ops:
        code, arg = int(code), int(arg)
        if code == native.OBSERVE:
            ch.observe([arg & 0xFFFFFFFF])
        elif code == native.SAMPLE:
            out.append(int(ch.sample(1)[0]))
        elif code == native.SAMPLE_BITS:
            out.append(ch.sample_bits(arg))
        elif code == native.GRIND:
            out.append(ch.grind(arg))
        elif code == native.CHECK_WITNESS:
            out.append(1 if ch.check_witness(arg & 63, arg >> 6) else 0)
        elif code == native.SAMPLE_EXT:
            out.extend(int(v) for v in ch.sample(4))
    return np.array(out, dtype=np.uint64)


def test_challenger_golden_transcripts(ctx, golden):
    for i in range(3):
        ch = ctx.challenger()
        got = _run_script(ch, golden[f"chal_{i}_ops"])
        assert got.shape == golden[f"chal_{i}_out"].shape and (got == golden[f"chal_{i}_out"]).all(), i
        ch.free()


def test_grind_from_every_input_length(ctx, golden, orc):
    """The GPU search returns the least passing witness (the oracle's), and the
    transcript continues identically, for 0..7 pending observations."""
    cb = golden["consts_bb"]
    rng = np.random.default_rng(21)
    for ell in range(8):
        prefix = [int(v) for v in rng.integers(0, BB_P, size=24 + ell)]
        bits = 7 + ell % 3
        ops = [(native.OBSERVE, v) for v in prefix] + [(native.GRIND, bits)] + [(native.SAMPLE, 0)] * 5
        want = orc.challenger_run(cb, RF, 13, ops)
        ch = ctx.challenger()
        ch.observe(prefix)
        got = [ch.grind(bits)] + [int(v) for v in ch.sample(5)]
        assert (np.array(got, dtype=np.uint64) == want).all(), ell
        ch.free()


def test_observe_slice_equals_single_observations(ctx):
    vals = np.arange(1, 20, dtype=np.uint32) * 123457
    a, b = ctx.challenger(), ctx.challenger()
    a.observe(vals)
    for v in vals:
        b.observe([v])
    assert (a.sample(11) == b.sample(11)).all()
    assert (a.state() == b.state()).all()


def test_clone_is_independent(ctx):
    a = ctx.challenger()
    a.observe(np.arange(13, dtype=np.uint32))
    b = a.clone()
    sa = a.sample(3)
    assert (b.sample(3) == sa).all()
    b.observe([5])
    a2 = a.sample(2)
    c = ctx.challenger()
    c.observe(np.arange(13, dtype=np.uint32))
    c.sample(3)
    assert (c.sample(2) == a2).all()


def test_observe_cap_equals_observing_the_root(ctx):
    m = ctx.upload(native.seeded(5, 64 * 12).reshape(64, 12))
    t = ctx.flat_commit(m)
    a, b = ctx.challenger(), ctx.challenger()
    a.observe([1, 2, 3])
    b.observe([1, 2, 3])
    a.observe_cap(t)
    b.observe(t.root())
    assert (a.sample(9) == b.sample(9)).all()


def test_fri_commit_with_challenger_golden(ctx, golden):
    for case in range(4):
        log_len, lb, lfpl, mla, pow_bits = (int(v) for v in golden[f"fchal_{case}_params"])
        ch = ctx.challenger()
        ch.observe(golden[f"fchal_{case}_pre"])
        r = ctx.fri_commit_phase_chal(ctx.upload(golden[f"fchal_{case}_cw"]), lb, lfpl, mla, ch, pow_bits)
        for k in ("roots", "betas", "witnesses", "final"):
            assert (r[k] == golden[f"fchal_{case}_{k}"]).all(), (case, k)
        assert (ch.sample(8) == golden[f"fchal_{case}_post"]).all(), case


def test_fri_commit_with_challenger_vs_oracle(ctx, golden, orc):
    cw = native.seeded(99, (1 << 14) * 4).reshape(-1, 4)
    for lb, lfpl, mla, pow_bits in ((1, 3, 1, 0), (2, 2, 3, 12)):
        want = orc.fri_commit_chal(golden["consts_bb"], RF, 13, cw, lb, lfpl, mla, pow_bits, [7, 8])
        ch = ctx.challenger()
        ch.observe([7, 8])
        r = ctx.fri_commit_phase_chal(ctx.upload(cw), lb, lfpl, mla, ch, pow_bits)
        for k in ("roots", "betas", "witnesses", "final"):
            assert (r[k] == want[k]).all(), (lb, lfpl, mla, pow_bits, k)
        assert (ch.sample(8) == want["post"]).all()


def test_challenger_errors(ctx):
    ch = ctx.challenger()
    with pytest.raises(ValueError):
        ch.grind(31)
    with pytest.raises(ValueError):
        ch.sample_bits(65)
    ch.observe(np.zeros(0, dtype=np.uint32))  # empty slice: allowed, no effect
    ch2 = ctx.challenger()
    assert (ch.sample(3) == ch2.sample(3)).all()


def test_grind_least_witness_beyond_the_first_wave():
    """19-bit grinds whose least witness lies past the first 2^20 witnesses
    (the first wave of 4096 chunks of 256): the minimum is found after later
    chunks are claimed and the remaining workgroups drain on the early exit.
    The witnesses are the oracle's (tests/test_oracle.py LATE_GRINDS)."""
    import os
    import sys

    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "plonky3-gpu_amd"))
    import workloads
    from test_oracle import LATE_GRINDS

    c = p3hip.Context(0)
    c.set_poseidon2(p3hip.BABYBEAR, np.array(workloads.p2_consts("bb"), dtype=np.uint32), RF, 13)
    for seed, want in LATE_GRINDS.items():
        prefix = native.seeded(seed, 11)
        ch = c.challenger()
        ch.observe(prefix)
        chk = ch.clone()
        assert ch.grind(19) == want, seed
        assert chk.check_witness(19, want)
        ch.free()
        chk.free()
    c.close()
