"""Benchmark: committed field elements/s for coset LDE + Poseidon2 Merkle commit.

Default workload (one MI355X): BabyBear trace of 2^20 rows x 256 columns,
log_blowup 1 (2^21 x 256 LDE, 2 GiB), shift 31, Poseidon2 width 16 / x^7 /
8 full + 13 partial rounds with seeded constants, padding-free sponge leaves
(rate 8, digest 8), 2-to-1 truncated-permutation tree, cap 1 root.
A step = coset LDE + leaf hashing + tree compression of one resident trace,
ending with the root copied to the host.  Inputs are resident in HBM before
the timed region.

    python bench.py [--steps K] [--warmup W] [--workload NAME]

Prints one JSON line.  `--gpus N` with N > 1 exits: this build runs on one GPU.
"""

from __future__ import annotations

import argparse
import importlib.util
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
BB_P = 0x78000001
GL_P = 0xFFFFFFFF00000001

# Published MI355X peaks used for roofline fractions.
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6  # vector and matrix FP64 are the same figure on MI355X
# The BabyBear leaf hash runs the exact binary64 Poseidon2 engine: one width-16
# permutation (8 full + 13 partial rounds, 132 of the 208 partial-round lane
# reductions kept) is 2142 add + 564 mul + 2255 fma binary64 operations =
# 7216 flops in 4961 FP64 VALU instructions (counted from poseidon2_f64.h);
# each digest lane's final canonical reduction adds 5 flops.
P2_F64_FLOPS_PER_PERM = 7216
P2_F64_FLOPS_PER_DIGEST = 8 * 5
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic.json")

WORKLOADS = {
    # name: (field, log_n, width, log_blowup, poseidon2 rounds_p)
    "bb_2^20x256_b1": ("bb", 20, 256, 1, 13),
    "bb_2^22x64_b1": ("bb", 22, 64, 1, 13),
    "bb_2^16x32_b1": ("bb", 16, 32, 1, 13),
    "gl_2^22x64_b2": ("gl", 22, 64, 2, 22),
}
DEFAULT = "bb_2^20x256_b1"
SPECIAL = ("batch64", "serve", "fri_bb4_2^23", "fri_bb4_2^23_chal", "grind_bb", "pcs_open_bb_2^20x256")


def load_p3hip():
    spec = importlib.util.spec_from_file_location("p3hip", os.path.join(ROOT, "plonky3-gpu_amd", "p3hip.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["p3hip"] = mod
    spec.loader.exec_module(mod)
    return mod


def p2_constants(field: str, rp: int, seed: int = 1) -> np.ndarray:
    rng = np.random.default_rng(seed)
    if field == "bb":
        return rng.integers(0, BB_P, size=8 * 16 + rp, dtype=np.uint32)
    return rng.integers(0, GL_P, size=8 * 8 + rp, dtype=np.uint64)


def host_cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(field: str, width: int, log_blowup: int, rp: int):
    """The reference's own CPU path (oracle/_ref, compiled from the reference
    sources in the build container) on a bounded sample of the same workload:
    2^15 rows x width, same blowup, LDE + single-matrix Poseidon2 commit, one
    thread.  Falls back to the oracle restatement when the reference build is
    absent."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import native  # test infrastructure; used here only as the CPU baseline

    lib = native.reference()
    kind = "reference"
    if lib is None:
        lib, kind = native.oracle(), "port"
    log_n = 15 if field == "bb" else 13
    n = 1 << log_n
    rng = np.random.default_rng(3)
    if field == "bb":
        trace = rng.integers(0, BB_P, size=(n, width), dtype=np.uint32)
    else:
        trace = rng.integers(0, GL_P, size=(n, width), dtype=np.uint64)
    consts = p2_constants(field, rp)
    t0 = time.perf_counter()
    lde = lib.lde(field, trace, log_blowup, 31 if field == "bb" else 7)
    lib.flat_commit(field, consts, 8, rp, lde)
    dt = time.perf_counter() - t0
    elems = (n << log_blowup) * width
    name = "BabyBear" if field == "bb" else "Goldilocks"
    return {"value": elems / dt, "unit": "committed field elems/s", "cores": 1, "kind": kind,
            "sample": f"{name} 2^{log_n} x {width}, log_blowup {log_blowup}: LDE + Poseidon2 flat commit, "
                      f"{dt:.2f} s on 1 host thread", "cpu_model": host_cpu_model()}


def cpu_baseline_threads(field: str, width: int, log_blowup: int, rp: int, threads: int = 0):
    """The oracle restatement on several host threads (columns split for the
    LDE, rows and tree nodes for the commit; the root equals the single-thread
    one), on a sample 4x the single-thread one.  Threads: 16, the host cores
    that go with one GPU of an eight-GPU MI355X node, or fewer if the host has
    fewer."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import native  # test infrastructure; used here only as the CPU baseline

    lib = native.oracle()
    threads = threads or min(16, os.cpu_count() or 1)
    log_n = 17 if field == "bb" else 15
    n = 1 << log_n
    rng = np.random.default_rng(5)
    if field == "bb":
        trace = rng.integers(0, BB_P, size=(n, width), dtype=np.uint32)
    else:
        trace = rng.integers(0, GL_P, size=(n, width), dtype=np.uint64)
    consts = p2_constants(field, rp)
    t0 = time.perf_counter()
    lib.lde_commit_mt(field, consts, 8, rp, trace, log_blowup, 31 if field == "bb" else 7, threads)
    dt = time.perf_counter() - t0
    elems = (n << log_blowup) * width
    name = "BabyBear" if field == "bb" else "Goldilocks"
    return {"value": elems / dt, "unit": "committed field elems/s", "cores": threads, "kind": "port",
            "sample": f"{name} 2^{log_n} x {width}, log_blowup {log_blowup}: LDE + Poseidon2 flat commit, "
                      f"{dt:.2f} s on {threads} host threads", "cpu_model": host_cpu_model()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default=DEFAULT, choices=sorted(WORKLOADS) + list(SPECIAL))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--hash-engine", type=int, default=1, help="BabyBear Poseidon2: 0 uint32, 1 binary64")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--tuning-file", default=None,
                    help="load the LDE pass shapes from this file if it exists, else autotune and save them there")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 1 GiB device-to-device copy (HBM counter calibration under rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is a design note only",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    if args.calibrate_copy:
        return calibrate_copy()
    p3hip = load_p3hip()
    if args.workload == "batch64":
        return bench_batch(p3hip, args)
    if args.workload == "serve":
        return bench_serve(p3hip, args)
    if args.workload == "fri_bb4_2^23":
        return bench_fri(p3hip, args)
    if args.workload == "fri_bb4_2^23_chal":
        return bench_fri(p3hip, args, challenger=True)
    if args.workload == "grind_bb":
        return bench_grind(p3hip, args)
    if args.workload == "pcs_open_bb_2^20x256":
        return bench_pcs_open(p3hip, args)
    field, log_n, width, lb, rp = WORKLOADS[args.workload]
    fid = p3hip.BABYBEAR if field == "bb" else p3hip.GOLDILOCKS
    n, N = 1 << log_n, 1 << (log_n + lb)
    shift = 31 if field == "bb" else 7
    ctx = p3hip.Context(0)
    ctx.set_poseidon2(fid, p2_constants(field, rp), 8, rp)
    ctx.set_hash_engine(args.hash_engine)
    rng = np.random.default_rng(2024)
    p = BB_P if field == "bb" else GL_P
    dtype = np.uint32 if field == "bb" else np.uint64
    trace = rng.integers(0, p, size=(n, width), dtype=dtype)
    m = ctx.upload(trace)
    if args.tuning_file and os.path.exists(args.tuning_file):
        ctx.tuning_load(args.tuning_file)
        tuned = {"loaded": args.tuning_file}
    else:
        tuned = None if args.no_autotune else ctx.autotune_lde(fid, log_n, width, lb, iters=2)
        if args.tuning_file and tuned is not None:
            ctx.tuning_save(args.tuning_file)

    def step():
        lde = ctx.coset_lde(m, lb, shift)
        tree = ctx.mmcs_commit([lde], 0)
        root = tree.root()
        tree.free()
        lde.free()
        return root

    for _ in range(args.warmup):
        step()
    ctx.sync()
    ctx.set_profiling(True)
    t0 = time.perf_counter()
    roots = [step() for _ in range(args.steps)]
    ctx.sync()
    dt = time.perf_counter() - t0
    phases = {ph: ctx.phase_stats(ph) for ph in ("lde", "leaf_hash", "compress")}
    ctx.set_profiling(False)
    assert all((r == roots[0]).all() for r in roots), "non-deterministic root"
    # host-handed input (the reference API's RowMajorMatrix): upload + step, outside the timed region
    t0 = time.perf_counter()
    m2 = ctx.upload(trace)
    upload_ms = (time.perf_counter() - t0) * 1e3
    lde = ctx.coset_lde(m2, lb, shift)
    tree = ctx.mmcs_commit([lde], 0)
    assert (tree.root() == roots[0]).all()
    host_ms = (time.perf_counter() - t0) * 1e3
    tree.free(), lde.free(), m2.free()
    del trace

    ms_per_step = dt * 1e3 / args.steps
    elems = N * width
    value = elems / (ms_per_step / 1e3)
    eb = 4 if field == "bb" else 8
    rate, dig = (8, 8) if field == "bb" else (4, 4)
    lde_ms = phases["lde"][1] / max(1, phases["lde"][0])
    hash_ms = phases["leaf_hash"][1] / max(1, phases["leaf_hash"][0])
    comp_ms = phases["compress"][1] / max(1, args.steps)
    # LDE phase: algorithmic bytes = eb * w * (3n + N)
    lde_bytes = eb * width * (3 * n + N)
    lde_gbs = lde_bytes / (lde_ms / 1e3) / 1e9
    # leaf hashing: one launch absorbs every LDE row
    perms = N * ((width + rate - 1) // rate)
    hash_bytes = eb * width * N + eb * dig * N
    traffic = load_traffic(args.workload)
    if field == "bb" and args.hash_engine == 1:
        flops = perms * P2_F64_FLOPS_PER_PERM + N * P2_F64_FLOPS_PER_DIGEST
        ach = flops / (hash_ms / 1e3) / 1e12
        roofline = {"bound": "valu", "kernel": "hash_rows_dense_f64_kernel (Poseidon2 leaf sponge, binary64)",
                    "achieved": ach, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": ach / FP64_PEAK_TFLOPS,
                    "traffic": traffic.get("hash_rows_dense_f64_kernel"), "algorithmic_flops": flops,
                    "algorithmic_bytes": hash_bytes, "hbm_GBps": hash_bytes / (hash_ms / 1e3) / 1e9,
                    "avg_launch_ms": hash_ms}
    else:
        roofline = {"bound": "valu", "kernel": "hash_rows_dense_kernel (Poseidon2 leaf sponge, integer)",
                    "achieved": perms / (hash_ms / 1e3) / 1e9, "unit": "Gperm/s", "peak": None, "frac": None,
                    "traffic": traffic.get("hash_rows_dense_kernel"), "algorithmic_bytes": hash_bytes,
                    "avg_launch_ms": hash_ms}
    out = {
        "metric": "committed field elems/s (LDE+Poseidon2 Merkle)",
        "value": value,
        "unit": "field elems/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32" if field == "bb" else "u64",
        "data": "synthetic (seeded uniform field elements, seeded Poseidon2 constants)",
        "config": {"workload": f"{'BabyBear' if field == 'bb' else 'Goldilocks'} trace 2^{log_n} x {width}, "
                               f"log_blowup {lb}, Poseidon2 W{16 if field == 'bb' else 8} RP{rp}, "
                               f"single-matrix MMCS root",
                   "rows": n, "width": width, "log_blowup": lb, "lde_rows": N},
        "phases_ms": {"lde": lde_ms, "leaf_hash": hash_ms, "compress": comp_ms},
        "roofline": roofline,
        "roofline_lde": {"bound": "hbm", "achieved": lde_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": lde_gbs / HBM_PEAK_GBS, "traffic": traffic.get("lde_phase"),
                         "algorithmic_bytes": lde_bytes, "avg_phase_ms": lde_ms},
        "pcie_inclusive": {"value": elems / (host_ms / 1e3), "unit": "field elems/s", "upload_ms": upload_ms,
                           "note": "trace handed over in pageable host memory: upload + LDE + commit + root"},
        "root": [int(x) for x in roots[0]],
        "hash_engine": "binary64" if args.hash_engine == 1 and field == "bb" else "uint32",
        "lde_tuning": tuned,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(field, width, lb, rp)
        out["cpu_baseline_threads"] = cpu_baseline_threads(field, width, lb, rp)
    print(json.dumps(out))
    ctx.close()


def load_traffic(workload: str) -> dict:
    """HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
    (each pass calibrated against a 1 GiB copy), kept as a profile summary."""
    try:
        with open(TRAFFIC_FILE) as f:
            return json.load(f).get(workload, {})
    except (OSError, ValueError):
        return {}


def calibrate_copy(nbytes: int = 1 << 30, reps: int = 3):
    import torch

    a = torch.randint(0, 1 << 30, (nbytes // 4,), dtype=torch.int32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        torch.add(a, 1, out=b)  # one elementwise kernel: reads nbytes, writes nbytes
    torch.cuda.synchronize()
    assert torch.equal(a + 1, b)
    print(json.dumps({"calibration": "elementwise copy (b = a + 1)", "bytes_read": nbytes, "bytes_written": nbytes,
                      "reps": reps}))


def batch_traces(seed: int = 7, count: int = 64):
    """64 BabyBear traces: heights 2^h, h uniform in [12, 20]; widths uniform in
    [8, 128]; log_blowup 1 or 2; uniform field elements; all from one seed."""
    rng = np.random.default_rng(seed)
    shapes = [(int(rng.integers(12, 21)), int(rng.integers(8, 129)), int(rng.integers(1, 3))) for _ in range(count)]
    traces = [rng.integers(0, BB_P, size=(1 << h, w), dtype=np.uint32) for h, w, _ in shapes]
    return shapes, traces


def bench_batch(p3hip, args):
    """64 BabyBear traces (heights 2^12..2^20, widths 8..128, blowup 2 or 4): one
    multi-height MMCS over 64 LDEs, planned and replayed as a
    HIP graph, against the plain per-trace loop (LDE + single-matrix commit +
    root per trace) on the same GPU and inputs."""
    ctx = p3hip.Context(0)
    ctx.set_poseidon2(p3hip.BABYBEAR, p2_constants("bb", 13), 8, 13)
    ctx.set_hash_engine(args.hash_engine)
    shapes, traces = batch_traces()
    mats = [ctx.upload(t) for t in traces]
    del traces
    ks = [k for _, _, k in shapes]
    elems = sum((1 << (h + k)) * w for h, w, k in shapes)
    # autotune the pass shape of the largest LDEs (they dominate the time)
    tuned = {}
    for h, w, k in sorted(shapes, key=lambda s: -(1 << s[0]) * s[1])[:6]:
        if (h, w, k) not in tuned:
            tuned[(h, w, k)] = ctx.autotune_lde(p3hip.BABYBEAR, h, w, k, iters=1)

    def plain_loop():
        roots = []
        for m, k in zip(mats, ks):
            lde = ctx.coset_lde(m, k, 31)
            t = ctx.flat_commit(lde)
            roots.append(t.root())
            t.free()
            lde.free()
        return roots

    for _ in range(max(1, args.warmup)):
        plain_loop()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plain_loop()
    ctx.sync()
    plain_ms = (time.perf_counter() - t0) * 1e3 / args.steps

    plan = ctx.batch_plan(mats, ks, 31, cap_height=0, use_graph=True)
    caps = [plan.run() for _ in range(max(1, args.warmup))]
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        caps.append(plan.run())
    ctx.sync()
    batch_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    assert all((c == caps[0]).all() for c in caps), "non-deterministic batched root"
    out = {
        "metric": "committed field elems/s (LDE+Poseidon2 Merkle), batched vs plain",
        "value": elems / (batch_ms / 1e3),
        "unit": "field elems/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": batch_ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
        "data": "synthetic (64 seeded traces, seeded Poseidon2 constants)",
        "config": {"workload": "64 BabyBear traces, h in [12,20], w in [8,128], log_blowup 1|2, one multi-height "
                               "MMCS, hipGraph replay", "committed_elems": elems},
        "plain_loop": {"value": elems / (plain_ms / 1e3), "ms_per_step": plain_ms},
        "batched_over_plain": plain_ms / batch_ms,
        "autotuned": {f"2^{h}x{w}_b{k}": v for (h, w, k), v in tuned.items()},
        "cap": [int(x) for x in caps[0][0]],
    }
    print(json.dumps(out))
    plan.free()
    ctx.close()


def bench_fri(p3hip, args, log_len: int = 23, log_blowup: int = 1, lfpl: int = 3, challenger: bool = False):
    """FRI commit phase on a random BabyBear^4 codeword of 2^23 elements: arity 2,
    log_blowup 1, final polynomial of 2^3 coefficients; each of the 19 layers
    (2^23 .. 2^5 elements) is committed with the single-matrix Poseidon2-16 MMCS
    (2 elements = 8 BabyBear words per leaf), then folded.  Challenges are
    seeded, or -- with `challenger` -- drawn from the device-resident duplex
    transcript in the loop (observe root, grind 0 bits, sample beta), as
    prove_fri runs it."""
    consts = p2_constants("bb", 13)
    ctx = p3hip.Context(0)
    ctx.set_poseidon2(p3hip.BABYBEAR, consts, 8, 13)
    ctx.set_hash_engine(args.hash_engine)
    rng = np.random.default_rng(5)
    cw = rng.integers(0, BB_P, size=(1 << log_len, 4), dtype=np.uint32)
    rounds = log_len - log_blowup - lfpl
    betas = rng.integers(0, BB_P, size=(rounds, 4), dtype=np.uint32)
    m = ctx.upload(cw)

    def step():
        if challenger:
            ch = ctx.challenger()
            ch.observe([1, 2, 3])
            r = ctx.fri_commit_phase_chal(m, log_blowup, lfpl, 1, ch, 0)
            ch.free()
            return r["roots"], r["final"]
        roots, final, _ = ctx.fri_commit_phase(m, log_blowup, lfpl, 1, betas)
        return roots, final

    for _ in range(args.warmup):
        step()
    ctx.sync()
    t0 = time.perf_counter()
    outs = [step() for _ in range(args.steps)]
    ctx.sync()
    dt = time.perf_counter() - t0
    # phase breakdown from a second pass with the timing events on: each event
    # costs GPU time between short launches, so they stay out of the timed loop
    ctx.set_profiling(True)
    for _ in range(args.steps):
        step()
    ctx.sync()
    phases = {ph: ctx.phase_stats(ph) for ph in ("fold", "leaf_hash", "compress")}
    ctx.set_profiling(False)
    assert all((o[0] == outs[0][0]).all() and (o[1] == outs[0][1]).all() for o in outs), "non-deterministic"
    ms = dt * 1e3 / args.steps
    layer_elems = sum(1 << l for l in range(lfpl + log_blowup + 1, log_len + 1))
    fold_ms = phases["fold"][1] / args.steps
    fold_bytes = sum(24 << l for l in range(lfpl + log_blowup + 1, log_len + 1))  # 16 B read + 8 B written per input
    out = {
        "metric": "committed BabyBear^4 elems/s (FRI commit phase)", "value": layer_elems / (ms / 1e3),
        "unit": "ext field elems/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
        "data": "synthetic (seeded codeword, challenges and Poseidon2 constants)",
        "config": {"workload": f"FRI commit phase, BabyBear^4 codeword 2^{log_len}, arity 2, log_blowup {log_blowup}, "
                               f"final poly 2^{lfpl}, {rounds} layer commits, "
                               + ("challenges from the duplex transcript" if challenger else "seeded challenges"),
                   "layer_elems": layer_elems},
        "phases_from": "a second pass of the same steps with phase timing events",
        "phases_ms": {"fold": fold_ms, "leaf_hash": phases["leaf_hash"][1] / args.steps,
                      "compress": phases["compress"][1] / args.steps},
        "roofline_fold": {"bound": "hbm", "achieved": fold_bytes / (fold_ms / 1e3) / 1e9, "peak": HBM_PEAK_GBS,
                          "unit": "GB/s", "frac": fold_bytes / (fold_ms / 1e3) / 1e9 / HBM_PEAK_GBS,
                          "algorithmic_bytes": fold_bytes, "launches": rounds},
        "roots_first_last": [[int(x) for x in outs[0][0][0]], [int(x) for x in outs[0][0][-1]]],
    }
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import native  # test infrastructure; used here only as the CPU baseline

        lib, kind = native.reference(), "reference"
        if lib is None:
            lib, kind = native.oracle(), "port"
        sl = 18
        t0 = time.perf_counter()
        lib.fri_commit(consts, 8, 13, cw[: 1 << sl], log_blowup, lfpl, 1, betas[: sl - log_blowup - lfpl])
        cdt = time.perf_counter() - t0
        se = sum(1 << l for l in range(lfpl + log_blowup + 1, sl + 1))
        out["cpu_baseline"] = {"value": se / cdt, "unit": "ext field elems/s", "cores": 1, "kind": kind,
                               "sample": f"codeword 2^{sl}, same parameters, {cdt:.2f} s on 1 host thread"}
    print(json.dumps(out))
    ctx.close()


def bench_grind(p3hip, args, bits: int = 24, cpu_bits: int = 13):
    """Proof-of-work search (DuplexChallenger::grind): from a seeded transcript,
    the least witness whose following sample has `bits` zero low bits; value =
    witnesses tried per second (the witness found + 1, over the wall time of
    the grind call)."""
    consts = p2_constants("bb", 13)
    ctx = p3hip.Context(0)
    ctx.set_poseidon2(p3hip.BABYBEAR, consts, 8, 13)
    prefix = np.arange(1, 12, dtype=np.uint32) * 7919
    runs = []
    for rep in range(args.warmup + args.steps):
        ch = ctx.challenger()
        ch.observe(prefix)
        ch.observe([rep])
        t0 = time.perf_counter()
        w = ch.grind(bits)
        dt = time.perf_counter() - t0
        ch.free()
        if rep >= args.warmup:
            runs.append((w, dt))
    tries = sum(w + 1 for w, _ in runs)
    total = sum(dt for _, dt in runs)
    out = {
        "metric": "proof-of-work witnesses tried/s (DuplexChallenger grind)", "value": tries / total,
        "unit": "witnesses/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": total * 1e3 / len(runs), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u32", "data": "synthetic (seeded transcript prefixes)",
        "config": {"workload": f"grind {bits} bits, BabyBear Poseidon2-16 duplex transcript (rate 8)",
                   "witnesses": [int(w) for w, _ in runs]},
    }
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import native  # test infrastructure; used here only as the CPU baseline

        lib, kind = native.reference(), "reference"
        if lib is None:
            lib, kind = native.oracle(), "port"
        ctries, ctime = 0, 0.0
        for rep in range(8):
            ops = [(native.OBSERVE, int(v)) for v in prefix] + [(native.OBSERVE, rep), (native.GRIND, cpu_bits)]
            t0 = time.perf_counter()
            (w,) = lib.challenger_run(consts, 8, 13, ops)
            ctime += time.perf_counter() - t0
            ctries += int(w) + 1
        out["cpu_baseline"] = {"value": ctries / ctime, "unit": "witnesses/s", "cores": 1, "kind": kind,
                               "sample": f"8 grinds of {cpu_bits} bits, {ctries} witnesses in {ctime:.2f} s"}
    print(json.dumps(out))
    ctx.close()


def bench_pcs_open(p3hip, args, log_n: int = 20, width: int = 256, n_points: int = 2, lfpl: int = 3,
                   num_queries: int = 100, query_pow_bits: int = 16):
    """A PCS round: TwoAdicFriPcs::commit of a 2^20 x 256 BabyBear trace
    (log_blowup 1) and ::open at 2 BabyBear^4 points -- barycentric opened
    values, quotient vector, prove_fri (arity 2, final poly 2^3), 16-bit query
    proof of work, 100 query indices and rows -- with the duplex transcript on
    the device.  value = committed-and-opened LDE elements per second."""
    consts = p2_constants("bb", 13)
    ctx = p3hip.Context(0)
    ctx.set_poseidon2(p3hip.BABYBEAR, consts, 8, 13)
    rng = np.random.default_rng(12)
    trace = ctx.upload(rng.integers(0, BB_P, size=(1 << log_n, width), dtype=np.uint32))
    pts = [rng.integers(0, BB_P, size=(n_points, 4), dtype=np.uint32)]
    if not args.no_autotune:
        ctx.autotune_lde(p3hip.BABYBEAR, log_n, width, 1, iters=2)

    def step():
        tree = ctx.pcs_commit([trace], [31], 1)
        ch = ctx.challenger()
        ch.observe_cap(tree)
        r = ctx.pcs_open(tree, [31], pts, ch, lfpl, 1, num_queries, 0, query_pow_bits)
        ch.free()
        tree.free()
        return r

    for _ in range(args.warmup):
        step()
    ctx.sync()
    t0 = time.perf_counter()
    outs = [step() for _ in range(args.steps)]
    ctx.sync()
    dt = time.perf_counter() - t0
    # phase breakdown from a second pass with the timing events on (kept out of the timed loop)
    ctx.set_profiling(True)
    for _ in range(args.steps):
        step()
    ctx.sync()
    phases = {ph: ctx.phase_stats(ph)[1] / args.steps
              for ph in ("lde", "leaf_hash", "compress", "pcs_open_values", "pcs_quotients", "fold", "fri_queries")}
    ctx.set_profiling(False)
    assert all((o["roots"] == outs[0]["roots"]).all() for o in outs), "non-deterministic"
    elems = (2 << log_n) * width
    ms = dt * 1e3 / args.steps
    out = {
        "metric": "PCS commit + open, LDE elems/s", "value": elems / (ms / 1e3), "unit": "field elems/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u32", "data": "synthetic (seeded trace and points)",
        "config": {"workload": f"TwoAdicFriPcs commit + open, BabyBear 2^{log_n} x {width}, log_blowup 1, "
                               f"{n_points} points, FRI arity 2, final poly 2^{lfpl}, {num_queries} queries, "
                               f"query PoW {query_pow_bits} bits"},
        "phases_from": "a second pass of the same steps with phase timing events",
        "phases_ms": phases,
        "rounds": len(outs[0]["roots"]),
    }
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import native  # test infrastructure; used here only as the CPU baseline

        lib, kind = native.reference(), "reference"
        if lib is None:
            lib, kind = native.oracle(), "port"
        sl = 12
        m = rng.integers(0, BB_P, size=(1 << sl, width), dtype=np.uint32)
        t0 = time.perf_counter()
        lib.pcs_open(consts, 8, 13, [m], [31], 1, pts, [], lfpl, 1, num_queries, 0, 8)
        cdt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": ((2 << sl) * width) / cdt, "unit": "field elems/s", "cores": 1,
                               "kind": kind,
                               "sample": f"commit + open of 2^{sl} x {width}, same parameters except 8-bit query "
                                         f"PoW, {cdt:.2f} s on 1 host thread"}
    print(json.dumps(out))
    ctx.close()


def _pct(xs, q):
    xs = sorted(xs)
    return xs[min(len(xs) - 1, int(round(q / 100.0 * (len(xs) - 1))))]


def bench_serve(p3hip, args, n_req: int = 96, pool: int = 24, budget_gib: int = 32):
    """Open-loop Poisson commit requests (BabyBear, log_blowup 1, heights
    2^[12,20], widths [8,128], 64 leaf openings each) at three arrival rates
    plus a saturating burst.  The served form (continuous batching over 4
    streams, admission under a fixed HBM budget, LDE and Merkle layers resident
    until the openings are taken) runs against the plain per-request path (one
    context, FIFO: upload, LDE, commit, root, openings, free) on the same
    arrivals.  Latency is scheduled arrival -> root on the host.  The request
    pool sits in page-locked host memory for both paths (uploads at the PCIe
    DMA rate); one burst from pageable copies of the pool is reported beside."""
    import threading

    consts = p2_constants("bb", 13)
    rng = np.random.default_rng(11)
    shapes = [(int(rng.integers(12, 21)), int(rng.integers(8, 129))) for _ in range(pool)]
    pageable = [rng.integers(0, BB_P, size=(1 << h, w), dtype=np.uint32) for h, w in shapes]
    pinned = []
    for t in pageable:
        pinned.append(p3hip.pinned_empty(t.shape, np.uint32))
        pinned[-1][...] = t
    traces = pinned
    order = rng.integers(0, pool, size=n_req)
    elems = sum((2 << shapes[i][0]) * shapes[i][1] for i in order)
    openings = [rng.integers(0, 2 << shapes[i][0], size=64) for i in order]

    ctx = p3hip.Context(0)
    ctx.set_poseidon2(p3hip.BABYBEAR, consts, 8, 13)

    def h2d_rate(a):  # best of 3 uploads, bytes/s
        best = float("inf")
        for _ in range(3):
            t0 = time.perf_counter()
            m = ctx.upload(a)
            best = min(best, time.perf_counter() - t0)
            m.free()
        return a.nbytes / best

    big = max(range(pool), key=lambda i: pageable[i].nbytes)
    h2d = {"bytes": int(pageable[big].nbytes), "pinned_GBps": h2d_rate(pinned[big]) / 1e9,
           "pageable_GBps": h2d_rate(pageable[big]) / 1e9}

    def plain_run(arrivals):
        lat, start = [], time.perf_counter()
        for k, (a, i) in enumerate(zip(arrivals, order)):
            now = time.perf_counter() - start
            if now < a:
                time.sleep(a - now)
            m = ctx.upload(traces[i])
            lde = ctx.coset_lde(m, 1, 31)
            tree = ctx.mmcs_commit([lde], 0)
            tree.root()
            lat.append((time.perf_counter() - start - a) * 1e3)
            tree.open(openings[k])
            tree.free(), lde.free(), m.free()
        return lat, time.perf_counter() - start

    def served_run(arrivals):
        with p3hip.Server(consts, 8, 13, hbm_budget=budget_gib << 30, n_slots=4, max_batch=8) as sv:
            ids, lag = [None] * n_req, [0.0] * n_req

            def submitter():
                for k, (a, i) in enumerate(zip(arrivals, order)):
                    now = time.perf_counter() - st
                    if now < a:
                        time.sleep(a - now)
                    lag[k] = (time.perf_counter() - st - a) * 1e3
                    ids[k] = sv.submit(traces[i], 1, 31)

            st = time.perf_counter()
            th = threading.Thread(target=submitter)
            th.start()
            lat = []
            for k in range(n_req):
                while ids[k] is None:
                    time.sleep(0.0002)
                _, server_ms = sv.wait(ids[k])
                lat.append(lag[k] + server_ms)
                h, w = shapes[order[k]]
                sv.open(ids[k], openings[k], w, h + 2)
                sv.release(ids[k])
            th.join()
            span = time.perf_counter() - st
            stats = sv.stats()
            assert stats["completed"] == n_req and stats["reserved_bytes"] == 0
            return lat, span, stats["batches"]

    plain_run(np.zeros(n_req))  # warm caches and tuning on both paths
    served_run(np.zeros(n_req))
    def burst():
        lat_p, span_p = plain_run(np.zeros(n_req))
        lat_s, span_s, nb = served_run(np.zeros(n_req))
        return span_p, {
            "served": {"elems_per_s": elems / span_s, "p50_ms": _pct(lat_s, 50), "p99_ms": _pct(lat_s, 99),
                       "batches": nb},
            "plain": {"elems_per_s": elems / span_p, "p50_ms": _pct(lat_p, 50), "p99_ms": _pct(lat_p, 99)}}

    traces = pageable
    _, burst_pageable = burst()
    traces = pinned
    span_p, burst_pinned = burst()
    res = {"budget_gib": budget_gib, "committed_elems": elems, "host_buffers": "pinned", "h2d": h2d,
           "uploaded_bytes": int(sum(pageable[i].nbytes for i in order)),
           "burst": burst_pinned, "burst_pageable": burst_pageable, "rates": []}
    svc = span_p / n_req  # plain per-request service time
    for factor in (0.5, 1.0, 1.6):
        lam = factor / svc
        arrivals = np.cumsum(rng.exponential(1.0 / lam, size=n_req))
        lat_p, span_p = plain_run(arrivals)
        lat_s, span_s, nb = served_run(arrivals)
        res["rates"].append({
            "arrival_rate_per_s": lam, "load_vs_plain_capacity": factor,
            "served": {"p50_ms": _pct(lat_s, 50), "p99_ms": _pct(lat_s, 99), "elems_per_s": elems / span_s,
                       "batches": nb},
            "plain": {"p50_ms": _pct(lat_p, 50), "p99_ms": _pct(lat_p, 99), "elems_per_s": elems / span_p},
        })
    top = res["rates"][-1]
    out = {
        "metric": "committed field elems/s (served, Poisson arrivals at 1.6x plain capacity)",
        "value": top["served"]["elems_per_s"], "unit": "field elems/s",
        "n_gpus": 1, "steps": n_req, "warmup": n_req, "ms_per_step": None, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u32",
        "data": "synthetic (seeded traces, Poisson arrivals)",
        "config": {"workload": f"{n_req} Poisson commit requests over {pool} BabyBear traces, h in [12,20], "
                               f"w in [8,128], log_blowup 1, 64 openings each; 4 slots, max batch 8, "
                               f"{budget_gib} GiB HBM budget; traces in pinned host memory"},
        "p50_ms": top["served"]["p50_ms"], "p99_ms": top["served"]["p99_ms"],
        "vs_plain": top["served"]["elems_per_s"] / top["plain"]["elems_per_s"],
        "serving": res,
    }
    print(json.dumps(out))
    ctx.close()


if __name__ == "__main__":
    main()
