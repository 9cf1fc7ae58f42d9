"""Write tests/golden/configs.npz: the oracle's commitments and openings for
every benchmark configuration at its full size.

Run in the build container (8 host threads, about 20 minutes, ~40 GiB of host
memory at the Goldilocks configuration):

    python tests/golden/gen_config_roots.py [--threads 8] [--only NAME ...]

Inputs are the synthetic workloads of plonky3-gpu_amd/workloads.py
(std::mt19937_64 streams); the oracle (oracle/liboracle.so) computes every
value on host threads.  Its threaded paths are tested equal to the
single-threaded restatement (tests/test_oracle.py), which the reference build's
golden vectors pin.

Contents (per configuration; DIG = 8 BabyBear / 4 Goldilocks words):
* single trace NAME (bb_2^16x32_b1, bb_2^20x256_b1, bb_2^22x64_b1, gl_2^22x64_b2):
  NAME/root [DIG], NAME/idx [64], NAME/vals [64, w] (LDE rows, natural order),
  NAME/proof [64, levels, DIG];
* fri_bb4: roots [19, 8], final [8, 4];
* batch64: cap [16, 8], idx [64], vals_sha [64, 32] (SHA-256 of each opening's
  concatenated rows as little-endian uint32), proof [64, levels, 8];
* serve16: roots [16, 8], idx [16, 64], vals_sha [16, 64, 32], proof_sha [16, 64, 32].
"""

from __future__ import annotations

import argparse
import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "plonky3-gpu_amd"))

import native  # noqa: E402
import workloads as W  # noqa: E402

OUT = os.path.join(HERE, "configs.npz")


def sha(a: np.ndarray) -> np.ndarray:
    return np.frombuffer(hashlib.sha256(np.ascontiguousarray(a).tobytes()).digest(), dtype=np.uint8).copy()


def consts(field):
    return np.array(W.p2_consts(field), dtype=native.DT[field])


def gen_single(orc, name, threads, out):
    c = W.SINGLE[name]
    f, lh, w, k = c["field"], c["log_n"], c["width"], c["added"]
    trace = native.seeded(c["seed"], (1 << lh) * w, f).reshape(1 << lh, w)
    idx = np.array(W.open_indices(c["seed"], 1 << (lh + k)), dtype=np.uint64)
    root, vals, proof = orc.lde_commit_open_mt(f, consts(f), W.RF, W.RP[f], trace, k, W.SHIFT[f], threads, idx)
    out.update({f"{name}/root": root, f"{name}/idx": idx, f"{name}/vals": vals, f"{name}/proof": proof})


def gen_fri(orc, threads, out):
    c = W.FRI_BB4
    cw = native.seeded(c["seed"], (1 << c["log_len"]) * 4).reshape(-1, 4)
    betas = native.seeded(c["seed"] + 1, 4 * c["rounds"]).reshape(-1, 4)
    roots, final = orc.fri_commit(consts("bb"), W.RF, W.RP["bb"], cw, c["log_blowup"], c["log_final_poly_len"],
                                  c["max_log_arity"], betas, threads=threads)
    assert roots.shape[0] == c["rounds"]
    out.update({"fri_bb4/roots": roots, "fri_bb4/final": final})


def gen_batch(orc, threads, out):
    shapes = W.batch_shapes()
    traces = [native.seeded(s, (1 << h) * w).reshape(1 << h, w) for h, w, _, s in shapes]
    ks = [k for _, _, k, _ in shapes]
    maxn = max(1 << (h + k) for h, _, k, _ in shapes)
    idx = np.array(W.open_indices(W.BATCH["seed"], maxn), dtype=np.uint64)
    cap, vals, proof = orc.batch_commit_mt(consts("bb"), W.RF, W.RP["bb"], traces, ks, W.SHIFT["bb"],
                                           W.BATCH["cap_height"], threads, idx)
    out.update({"batch64/cap": cap, "batch64/idx": idx, "batch64/vals_sha": np.stack([sha(v) for v in vals]),
                "batch64/proof": proof})


def gen_serve(orc, threads, out):
    shapes = W.serve_shapes()
    roots, idxs, vsha, psha = [], [], [], []
    for h, w, k, s in shapes:
        trace = native.seeded(s, (1 << h) * w).reshape(1 << h, w)
        idx = np.array(W.open_indices(s, 1 << (h + k)), dtype=np.uint64)
        root, vals, proof = orc.lde_commit_open_mt("bb", consts("bb"), W.RF, W.RP["bb"], trace, k, W.SHIFT["bb"],
                                                   threads, idx)
        roots.append(root)
        idxs.append(idx)
        vsha.append(np.stack([sha(v) for v in vals]))
        psha.append(np.stack([sha(p) for p in proof]))
    out.update({"serve16/roots": np.stack(roots), "serve16/idx": np.stack(idxs), "serve16/vals_sha": np.stack(vsha),
                "serve16/proof_sha": np.stack(psha)})


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=min(8, os.cpu_count() or 1))
    ap.add_argument("--only", nargs="*", default=None)
    args = ap.parse_args()
    orc = native.oracle()
    out = {}
    if os.path.exists(OUT):
        with np.load(OUT, allow_pickle=False) as old:
            out.update({k: old[k] for k in old.files})
    jobs = [(n, (lambda n=n: gen_single(orc, n, args.threads, out))) for n in W.SINGLE]
    jobs += [("fri_bb4", lambda: gen_fri(orc, args.threads, out)),
             ("batch64", lambda: gen_batch(orc, args.threads, out)),
             ("serve16", lambda: gen_serve(orc, args.threads, out))]
    for name, job in jobs:
        if args.only is not None and name not in args.only:
            continue
        t0 = time.time()
        job()
        print(f"{name}: {time.time() - t0:.1f} s", flush=True)
        np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes)")


if __name__ == "__main__":
    main()
