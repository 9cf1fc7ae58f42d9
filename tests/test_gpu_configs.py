"""Every benchmark configuration at its full size, bit for bit against the
oracle: commitments, openings and FRI layer roots of the same std::mt19937_64
inputs (tests/golden/configs.npz, written by tests/golden/gen_config_roots.py
with the oracle on host threads).

* BabyBear 2^16 x 32, 2^20 x 256, 2^22 x 64 (log_blowup 1) and Goldilocks
  2^22 x 64 (log_blowup 2): LDE + single-matrix commit root, 64 openings (rows
  and authentication paths), and the device verifier accepting them;
* FRI commit phase on a BabyBear^4 codeword of 2^23: 19 layer roots and the
  final polynomial;
* 64 traces (heights 2^12..2^20, widths 8..128, log_blowup 1|2) as one
  multi-height MMCS with a cap of 16, through the per-call batch API, the
  planned HIP-graph replay, and the unbatched LDE + MMCS calls: cap and 64
  openings;
* 16 requests served concurrently (3 slots, continuous batching): each root and
  64 openings equal the oracle's commit of that request alone.

Inputs are regenerated here from the seeds of plonky3-gpu_amd/workloads.py.
"""

import hashlib
import os
import sys

import numpy as np
import pytest

import native
import p3hip

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "plonky3-gpu_amd"))
import workloads as W  # noqa: E402

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "configs.npz")
FID = {"bb": p3hip.BABYBEAR, "gl": p3hip.GOLDILOCKS}


def say(msg):
    print(f"[configs] {msg}", flush=True)


def sha(a):
    return np.frombuffer(hashlib.sha256(np.ascontiguousarray(a).tobytes()).digest(), dtype=np.uint8)


@pytest.fixture(scope="module")
def fx():
    with np.load(FIXTURE, allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


@pytest.fixture(scope="module")
def wctx():
    c = p3hip.Context(0)
    c.set_poseidon2(p3hip.BABYBEAR, np.array(W.p2_consts("bb"), dtype=np.uint32), W.RF, W.RP["bb"])
    c.set_poseidon2(p3hip.GOLDILOCKS, np.array(W.p2_consts("gl"), dtype=np.uint64), W.RF, W.RP["gl"])
    yield c
    c.close()


@pytest.mark.parametrize("name", list(W.SINGLE))
def test_single_trace_config_full_size(wctx, fx, name):
    c = W.SINGLE[name]
    f, lh, w, k = c["field"], c["log_n"], c["width"], c["added"]
    say(f"{name}: seed {c['seed']:#x}, trace 2^{lh} x {w}, added bits {k}")
    trace = native.seeded(c["seed"], (1 << lh) * w, f).reshape(1 << lh, w)
    m = wctx.upload(trace)
    del trace
    lde = wctx.coset_lde(m, k, W.SHIFT[f])
    tree = wctx.mmcs_commit([lde], 0)
    assert (tree.root() == fx[f"{name}/root"]).all(), name
    idx = fx[f"{name}/idx"]
    assert (idx == np.array(W.open_indices(c["seed"], 1 << (lh + k)), dtype=np.uint64)).all()
    vals, proof = tree.open(idx)
    assert (vals == fx[f"{name}/vals"]).all(), name
    assert (proof == fx[f"{name}/proof"]).all(), name
    flat = wctx.flat_commit(lde)
    assert (flat.root() == fx[f"{name}/root"]).all(), name
    dims = [(1 << (lh + k), w)]
    for q in (0, 17, 63):
        assert wctx.mmcs_verify(FID[f], tree.cap(), dims, int(idx[q]), vals[q], proof[q])
    for h in (flat, tree, lde, m):
        h.free()
    wctx.trim()


def test_fri_config_full_size(wctx, fx):
    c = W.FRI_BB4
    say(f"fri_bb4: seed {c['seed']:#x}, codeword 2^{c['log_len']}")
    cw = native.seeded(c["seed"], (1 << c["log_len"]) * 4).reshape(-1, 4)
    betas = native.seeded(c["seed"] + 1, 4 * c["rounds"]).reshape(-1, 4)
    m = wctx.upload(cw)
    roots, final, _ = wctx.fri_commit_phase(m, c["log_blowup"], c["log_final_poly_len"], c["max_log_arity"], betas)
    assert roots.shape == (c["rounds"], 8)
    assert (roots == fx["fri_bb4/roots"]).all()
    assert (final == fx["fri_bb4/final"]).all()
    m.free()
    wctx.trim()


def _check_batch_openings(tree, fx):
    vals, proof = tree.open(fx["batch64/idx"])
    for q in range(vals.shape[0]):
        assert (sha(vals[q]) == fx["batch64/vals_sha"][q]).all(), q
    assert (proof == fx["batch64/proof"]).all()


def test_batch_config_full_size(wctx, fx):
    shapes = W.batch_shapes()
    say(f"batch64: seed {W.BATCH['seed']:#x}, {len(shapes)} traces, {W.committed_elems(shapes)} committed elems")
    mats = []
    for h, w, _, s in shapes:
        mats.append(wctx.upload(native.seeded(s, (1 << h) * w).reshape(1 << h, w)))
    ks = [k for _, _, k, _ in shapes]
    ch = W.BATCH["cap_height"]
    want = fx["batch64/cap"]
    say("batch64: per-call batched commit")
    t = wctx.batch_commit(mats, ks, W.SHIFT["bb"], ch)
    assert (t.cap() == want).all()
    _check_batch_openings(t, fx)
    t.free()
    say("batch64: planned HIP-graph replays, LDE as launches grouped by pass shape")
    plan = wctx.batch_plan(mats, ks, W.SHIFT["bb"], cap_height=ch, use_graph=True)
    assert plan.lde_launches() > 1
    for _ in range(2):
        assert (plan.run() == want).all()
    _check_batch_openings(plan.tree(), fx)
    plan.free()
    say("batch64: planned HIP-graph replays, LDE as one work-queue launch")
    plan = wctx.batch_plan(mats, ks, W.SHIFT["bb"], cap_height=ch, use_graph=True, lde_queue=True)
    assert plan.lde_launches() == 1
    for _ in range(2):
        assert (plan.run() == want).all()
    plan.free()
    say("batch64: unbatched LDE calls + one MMCS commit")
    ldes = [wctx.coset_lde(m, k, W.SHIFT["bb"]) for m, k in zip(mats, ks)]
    t = wctx.mmcs_commit(ldes, ch)
    assert (t.cap() == want).all()
    t.free()
    for x in ldes + mats:
        x.free()
    wctx.trim()


def test_serve_mix_full_size(fx):
    shapes = W.serve_shapes()
    say(f"serve16: seed {W.SERVE_MIX['seed']:#x}, {len(shapes)} requests, 3 slots")
    traces = [native.seeded(s, (1 << h) * w).reshape(1 << h, w) for h, w, _, s in shapes]
    consts = np.array(W.p2_consts("bb"), dtype=np.uint32)
    with p3hip.Server(consts, W.RF, W.RP["bb"], n_slots=3, max_batch=4) as sv:
        ids = [sv.submit(t, 1, W.SHIFT["bb"]) for t in traces]
        for i, (rid, (h, w, k, _)) in enumerate(zip(ids, shapes)):
            root, _ = sv.wait(rid)
            assert (root == fx["serve16/roots"][i]).all(), i
            vals, proof = sv.open(rid, fx["serve16/idx"][i], w, h + k + 1)
            for q in range(vals.shape[0]):
                assert (sha(vals[q]) == fx["serve16/vals_sha"][i][q]).all(), (i, q)
                assert (sha(proof[q]) == fx["serve16/proof_sha"][i][q]).all(), (i, q)
            sv.release(rid)
        assert sv.stats()["reserved_bytes"] == 0
