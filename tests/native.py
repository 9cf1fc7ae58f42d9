"""ctypes wrappers over the CPU oracle (oracle/liboracle.so) and, when it has
been built in this container, the reference build (oracle/_ref/libp3ref.so).

Test infrastructure only.  Both libraries share one calling convention, so
``Native(lib, prefix)`` drives either: prefix "orc" for the oracle, "ref" for
the reference build.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")
REF_SO = os.path.join(ORACLE_DIR, "_ref", "libp3ref.so")

BB_P = 0x78000001
GL_P = 0xFFFFFFFF00000001
U32, U64 = np.uint32, np.uint64
DT = {"bb": U32, "gl": U64}
DIG = {"bb": 8, "gl": 4}
WIDTH = {"bb": 16, "gl": 8}


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def build_oracle():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


class Native:
    def __init__(self, path: str, prefix: str):
        self.lib = C.CDLL(path)
        self.pre = prefix

    def fn(self, name):
        return getattr(self.lib, f"{self.pre}_{name}")

    def _rc(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{self.pre}_{what} failed ({rc})")

    # Poseidon2 ------------------------------------------------------------
    def permute(self, f, consts, rf, rp, states):
        s = np.ascontiguousarray(states, dtype=DT[f]).copy()
        c = np.ascontiguousarray(consts, dtype=DT[f])
        self._rc(self.fn(f"{f}_permute")(_p(c), C.c_int(rf), C.c_int(rp), _p(s), C.c_size_t(s.size // WIDTH[f])),
                 "permute")
        return s

    def sponge(self, f, consts, rf, rp, data):
        d = np.ascontiguousarray(data, dtype=DT[f])
        c = np.ascontiguousarray(consts, dtype=DT[f])
        out = np.zeros(DIG[f], dtype=DT[f])
        self._rc(self.fn(f"{f}_sponge")(_p(c), C.c_int(rf), C.c_int(rp), _p(d), C.c_size_t(d.size), _p(out)),
                 "sponge")
        return out

    def compress(self, f, consts, rf, rp, pair):
        d = np.ascontiguousarray(pair, dtype=DT[f]).reshape(-1)
        c = np.ascontiguousarray(consts, dtype=DT[f])
        out = np.zeros(DIG[f], dtype=DT[f])
        self._rc(self.fn(f"{f}_compress")(_p(c), C.c_int(rf), C.c_int(rp), _p(d), _p(out)), "compress")
        return out

    # DFT ------------------------------------------------------------------
    def dft(self, f, kind, mat, shift=1):
        m = np.ascontiguousarray(mat, dtype=DT[f])
        h, w = m.shape
        out = np.zeros_like(m)
        sh = C.c_uint32(shift) if f == "bb" else C.c_uint64(shift)
        self._rc(self.fn(f"{f}_dft")(C.c_int(kind), _p(m), C.c_size_t(h.bit_length() - 1), C.c_size_t(w), sh, _p(out)),
                 "dft")
        return out

    def lde(self, f, mat, added, shift):
        m = np.ascontiguousarray(mat, dtype=DT[f])
        h, w = m.shape
        out = np.zeros((h << added, w), dtype=DT[f])
        sh = C.c_uint32(shift) if f == "bb" else C.c_uint64(shift)
        self._rc(self.fn(f"{f}_lde")(_p(m), C.c_size_t(h.bit_length() - 1), C.c_size_t(w), C.c_size_t(added), sh,
                                     _p(out)), "lde")
        return out

    def lde_brute(self, f, mat, added, shift):
        m = np.ascontiguousarray(mat, dtype=DT[f])
        h, w = m.shape
        out = np.zeros((h << added, w), dtype=DT[f])
        sh = C.c_uint32(shift) if f == "bb" else C.c_uint64(shift)
        self._rc(self.fn(f"{f}_lde_brute")(_p(m), C.c_size_t(h.bit_length() - 1), C.c_size_t(w), C.c_size_t(added),
                                           sh, _p(out)), "lde_brute")
        return out

    # MMCS -----------------------------------------------------------------
    def flat_commit(self, f, consts, rf, rp, mat, with_tree=False):
        m = np.ascontiguousarray(mat, dtype=DT[f])
        h, w = m.shape
        c = np.ascontiguousarray(consts, dtype=DT[f])
        root = np.zeros(DIG[f], dtype=DT[f])
        tree = None
        if with_tree:
            total, cnt = 0, h
            while True:
                total += cnt
                if cnt == 1:
                    break
                cnt = (cnt + 1) // 2
            tree = np.zeros(total * DIG[f], dtype=DT[f])
        self._rc(self.fn(f"{f}_flat_commit")(_p(c), C.c_int(rf), C.c_int(rp), _p(m), C.c_size_t(h), C.c_size_t(w),
                                             _p(root), _p(tree) if with_tree else None), "flat_commit")
        if not with_tree:
            return root
        layers, off, cnt = [], 0, h
        while True:
            layers.append(tree[off:off + cnt * DIG[f]].reshape(cnt, DIG[f]))
            off += cnt * DIG[f]
            if cnt == 1:
                break
            cnt = (cnt + 1) // 2
        return root, layers

    def lde_commit_mt(self, f, consts, rf, rp, mat, added, shift, threads):
        """Root of the flat commit of the LDE, computed on `threads` host threads."""
        m = np.ascontiguousarray(mat, dtype=DT[f])
        h, w = m.shape
        c = np.ascontiguousarray(consts, dtype=DT[f])
        root = np.zeros(DIG[f], dtype=DT[f])
        sh = C.c_uint32(shift) if f == "bb" else C.c_uint64(shift)
        self._rc(self.fn(f"{f}_lde_commit_mt")(_p(c), C.c_int(rf), C.c_int(rp), _p(m), C.c_size_t(h.bit_length() - 1),
                                               C.c_size_t(w), C.c_size_t(added), sh, C.c_int(threads), _p(root)),
                 "lde_commit_mt")
        return root

    def lde_commit_open_mt(self, f, consts, rf, rp, mat, added, shift, threads, indices):
        """LDE + flat commit on `threads` host threads, with openings: (root,
        values n_idx x w, proofs n_idx x levels x DIG)."""
        m = np.ascontiguousarray(mat, dtype=DT[f])
        h, w = m.shape
        c = np.ascontiguousarray(consts, dtype=DT[f])
        idx = np.ascontiguousarray(indices, dtype=U64)
        levels = (h << added).bit_length() - 1
        root = np.zeros(DIG[f], dtype=DT[f])
        vals = np.zeros(max(1, idx.size * w), dtype=DT[f])
        proof = np.zeros(max(1, idx.size * levels * DIG[f]), dtype=DT[f])
        sh = C.c_uint32(shift) if f == "bb" else C.c_uint64(shift)
        self._rc(self.fn(f"{f}_lde_commit_open_mt")(_p(c), C.c_int(rf), C.c_int(rp), _p(m),
                                                    C.c_size_t(h.bit_length() - 1), C.c_size_t(w),
                                                    C.c_size_t(added), sh, C.c_int(threads), _p(idx),
                                                    C.c_size_t(idx.size), _p(root), _p(vals), _p(proof)),
                 "lde_commit_open_mt")
        return (root, vals[: idx.size * w].reshape(idx.size, w),
                proof[: idx.size * levels * DIG[f]].reshape(idx.size, levels, DIG[f]))

    def batch_commit_mt(self, consts, rf, rp, traces, adds, shift, cap_height, threads, indices):
        """LDE of every BabyBear trace, one multi-height MMCS over them, with
        openings: (cap, values n_idx x total_width, proofs n_idx x levels x 8)."""
        ts = [np.ascontiguousarray(t, dtype=U32) for t in traces]
        n = len(ts)
        lhs = (C.c_size_t * n)(*[t.shape[0].bit_length() - 1 for t in ts])
        ws = (C.c_size_t * n)(*[t.shape[1] for t in ts])
        ks = (C.c_size_t * n)(*[int(k) for k in adds])
        ptrs = (C.c_void_p * n)(*[t.ctypes.data for t in ts])
        c = np.ascontiguousarray(consts, dtype=U32)
        idx = np.ascontiguousarray(indices, dtype=U64)
        tw = sum(t.shape[1] for t in ts)
        maxn = max(t.shape[0] << int(k) for t, k in zip(ts, adds))
        levels = maxn.bit_length() - 1 - cap_height
        cap = np.zeros((1 << cap_height) * 8, dtype=U32)
        vals = np.zeros(max(1, idx.size * tw), dtype=U32)
        proof = np.zeros(max(1, idx.size * levels * 8), dtype=U32)
        plen = C.c_size_t(0)
        self._rc(self.fn("bb_batch_commit_mt")(_p(c), C.c_int(rf), C.c_int(rp), C.c_size_t(n), lhs, ws, ks, ptrs,
                                               C.c_uint32(shift), C.c_size_t(cap_height), C.c_int(threads), _p(idx),
                                               C.c_size_t(idx.size), _p(cap), _p(vals), _p(proof), C.byref(plen)),
                 "batch_commit_mt")
        assert plen.value == levels
        return (cap.reshape(-1, 8), vals[: idx.size * tw].reshape(idx.size, tw),
                proof[: idx.size * levels * 8].reshape(idx.size, levels, 8))

    def mmcs(self, f, consts, rf, rp, mats, cap_height=0, open_index=-1):
        ms = [np.ascontiguousarray(m, dtype=DT[f]) for m in mats]
        n = len(ms)
        hs = (C.c_size_t * n)(*[m.shape[0] for m in ms])
        ws = (C.c_size_t * n)(*[m.shape[1] for m in ms])
        ptrs = (C.c_void_p * n)(*[m.ctypes.data for m in ms])
        c = np.ascontiguousarray(consts, dtype=DT[f])
        cap = np.zeros((1 << cap_height) * DIG[f], dtype=DT[f])
        total_w = sum(m.shape[1] for m in ms)
        vals = np.zeros(max(1, total_w), dtype=DT[f])
        proof = np.zeros(64 * DIG[f], dtype=DT[f])
        plen = C.c_size_t(0)
        self._rc(self.fn(f"{f}_mmcs")(_p(c), C.c_int(rf), C.c_int(rp), C.c_size_t(n), hs, ws, ptrs,
                                      C.c_size_t(cap_height), _p(cap), C.c_longlong(open_index), _p(vals), _p(proof),
                                      C.byref(plen)), "mmcs")
        cap = cap.reshape(-1, DIG[f])
        if open_index < 0:
            return cap
        return cap, vals[:total_w], proof[: plen.value * DIG[f]].reshape(-1, DIG[f])

    def mmcs_verify(self, consts, rf, rp, dims, cap_height, cap, index, values, proof):
        n = len(dims)
        hs = (C.c_size_t * n)(*[d[0] for d in dims])
        ws = (C.c_size_t * n)(*[d[1] for d in dims])
        c = np.ascontiguousarray(consts, dtype=U32)
        cp = np.ascontiguousarray(cap, dtype=U32)
        v = np.ascontiguousarray(values, dtype=U32)
        pr = np.ascontiguousarray(proof, dtype=U32)
        return self.fn("bb_mmcs_verify")(_p(c), C.c_int(rf), C.c_int(rp), C.c_size_t(n), hs, ws,
                                         C.c_size_t(cap_height), _p(cp), C.c_size_t(index), _p(v), _p(pr),
                                         C.c_size_t(pr.size // 8))

    def flat_open(self, consts, rf, rp, mat, index):
        m = np.ascontiguousarray(mat, dtype=U32)
        h, w = m.shape
        c = np.ascontiguousarray(consts, dtype=U32)
        row = np.zeros(w, dtype=U32)
        steps, cnt = 0, h
        while cnt > 1:
            cnt = (cnt + 1) // 2
            steps += 1
        proof = np.zeros(max(1, steps) * 8, dtype=U32)
        self._rc(self.fn("bb_flat_open")(_p(c), C.c_int(rf), C.c_int(rp), _p(m), C.c_size_t(h), C.c_size_t(w),
                                         C.c_size_t(index), _p(row), _p(proof)), "flat_open")
        return row, proof[: steps * 8].reshape(-1, 8)

    # FRI ------------------------------------------------------------------
    def bb4_mul(self, a, b):
        x = np.ascontiguousarray(a, dtype=U32)
        y = np.ascontiguousarray(b, dtype=U32)
        out = np.zeros(4, dtype=U32)
        r = self.fn("bb4_mul")(_p(x), _p(y), _p(out))
        if self.pre == "ref":
            self._rc(r, "bb4_mul")
        return out

    def fold(self, log_height, log_arity, beta, cur):
        c = np.ascontiguousarray(cur, dtype=U32).reshape(-1)
        b = np.ascontiguousarray(beta, dtype=U32)
        out = np.zeros(c.size >> log_arity, dtype=U32)
        self._rc(self.fn("bb4_fold")(C.c_size_t(log_height), C.c_size_t(log_arity), _p(b), _p(c),
                                     C.c_size_t(c.size // 4), _p(out)), "fold")
        return out.reshape(-1, 4)

    def fri_commit(self, consts, rf, rp, codeword, log_blowup, log_final_poly_len, max_log_arity, betas, threads=1):
        cw = np.ascontiguousarray(codeword, dtype=U32).reshape(-1)
        c = np.ascontiguousarray(consts, dtype=U32)
        b = np.ascontiguousarray(betas, dtype=U32).reshape(-1)
        log_len = (cw.size // 4).bit_length() - 1
        roots = np.zeros(64 * 8, dtype=U32)
        final = np.zeros(4 << log_final_poly_len, dtype=U32)
        nr = C.c_size_t(0)
        if threads > 1:
            self._rc(self.fn("bb4_fri_commit_mt")(_p(c), C.c_int(rf), C.c_int(rp), _p(cw), C.c_size_t(log_len),
                                                  C.c_size_t(log_blowup), C.c_size_t(log_final_poly_len),
                                                  C.c_size_t(max_log_arity), _p(b), C.c_int(threads), _p(roots),
                                                  _p(final), C.byref(nr)), "fri_commit_mt")
        else:
            self._rc(self.fn("bb4_fri_commit")(_p(c), C.c_int(rf), C.c_int(rp), _p(cw), C.c_size_t(log_len),
                                               C.c_size_t(log_blowup), C.c_size_t(log_final_poly_len),
                                               C.c_size_t(max_log_arity), _p(b), _p(roots), _p(final), C.byref(nr)),
                     "fri_commit")
        return roots[: nr.value * 8].reshape(-1, 8), final.reshape(-1, 4)

    # Challenger ------------------------------------------------------------
    def challenger_run(self, consts, rf, rp, ops, max_out=4096):
        """ops: list of (code, arg); see op_* helpers.  Returns the results."""
        o = np.array([(int(c) & 0xFF) | (int(a) << 8) for c, a in ops], dtype=U64)
        c = np.ascontiguousarray(consts, dtype=U32)
        out = np.zeros(max_out, dtype=U64)
        n = C.c_size_t(0)
        self._rc(self.fn("bb_challenger_run")(_p(c), C.c_int(rf), C.c_int(rp), _p(o), C.c_size_t(o.size), _p(out),
                                              C.byref(n)), "challenger_run")
        return out[: n.value]

    def fri_commit_chal(self, consts, rf, rp, codeword, log_blowup, log_final_poly_len, max_log_arity, pow_bits,
                        pre=()):
        cw = np.ascontiguousarray(codeword, dtype=U32).reshape(-1)
        c = np.ascontiguousarray(consts, dtype=U32)
        pr = np.ascontiguousarray(np.array(pre, dtype=U32).reshape(-1)) if len(pre) else np.zeros(1, dtype=U32)
        log_len = (cw.size // 4).bit_length() - 1
        roots = np.zeros(64 * 8, dtype=U32)
        betas = np.zeros(64 * 4, dtype=U32)
        wit = np.zeros(64, dtype=U64)
        final = np.zeros(4 << log_final_poly_len, dtype=U32)
        post = np.zeros(8, dtype=U32)
        nr = C.c_size_t(0)
        self._rc(self.fn("bb4_fri_commit_chal")(_p(c), C.c_int(rf), C.c_int(rp), _p(cw), C.c_size_t(log_len),
                                                C.c_size_t(log_blowup), C.c_size_t(log_final_poly_len),
                                                C.c_size_t(max_log_arity), C.c_size_t(pow_bits), _p(pr),
                                                C.c_size_t(len(pre)), _p(roots), _p(betas), _p(wit), _p(final),
                                                _p(post), C.byref(nr)), "fri_commit_chal")
        r = nr.value
        return {"roots": roots[: 8 * r].reshape(-1, 8), "betas": betas[: 4 * r].reshape(-1, 4), "witnesses": wit[:r],
                "final": final.reshape(-1, 4), "post": post}

    def pcs_open(self, consts, rf, rp, mats, shifts, log_blowup, points, pre, lfpl, mla, num_queries,
                 commit_pow_bits, query_pow_bits):
        """TwoAdicFriPcs commit + open (one batch); points: list per matrix of
        BabyBear^4 values (4 words each).  Returns a dict of the proof parts."""
        ms = [np.ascontiguousarray(m, dtype=U32) for m in mats]
        n = len(ms)
        lhs = (C.c_size_t * n)(*[m.shape[0].bit_length() - 1 for m in ms])
        ws = (C.c_size_t * n)(*[m.shape[1] for m in ms])
        npts = (C.c_size_t * n)(*[len(p) for p in points])
        ptrs = (C.c_void_p * n)(*[m.ctypes.data for m in ms])
        sh = np.ascontiguousarray(shifts, dtype=U32)
        c = np.ascontiguousarray(consts, dtype=U32)
        flat_pts = np.ascontiguousarray(np.concatenate([np.array(p, dtype=U32).reshape(-1) for p in points]
                                                       + [np.zeros(4, dtype=U32)]))
        pr = np.ascontiguousarray(np.array(pre, dtype=U32).reshape(-1)) if len(pre) else np.zeros(1, dtype=U32)
        n_open = sum(len(p) * m.shape[1] for p, m in zip(points, ms))
        root = np.zeros(8, dtype=U32)
        opened = np.zeros(max(1, 4 * n_open), dtype=U32)
        roots = np.zeros(64 * 8, dtype=U32)
        cwit = np.zeros(64, dtype=U64)
        final = np.zeros(4 << lfpl, dtype=U32)
        qwit = np.zeros(1, dtype=U64)
        qidx = np.zeros(max(1, num_queries), dtype=U64)
        tw = sum(m.shape[1] for m in ms)
        qrows = np.zeros(max(1, num_queries * tw), dtype=U32)
        post = np.zeros(8, dtype=U32)
        nr = C.c_size_t(0)
        self._rc(self.fn("bb_pcs_open")(_p(c), C.c_int(rf), C.c_int(rp), C.c_size_t(n), lhs, ws, _p(sh), ptrs,
                                        C.c_size_t(log_blowup), npts, _p(flat_pts), _p(pr), C.c_size_t(len(pre)),
                                        C.c_size_t(lfpl), C.c_size_t(mla), C.c_size_t(num_queries),
                                        C.c_size_t(commit_pow_bits), C.c_size_t(query_pow_bits), _p(root),
                                        _p(opened), _p(roots), _p(cwit), _p(final), _p(qwit), _p(qidx), _p(qrows),
                                        _p(post), C.byref(nr)), "pcs_open")
        r = nr.value
        return {"root": root, "opened": opened[: 4 * n_open].reshape(-1, 4), "roots": roots[: 8 * r].reshape(-1, 8),
                "commit_witnesses": cwit[:r], "final": final.reshape(-1, 4), "query_witness": qwit,
                "query_indices": qidx[:num_queries], "query_rows": qrows[: num_queries * tw].reshape(num_queries, tw),
                "post": post}

    def pcs_commit(self, consts, rf, rp, mats, shifts, log_blowup):
        ms = [np.ascontiguousarray(m, dtype=U32) for m in mats]
        n = len(ms)
        lh = (C.c_size_t * n)(*[m.shape[0].bit_length() - 1 for m in ms])
        ws = (C.c_size_t * n)(*[m.shape[1] for m in ms])
        sh = np.ascontiguousarray(shifts, dtype=U32)
        ptrs = (C.c_void_p * n)(*[m.ctypes.data for m in ms])
        c = np.ascontiguousarray(consts, dtype=U32)
        root = np.zeros(8, dtype=U32)
        self._rc(self.fn("bb_pcs_commit")(_p(c), C.c_int(rf), C.c_int(rp), C.c_size_t(n), lh, ws, _p(sh), ptrs,
                                          C.c_size_t(log_blowup), _p(root)), "pcs_commit")
        return root


# challenger script op codes (oracle/oracle.cpp run_script)
OBSERVE, SAMPLE, SAMPLE_BITS, GRIND, CHECK_WITNESS, SAMPLE_EXT = range(6)


def check_witness_op(bits: int, witness: int):
    return (CHECK_WITNESS, bits | (witness << 6))


def oracle() -> Native:
    if not os.path.exists(ORACLE_SO):
        build_oracle()
    lib = Native(ORACLE_SO, "orc")
    f = lib.lib
    f.orc_bb_mul.restype = C.c_uint32
    f.orc_bb_mul.argtypes = [C.c_uint32, C.c_uint32]
    f.orc_bb_pow.restype = C.c_uint32
    f.orc_bb_pow.argtypes = [C.c_uint32, C.c_uint64]
    f.orc_bb_two_adic_generator.restype = C.c_uint32
    f.orc_bb_two_adic_generator.argtypes = [C.c_uint32]
    f.orc_gl_mul.restype = C.c_uint64
    f.orc_gl_mul.argtypes = [C.c_uint64, C.c_uint64]
    f.orc_gl_pow.restype = C.c_uint64
    f.orc_gl_pow.argtypes = [C.c_uint64, C.c_uint64]
    f.orc_gl_two_adic_generator.restype = C.c_uint64
    f.orc_gl_two_adic_generator.argtypes = [C.c_uint32]
    return lib


def reference():
    """The reference build, or None when it has not been built here."""
    if not os.path.exists(REF_SO):
        return None
    return Native(REF_SO, "ref")


def seeded(seed: int, n: int, field: str = "bb") -> np.ndarray:
    """std::mt19937_64(seed), value = rng() % p (the oracle's generator)."""
    lib = oracle().lib
    if field == "bb":
        out = np.empty(n, dtype=U32)
        lib.orc_bb_fill(C.c_uint64(seed), _p(out), C.c_size_t(n))
    else:
        out = np.empty(n, dtype=U64)
        lib.orc_gl_fill(C.c_uint64(seed), _p(out), C.c_size_t(n))
    return out


def p2_consts(seed: int, field: str = "bb", rf: int = 8, rp: int = 13) -> np.ndarray:
    """Poseidon2 constants from an independent stream (seed ^ tag):
    initial[rf/2][W], terminal[rf/2][W], internal[rp]."""
    return seeded(seed ^ 0x5EED_C0DE_0000_0002, rf * WIDTH[field] + rp, field)
