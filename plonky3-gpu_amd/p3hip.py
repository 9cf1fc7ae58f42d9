"""Python host binding of libp3hip (include/p3hip.h) over ctypes.

The classes mirror the reference's C++ API for the commit hot path
(dloghin/plonky3-gpu, src/include/):

* ``Radix2DitHip``      -- Radix2Dit / NttCuda: dft_batch, idft_batch,
                           coset_dft_batch, coset_idft_batch, coset_lde_batch
                           (radix2_dit.hpp:82-233, ntt_cuda.hpp:310-378)
* ``MerkleTreeMmcsHip`` -- p3_merkle_tree::MerkleTreeMmcs: commit, open_batch
                           (merkle_tree_mmcs.hpp:100-152)
* ``FlatMerkleTreeMmcsHip`` -- p3_merkle::MerkleTreeMmcs: commit_matrix,
                           open_row (fri_merkle_tree_mmcs.hpp:55-141)
* ``fri_fold`` / ``fri_commit_phase`` -- fold_matrix_cuda and the prove_fri
                           commit loop (fri_fold_cuda.hpp:306-348,
                           fri_prover.hpp:83-178)
* ``pcs_commit``        -- TwoAdicFriPcs::commit (two_adic_fri_pcs.hpp:207-266)

Matrices are numpy arrays of canonical field elements (uint32 BabyBear,
uint64 Goldilocks), shape (height, width).  Errors are raised as ValueError
(the reference's std::invalid_argument) or RuntimeError (std::runtime_error).
The library must be built (``python __graft_entry__.py`` or
``make -C plonky3-gpu_amd/csrc``); there is no CPU fallback.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

BABYBEAR = 0
GOLDILOCKS = 1
# p3h_batch_plan flags
BATCH_GRAPH = 1
BATCH_QUEUE_LDE = 2
BB_P = 0x78000001
GL_P = 0xFFFFFFFF00000001
DTYPE = {BABYBEAR: np.uint32, GOLDILOCKS: np.uint64}
DIGEST = {BABYBEAR: 8, GOLDILOCKS: 4}
P2_WIDTH = {BABYBEAR: 16, GOLDILOCKS: 8}

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libp3hip.so")

_OK, _EINVAL, _ENOMEM, _EDEVICE, _EBUDGET, _ESTATE = range(6)


class BudgetError(RuntimeError):
    """A request exceeded the context's HBM budget."""


_lib = None

_vp = C.c_void_p
_u32, _u64, _i32 = C.c_uint32, C.c_uint64, C.c_int
_SIGS = {
    "p3h_ctx_create": (_i32, [_i32, _u64, C.POINTER(_vp)]),
    "p3h_ctx_destroy": (_i32, [_vp]),
    "p3h_last_error": (C.c_char_p, [_vp]),
    "p3h_ctx_sync": (_i32, [_vp]),
    "p3h_ctx_memory": (_i32, [_vp, C.POINTER(_u64), C.POINTER(_u64), C.POINTER(_u64)]),
    "p3h_ctx_trim": (_i32, [_vp]),
    "p3h_ctx_table_bytes": (_i32, [_vp, C.POINTER(_u64)]),
    "p3h_probe_butterflies": (_i32, [_vp, _i32, _u32, _u32, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "p3h_ctx_set_profiling": (_i32, [_vp, _i32]),
    "p3h_phase_stats": (_i32, [_vp, C.c_char_p, C.POINTER(_u64), C.POINTER(C.c_double)]),
    "p3h_ctx_set_ntt_tuning": (_i32, [_vp, _u32, _u32, _u32]),
    "p3h_ctx_set_hash_engine": (_i32, [_vp, _i32]),
    "p3h_host_alloc": (_i32, [_u64, C.POINTER(_vp)]),
    "p3h_host_free": (_i32, [_vp]),
    "p3h_seeded_fill": (_i32, [_i32, _u64, _vp, _u64]),
    "p3h_poseidon2_set": (_i32, [_vp, _i32, _u32, _u32, _u32, _vp]),
    "p3h_poseidon2_permute": (_i32, [_vp, _i32, _vp, _u64]),
    "p3h_mat_upload": (_i32, [_vp, _i32, _vp, _u64, _u32, C.POINTER(_vp)]),
    "p3h_mat_alloc": (_i32, [_vp, _i32, _u64, _u32, C.POINTER(_vp)]),
    "p3h_mat_download": (_i32, [_vp, _vp, _vp]),
    "p3h_mat_download_rows": (_i32, [_vp, _vp, _u64, _u64, _vp]),
    "p3h_mat_info": (_i32, [_vp, C.POINTER(_i32), C.POINTER(_u64), C.POINTER(_u32)]),
    "p3h_mat_free": (_i32, [_vp]),
    "p3h_mat_device_ptr": (_i32, [_vp, C.POINTER(_vp)]),
    "p3h_fri_fold_device": (_i32, [_vp, _vp, _vp, _vp, _u32, _u32, _u64]),
    "p3h_fri_fold_device_noalloc": (_i32, [_vp, _vp, _vp, _vp, _u32, _u32, _u64]),
    "p3h_dft": (_i32, [_vp, _i32, _vp, _u64, C.POINTER(_vp)]),
    "p3h_coset_lde": (_i32, [_vp, _vp, _u32, _u64, C.POINTER(_vp)]),
    "p3h_mmcs_commit": (_i32, [_vp, C.POINTER(_vp), _u32, _u32, C.POINTER(_vp)]),
    "p3h_flat_commit": (_i32, [_vp, _vp, C.POINTER(_vp)]),
    "p3h_tree_cap": (_i32, [_vp, _vp, _vp, C.POINTER(_u32)]),
    "p3h_tree_info": (_i32, [_vp, C.POINTER(_u32), C.POINTER(_u64), C.POINTER(_u32)]),
    "p3h_tree_matrices": (_i32, [_vp, C.POINTER(_u32), _vp, _vp]),
    "p3h_tree_layer": (_i32, [_vp, _vp, _u32, _vp]),
    "p3h_tree_open": (_i32, [_vp, _vp, _vp, _u32, _vp, _vp]),
    "p3h_tree_free": (_i32, [_vp]),
    "p3h_mmcs_verify": (_i32, [_vp, _i32, _vp, _u32, _vp, _vp, _u32, _u64, _vp, _vp, _u32, C.POINTER(C.c_int)]),
    "p3h_flat_verify": (_i32, [_vp, _i32, _vp, _u32, _u64, _vp, _u32, _vp, _u32, C.POINTER(C.c_int)]),
    "p3h_fri_fold": (_i32, [_vp, _vp, _u32, _vp, C.POINTER(_vp)]),
    "p3h_fri_commit_phase": (_i32, [_vp, _vp, _u32, _u32, _u32, _vp, _vp, _vp, C.POINTER(_u32), C.POINTER(_vp)]),
    "p3h_fri_commit_phase_chal": (_i32, [_vp, _vp, _u32, _u32, _u32, _vp, _u32, _vp, _vp, _vp, _vp, C.POINTER(_u32),
                                         C.POINTER(_vp)]),
    "p3h_fri_layer_tree": (_i32, [_vp, _u32, C.POINTER(_vp)]),
    "p3h_challenger_create": (_i32, [_vp, C.POINTER(_vp)]),
    "p3h_challenger_clone": (_i32, [_vp, C.POINTER(_vp)]),
    "p3h_challenger_free": (_i32, [_vp]),
    "p3h_challenger_observe": (_i32, [_vp, _vp, _u32]),
    "p3h_challenger_observe_cap": (_i32, [_vp, _vp]),
    "p3h_challenger_sample": (_i32, [_vp, _u32, _vp]),
    "p3h_challenger_sample_bits": (_i32, [_vp, _u32, C.POINTER(_u64)]),
    "p3h_challenger_grind": (_i32, [_vp, _u32, C.POINTER(_u64)]),
    "p3h_challenger_check_witness": (_i32, [_vp, _u32, _u64, C.POINTER(_i32)]),
    "p3h_challenger_state": (_i32, [_vp, _vp]),
    "p3h_fri_free": (_i32, [_vp]),
    "p3h_pcs_commit": (_i32, [_vp, C.POINTER(_vp), _vp, _u32, _u32, C.POINTER(_vp)]),
    "p3h_pcs_open": (_i32, [_vp, _vp, _vp, _vp, _vp, _vp, _u32, _u32, _u32, _u32, _u32, _vp, _vp, _vp, _vp, _vp,
                            _vp, _vp, C.POINTER(_u32), C.POINTER(_vp)]),
    "p3h_batch_commit": (_i32, [_vp, C.POINTER(_vp), _vp, _u32, _u64, _u32, C.POINTER(_vp)]),
    "p3h_mat_write": (_i32, [_vp, _vp, _vp]),
    "p3h_batch_plan": (_i32, [_vp, C.POINTER(_vp), _vp, _u32, _u64, _u32, _i32, C.POINTER(_vp)]),
    "p3h_batch_run": (_i32, [_vp, _vp]),
    "p3h_batch_tree": (_i32, [_vp, C.POINTER(_vp)]),
    "p3h_batch_lde_launches": (_i32, [_vp, C.POINTER(_u32)]),
    "p3h_batch_free": (_i32, [_vp]),
    "p3h_autotune_lde": (_i32, [_vp, _i32, _u32, _u32, _u32, _u32, C.POINTER(_u32), C.POINTER(_u32),
                                C.POINTER(_u32), C.POINTER(C.c_double)]),
    "p3h_tuning_save": (_i32, [_vp, C.c_char_p]),
    "p3h_tuning_load": (_i32, [_vp, C.c_char_p]),
    "p3h_serve_create": (_i32, [_i32, _u64, _u32, _u32, _vp, _u32, _u32, C.POINTER(_vp)]),
    "p3h_serve_submit": (_i32, [_vp, _vp, _u64, _u32, _u32, _u64, C.POINTER(_u64)]),
    "p3h_serve_wait": (_i32, [_vp, _u64, _vp, C.POINTER(C.c_double)]),
    "p3h_serve_poll": (_i32, [_vp, _u64, C.POINTER(_i32)]),
    "p3h_serve_open": (_i32, [_vp, _u64, _vp, _u32, _vp, _vp]),
    "p3h_serve_release": (_i32, [_vp, _u64]),
    "p3h_serve_stats": (_i32, [_vp, C.POINTER(_u64), C.POINTER(_u64), C.POINTER(_u64), C.POINTER(_u64)]),
    "p3h_serve_last_error": (C.c_char_p, [_vp]),
    "p3h_serve_destroy": (_i32, [_vp]),
}


def exported_symbols():
    """Names the C ABI declares (include/p3hip.h)."""
    return list(_SIGS)


def load_library(path: str = LIB_PATH):
    """Load libp3hip.so and attach signatures; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"libp3hip.so not built at {path}; run `python __graft_entry__.py`")
    lib = C.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class PinnedArray(np.ndarray):
    """A numpy array whose storage is page-locked host memory (p3h_host_alloc);
    the memory is freed when the last view of it goes away."""


class _PinnedBlock:
    def __init__(self, lib, addr):
        self.lib, self.addr = lib, addr

    def __del__(self):
        if self.addr:
            self.lib.p3h_host_free(self.addr)
            self.addr = None


def pinned_empty(shape, dtype) -> PinnedArray:
    """An uninitialised array in page-locked host memory: uploads from it, and
    serving submissions of it, run at the full PCIe DMA rate."""
    lib = load_library()
    dtype = np.dtype(dtype)
    nbytes = int(np.prod(shape, dtype=np.int64)) * dtype.itemsize
    p = _vp()
    rc = lib.p3h_host_alloc(max(nbytes, 1), C.byref(p))
    if rc != _OK:
        raise MemoryError(f"p3h_host_alloc({nbytes}) failed with status {rc}")
    buf = (C.c_char * max(nbytes, 1)).from_address(p.value)
    # every numpy view of the memory keeps `buf` alive, and `buf` the block
    buf._p3h_owner = _PinnedBlock(lib, p.value)
    return np.frombuffer(buf, dtype=dtype, count=nbytes // dtype.itemsize).reshape(shape).view(PinnedArray)


def seeded_fill(field: int, seed: int, n: int, out: np.ndarray | None = None) -> np.ndarray:
    """n canonical field elements, element i = the i-th std::mt19937_64(seed)
    output mod p (p3h_seeded_fill; host only, no device work)."""
    lib = load_library()
    dt = DTYPE[field]
    if out is None:
        out = np.empty(n, dtype=dt)
    if out.dtype != dt or out.size < n or not out.flags.c_contiguous:
        raise ValueError("seeded_fill: output must be a contiguous array of the field's dtype")
    rc = lib.p3h_seeded_fill(field, seed, out.ctypes.data_as(_vp), n)
    if rc != _OK:
        raise ValueError(f"p3h_seeded_fill failed with status {rc}")
    return out


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_vp)


def _field_of(dtype) -> int:
    if dtype == np.uint32:
        return BABYBEAR
    if dtype == np.uint64:
        return GOLDILOCKS
    raise ValueError(f"unsupported dtype {dtype}")


class Context:
    """One device context (one HIP stream, a caching allocator, an HBM budget)."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        self.lib = load_library()
        h = _vp()
        rc = self.lib.p3h_ctx_create(device, hbm_budget, C.byref(h))
        if rc != _OK:
            raise RuntimeError(f"p3h_ctx_create failed with status {rc}")
        self.h = h

    # -- plumbing ---------------------------------------------------------
    def _check(self, rc: int):
        if rc == _OK:
            return
        msg = self.lib.p3h_last_error(self.h).decode()
        if rc == _EINVAL:
            raise ValueError(msg)
        if rc == _EBUDGET:
            raise BudgetError(msg)
        raise RuntimeError(f"[status {rc}] {msg}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.p3h_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def sync(self):
        self._check(self.lib.p3h_ctx_sync(self.h))

    def memory(self):
        a, b, c = _u64(), _u64(), _u64()
        self._check(self.lib.p3h_ctx_memory(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return {"in_use": a.value, "budget": b.value, "cached": c.value}

    def probe_butterflies(self, field: int, blocks: int = 16384, iters: int = 200):
        """(butterflies per second, kernel ms): the NTT butterflies' VALU
        ceiling (p3h_probe_butterflies)."""
        ms, n = C.c_double(), C.c_double()
        self._check(self.lib.p3h_probe_butterflies(self.h, field, blocks, iters, C.byref(ms), C.byref(n)))
        return n.value / (ms.value / 1e3), ms.value

    def table_bytes(self) -> int:
        """Bytes of in_use held by the context's table caches."""
        v = _u64()
        self._check(self.lib.p3h_ctx_table_bytes(self.h, C.byref(v)))
        return v.value

    def trim(self):
        self._check(self.lib.p3h_ctx_trim(self.h))

    def set_profiling(self, on: bool):
        self._check(self.lib.p3h_ctx_set_profiling(self.h, 1 if on else 0))

    def phase_stats(self, phase: str):
        n, ms = _u64(), C.c_double()
        self._check(self.lib.p3h_phase_stats(self.h, phase.encode(), C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def set_hash_engine(self, engine: int):
        """BabyBear Poseidon2 engine: 0 = uint32 Montgomery, 1 = exact binary64."""
        self._check(self.lib.p3h_ctx_set_hash_engine(self.h, engine))

    def set_ntt_tuning(self, log_cols: int = 0, max_stages: int = 0, log_slab: int = 0):
        self._check(self.lib.p3h_ctx_set_ntt_tuning(self.h, log_cols, max_stages, log_slab))

    # -- Poseidon2 --------------------------------------------------------
    def set_poseidon2(self, field: int, consts: np.ndarray, rounds_f: int = 8, rounds_p: int = 13):
        width = P2_WIDTH[field]
        c = np.ascontiguousarray(consts, dtype=DTYPE[field])
        if c.size != rounds_f * width + rounds_p:
            raise ValueError("constant count does not match rounds_f * width + rounds_p")
        self._check(self.lib.p3h_poseidon2_set(self.h, field, width, rounds_f, rounds_p, _ptr(c)))

    def permute(self, field: int, states: np.ndarray) -> np.ndarray:
        s = np.ascontiguousarray(states, dtype=DTYPE[field]).copy()
        n = s.size // P2_WIDTH[field]
        self._check(self.lib.p3h_poseidon2_permute(self.h, field, _ptr(s), n))
        return s

    # -- matrices ---------------------------------------------------------
    def upload(self, a: np.ndarray) -> "Mat":
        a = np.ascontiguousarray(a)
        if a.ndim == 1:
            a = a.reshape(-1, 1)
        field = _field_of(a.dtype)
        h = _vp()
        self._check(self.lib.p3h_mat_upload(self.h, field, _ptr(a), a.shape[0], a.shape[1], C.byref(h)))
        return Mat(self, h)

    def alloc(self, field: int, height: int, width: int) -> "Mat":
        h = _vp()
        self._check(self.lib.p3h_mat_alloc(self.h, field, height, width, C.byref(h)))
        return Mat(self, h)

    # -- operations -------------------------------------------------------
    def dft(self, kind: int, m: "Mat", shift: int = 1) -> "Mat":
        h = _vp()
        self._check(self.lib.p3h_dft(self.h, kind, m.h, shift, C.byref(h)))
        return Mat(self, h)

    def coset_lde(self, m: "Mat", added_bits: int, shift: int) -> "Mat":
        h = _vp()
        self._check(self.lib.p3h_coset_lde(self.h, m.h, added_bits, shift, C.byref(h)))
        return Mat(self, h)

    def mmcs_commit(self, mats, cap_height: int = 0) -> "Tree":
        arr = (_vp * len(mats))(*[m.h for m in mats])
        h = _vp()
        self._check(self.lib.p3h_mmcs_commit(self.h, arr, len(mats), cap_height, C.byref(h)))
        return Tree(self, h, mats[0].field)

    def flat_commit(self, m: "Mat") -> "Tree":
        h = _vp()
        self._check(self.lib.p3h_flat_commit(self.h, m.h, C.byref(h)))
        return Tree(self, h, m.field)

    def mmcs_verify(self, field: int, cap, dims, index: int, values, proof) -> bool:
        """MerkleTreeMmcs::verify_batch (merkle_tree_mmcs.hpp:168-254).  dims:
        (height, width) per matrix in commit order; values: their opened rows
        concatenated in that order; proof: sibling digests, leaf level first;
        cap: the commitment's digests."""
        dt, dg = DTYPE[field], DIGEST[field]
        cap = np.ascontiguousarray(cap, dtype=dt).reshape(-1, dg)
        hs = np.ascontiguousarray([int(h) for h, _ in dims], dtype=np.uint64)
        ws = np.ascontiguousarray([int(w) for _, w in dims], dtype=np.uint32)
        vals = np.ascontiguousarray(values, dtype=dt).reshape(-1)
        if vals.size != int(ws.sum()):
            raise ValueError("opened values do not match the matrix widths")
        pr = np.ascontiguousarray(proof, dtype=dt).reshape(-1, dg)
        ok = C.c_int(0)
        self._check(self.lib.p3h_mmcs_verify(self.h, field, _ptr(cap), cap.shape[0], _ptr(hs), _ptr(ws), len(dims),
                                             index, _ptr(vals), _ptr(pr), pr.shape[0], C.byref(ok)))
        return bool(ok.value)

    def flat_verify(self, field: int, root, index: int, row, proof) -> bool:
        """p3_merkle::MerkleTreeMmcs::verify_row (fri_merkle_tree_mmcs.hpp:146-168)."""
        dt, dg = DTYPE[field], DIGEST[field]
        root = np.ascontiguousarray(root, dtype=dt).reshape(-1, dg)
        r = np.ascontiguousarray(row, dtype=dt).reshape(-1)
        pr = np.ascontiguousarray(proof, dtype=dt).reshape(-1, dg)
        ok = C.c_int(0)
        self._check(self.lib.p3h_flat_verify(self.h, field, _ptr(root), root.shape[0], index, _ptr(r), r.size,
                                             _ptr(pr), pr.shape[0], C.byref(ok)))
        return bool(ok.value)

    def pcs_commit(self, mats, shifts, log_blowup: int) -> "Tree":
        arr = (_vp * len(mats))(*[m.h for m in mats])
        sh = np.ascontiguousarray(shifts, dtype=np.uint64)
        h = _vp()
        self._check(self.lib.p3h_pcs_commit(self.h, arr, _ptr(sh), len(mats), log_blowup, C.byref(h)))
        t = Tree(self, h, BABYBEAR)
        t.mat_widths = [m.width for m in mats]  # the committed matrices, for pcs_open's output sizes
        return t

    def pcs_open(self, tree: "Tree", shifts, points, challenger: "Challenger", log_final_poly_len: int,
                 max_log_arity: int, num_queries: int, commit_pow_bits: int = 0, query_pow_bits: int = 0,
                 keep: bool = False):
        """TwoAdicFriPcs::open over the LDEs of a p3h_pcs_commit tree; points: per
        matrix an (n x 4) array of BabyBear^4 values.  Returns a dict."""
        n = len(points)
        sh = np.ascontiguousarray(shifts, dtype=np.uint64)
        npts = np.array([len(p) for p in points], dtype=np.uint32)
        flat = np.ascontiguousarray(np.concatenate([np.asarray(p, dtype=np.uint32).reshape(-1) for p in points]
                                                   + [np.zeros(4, dtype=np.uint32)]))
        widths = tree.mat_widths
        n_open = int(sum(int(k) * w for k, w in zip(npts, widths)))
        tw = int(sum(widths))
        opened = np.zeros((max(1, n_open), 4), dtype=np.uint32)
        roots = np.zeros((32, 8), dtype=np.uint32)
        cwit = np.zeros(32, dtype=np.uint64)
        final = np.zeros((1 << log_final_poly_len, 4), dtype=np.uint32)
        qwit = np.zeros(1, dtype=np.uint64)
        qidx = np.zeros(max(1, num_queries), dtype=np.uint64)
        qrows = np.zeros((max(1, num_queries), tw), dtype=np.uint32)
        nr = _u32()
        kp = _vp()
        assert n == len(widths), "one point list per committed matrix"
        self._check(self.lib.p3h_pcs_open(self.h, tree.h, _ptr(sh), _ptr(npts), _ptr(flat), challenger.h,
                                          log_final_poly_len, max_log_arity, num_queries, commit_pow_bits,
                                          query_pow_bits, _ptr(opened), _ptr(roots), _ptr(cwit), _ptr(final),
                                          _ptr(qwit), _ptr(qidx), _ptr(qrows), C.byref(nr),
                                          C.byref(kp) if keep else None))
        r = nr.value
        out = {"opened": opened[:n_open], "roots": roots[:r], "commit_witnesses": cwit[:r], "final": final,
               "query_witness": qwit, "query_indices": qidx[:num_queries], "query_rows": qrows[:num_queries]}
        if keep:
            out["fri"] = Fri(self, kp)
        return out

    def batch_commit(self, traces, added_bits, shift: int, cap_height: int = 0) -> "Tree":
        arr = (_vp * len(traces))(*[m.h for m in traces])
        ab = np.ascontiguousarray(added_bits, dtype=np.uint32)
        h = _vp()
        self._check(self.lib.p3h_batch_commit(self.h, arr, _ptr(ab), len(traces), shift, cap_height, C.byref(h)))
        return Tree(self, h, BABYBEAR)

    def batch_plan(self, traces, added_bits, shift: int, cap_height: int = 0, use_graph: bool = True,
                   lde_queue: bool = False) -> "BatchPlan":
        """Planned batched commit (p3h_batch_plan).  lde_queue: the LDE passes of
        all traces as one work-queue launch instead of launches grouped by pass
        shape."""
        arr = (_vp * len(traces))(*[m.h for m in traces])
        ab = np.ascontiguousarray(added_bits, dtype=np.uint32)
        h = _vp()
        flags = (BATCH_GRAPH if use_graph else 0) | (BATCH_QUEUE_LDE if lde_queue else 0)
        self._check(self.lib.p3h_batch_plan(self.h, arr, _ptr(ab), len(traces), shift, cap_height, flags,
                                            C.byref(h)))
        return BatchPlan(self, h, traces)

    def autotune_lde(self, field: int, log_n: int, width: int, added_bits: int, iters: int = 3):
        lc, ms, sl, t = _u32(), _u32(), _u32(), C.c_double()
        self._check(self.lib.p3h_autotune_lde(self.h, field, log_n, width, added_bits, iters, C.byref(lc),
                                              C.byref(ms), C.byref(sl), C.byref(t)))
        return {"log_cols": lc.value, "max_stages": ms.value, "log_slab": sl.value, "ms": t.value}

    def tuning_save(self, path: str):
        self._check(self.lib.p3h_tuning_save(self.h, path.encode()))

    def tuning_load(self, path: str):
        self._check(self.lib.p3h_tuning_load(self.h, path.encode()))

    def fri_fold(self, cur: "Mat", log_arity: int, beta) -> "Mat":
        b = np.ascontiguousarray(beta, dtype=np.uint32)
        h = _vp()
        self._check(self.lib.p3h_fri_fold(self.h, cur.h, log_arity, _ptr(b), C.byref(h)))
        return Mat(self, h)

    def fri_fold_device(self, d_in: int, d_out: int, beta, log_arity: int, log_height: int, n: int):
        """fold_matrix_cuda_device: fold n rows from device address d_in into
        device address d_out (BabyBear^4, 4 words per value)."""
        b = np.ascontiguousarray(beta, dtype=np.uint32).reshape(4)
        self._check(self.lib.p3h_fri_fold_device(self.h, d_in, d_out, _ptr(b), log_arity, log_height, n))

    def fri_commit_phase(self, codeword: "Mat", log_blowup: int, log_final_poly_len: int, max_log_arity: int,
                         betas: np.ndarray, keep: bool = False):
        """Returns (roots [rounds x 8], final_poly [2^lfpl x 4], Fri or None)."""
        b = np.ascontiguousarray(betas, dtype=np.uint32).reshape(-1)
        max_rounds = max(1, b.size // 4)
        roots = np.zeros((max_rounds, 8), dtype=np.uint32)
        final = np.zeros((1 << log_final_poly_len, 4), dtype=np.uint32)
        nr = _u32()
        kp = _vp()
        self._check(self.lib.p3h_fri_commit_phase(self.h, codeword.h, log_blowup, log_final_poly_len, max_log_arity,
                                                  _ptr(b), _ptr(roots), _ptr(final), C.byref(nr),
                                                  C.byref(kp) if keep else None))
        fri = Fri(self, kp) if keep else None
        return roots[: nr.value], final, fri

    def challenger(self) -> "Challenger":
        return Challenger(self)

    def fri_commit_phase_chal(self, codeword: "Mat", log_blowup: int, log_final_poly_len: int, max_log_arity: int,
                              challenger: "Challenger", pow_bits: int = 0, keep: bool = False):
        """prove_fri's commit loop with the transcript in the loop.  Returns a dict
        with roots, betas, witnesses, final (and fri when keep)."""
        roots = np.zeros((32, 8), dtype=np.uint32)
        betas = np.zeros((32, 4), dtype=np.uint32)
        wit = np.zeros(32, dtype=np.uint64)
        final = np.zeros((1 << log_final_poly_len, 4), dtype=np.uint32)
        nr = _u32()
        kp = _vp()
        self._check(self.lib.p3h_fri_commit_phase_chal(self.h, codeword.h, log_blowup, log_final_poly_len,
                                                       max_log_arity, challenger.h, pow_bits, _ptr(roots),
                                                       _ptr(betas), _ptr(wit), _ptr(final), C.byref(nr),
                                                       C.byref(kp) if keep else None))
        r = nr.value
        out = {"roots": roots[:r], "betas": betas[:r], "witnesses": wit[:r], "final": final}
        if keep:
            out["fri"] = Fri(self, kp)
        return out


class Mat:
    """Device-resident matrix handle."""

    def __init__(self, ctx: Context, h):
        self.ctx, self.h = ctx, h
        f, hh, w = _i32(), _u64(), _u32()
        ctx._check(ctx.lib.p3h_mat_info(h, C.byref(f), C.byref(hh), C.byref(w)))
        self.field, self.height, self.width = f.value, hh.value, w.value

    @property
    def shape(self):
        return (self.height, self.width)

    def download(self) -> np.ndarray:
        out = np.empty((self.height, self.width), dtype=DTYPE[self.field])
        self.ctx._check(self.ctx.lib.p3h_mat_download(self.ctx.h, self.h, _ptr(out)))
        return out

    def write(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=DTYPE[self.field])
        if a.size != self.height * self.width:
            raise ValueError("shape mismatch")
        self.ctx._check(self.ctx.lib.p3h_mat_write(self.ctx.h, self.h, _ptr(a)))

    def device_ptr(self) -> int:
        """Device address of the first element (p3h_mat_device_ptr)."""
        p = _vp()
        self.ctx._check(self.ctx.lib.p3h_mat_device_ptr(self.h, C.byref(p)))
        return p.value

    def download_rows(self, row0: int, nrows: int) -> np.ndarray:
        out = np.empty((nrows, self.width), dtype=DTYPE[self.field])
        self.ctx._check(self.ctx.lib.p3h_mat_download_rows(self.ctx.h, self.h, row0, nrows, _ptr(out)))
        return out

    def free(self):
        if self.h:
            self.ctx.lib.p3h_mat_free(self.h)
            self.h = None

    def __del__(self):
        try:
            if self.h and self.ctx.h:
                self.free()
        except Exception:
            pass


class Tree:
    """Device-resident Merkle tree (prover data)."""

    def __init__(self, ctx: Context, h, field: int, owned: bool = True):
        self.ctx, self.h, self.field, self.owned = ctx, h, field, owned
        nl, nleaf, tw = _u32(), _u64(), _u32()
        ctx._check(ctx.lib.p3h_tree_info(h, C.byref(nl), C.byref(nleaf), C.byref(tw)))
        self.n_layers, self.n_leaves, self.total_width = nl.value, nleaf.value, tw.value

    def cap(self) -> np.ndarray:
        n = _u32()
        self.ctx._check(self.ctx.lib.p3h_tree_cap(self.ctx.h, self.h, None, C.byref(n)))
        out = np.empty((n.value, DIGEST[self.field]), dtype=DTYPE[self.field])
        self.ctx._check(self.ctx.lib.p3h_tree_cap(self.ctx.h, self.h, _ptr(out), C.byref(n)))
        return out

    def root(self) -> np.ndarray:
        return self.cap()[0]

    def dims(self):
        """(height, width) of each committed matrix in commit order (p3h_tree_matrices)."""
        n = _u32()
        self.ctx._check(self.ctx.lib.p3h_tree_matrices(self.h, C.byref(n), None, None))
        hs = np.empty(n.value, dtype=np.uint64)
        ws = np.empty(n.value, dtype=np.uint32)
        self.ctx._check(self.ctx.lib.p3h_tree_matrices(self.h, C.byref(n), _ptr(hs), _ptr(ws)))
        return [(int(h), int(w)) for h, w in zip(hs, ws)]

    def layer(self, i: int) -> np.ndarray:
        cnt = self.n_leaves
        # counts per layer: computed by halving (flat trees round up)
        for _ in range(i):
            cnt = (cnt + 1) // 2
        out = np.empty((cnt, DIGEST[self.field]), dtype=DTYPE[self.field])
        self.ctx._check(self.ctx.lib.p3h_tree_layer(self.ctx.h, self.h, i, _ptr(out)))
        return out

    def open(self, indices):
        """Returns (values [n x total_width], proof [n x (n_layers-1) x DIGEST])."""
        idx = np.ascontiguousarray(indices, dtype=np.uint64).reshape(-1)
        vals = np.empty((idx.size, self.total_width), dtype=DTYPE[self.field])
        proof = np.empty((idx.size, self.n_layers - 1, DIGEST[self.field]), dtype=DTYPE[self.field])
        self.ctx._check(self.ctx.lib.p3h_tree_open(self.ctx.h, self.h, _ptr(idx), idx.size, _ptr(vals), _ptr(proof)))
        return vals, proof

    def free(self):
        if self.h and self.owned:
            self.ctx.lib.p3h_tree_free(self.h)
        self.h = None

    def __del__(self):
        try:
            if self.h and self.ctx.h:
                self.free()
        except Exception:
            pass


class BatchPlan:
    """A planned batched commit (optionally a captured HIP graph)."""

    def __init__(self, ctx: Context, h, traces):
        self.ctx, self.h, self.traces = ctx, h, traces

    def run(self, want_cap: bool = True):
        if not want_cap:
            self.ctx._check(self.ctx.lib.p3h_batch_run(self.h, None))
            return None
        n = _u32()
        tr = _vp()
        self.ctx._check(self.ctx.lib.p3h_batch_tree(self.h, C.byref(tr)))
        self.ctx._check(self.ctx.lib.p3h_tree_cap(self.ctx.h, tr, None, C.byref(n)))
        out = np.empty((n.value, 8), dtype=np.uint32)
        self.ctx._check(self.ctx.lib.p3h_batch_run(self.h, _ptr(out)))
        return out

    def lde_launches(self) -> int:
        """Kernel launches of the LDE phase per run (1: the queue launch)."""
        n = _u32()
        self.ctx._check(self.ctx.lib.p3h_batch_lde_launches(self.h, C.byref(n)))
        return n.value

    def tree(self) -> "Tree":
        tr = _vp()
        self.ctx._check(self.ctx.lib.p3h_batch_tree(self.h, C.byref(tr)))
        return Tree(self.ctx, tr, BABYBEAR, owned=False)

    def free(self):
        if self.h:
            self.ctx.lib.p3h_batch_free(self.h)
            self.h = None

    def __del__(self):
        try:
            if self.h and self.ctx.h:
                self.free()
        except Exception:
            pass


class Challenger:
    """DuplexChallenger<BabyBear, Poseidon2 W16, 16, 8> with its state on the
    device (p3h_challenger_*), using the context's BabyBear Poseidon2 constants."""

    def __init__(self, ctx: "Context", h=None):
        self.ctx = ctx
        if h is None:
            h = _vp()
            ctx._check(ctx.lib.p3h_challenger_create(ctx.h, C.byref(h)))
        self.h = h

    def clone(self) -> "Challenger":
        h = _vp()
        self.ctx._check(self.ctx.lib.p3h_challenger_clone(self.h, C.byref(h)))
        return Challenger(self.ctx, h)

    def observe(self, values):
        v = np.ascontiguousarray(np.atleast_1d(values), dtype=np.uint32).reshape(-1)
        self.ctx._check(self.ctx.lib.p3h_challenger_observe(self.h, _ptr(v), v.size))

    def observe_cap(self, tree: "Tree"):
        self.ctx._check(self.ctx.lib.p3h_challenger_observe_cap(self.h, tree.h))

    def sample(self, n: int = 1) -> np.ndarray:
        out = np.zeros(max(1, n), dtype=np.uint32)
        self.ctx._check(self.ctx.lib.p3h_challenger_sample(self.h, n, _ptr(out)))
        return out[:n]

    def sample_ext(self) -> np.ndarray:
        return self.sample(4)

    def sample_bits(self, bits: int) -> int:
        r = _u64()
        self.ctx._check(self.ctx.lib.p3h_challenger_sample_bits(self.h, bits, C.byref(r)))
        return r.value

    def grind(self, bits: int) -> int:
        w = _u64()
        self.ctx._check(self.ctx.lib.p3h_challenger_grind(self.h, bits, C.byref(w)))
        return w.value

    def check_witness(self, bits: int, witness: int) -> bool:
        ok = _i32()
        self.ctx._check(self.ctx.lib.p3h_challenger_check_witness(self.h, bits, witness, C.byref(ok)))
        return bool(ok.value)

    def state(self) -> np.ndarray:
        out = np.zeros(34, dtype=np.uint32)
        self.ctx._check(self.ctx.lib.p3h_challenger_state(self.h, _ptr(out)))
        return out

    def free(self):
        if self.h:
            self.ctx.lib.p3h_challenger_free(self.h)
            self.h = None

    def __del__(self):
        try:
            if self.h and self.ctx.h:
                self.free()
        except Exception:
            pass


class Fri:
    """Kept FRI commit-phase layers."""

    def __init__(self, ctx: Context, h):
        self.ctx, self.h = ctx, h

    def layer_tree(self, r: int) -> Tree:
        t = _vp()
        self.ctx._check(self.ctx.lib.p3h_fri_layer_tree(self.h, r, C.byref(t)))
        return Tree(self.ctx, t, BABYBEAR, owned=False)

    def free(self):
        if self.h:
            self.ctx.lib.p3h_fri_free(self.h)
            self.h = None

    def __del__(self):
        try:
            if self.h and self.ctx.h:
                self.free()
        except Exception:
            pass


class Server:
    """Commit server (p3h_serve_*): continuous batching over n_slots streams,
    HBM-budget admission, resident results for openings."""

    def __init__(self, consts, rounds_f: int = 8, rounds_p: int = 13, device: int = 0, hbm_budget: int = 0,
                 n_slots: int = 4, max_batch: int = 8):
        self.lib = load_library()
        c = np.ascontiguousarray(consts, dtype=np.uint32)
        h = _vp()
        rc = self.lib.p3h_serve_create(device, hbm_budget, n_slots, max_batch, _ptr(c), rounds_f, rounds_p,
                                       C.byref(h))
        if rc != _OK:
            raise RuntimeError(f"p3h_serve_create failed with status {rc}")
        self.h = h
        self._held = {}  # request id -> host trace the server reads until done

    def _check(self, rc: int):
        if rc == _OK:
            return
        msg = self.lib.p3h_serve_last_error(self.h).decode()
        if rc == _EINVAL:
            raise ValueError(msg)
        if rc == _EBUDGET:
            raise BudgetError(msg)
        raise RuntimeError(f"[status {rc}] {msg}")

    def submit(self, trace: np.ndarray, added_bits: int = 1, shift: int = 31) -> int:
        t = np.ascontiguousarray(trace, dtype=np.uint32)
        rid = _u64()
        self._check(self.lib.p3h_serve_submit(self.h, _ptr(t), t.shape[0], t.shape[1], added_bits, shift,
                                              C.byref(rid)))
        self._held[rid.value] = t
        return rid.value

    def wait(self, rid: int):
        """(root [8], latency_ms)"""
        root = np.empty(8, dtype=np.uint32)
        lat = C.c_double()
        self._check(self.lib.p3h_serve_wait(self.h, rid, _ptr(root), C.byref(lat)))
        self._held.pop(rid, None)
        return root, lat.value

    def done(self, rid: int) -> bool:
        d = _i32()
        self._check(self.lib.p3h_serve_poll(self.h, rid, C.byref(d)))
        if d.value:
            self._held.pop(rid, None)
        return bool(d.value)

    def open(self, rid: int, indices, width: int, n_layers: int):
        idx = np.ascontiguousarray(indices, dtype=np.uint64).reshape(-1)
        vals = np.empty((idx.size, width), dtype=np.uint32)
        proof = np.empty((idx.size, n_layers - 1, 8), dtype=np.uint32)
        self._check(self.lib.p3h_serve_open(self.h, rid, _ptr(idx), idx.size, _ptr(vals), _ptr(proof)))
        return vals, proof

    def release(self, rid: int):
        self._check(self.lib.p3h_serve_release(self.h, rid))
        self._held.pop(rid, None)

    def stats(self):
        a, b, c, d = _u64(), _u64(), _u64(), _u64()
        self._check(self.lib.p3h_serve_stats(self.h, C.byref(a), C.byref(b), C.byref(c), C.byref(d)))
        return {"completed": a.value, "batches": b.value, "reserved_bytes": c.value, "committed_elems": d.value}

    def close(self):
        """Destroy the server: queued requests fail, running batches finish,
        calls in progress on other threads return first."""
        h = getattr(self, "h", None)
        if h:
            self.h = None
            self.lib.p3h_serve_destroy(h)
            self._held.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


# ----------------------------------------------------------- reference API --
class Radix2DitHip:
    """Drop-in for Radix2Dit<F> / NttCuda<F> on numpy matrices (host in, host out)."""

    def __init__(self, ctx: Context):
        self.ctx = ctx

    def _run(self, kind, mat, shift=1):
        m = self.ctx.upload(mat)
        try:
            if m.height <= 1:
                return np.array(mat, copy=True)
            out = self.ctx.dft(kind, m, shift)
            try:
                return out.download()
            finally:
                out.free()
        finally:
            m.free()

    def dft_batch(self, mat):
        return self._run(0, mat)

    def idft_batch(self, mat):
        return self._run(1, mat)

    def coset_dft_batch(self, mat, shift):
        return self._run(2, mat, shift)

    def coset_idft_batch(self, mat, shift):
        return self._run(3, mat, shift)

    def coset_lde_batch(self, mat, added_bits, shift):
        m = self.ctx.upload(mat)
        try:
            out = self.ctx.coset_lde(m, added_bits, shift)
            try:
                return out.download()
            finally:
                out.free()
        finally:
            m.free()


class MerkleTreeMmcsHip:
    """Drop-in for p3_merkle_tree::MerkleTreeMmcs (multi-height, cap)."""

    def __init__(self, ctx: Context, cap_height: int = 0):
        self.ctx, self.cap_height = ctx, cap_height

    def commit(self, matrices):
        mats = [self.ctx.upload(m) for m in matrices]
        tree = self.ctx.mmcs_commit(mats, self.cap_height)
        for m in mats:
            m.free()  # the tree keeps its own references
        return tree.cap(), tree

    def open_batch(self, index: int, tree: Tree):
        """(opened_values: one row per matrix in commit order, proof digests)."""
        vals, proof = tree.open([index])
        cuts = np.cumsum([w for _, w in tree.dims()])[:-1]
        return list(np.split(vals[0], cuts)), proof[0]

    def verify_batch(self, cap, dims, index: int, opened_values, proof) -> bool:
        """verify_batch (merkle_tree_mmcs.hpp:168-254); dims: (height, width) per
        matrix in commit order."""
        if len(dims) != len(opened_values) or any(len(v) != w for v, (_, w) in zip(opened_values, dims)):
            return False
        field = _field_of(np.asarray(cap).dtype)
        vals = np.concatenate([np.asarray(v).reshape(-1) for v in opened_values]) if dims else np.empty(0)
        return self.ctx.mmcs_verify(field, cap, dims, index, vals, proof)


class FlatMerkleTreeMmcsHip:
    """Drop-in for p3_merkle::MerkleTreeMmcs (single matrix, zero-padded odd layers)."""

    def __init__(self, ctx: Context):
        self.ctx = ctx

    def commit_matrix(self, data, width: int):
        arr = np.ascontiguousarray(data)
        if width == 0 or arr.size % width != 0:
            raise ValueError("commit_matrix: bad dimensions")
        if arr.size == 0:
            raise ValueError("commit_matrix: empty matrix")
        m = self.ctx.upload(arr.reshape(-1, width))
        tree = self.ctx.flat_commit(m)
        m.free()
        return tree.root(), tree

    def open_row(self, tree: Tree, row_index: int):
        vals, proof = tree.open([row_index])
        return vals[0], proof[0]

    def verify_row(self, commit, row_index: int, row_vals, proof) -> bool:
        """verify_row (fri_merkle_tree_mmcs.hpp:146-168)."""
        return self.ctx.flat_verify(_field_of(np.asarray(commit).dtype), commit, row_index, row_vals, proof)
