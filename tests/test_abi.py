"""The C-ABI library loads and exports every entry point include/p3hip.h declares
(CPU only: no device calls)."""

import ctypes
import os
import re

import p3hip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "p3hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(p3h_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "p3h_coset_lde" in names and "p3h_mmcs_commit" in names and "p3h_fri_commit_phase" in names
    assert len(names) >= 30


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(p3hip.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    assert sorted(p3hip.exported_symbols()) == declared()


def test_host_alloc_rejects_bad_arguments_before_any_device_call():
    """Argument checks of the pinned-buffer calls return status codes before any
    HIP call, so they run without a device."""
    lib = p3hip.load_library()
    out = ctypes.c_void_p(12345)
    assert lib.p3h_host_alloc(0, ctypes.byref(out)) == 1  # P3H_EINVAL
    assert out.value is None  # cleared on failure
    assert lib.p3h_host_alloc(64, None) == 1
    assert lib.p3h_host_free(None) == 0  # freeing nothing is a no-op


def test_seeded_fill_matches_the_oracle_stream():
    """p3h_seeded_fill is host-only: the benchmark's synthetic inputs equal the
    oracle's std::mt19937_64 stream element for element."""
    import native
    import numpy as np

    for field, name in ((p3hip.BABYBEAR, "bb"), (p3hip.GOLDILOCKS, "gl")):
        got = p3hip.seeded_fill(field, 0x2A, 5000)
        assert (got == native.seeded(0x2A, 5000, name)).all()
    lib = p3hip.load_library()
    assert lib.p3h_seeded_fill(7, 1, None, 0) == 1  # unknown field: P3H_EINVAL
    assert lib.p3h_seeded_fill(0, 1, None, 4) == 1  # null output
    assert np.asarray(p3hip.seeded_fill(p3hip.BABYBEAR, 1, 0)).size == 0
