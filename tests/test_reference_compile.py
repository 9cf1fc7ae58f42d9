"""The reference's own PCS, FRI prover/verifier and MMCS call patterns compile
unchanged over this library's C++ adapters (tests/cpp/reference_pcs_over_adapters.cpp),
against the reference headers where they lie.  CPU only, compile only
(-fsyntax-only: every template the translation unit uses is instantiated);
skipped when the reference sources are not present."""

import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_INC = "/root/reference/src/include"
TU = os.path.join(ROOT, "tests", "cpp", "reference_pcs_over_adapters.cpp")


def _compile(adapter_dir):
    return subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I" + REF_INC, "-I" + os.path.join(ROOT, "include"),
                           "-I" + adapter_dir, TU], capture_output=True, text=True)


@pytest.mark.skipif(not os.path.isdir(REF_INC), reason="reference headers not present")
def test_reference_pcs_compiles_over_adapters():
    r = _compile(os.path.join(ROOT, "plonky3-gpu_amd", "cpp"))
    assert r.returncode == 0, r.stderr[-4000:]


@pytest.mark.skipif(not os.path.isdir(REF_INC), reason="reference headers not present")
def test_check_fails_without_get_row_as_vals(tmp_path):
    """The check has teeth: with HipFlatMmcs::get_row_as_vals renamed away,
    the reference's ExtensionMmcs::get_row (fri_extension_mmcs.hpp:96) no
    longer compiles."""
    src = open(os.path.join(ROOT, "plonky3-gpu_amd", "cpp", "p3hip.hpp")).read()
    broken = src.replace("get_row_as_vals(", "get_row_as_vals_renamed(")
    assert broken != src
    (tmp_path / "p3hip.hpp").write_text(broken)
    r = _compile(str(tmp_path))
    assert r.returncode != 0 and "get_row_as_vals" in r.stderr
