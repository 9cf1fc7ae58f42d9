"""The C++ host adapters (plonky3-gpu_amd/cpp/p3hip.hpp), driven like the
reference's callers drive Radix2Dit / MerkleTreeMmcs / fold_matrix_cuda,
checked against the oracle on the same seeded inputs."""

import os
import subprocess

import numpy as np
import pytest

import native

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "adapter_check")


def build_adapter_check():
    subprocess.run(["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
                    "-I" + os.path.join(ROOT, "plonky3-gpu_amd", "cpp"), os.path.join(ROOT, "tests", "cpp",
                                                                                     "adapter_check.cpp"),
                    "-L" + os.path.join(ROOT, "plonky3-gpu_amd"), "-lp3hip",
                    "-Wl,-rpath," + os.path.join(ROOT, "plonky3-gpu_amd"), "-o", BIN], check=True)


def test_cpp_adapters_match_oracle(orc):
    build_adapter_check()  # always: the adapters are header-only
    out = subprocess.run([BIN], check=True, capture_output=True, text=True, timeout=120).stdout
    lines = {l.split()[0]: l.split()[1:] for l in out.strip().splitlines()}
    consts = native.p2_consts(1, "bb", 8, 13)
    trace = native.seeded(5, 1024 * 8).reshape(1024, 8)
    lde = orc.lde("bb", trace, 1, 31)
    assert int(lines["lde_rows"][0]) == 2048
    assert [int(x) for x in lines["lde_row3"]] == [int(x) for x in lde[3]]
    d = orc.dft("bb", 0, native.seeded(6, 128).reshape(64, 2))
    assert [int(x) for x in lines["dft_row1"]] == [int(x) for x in d[1]]
    assert [int(x) for x in lines["flat_root"]] == [int(x) for x in orc.flat_commit("bb", consts, 8, 13, lde)]
    _, proof = orc.flat_open(consts, 8, 13, lde, 77)
    assert int(lines["flat_proof_len"][0]) == len(proof) and int(lines["flat_proof_len"][2]) == int(proof[0][0])
    cap = orc.mmcs("bb", consts, 8, 13, [lde, trace], 1)
    assert [int(x) for x in lines["mmcs_cap"]] == [int(x) for x in cap.reshape(-1)]
    # one opened row per matrix, in commit order, as the reference's open_batch
    assert lines["mmcs_open_len"] == ["2", "8", "8", "proof_len", "10"]
    # verify_batch: the honest opening passes; a changed value, a changed sibling
    # or another index fails
    assert lines["mmcs_verify"] == ["1", "0", "0", "0"]
    assert lines["flat_verify"] == ["1", "0"]
    cw = native.seeded(7, 64 * 4).reshape(64, 4)
    f = orc.fold(6, 1, np.array([3, 1, 4, 1], dtype=np.uint32), cw)
    assert [int(x) for x in lines["fold0"]] == [int(x) for x in f[0]]
    ops = [(native.OBSERVE, int(v)) for v in native.seeded(9, 11)]
    ops += [(native.OBSERVE, int(v)) for v in orc.flat_commit("bb", consts, 8, 13, lde)]
    ops += [(native.SAMPLE_EXT, 0), (native.GRIND, 6), (native.SAMPLE, 0)]
    want = orc.challenger_run(consts, 8, 13, ops)
    assert [int(x) for x in lines["chal_ext"]] == [int(x) for x in want[:4]]
    assert [int(x) for x in lines["chal_grind"][:2]] == [int(want[4]), int(want[5])]
    # the copy continues from the transcript as it was before grinding
    fresh = orc.challenger_run(consts, 8, 13, ops[:-2] + [(native.SAMPLE, 0)])
    assert int(lines["chal_grind"][2]) == int(fresh[4])
    assert "invalid_argument_raised" in lines
