"""Sanitizer builds (CPU only): the oracle and the host-only planners of
libp3hip under AddressSanitizer + UndefinedBehaviorSanitizer.

* tests/cpp/oracle_sanitize.cpp drives every oracle entry point on small and
  edge-case inputs (g++ -fsanitize=address,undefined);
* tests/cpp/planner_sanitize.cpp drives plan_lde_batch (csrc/ntt_impl.h) and the
  serving footprint and admission rule (csrc/host_plan.h); hipcc applies the
  sanitizers to host code only, and the device code, built at -O0 to keep the
  build short, is never launched.

Any sanitizer report aborts the driver (-fno-sanitize-recover=all)."""

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all"]


def _run(exe):
    env = dict(os.environ, ASAN_OPTIONS="abort_on_error=1:detect_leaks=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), (r.stdout + r.stderr)[-4000:]


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not present")
def test_oracle_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "oracle_san")
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", *SAN, "-pthread", "-I" + os.path.join(ROOT, "oracle"),
                    os.path.join(ROOT, "tests", "cpp", "oracle_sanitize.cpp"), os.path.join(ROOT, "oracle", "oracle.cpp"),
                    "-o", exe], check=True)
    _run(exe)


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not present")
def test_host_planners_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "planner_san")
    host_san = [f for s in SAN for f in ("-Xarch_host", s)]
    subprocess.run([HIPCC, "-O0", "-g", "-std=c++17", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
                    "-packed-fp32-ops", *host_san, "-I" + os.path.join(ROOT, "include"),
                    "-I" + os.path.join(ROOT, "plonky3-gpu_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpp", "planner_sanitize.cpp"),
                    os.path.join(ROOT, "plonky3-gpu_amd", "csrc", "ntt_bb.hip"),
                    os.path.join(ROOT, "plonky3-gpu_amd", "csrc", "ntt_gl.hip"),
                    os.path.join(ROOT, "plonky3-gpu_amd", "csrc", "ntt_queue.hip"), "-o", exe], check=True)
    _run(exe)
