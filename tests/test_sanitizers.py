"""The CPU oracle and the host build of the device SA math run clean under
AddressSanitizer and UndefinedBehaviorSanitizer (small drivers in
tests/native/sanitize_*.{c,cpp}: a whole oracle solve + post-processing + the
region hook, and the node/region math of ebem_sa.hpp over random pairs)."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

SAN = ["-fsanitize=address", "-fsanitize=undefined", "-fno-sanitize-recover=undefined"]


def _run(exe):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    env.pop("LD_PRELOAD", None)
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
    return r.stdout


def test_oracle_under_asan_ubsan(tmp_path):
    gcc = shutil.which("gcc")
    if not gcc:
        pytest.skip("gcc not available")
    exe = str(tmp_path / "san_oracle")
    subprocess.run([gcc, "-O1", "-g", "-fno-omit-frame-pointer", "-std=c11", "-ffp-contract=off",
                    "-Wno-unknown-pragmas", *SAN, "-o", exe,
                    os.path.join(ROOT, "tests", "native", "sanitize_oracle.c"),
                    os.path.join(ROOT, "oracle", "ebem_oracle.c"), "-lm"], check=True)
    assert "oracle ok" in _run(exe)


def test_device_math_host_build_under_asan_ubsan(tmp_path):
    hipcc = "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    exe = str(tmp_path / "san_sa")
    # the sanitizers apply to the host compilation only
    host_san = [f for s in SAN for f in ("-Xarch_host", s)]
    subprocess.run([hipcc, "--offload-arch=gfx950", "--offload-host-only", "-O1", "-g", "-std=c++17",
                    "-fno-omit-frame-pointer", "-ffp-contract=fast-honor-pragmas", *host_san, "-o", exe,
                    os.path.join(ROOT, "tests", "native", "sanitize_sa.cpp")], check=True)
    assert "device math ok" in _run(exe)
