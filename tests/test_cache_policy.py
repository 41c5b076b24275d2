"""Reuse rules of the library's caching device allocator, checked on the host.

The bookkeeping lives in rte-rrtmgp-cpp_amd/csrc/cache_policy.h without HIP calls; this test
compiles tests/native/cache_policy_check.cpp against it with g++ and runs it.  It covers the
cross-stream case: a block freed on another stream than the one it was allocated on must come
back fenced, never as a plain same-stream reuse.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cache_policy(tmp_path):
    src = os.path.join(ROOT, "tests", "native", "cache_policy_check.cpp")
    exe = str(tmp_path / "cache_policy_check")
    inc = os.path.join(ROOT, "rte-rrtmgp-cpp_amd", "csrc")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-I", inc, src, "-o", exe],
                       capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail(r.stderr)
    out = subprocess.run([exe], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip() == "ok"
