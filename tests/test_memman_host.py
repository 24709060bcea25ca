"""The device-wide memory manager's policies on the host (no GPU): its
source is built with host stand-ins for the HIP calls it makes and
AddressSanitizer, and a C++ driver checks budgets, cross-volume LRU
eviction, this frame's slots, slots being filled, chunk reclaim across slot
sizes, re-registration, the host brick cache and the worker pool."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HERE = os.path.join(ROOT, "tests", "memman_host")


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.isdir("/opt/rocm/include"),
                    reason="needs g++ and the ROCm headers")
def test_memman_policies(tmp_path):
    exe = str(tmp_path / "memman_policy_test")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-D__HIP_PLATFORM_AMD__",
           "-I/opt/rocm/include", os.path.join(HERE, "memman_policy_test.cpp"),
           os.path.join(ROOT, "imagevis3d_amd", "csrc", "memman.cpp"), os.path.join(HERE, "hip_stub.cpp"),
           "-lpthread", "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120,
                       env={**os.environ, "ASAN_OPTIONS": "detect_leaks=0"})
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "ALL OK" in r.stdout, r.stderr
