"""Host-side check of the projection basis restart (no GPU needed): tests/cpp/proj_restart_check.cpp
drives proj_restart_basis (ho-cxx_amd/csrc/hoc_host.cpp), which the solver calls when a full
basis restarts from the span of its newest solutions: the combinations it returns must be
orthonormal in the measured Gram inner product, span the newest solutions' coefficient columns,
give no weight to a vector that takes no part, and drop a repeated column."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "ho-cxx_amd", "csrc")


def test_projection_restart_basis(tmp_path):
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    exe = tmp_path / "proj_restart_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC, os.path.join(ROOT, "tests", "cpp", "proj_restart_check.cpp"),
                    os.path.join(CSRC, "hoc_host.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(out.stdout)
    assert out.returncode == 0 and out.stdout.strip().endswith("OK"), out.stdout
