"""Host-side checks of the sparse hierarchy the device V-cycle runs on (no GPU needed):
ELL packing keeps every column index inside the matrix (padding of the rectangular
prolongation included), ELL and CSR products agree, and the V-cycle's sequence of sparse
steps (tests/cpp/vcycle_check.cpp, replaying csr_vcycle with bounds-checked vectors) is a
fixed linear and symmetric operator, as PCG requires of its preconditioner."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "ho-cxx_amd", "csrc")


def test_vcycle_levels(tmp_path):
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    exe = tmp_path / "vcycle_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC, os.path.join(ROOT, "tests", "cpp", "vcycle_check.cpp"),
                    os.path.join(CSRC, "hoc_host.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(out.stdout)
    assert out.returncode == 0 and out.stdout.strip().endswith("OK"), out.stdout
