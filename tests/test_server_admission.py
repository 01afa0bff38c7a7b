"""CPU tests of the assembly server's admission, eviction and cancellation
bookkeeping (finetools.jl_amd/csrc/sched.hpp, device-free): a C++ check
program compiled with the host compiler and run here."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "admission_check.cpp")
INC = os.path.join(ROOT, "finetools.jl_amd", "csrc")


def test_admission_refusal_eviction_and_cancellation(tmp_path):
    cxx = shutil.which("g++") or shutil.which("c++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    exe = str(tmp_path / "admission_check")
    subprocess.check_call([cxx, "-std=c++17", "-O1", "-Wall", "-I", INC, SRC, "-o", exe])
    out = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip() == "ok"
