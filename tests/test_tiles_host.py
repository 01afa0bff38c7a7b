"""CPU test of the host tile builder (finetools.jl_amd/csrc/tiles.cpp), the
reference that tests/test_gpu_tiles.py compares the device builder with byte
for byte: tests/cpp/tiles_check.cpp restates the Morton box-tree partition
with sets and checks every tile against the mesh (owned columns, halo lists,
block-task items and flags, each node block produced exactly once,
coordinates, layout) and that the bytes do not depend on the host thread
count.  Compiled with hipcc (the sources include the HIP runtime headers);
nothing runs on a device."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "finetools.jl_amd", "csrc")


def test_host_tile_builder(tmp_path):
    hipcc = shutil.which("hipcc") or ("/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else None)
    if hipcc is None:
        pytest.skip("no hipcc")
    exe = str(tmp_path / "tiles_check")
    srcs = [os.path.join(ROOT, "tests", "cpp", "tiles_check.cpp")] + \
        [os.path.join(CSRC, f) for f in ("tiles.cpp", "meshgen.cpp", "pool.cpp")]
    subprocess.check_call([hipcc, "-O2", "-std=c++17", "-I", CSRC, "-I", os.path.join(ROOT, "include")] + srcs +
                          ["-o", exe, "-lpthread"], cwd=str(tmp_path))
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr[-4000:]
    assert out.stdout.strip().splitlines()[-1] == "ok"
