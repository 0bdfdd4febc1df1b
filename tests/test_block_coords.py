"""The block-coordinate arithmetic every list-walking kernel uses (bid -> bx, by, bz
by multiplication with precomputed division constants) equals integer division,
checked on the host with the device header itself."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc is needed to compile the device header")
def test_block_coords_equal_integer_division(tmp_path):
    exe = tmp_path / "block_coords_check"
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(ROOT, "adaptive-pf-flip_amd", "csrc"),
                    os.path.join(ROOT, "tests", "host", "block_coords_check.cpp"), "-o", str(exe)],
                   check=True, capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert " 0 bad" in out.stdout
