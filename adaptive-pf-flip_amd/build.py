"""Build the HIP library of the MI355X PF-FLIP step (gfx950) in-tree.

Produces ``adaptive-pf-flip_amd/pfflip_amd/libpfflip_hip.so`` from
``csrc/pf_api.hip`` (one translation unit that includes the kernel files) and the
C++ host-class demo binary ``host/sim_demo`` when requested.

Usage: python adaptive-pf-flip_amd/build.py [--verbose]
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OUT_DIR = os.path.join(PKG, "pfflip_amd")
LIB = os.path.join(OUT_DIR, "libpfflip_hip.so")
HOST = os.path.join(PKG, "host")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
ARCH = os.environ.get("PF_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    f"--offload-arch={ARCH}",
    # keep packed FP32 instructions out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]


# translation units of the library (pf_api.hip includes the kernel files it launches)
LIB_SOURCES = ["pf_api.hip", "pf_noalloc.hip"]


def build_lib(verbose: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + HIP_FLAGS + [os.path.join(CSRC, f) for f in LIB_SOURCES] + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


def build_host_demo(verbose: bool = False) -> str:
    """C++17 drop-in host class demo, linked against the HIP library."""
    out = os.path.join(HOST, "sim_demo")
    cmd = ["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-I", HOST,
           os.path.join(HOST, "sim_demo.cpp"), os.path.join(HOST, "sparse_sim3d_tp_gpu.cpp"),
           os.path.join(HOST, "sparse_sim3d_gpu.cpp"), os.path.join(HOST, "multires_sim3d_tp_gpu.cpp"),
           os.path.join(HOST, "render_cache3d_gpu.cpp"),
           "-o", out, "-L", OUT_DIR, "-lpfflip_hip", f"-Wl,-rpath,{OUT_DIR}"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return out


def main() -> None:
    verbose = "--verbose" in sys.argv
    print(build_lib(verbose))
    if os.path.exists(os.path.join(HOST, "sim_demo.cpp")):
        print(build_host_demo(verbose))


if __name__ == "__main__":
    main()
