"""Build step for the native parts of vkdispatch_amd.

``build_native()`` compiles ``csrc/runtime.hip`` (the C-ABI runtime declared in
``include/vkdfft.h``) into ``_lib/libvkdfft.so`` with hipcc for gfx950.
FFT kernels themselves are instantiated per plan and compiled with hiprtc at
run time (see ``fft/kernels.py``); ``build()`` in ``__graft_entry__`` also
pre-compiles the kernels of the standard workloads into the on-disk kernel
cache so that a fresh GPU host does not pay the JIT latency.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
LIB_DIR = os.path.join(PKG_DIR, "_lib")
LIB_PATH = os.path.join(LIB_DIR, "libvkdfft.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = os.environ.get("VKX_OFFLOAD_ARCH", "gfx950")

# Packed FP32 (v_pk_*_f32) is disabled for every HIP translation unit.
HIP_FLAGS = [
    "-O3",
    "-std=c++17",
    f"--offload-arch={ARCH}",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    path = os.path.join(ROCM, "bin", "hipcc")
    if os.path.exists(path):
        return path
    found = shutil.which("hipcc")
    if not found:
        raise RuntimeError("hipcc not found; set ROCM_PATH")
    return found


def build_native(verbose: bool = False) -> str:
    """Compile the runtime library; returns the path of the built .so."""
    os.makedirs(LIB_DIR, exist_ok=True)
    srcs = [os.path.join(PKG_DIR, "csrc", f) for f in ("runtime.hip", "ragged.hip")]
    tmp = LIB_PATH + ".tmp"
    cmd = [hipcc(), *HIP_FLAGS, "-shared", "-I", os.path.join(REPO_DIR, "include"),
           "-I", os.path.join(PKG_DIR, "csrc"), *srcs,
           "-o", tmp, f"-L{ROCM}/lib", "-lhiprtc", f"-Wl,-rpath,{ROCM}/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed ({proc.returncode}):\n{proc.stdout}\n{proc.stderr}")
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build_native(verbose="-v" in sys.argv))
