"""Builds phoenix_amd/libphx.so (the C ABI of include/phx.h) in-tree.

hipcc compiles the device TU for gfx950 only; g++ compiles the host TU; hipcc
links both into one shared library next to this file.  Packed FP32 VALU ops are
disabled (``-Xclang -target-feature -Xclang -packed-fp32-ops``): some of their
forms have produced wrong results on the MI355X boxes this runs on.  FP32
division and square root use the hardware reciprocal / square-root sequences
(``-fno-hip-fp32-correctly-rounded-divide-sqrt``, ~1 ulp instead of correctly
rounded), and FP32 denormals are flushed to zero (``-fgpu-flush-denormals-to-zero``), which
lets a division be one reciprocal and one multiply; FP64 is unaffected by both.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libphx.so")
CLI = os.path.join(HERE, "phoenix")
BUILD = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = ["-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-Wall", "-Wno-unused-function",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-munsafe-fp-atomics",
             "-fno-hip-fp32-correctly-rounded-divide-sqrt", "-fgpu-flush-denormals-to-zero",
             f"-I{ROOT}/include"]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra", f"-I{ROOT}/include"]


def _run(cmd):
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode:
        sys.stderr.write(" ".join(cmd) + "\n" + p.stdout + p.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} (exit {p.returncode})")
    return p


def build(verbose: bool = False, out: str = LIB, extra_hip_flags=(), device_src: str | None = None) -> str:
    os.makedirs(BUILD, exist_ok=True)
    dev_o = os.path.join(BUILD, "phx_device.o")
    host_o = os.path.join(BUILD, "phx_host.o")
    _run([HIPCC, *HIP_FLAGS, *extra_hip_flags, "-c", device_src or os.path.join(CSRC, "phx_device.hip"), "-o", dev_o])
    _run([os.environ.get("CXX", "g++"), *CXX_FLAGS, "-c", os.path.join(CSRC, "phx_host.cpp"), "-o", host_o])
    tmp = out + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp, dev_o, host_o, "-ldl"])
    os.replace(tmp, out)
    if out == LIB:  # the command-line driver, linked against the library next to it
        _run([os.environ.get("CXX", "g++"), *CXX_FLAGS, os.path.join(CSRC, "phx_cli.cpp"), "-o", CLI + ".tmp",
              f"-L{HERE}", "-lphx", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib"])
        os.replace(CLI + ".tmp", CLI)
    if verbose:
        print("built", out)
    return out


def clean() -> None:
    shutil.rmtree(BUILD, ignore_errors=True)


if __name__ == "__main__":
    build(verbose=True)
