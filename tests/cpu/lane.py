"""Runs the CPU build of the march's per-ray code (tests/cpu/lane_host.cpp) on a golden case.

The model and block tables come from the library's host layer (phx_load_config / grid / dump,
no GPU involved), the decoded primitives from the oracle's host decode (the reference's
convert2prim).  Test infrastructure only.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")
MAGIC = 0x50485831


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def binary(batched: int, sanitized: bool) -> str:
    return os.path.join(BUILD, f"lane_host{'_asan' if sanitized else ''}_b{batched}")


def write_input(path: str, model, prim: np.ndarray, precision: int, pix: np.ndarray, escape: int = 1,
                field_free: int = 1) -> None:
    from phoenix_amd import capi

    m = model
    with open(path, "wb") as f:
        f.write(np.array([MAGIC, ctypes.sizeof(capi.Model), precision, escape, pix.size, field_free],
                         dtype=np.int32).tobytes())
        f.write(bytes(m))
        f.write(np.ctypeslib.as_array(m.lb, shape=(m.nleafs * 3,)).astype(np.float64).tobytes())
        f.write(np.ctypeslib.as_array(m.dxc, shape=(m.nleafs * 3,)).astype(np.float64).tobytes())
        for L in range(m.nlevels):
            n = m.level_dims[L][0] * m.level_dims[L][1] * m.level_dims[L][2]
            f.write(np.ctypeslib.as_array(m.level_map[L], shape=(n,)).astype(np.int32).tobytes())
        f.write(np.ascontiguousarray(prim, dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(pix, dtype=np.int32).tobytes())


def run(model, prim: np.ndarray, precision: int, pix, batched: int = 1, sanitized: bool = False, escape: int = 1,
        field_free: int = 1, timeout: float = 900.0, two_phase: bool = False, env: dict | None = None, points=None):
    """Returns (res [npix, nf, 6], stats [steps, nplasma, bad, fp64_steps]).  two_phase: the recording
    march and the transfer of every frequency over the records instead of the in-place march.
    points: a parameter sweep through model's camera (models with their own units and the same
    frequency count); res is then [len(points), npix, nf, 6]."""
    pix = np.ascontiguousarray(pix, dtype=np.int32)
    with tempfile.TemporaryDirectory() as td:
        inp, out = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        write_input(inp, model, prim, precision, pix, escape, field_free)
        extra = ["two"] if two_phase else []
        if points is not None:
            pts = os.path.join(td, "points.bin")
            with open(pts, "wb") as f:
                f.write(np.array([len(points)], dtype=np.int32).tobytes())
                for p in points:
                    f.write(bytes(p))
            extra = ["sweep", pts]
        env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
                   UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", **(env or {}))
        p = subprocess.run([binary(batched, sanitized), inp, out] + extra, capture_output=True, text=True, env=env,
                           timeout=timeout)
        if p.returncode:
            raise RuntimeError(f"lane_host failed ({p.returncode}):\n{p.stderr[-4000:]}")
        with open(out, "rb") as f:
            st = np.frombuffer(f.read(32), dtype=np.int64).copy()
            res = np.frombuffer(f.read(), dtype=np.float64).copy()
    if points is not None:
        return res.reshape(len(points), pix.size, model.nfreqs, 6), st
    res = res.reshape(pix.size, model.nfreqs, 6)
    return res, st
