"""ctypes wrapper of the oracle (liboracle.so, the CPU restatement in phx_oracle.cpp).

Test infrastructure only: imported by tests/, by ``__graft_entry__.smoke()`` and by
``bench.py``'s CPU-baseline leg, always as the checker.  The product (phoenix_amd)
never imports this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")
MAXF = 32


class Params(ctypes.Structure):
    _fields_ = [
        ("inclination", ctypes.c_double), ("fov_x", ctypes.c_double), ("fov_y", ctypes.c_double),
        ("r_cam", ctypes.c_double), ("t_init", ctypes.c_double), ("pa", ctypes.c_double),
        ("phi_cam", ctypes.c_double),
        ("img_width", ctypes.c_int), ("img_height", ctypes.c_int), ("nfreqs", ctypes.c_int),
        ("freqs", ctypes.c_double * MAXF),
        ("mass", ctypes.c_double), ("distance", ctypes.c_double),
        ("stepsize", ctypes.c_double), ("max_steps", ctypes.c_int),
        ("horizon_margin", ctypes.c_double), ("max_order", ctypes.c_int),
        ("m_unit", ctypes.c_double), ("r_low", ctypes.c_double), ("r_high", ctypes.c_double),
        ("beta_c", ctypes.c_double), ("sigma_cut", ctypes.c_double), ("thetae_min", ctypes.c_double),
        ("thetae_max", ctypes.c_double), ("rt_outer", ctypes.c_double),
        ("nspin", ctypes.c_int), ("bpol", ctypes.c_int), ("nlevels", ctypes.c_int), ("sfc", ctypes.c_int),
    ]


DEFAULTS = dict(inclination=10.0, fov_x=40.0, fov_y=40.0, r_cam=1e4, t_init=0.0, pa=0.0, phi_cam=0.0,
                img_width=128, img_height=128, freqs=(230e9,), mass=4.14e6, distance=8.127,
                stepsize=0.01, max_steps=10000, horizon_margin=0.01, max_order=100, m_unit=1.2e18,
                r_low=1.0, r_high=3.0, beta_c=1.0, sigma_cut=1.0, thetae_min=1e-3, thetae_max=100.0,
                rt_outer=50.0, nspin=3, bpol=1, nlevels=-1, sfc=1)


def make_params(**kw) -> Params:
    d = dict(DEFAULTS)
    d.update(kw)
    p = Params()
    for k, v in d.items():
        if k == "freqs":
            p.nfreqs = len(v)
            for i, f in enumerate(v):
                p.freqs[i] = f
        else:
            setattr(p, k, v)
    return p


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)
    if os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-s", "-C", HERE, "ref", "ref_output"], check=True)
        if os.path.exists(os.path.join(HERE, "..", "phoenix_amd", "libphx.so")):
            subprocess.run(["make", "-s", "-C", HERE, "ref_main_phx"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        dp = ctypes.POINTER(ctypes.c_double)
        L.orc_load.restype = ctypes.c_void_p
        L.orc_load.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(Params), ctypes.c_int]
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_free.argtypes = [ctypes.c_void_p]
        L.orc_render.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_long), ctypes.c_long, dp,
                                 ctypes.POINTER(ctypes.c_longlong), ctypes.c_int]
        L.orc_counts.argtypes = [ctypes.POINTER(ctypes.c_longlong)]
        L.orc_model_info.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), dp]
        L.orc_tables.argtypes = [ctypes.c_void_p, dp, dp, dp, ctypes.POINTER(ctypes.c_int),
                                 ctypes.POINTER(ctypes.c_int)]
        L.orc_find_leaf.argtypes = [ctypes.c_void_p, dp, ctypes.c_long, ctypes.POINTER(ctypes.c_int),
                                    ctypes.POINTER(ctypes.c_int)]
        L.orc_fluid.argtypes = [ctypes.c_void_p, dp, ctypes.c_long, dp]
        L.orc_metric.argtypes = [ctypes.c_double, ctypes.c_double, dp, dp, dp]
        L.orc_metric_split.argtypes = [ctypes.c_double, ctypes.c_double, dp, dp, dp]
        L.orc_christoffel.argtypes = [ctypes.c_double, ctypes.c_double, dp, dp, dp, dp]
        L.orc_bessel_ratios.argtypes = [dp, ctypes.c_long, dp, dp, ctypes.c_int]
        L.orc_bessel_scaled.argtypes = [ctypes.c_double, dp]
        L.orc_thermal.argtypes = [dp, dp]
        L.orc_invariant.argtypes = [dp, ctypes.c_double, ctypes.c_double, dp]
        L.orc_segment_op.argtypes = [dp, ctypes.c_double, dp, dp]
        L.orc_invert4.argtypes = [dp, dp]
        L.orc_camera.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.POINTER(Params), ctypes.c_double,
                                 ctypes.c_double, dp, dp]
        L.orc_geodesic.argtypes = [ctypes.c_double, ctypes.c_double, dp, ctypes.c_double, ctypes.c_int, dp,
                                   ctypes.c_int]
        L.orc_render_ham.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_long), ctypes.c_long, dp,
                                     ctypes.POINTER(ctypes.c_longlong), ctypes.c_int]
        L.orc_wp_kappa.argtypes = [ctypes.c_double, ctypes.c_double, dp, dp, dp, dp]
        L.orc_ham_rhs.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, dp, dp, dp]
        L.orc_step_len.restype = ctypes.c_double
        L.orc_step_len.argtypes = [ctypes.c_double, dp, dp, dp]
        _lib = L
    return _lib


def ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def iptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


class Oracle:
    """A loaded snapshot + run parameters at one precision (32, 64, 132, 164)."""

    def __init__(self, dump: str, grid: str, precision: int = 64, **params):
        self.params = make_params(**params)
        self.precision = precision
        self.h = lib().orc_load(dump.encode(), grid.encode(), ctypes.byref(self.params), precision)
        if not self.h:
            raise RuntimeError(lib().orc_last_error().decode())
        self.W, self.H, self.nf = self.params.img_width, self.params.img_height, self.params.nfreqs

    def close(self):
        if self.h:
            lib().orc_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def render(self, pixels=None, threads: int = 0):
        """Returns (res [npix, nf, 6], stats [steps, nplasma, bad])."""
        if pixels is None:
            pixels = np.arange(self.W * self.H, dtype=np.int64)
        pixels = np.ascontiguousarray(pixels, dtype=np.int64)
        res = np.zeros((pixels.size, self.nf, 6))
        st = (ctypes.c_longlong * 3)()
        lib().orc_render(self.h, pixels.ctypes.data_as(ctypes.POINTER(ctypes.c_long)), pixels.size, ptr(res), st,
                         threads)
        return res, np.array(list(st), dtype=np.int64)

    def render_ham(self, pixels=None, threads: int = 0):
        """The Hamiltonian twin (ham_twin.cpp: Hamiltonian geodesic, Walker-Penrose polarization)
        on a handle of precision 64 or 164.  Returns (res [npix, nf, 6], stats)."""
        if pixels is None:
            pixels = np.arange(self.W * self.H, dtype=np.int64)
        pixels = np.ascontiguousarray(pixels, dtype=np.int64)
        res = np.zeros((pixels.size, self.nf, 6))
        st = (ctypes.c_longlong * 3)()
        rc = lib().orc_render_ham(self.h, pixels.ctypes.data_as(ctypes.POINTER(ctypes.c_long)), pixels.size,
                                  ptr(res), st, threads)
        if rc:
            raise RuntimeError("the Hamiltonian twin runs at precision 64 or 164")
        return res, np.array(list(st), dtype=np.int64)

    def info(self):
        iv = np.zeros(9, dtype=np.int32)
        dv = np.zeros(7)
        lib().orc_model_info(self.h, iptr(iv), ptr(dv))
        return {"nleafs": int(iv[0]), "cells": int(iv[1]), "nlevels": int(iv[2]), "nx": iv[3:6].tolist(),
                "ng": iv[6:9].tolist(), "a": dv[0], "hslope": dv[1], "L": dv[2], "cut_in": dv[3],
                "r_out": dv[4], "B_unit": dv[5], "Ne_unit": dv[6]}

    def tables(self):
        i = self.info()
        n, c = i["nleafs"], i["cells"]
        prim = np.zeros((n, c, 8))
        lb = np.zeros((n, 3))
        dxc = np.zeros((n, 3))
        lev = np.zeros(n, dtype=np.int32)
        ind = np.zeros((n, 3), dtype=np.int32)
        lib().orc_tables(self.h, ptr(prim), ptr(lb), ptr(dxc), iptr(lev), iptr(ind))
        return prim, lb, dxc, lev, ind

    def find_leaf(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 4)
        leaf = np.zeros(len(x), dtype=np.int32)
        cand = np.zeros(len(x), dtype=np.int32)
        lib().orc_find_leaf(self.h, ptr(x), len(x), iptr(leaf), iptr(cand))
        return leaf, cand

    def fluid(self, X):
        X = np.ascontiguousarray(X, dtype=np.float64).reshape(-1, 4)
        out = np.zeros((len(X), 14))
        lib().orc_fluid(self.h, ptr(X), len(X), ptr(out))
        return out


def wp_kappa(a: float, hs: float, X, k, f) -> np.ndarray:
    """Walker-Penrose constant (kappa_1, kappa_2) of (k, f) at X (MKS-BHAC Kerr-Schild components)."""
    X, k, f = (np.ascontiguousarray(v, dtype=np.float64) for v in (X, k, f))
    q = np.zeros(2)
    lib().orc_wp_kappa(a, hs, ptr(X), ptr(k), ptr(f), ptr(q))
    return q


def ham_rhs(a: float, hs: float, pt: float, pp: float, y):
    """Hamilton's equations at y = (x1, x2, x3, p_1, p_2): returns (dy/dlambda, k^mu)."""
    y = np.ascontiguousarray(y, dtype=np.float64)
    d, k = np.zeros(5), np.zeros(4)
    lib().orc_ham_rhs(a, hs, pt, pp, ptr(y), ptr(d), ptr(k))
    return d, k


def op_counts() -> np.ndarray:
    out = (ctypes.c_longlong * 24)()
    lib().orc_counts(out)
    return np.array(list(out), dtype=np.int64).reshape(4, 6)


def op_counts_reset() -> None:
    lib().orc_counts_reset()


# ----------------------------------------------------------------- reference harness
def ref_binary(precision: int) -> str:
    return os.path.join(REF_DIR, f"phx_ref_fp{precision}")


def have_ref() -> bool:
    return os.path.exists(ref_binary(32)) and os.path.exists(ref_binary(64))


def run_ref(precision: int, config: str, out: str, pixels=None, threads: int = 8, overrides=()):
    """Runs the reference CPU build (oracle/_ref) and returns (res [W*H, nf, 6], stats, seconds)."""
    cmd = [ref_binary(precision), config, out]
    if pixels is not None:
        cmd += ["--pixels", ",".join(str(int(v)) for v in pixels)]
    cmd += list(overrides)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    p = subprocess.run(cmd, env=env, capture_output=True, text=True)
    if p.returncode:
        raise RuntimeError(p.stderr)
    return read_ref_output(out)


def read_ref_output(path: str):
    with open(path, "rb") as f:
        hdr = np.frombuffer(f.read(16), dtype=np.int32)
        st = np.frombuffer(f.read(24), dtype=np.int64)
        secs = float(np.frombuffer(f.read(8), dtype=np.float64)[0])
        W, H, nf = int(hdr[0]), int(hdr[1]), int(hdr[2])
        res = np.frombuffer(f.read(), dtype=np.float64).reshape(W * H, nf, 6)
    return res, st.copy(), secs
