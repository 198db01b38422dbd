"""Image comparison against the FP64 judge with the acceptance gates of the path:

* per pixel, on pixels whose |I| is above 1e-6 of the image peak:
  |I - I_ref| / |I_ref| <= 1e-3 and |X - X_ref| / |I_ref| <= 1e-3 for X in Q, U, V;
* total flux: |F_I - F_I,ref| / |F_I,ref| <= 1e-4 and |F_X - F_X,ref| / |F_I,ref| <= 1e-4.

A handful of photon-ring pixels are chaotic (a 1-ulp change flips a theta
turning point or the step on which a ray ends); the reference's own FP32 build
misses the per-pixel gate on 0-2 pixels per image against its FP64 build.  The
allowance for such pixels is max(2, 0.1% of the gated pixels).
"""
from __future__ import annotations

import math

import numpy as np

PIX_TOL = 1e-3
FLUX_TOL = 1e-4
FLOOR = 1e-6


def compare(res: np.ndarray, ref: np.ndarray, pix_tol=PIX_TOL, flux_tol=FLUX_TOL, floor=FLOOR) -> list:
    """res, ref: [npix, nf, 6].  Returns one dict of error figures per frequency."""
    out = []
    for f in range(ref.shape[1]):
        a, b = res[:, f, :4], ref[:, f, :4]
        fin = np.all(np.isfinite(a), axis=1) & np.all(np.isfinite(b), axis=1)
        I = b[:, 0]
        peak = np.max(np.abs(I[fin])) if fin.any() else 0.0
        m = fin & (np.abs(I) > floor * peak) if peak > 0 else fin & False
        with np.errstate(divide="ignore", invalid="ignore"):
            rel = np.abs(a[m, 0] - I[m]) / np.abs(I[m])
            quv = np.max(np.abs(a[m, 1:] - b[m, 1:]) / np.abs(I[m])[:, None], axis=1) if m.any() else np.zeros(0)
        nbad = int(np.sum((rel > pix_tol) | (quv > pix_tol)))
        Fa = a[np.all(np.isfinite(a), axis=1)].sum(0)
        Fb = b[np.all(np.isfinite(b), axis=1)].sum(0)
        dI = abs(Fa[0] - Fb[0]) / abs(Fb[0]) if Fb[0] != 0 else abs(Fa[0])
        dQUV = float(np.max(np.abs(Fa[1:] - Fb[1:]))) / abs(Fb[0]) if Fb[0] != 0 else float(np.max(np.abs(Fa[1:])))
        out.append({
            "f": f, "gated": int(m.sum()), "nbad": nbad, "allowed": max(2, math.ceil(1e-3 * int(m.sum()))),
            "rel_I_max": float(rel.max()) if rel.size else 0.0,
            "rel_I_p999": float(np.percentile(rel, 99.9)) if rel.size else 0.0,
            "quv_max": float(quv.max()) if quv.size else 0.0,
            "flux_dI": float(dI), "flux_dQUV": float(dQUV),
            "nonfinite": int((~np.all(np.isfinite(a), axis=1)).sum()),
            "zero_mismatch": int(np.sum((a[:, 0] == 0) != (b[:, 0] == 0))),
            "flux": Fa.tolist(), "flux_ref": Fb.tolist(),
        })
    return out


def passes(figs: list, flux_tol=FLUX_TOL) -> bool:
    return all(d["nbad"] <= d["allowed"] and d["flux_dI"] <= flux_tol and d["flux_dQUV"] <= flux_tol
               and d["nonfinite"] == 0 for d in figs)


def summary(figs: list) -> str:
    return "; ".join(
        f"f{d['f']}: gated {d['gated']} bad {d['nbad']}/{d['allowed']} relI max {d['rel_I_max']:.2e} "
        f"p99.9 {d['rel_I_p999']:.2e} quv {d['quv_max']:.2e} flux dI {d['flux_dI']:.1e} dQUV {d['flux_dQUV']:.1e} "
        f"zero-mismatch {d['zero_mismatch']}" for d in figs)
