"""Writes the golden fixtures of tests/golden from the reference's own CPU build.

Runs oracle/_ref/phx_ref_fp{32,64} (built by oracle/Makefile from the reference's
sources) on the seeded synthetic snapshots and stores, per case, the FreqOut
array (I, Q, U, V, tau, tauF per pixel and frequency), the run statistics
(steps, plasma segments, dropped segments) and the snapshot's sha256.

    python tests/golden/make_golden.py            # all cases
    python tests/golden/make_golden.py c1_i10_128  # one case
"""
from __future__ import annotations

import json
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from phoenix_amd import snapshot as S  # noqa: E402
from phoenix_amd.configs import write_config  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
FREQS8 = "[" + ", ".join(repr(float(f)) for f in np.geomspace(86e9, 690e9, 8)) + "]"

# name -> (snapshot preset, config overrides as section__key: TOML text)
CASES = {
    "c1_i10_128": ("c1", {}),
    "c1_i60_48": ("c1", {"camera__inclination": "60", "camera__img_width": "48", "camera__img_height": "48"}),
    "c1_i163_48": ("c1", {"camera__inclination": "163", "camera__img_width": "48", "camera__img_height": "48"}),
    "c1_3freq_48": ("c1", {"camera__frequency_list": "[86e9, 230e9, 345e9]", "camera__img_width": "48",
                           "camera__img_height": "48"}),
    "c1_pa30_phi45_40x48": ("c1", {"camera__pa": "30", "camera__phi": "45", "camera__img_width": "40",
                                   "camera__img_height": "48", "camera__fov_x": "36"}),
    "c1_nlev1_32": ("c1", {"grmhd__nlevels": "1", "camera__img_width": "32", "camera__img_height": "32"}),
    "c1_rhigh40_i90_32": ("c1", {"plasma__r_high": "40", "camera__inclination": "90", "camera__img_width": "32",
                                 "camera__img_height": "32", "plasma__m_unit": "5e18"}),
    "c1a05_i10_48": ("c1_a05", {"camera__img_width": "48", "camera__img_height": "48"}),
    "c2_i10_64": ("c2", {"camera__img_width": "64", "camera__img_height": "64"}),
    # the precision workload's shape: 8 frequencies log-spaced over 86-690 GHz, both spins, 10 / 60 / 163 deg
    "c2_i163_8f_40": ("c2", {"camera__inclination": "163", "camera__img_width": "40", "camera__img_height": "40",
                             "camera__frequency_list": FREQS8}),
    "c2a05_i10_8f_40": ("c2_a05", {"camera__img_width": "40", "camera__img_height": "40",
                                   "camera__frequency_list": FREQS8}),
    "c2a05_i60_8f_40": ("c2_a05", {"camera__inclination": "60", "camera__img_width": "40",
                                   "camera__img_height": "40", "camera__frequency_list": FREQS8}),
    "c2a05_i163_8f_40": ("c2_a05", {"camera__inclination": "163", "camera__img_width": "40",
                                    "camera__img_height": "40", "camera__frequency_list": FREQS8}),
    # the serving snapshots: 16^3 blocks, 3 AMR levels, spins 0.9375 and 0.5; request shapes of the mix
    "c5_i10_48": ("c5", {"camera__img_width": "48", "camera__img_height": "48"}),
    "c5_i120_32_2f": ("c5", {"camera__inclination": "120", "camera__img_width": "32", "camera__img_height": "32",
                             "camera__frequency_list": "[150e9, 460e9]", "plasma__m_unit": "3e18",
                             "plasma__r_high": "25"}),
    "c5a05_i60_40_4f": ("c5_a05", {"camera__inclination": "60", "camera__img_width": "40",
                                   "camera__img_height": "40", "camera__frequency_list": "[86e9, 230e9, 345e9, 690e9]",
                                   "plasma__m_unit": "7e17", "plasma__r_high": "12"}),
    "c5a05_i170_32": ("c5_a05", {"camera__inclination": "170", "camera__img_width": "32",
                                 "camera__img_height": "32", "camera__frequency": "345e9"}),
    # base blocks in k, j, i order (grmhd.sfc = false), 8^3 blocks on a 6x3x3 base
    "c1sfc0_i30_40": ("c1_sfc0", {"camera__inclination": "30", "camera__img_width": "40", "camera__img_height": "40",
                                  "grmhd__sfc": "false"}),
}


def fixture_path(case: str, precision: int) -> str:
    return os.path.join(HERE, f"{case}_fp{precision}.npz")


def make(case: str, snapdir: str) -> None:
    preset, over = CASES[case]
    dump, grid, info = S.ensure_snapshot(preset, snapdir)
    with tempfile.TemporaryDirectory() as td:
        cfg = write_config(os.path.join(td, "c.toml"), dump, grid, **over)
        for prec in (64, 32):
            res, st, secs = O.run_ref(prec, cfg, os.path.join(td, f"o{prec}.bin"))
            np.savez_compressed(fixture_path(case, prec), res=res, stats=st,
                                meta=np.array(json.dumps({"case": case, "preset": preset, "overrides": over,
                                                          "sha256": info["sha256"], "precision": prec,
                                                          "seconds_8threads": secs})))
            print(case, prec, st.tolist(), f"{secs:.2f}s", os.path.getsize(fixture_path(case, prec)) // 1024, "KiB")


def load(case: str, precision: int):
    z = np.load(fixture_path(case, precision), allow_pickle=False)
    return z["res"], z["stats"], json.loads(str(z["meta"]))


if __name__ == "__main__":
    if not O.have_ref():
        O.build()
    names = sys.argv[1:] or list(CASES)
    for n in names:
        make(n, S.default_dir())
