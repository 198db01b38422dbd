"""The oracle (oracle/phx_oracle.cpp) against the reference's own CPU build.

The fixtures in tests/golden were written by tests/golden/make_golden.py from
oracle/_ref (the reference's sources compiled as they lie).  On the same
snapshot bytes the oracle must agree bit for bit at both precisions, including
the step / plasma-segment / dropped-segment counts.
"""
import numpy as np
import pytest

from tests.golden import make_golden as G

CASES = sorted(G.CASES)


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("case", CASES)
def test_oracle_matches_reference(case, precision, snapdir, oracle_lib):
    O = oracle_lib
    res_ref, st_ref, meta = G.load(case, precision)
    preset, over = G.CASES[case]
    from phoenix_amd import snapshot as S
    dump, grid, info = S.ensure_snapshot(preset, snapdir)
    params = params_from_overrides(over)
    o = O.Oracle(dump, grid, precision, **params)
    res, st = o.render(threads=8)
    o.close()
    assert res.shape == res_ref.shape
    if info["sha256"] != meta["sha256"]:
        pytest.skip("snapshot bytes differ from the fixture's (different generator numerics)")
    assert st.tolist() == st_ref.tolist()
    same = np.all((res == res_ref) | (np.isnan(res) & np.isnan(res_ref)), axis=(1, 2))
    assert same.all(), f"{(~same).sum()} pixels differ"


def params_from_overrides(over):
    """Oracle keyword parameters for a fixture's config overrides."""
    p = {}
    names = {"inclination": "inclination", "img_width": "img_width", "img_height": "img_height",
             "fov_x": "fov_x", "fov_y": "fov_y", "pa": "pa", "phi": "phi_cam", "r_high": "r_high",
             "m_unit": "m_unit", "nlevels": "nlevels"}
    for k, v in over.items():
        sec, key = k.split("__", 1)
        if key == "frequency_list":
            p["freqs"] = tuple(float(x) for x in v.strip("[]").split(","))
        elif key == "frequency":
            p["freqs"] = (float(v),)
        elif key == "sfc":
            p["sfc"] = int(v == "true")
        elif key in ("img_width", "img_height", "nlevels"):
            p[names[key]] = int(float(v))
        else:
            p[names[key]] = float(v)
    return p
