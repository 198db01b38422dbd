"""The march's per-ray code (phoenix_amd/csrc/phx_lane.h, phx_march.h) built for the CPU
(tests/cpu/lane_host.cpp) and run under the address and undefined-behaviour sanitizers:

* the FP64 march's two memory forms (short-circuit box tests and interleaved gathers, or the
  batched forms the FP32-cell marches use) give bit-identical images and counts;
* the FP64 march against the reference's FP64 build (the golden fixtures) per pixel, and against
  the oracle on the same pixels for the step / segment / drop counts;
* the FP32 and mixed-precision marches against the reference's FP64 build with the parity gates;
* escaping rays ended at rt_outer change nothing, and with rt_outer inside the photon orbits the
  early end is not taken.

The kernel runs the same Lane code; this exercises it without a GPU, one lane at a time.
"""
import numpy as np
import pytest

from phoenix_amd import capi, configs, snapshot as S
from tests import parity
from tests.cpu import lane
from tests.golden import make_golden as G
from tests.test_oracle_golden import params_from_overrides

CASES = ["c1_i60_48", "c1_i163_48", "c1_nlev1_32", "c1sfc0_i30_40", "c2_i163_8f_40", "c5a05_i60_40_4f"]


@pytest.fixture(scope="module", autouse=True)
def built():
    lane.build()


_cache = {}


def _setup(case, snapdir, tmp_path, oracle_lib):
    if case not in _cache:
        preset, over = G.CASES[case]
        dump, grid, _ = S.ensure_snapshot(preset, snapdir)
        run = capi.Run(configs.write_config(str(tmp_path / f"{case}.toml"), dump, grid, **over))
        o = oracle_lib.Oracle(dump, grid, 64, **params_from_overrides(over))
        prim = o.tables()[0]
        _cache[case] = (run, o, prim)
    return _cache[case]


def _pixels(run, stride):
    n = run.model.img_width * run.model.img_height
    return np.arange(n // (2 * stride), n, stride, dtype=np.int32)


@pytest.mark.parametrize("case", CASES)
def test_fp64_forms_identical_and_match_reference(case, snapdir, tmp_path, oracle_lib):
    run, o, prim = _setup(case, snapdir, tmp_path, oracle_lib)
    pix = _pixels(run, 11)
    r0, s0 = lane.run(run.model, prim, capi.FP64, pix, batched=0, sanitized=True)
    r1, s1 = lane.run(run.model, prim, capi.FP64, pix, batched=1, sanitized=True)
    assert np.array_equal(r0, r1) and np.array_equal(s0, s1), "the FP64 memory forms differ"
    ref, _, _ = G.load(case, 64)
    ref = ref[pix]
    m = ref[..., 0] != 0
    rel = np.abs(r1[..., 0][m] - ref[..., 0][m]) / np.abs(ref[..., 0][m])
    quv = np.max(np.abs(r1[..., 1:4] - ref[..., 1:4])[m] / np.abs(ref[..., 0][m])[:, None], axis=1)
    print(f"\n{case}: {pix.size} pixels, FP64 lane vs reference FP64: max |dI|/I {rel.max():.2e}, "
          f"|dQUV|/I {quv.max():.2e}; counts {s1[:3].tolist()}")
    assert np.array_equal(ref[..., 0] == 0, r1[..., 0] == 0)
    assert rel.max() <= 1e-9 and quv.max() <= 1e-9
    _, st = o.render(pixels=pix.astype(np.int64), threads=8)
    assert np.all(np.abs(s1[:3] - st) <= 1), (s1[:3], st)


@pytest.mark.parametrize("precision", [capi.FP32, capi.MIXED])
@pytest.mark.parametrize("case", CASES)
def test_fp32_and_mixed_lanes_within_gates(case, precision, snapdir, tmp_path, oracle_lib):
    run, _, prim = _setup(case, snapdir, tmp_path, oracle_lib)
    pix = _pixels(run, 7)
    res, st = lane.run(run.model, prim, precision, pix, sanitized=True)
    ref, _, _ = G.load(case, 64)
    figs = parity.compare(res, ref[pix], preset=G.CASES[case][0])
    print(f"\n{case} fp{precision}: {parity.summary(figs[:2])}; counts {st.tolist()}")
    assert parity.passes(figs), parity.summary(figs)
    if precision == capi.MIXED:
        assert st[3] > 0  # steps near the horizon or the poles ran in FP64


@pytest.mark.parametrize("precision", [capi.FP32, capi.FP64, capi.MIXED])
def test_escape_changes_nothing_and_is_off_inside_the_photon_orbits(precision, snapdir, tmp_path, oracle_lib):
    run, _, prim = _setup("c1_i163_48", snapdir, tmp_path, oracle_lib)
    pix = _pixels(run, 5)
    full, s_full = lane.run(run.model, prim, precision, pix, escape=1)
    cut, s_cut = lane.run(run.model, prim, precision, pix, escape=2)
    assert np.array_equal(full, cut) and s_cut[0] < 0.9 * s_full[0]
    assert np.array_equal(s_full[1:3], s_cut[1:3])
    # rt_outer = 3 lies inside the retrograde photon orbit (3.93 at spin 0.9375): rays moving outward
    # there may still turn back, so escape 2 must march them on exactly as escape 1 does
    rq = run.request(rt_outer=3.0)
    full3, s3 = lane.run(rq, prim, precision, pix, escape=1)
    cut3, c3 = lane.run(rq, prim, precision, pix, escape=2)
    assert np.array_equal(full3, cut3) and np.array_equal(s3, c3)


@pytest.mark.parametrize("precision", [capi.MIXED, capi.FP32, capi.FP64])
def test_two_phase_equals_in_place_march(precision, snapdir, tmp_path, oracle_lib):
    """The two-phase render (the recording march, then every frequency folded over the records) gives
    the in-place march's results and counts bit for bit on the CPU, also when the record pool is too
    small for the rays (8 blocks: the render runs again with a larger pool), under the sanitizers."""
    case = "c2_i163_8f_40"
    run, _, prim = _setup(case, snapdir, tmp_path, oracle_lib)
    pix = _pixels(run, 13)
    r0, s0 = lane.run(run.model, prim, precision, pix, sanitized=True)
    r1, s1 = lane.run(run.model, prim, precision, pix, sanitized=True, two_phase=True)
    r2, s2 = lane.run(run.model, prim, precision, pix, sanitized=True, two_phase=True, env={"LANE_REC_BLOCKS": "8"})
    print(f"\n{case} fp{precision} nf={run.model.nfreqs}: counts {s0.tolist()}")
    assert np.array_equal(r0, r1) and np.array_equal(s0, s1)
    assert np.array_equal(r1, r2) and np.array_equal(s1, s2)


@pytest.mark.parametrize("precision", [capi.MIXED, capi.FP32, capi.FP64])
def test_sweep_equals_each_point_alone(precision, snapdir, tmp_path, oracle_lib):
    """A parameter sweep (the rays marched once and recorded without plasma units, every point's
    n_e, |B|, Theta_e, Theta_e cut and transfer from the records) gives each point's own in-place
    render bit for bit on the CPU, under the sanitizers."""
    case = "c2_i163_8f_40"
    run, _, prim = _setup(case, snapdir, tmp_path, oracle_lib)
    pix = _pixels(run, 17)
    points = [run.request(m_unit=3e18, r_high=20.0), run.request(m_unit=6e17, r_high=1.5),
              run.request(m_unit=1.2e18, r_high=40.0)]
    res, st = lane.run(run.model, prim, precision, pix, sanitized=True, points=points)
    for k, p in enumerate(points):
        alone, sa = lane.run(p, prim, precision, pix)
        print(f"\n{case} fp{precision} point {k}: identical {np.array_equal(res[k], alone)}")
        assert np.array_equal(res[k], alone), f"sweep point {k} differs from its own render"
    assert st[0] == sa[0]  # one march: the steps of one render
