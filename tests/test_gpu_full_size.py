"""The measured workloads at their full sizes, through properties that hold at any size:

* the headline image (c2 snapshot, 1024 x 1024, 230 GHz, i = 10 deg) in mixed precision and in
  plain FP32 against the FP64 march: the per-pixel and total-flux gates;
* the same image at the 8 frequencies of the precision form: two-phase equals in place bit for bit;
* the 16-point sweep of that image: every point equals its own render bit for bit;
* the 64-image batch of uneven size (13.9M rays): every image equals its own render bit for bit;
* two 2048 x 2048 x 8-frequency images of the precision form (c2 at i = 60 deg, c2_a05 at i = 163 deg)
  against the FP64 march;
* requests of the serving workload's kind served together: each equals its render alone.
"""
import os

import numpy as np
import pytest

from bench import FREQ, W, H, c4_requests
from phoenix_amd import capi, configs, snapshot as S
from tests import parity

pytestmark = pytest.mark.gpu
FREQS8 = tuple(float(f) for f in np.geomspace(86e9, 690e9, 8))


@pytest.fixture(scope="module")
def c2(snapdir, tmp_path_factory):
    dump, grid, _ = S.ensure_snapshot("c2", snapdir)
    cfg = configs.write_config(str(tmp_path_factory.mktemp("full") / "c2.toml"), dump, grid,
                               camera__img_width=str(W), camera__img_height=str(H), camera__frequency=str(FREQ),
                               camera__inclination="10", plasma__m_unit="1.2e18", plasma__r_high="3")
    run = capi.Run(cfg)
    snap = capi.Snapshot(run, capi.TABLE_FP32 | capi.TABLE_FP64)
    yield run, snap
    snap.close()
    run.close()


def test_headline_image_within_the_gates_of_fp64(c2):
    run, snap = c2
    ref64, _ = snap.render(None, capi.FP64)
    for name, prec, tune in (("mixed", capi.MIXED, None), ("plain fp32", capi.FP32, {"queue": 0, "refill": 64, "escape": 1})):
        res, info = snap.render(None, prec, tune=tune)
        figs = parity.compare(res, ref64, preset="c2")
        print(f"\n1024^2 {name}: {parity.summary(figs)}; kernel {info['kernel_s'] * 1e3:.1f} ms")
        assert parity.passes(figs), parity.summary(figs)


def test_headline_8_frequencies_two_phase_equals_in_place(c2):
    run, snap = c2
    rq = run.request(freqs=FREQS8)
    os.environ["PHX_TWO_PHASE_MIN_NF"] = "0"
    try:
        ref, iref = snap.render(rq, capi.MIXED)
    finally:
        os.environ.pop("PHX_TWO_PHASE_MIN_NF", None)
    res, info = snap.render(rq, capi.MIXED)
    print(f"\n1024^2 x 8: two-phase {info['kernel_s'] * 1e3:.1f} ms, in place {iref['kernel_s'] * 1e3:.1f} ms")
    assert np.array_equal(res, ref)
    assert (info["steps"], info["nplasma"], info["bad"]) == (iref["steps"], iref["nplasma"], iref["bad"])


def test_sweep_of_16_points_equals_each_point_alone(c2):
    run, snap = c2
    pts = [run.request(m_unit=float(m), r_high=float(r)) for m in np.geomspace(5e17, 5e18, 4)
           for r in np.linspace(1.0, 40.0, 4)]
    got, info = snap.render_batch(pts, capi.MIXED)
    assert info["rays"] == 16 * W * H
    for k, p in enumerate(pts):
        alone, _ = snap.render(p, capi.MIXED)
        assert np.array_equal(got[k], alone), f"sweep point {k} differs from its own render"


def test_c4_batch_equals_each_image_alone(c2):
    run, snap = c2
    reqs = c4_requests(run)
    got, info = snap.render_batch(reqs, capi.MIXED)
    assert info["rays"] == 13926400
    for k, rq in enumerate(reqs):
        alone, _ = snap.render(rq, capi.MIXED)
        assert np.array_equal(got[k], alone), f"batch image {k} ({rq.img_width}^2) differs from its own render"


def test_precision_form_2048_8_frequencies_within_the_gates_of_fp64(c2):
    run, snap = c2
    rq = run.request(img_width=2048, img_height=2048, inclination=60.0, freqs=FREQS8)
    ref64, _ = snap.render(rq, capi.FP64)
    res, info = snap.render(rq, capi.MIXED)
    figs = parity.compare(res, ref64, preset="c2")
    print(f"\n2048^2 x 8 mixed: {parity.summary(figs)}; kernel {info['kernel_s'] * 1e3:.1f} ms")
    assert parity.passes(figs), parity.summary(figs)


def test_precision_form_2048_spin_half_south_within_the_gates_of_fp64(snapdir, tmp_path):
    """The second C3 snapshot (spin 0.5) from below the equator (i = 163 deg), 2048 x 2048 x 8."""
    dump, grid, _ = S.ensure_snapshot("c2_a05", snapdir)
    run = capi.Run(configs.write_config(str(tmp_path / "a05.toml"), dump, grid))
    snap = capi.Snapshot(run, capi.TABLE_FP32 | capi.TABLE_FP64)
    try:
        rq = run.request(img_width=2048, img_height=2048, inclination=163.0, freqs=FREQS8)
        ref64, _ = snap.render(rq, capi.FP64)
        res, info = snap.render(rq, capi.MIXED)
        figs = parity.compare(res, ref64, preset="c2_a05")
        print(f"\n2048^2 x 8 spin 0.5 i=163 mixed: {parity.summary(figs[:2])}")
        assert parity.passes(figs), parity.summary(figs)
    finally:
        snap.close()
        run.close()


def test_served_c5_requests_equal_their_renders_alone(snapdir, tmp_path):
    """Requests of the serving workload's kind (the two 1.9M / 1.4M-cell c5 snapshots, 128^2 to
    1024^2 pixels, 1 to 4 frequencies, the units of each drawn at random) submitted together to one
    server: every result equals the same request rendered alone, bit for bit."""
    from bench import c5_request

    runs, snaps = [], []
    for name in ("c5", "c5_a05"):
        dump, grid, _ = S.ensure_snapshot(name, snapdir)
        r = capi.Run(configs.write_config(str(tmp_path / f"{name}.toml"), dump, grid))
        runs.append(r)
        snaps.append(capi.Snapshot(r, capi.TABLE_FP32))
    rng = np.random.default_rng(11)
    reqs = [(k % 2, c5_request(runs[k % 2], rng)) for k in range(8)]
    srv = capi.Server(snaps, capi.MIXED)
    try:
        got = [srv.wait(t)[0] for t in [srv.submit(k, rq) for k, rq in reqs]]
    finally:
        srv.stop()
    try:
        for (k, rq), g in zip(reqs, got):
            alone, _ = snaps[k].render(rq, capi.MIXED)
            print(f"\nsnapshot {k}: {rq.img_width}^2 x {rq.nfreqs}: served == alone {np.array_equal(g, alone)}")
            assert np.array_equal(g, alone)
    finally:
        for s in snaps:
            s.close()
        for r in runs:
            r.close()
