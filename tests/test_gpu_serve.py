"""The serving loop (phx_server_*): requests that arrive while others are on the GPU share its
launches (continuous batching), and each result equals the same request rendered alone.

A request's rays go through the same per-ray code in or out of the mix, and launches take whole
waves of 64 of a request's jobs, so that a wave marches the same rays as in the alone render (in
mixed precision they share their FP64 steps).  A result is then bit-identical to the alone render
whenever the launches that marched it used the same kernel instantiation (all requests of the mix
have one frequency chunk of the same width, or render in two phases); otherwise it is held to the
parity gates.
"""
import threading
import time

import numpy as np
import pytest

from phoenix_amd import capi, configs, snapshot as S
from tests import parity

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env(snapdir, tmp_path_factory):
    dump, grid, _ = S.ensure_snapshot("c1", snapdir)
    d2, g2, _ = S.ensure_snapshot("c1_a05", snapdir)
    td = tmp_path_factory.mktemp("serve")
    run = capi.Run(configs.write_config(str(td / "a.toml"), dump, grid))
    run2 = capi.Run(configs.write_config(str(td / "b.toml"), d2, g2))
    snaps = [capi.Snapshot(run, capi.TABLE_FP32 | capi.TABLE_FP64), capi.Snapshot(run2, capi.TABLE_FP32)]
    yield run, run2, snaps
    for s in snaps:
        s.close()
    run.close()
    run2.close()


def _rq(run, w, h, freqs, incl, m_unit=1.2e18, r_high=3.0):
    return run.request(img_width=w, img_height=h, freqs=freqs, inclination=incl, m_unit=m_unit, r_high=r_high)


@pytest.mark.parametrize("order", [0, 1])
def test_mix_equals_alone_bitwise(env, order):
    """Oldest-request-first (order 0) and least-attained-service (order 1) job selection."""
    run, run2, snaps = env
    reqs = [(0, _rq(run, 16, 16, (230e9,), 10.0)), (1, _rq(run2, 24, 16, (230e9,), 60.0, 3e18, 20.0)),
            (0, _rq(run, 32, 32, (345e9,), 163.0, 5e17, 40.0)), (1, _rq(run2, 8, 8, (86e9,), 120.0))]
    # several launches per request, requests overlap
    srv = capi.Server(snaps, capi.FP32, quantum_jobs=300, order=order)
    tickets = [srv.submit(k, r) for k, r in reqs]
    outs = [srv.wait(t) for t in tickets]
    st = srv.stats()
    srv.stop()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["launches"] > len(reqs)
    assert 0 < st["p50_s"] <= st["p99_s"] <= st["max_s"]
    for (k, r), (res, lat) in zip(reqs, outs):
        alone, _ = snaps[k].render(r, capi.FP32)
        assert lat > 0
        assert np.array_equal(res, alone), f"request {r.img_width}x{r.img_height} on snapshot {k}: mix != alone"


@pytest.mark.parametrize("precision", [capi.FP32, capi.MIXED])
def test_mix_of_frequency_counts_within_gates(env, precision):
    run, _, snaps = env
    reqs = [_rq(run, 16, 16, (230e9, 345e9, 86e9), 30.0), _rq(run, 20, 12, (230e9,), 80.0, 2e18, 10.0),
            _rq(run, 8, 8, tuple(np.geomspace(86e9, 690e9, 6)), 150.0)]
    srv = capi.Server(snaps[:1], precision, quantum_jobs=128)
    tickets = [srv.submit(0, r) for r in reqs]
    outs = [srv.wait(t)[0] for t in tickets]
    srv.stop()
    for r, res in zip(reqs, outs):
        alone, _ = snaps[0].render(r, precision)
        figs = parity.compare(res, alone, preset="c1")
        print(f"\n{r.img_width}x{r.img_height} nf={r.nfreqs}: exact={np.array_equal(res, alone)}; "
              f"{parity.summary(figs)}")
        assert parity.passes(figs), parity.summary(figs)


def test_budget_admission_and_threaded_submitters(env):
    run, _, snaps = env
    one = 16 * 16 * 48  # bytes of one 16x16 single-frequency result
    srv = capi.Server(snaps[:1], capi.FP32, quantum_jobs=64, hbm_budget=2 * one)
    reqs = [_rq(run, 16, 16, (230e9,), 10.0 + 15.0 * i) for i in range(6)]
    tickets = [None] * len(reqs)

    def submit(i):
        tickets[i] = srv.submit(0, reqs[i])

    th = [threading.Thread(target=submit, args=(i,)) for i in range(len(reqs))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    outs = [srv.wait(t)[0] for t in tickets]
    st = srv.stats()
    srv.stop()
    assert st["completed"] == len(reqs)
    assert st["peak_bytes"] <= 2 * one  # never more than the budget admitted at once
    for r, res in zip(reqs, outs):
        alone, _ = snaps[0].render(r, capi.FP32)
        assert np.array_equal(res, alone)


def test_stop_fails_unfinished_requests_and_wakes_waiters(env):
    """phx_server_stop with requests still queued: they fail with "server stopped", a thread blocked
    in phx_server_wait on one of them returns (with the error), and the stop returns after it."""
    run, _, snaps = env
    srv = capi.Server(snaps[:1], capi.FP32, quantum_jobs=64, max_active=1)
    tickets = [srv.submit(0, _rq(run, 48, 48, (230e9,), 10.0 + 20.0 * i)) for i in range(6)]
    out = {}

    def waiter():
        try:
            out["res"] = srv.wait(tickets[-1])
        except capi.PhxError as e:
            out["err"] = str(e)

    th = threading.Thread(target=waiter)
    th.start()
    time.sleep(0.5)  # the waiter is inside phx_server_wait
    srv.stop()
    th.join(timeout=60)
    assert not th.is_alive()
    assert "server stopped" in out.get("err", "") or out.get("res") is not None


def test_server_rejects_bad_requests(env):
    run, _, snaps = env
    srv = capi.Server(snaps[:1], capi.FP32, hbm_budget=1000)
    with pytest.raises(capi.PhxError):
        srv.submit(0, _rq(run, 64, 64, (230e9,), 10.0))  # larger than the budget
    with pytest.raises(capi.PhxError):
        srv.submit(3, _rq(run, 4, 4, (230e9,), 10.0))  # no such snapshot
    srv.stop()


def test_device_flux_matches_compute_spec(env):
    """The flux reduced on the device (phx_snapshot_flux, and a served request's spectrum with no
    image copied back) against the host's phx_compute_spec of the same image, to 1e-12 of I."""
    run, _, snaps = env
    rq = _rq(run, 40, 32, (86e9, 230e9, 690e9), 60.0, 2e18, 10.0)
    res, _ = snaps[0].render(rq, capi.FP32)
    host = run.flux(res, rq)
    dev = snaps[0].flux(rq.nfreqs)
    print(f"\ndevice flux {dev[:, 0]} Jy, host {host[:, 0]} Jy")
    assert np.all(np.abs(dev - host) <= 1e-12 * np.abs(host[:, :1]))
    srv = capi.Server(snaps[:1], capi.FP32, quantum_jobs=500)
    t_img = srv.submit(0, rq)
    t_spec = srv.submit(0, rq, image=False, spec=True)
    img, _ = srv.wait(t_img)
    none, spec, _ = srv.wait_spec(t_spec)
    srv.stop()
    assert none is None and np.array_equal(img, res)
    assert np.all(np.abs(spec - host) <= 1e-12 * np.abs(host[:, :1]))


@pytest.mark.parametrize("precision", [capi.MIXED, capi.FP32])
def test_served_sweep_points_share_the_march(env, precision):
    """Sweep points submitted together (one camera; m_unit, R_high and frequencies differ) are
    marched once per launch and folded point by point; each result equals the point rendered alone
    bit for bit, beside a request through another camera, and with several launches per request."""
    run, run2, snaps = env
    cam = dict(img_width=32, img_height=24, inclination=45.0)
    points = [run.request(**cam, freqs=(230e9,)),
              run.request(**cam, freqs=(86e9, 345e9), m_unit=3e18, r_high=20.0),
              run.request(**cam, freqs=(230e9,), m_unit=6e17, r_high=1.5),
              run.request(**cam, freqs=tuple(np.geomspace(86e9, 690e9, 3)), m_unit=2e18, r_high=40.0)]
    other = _rq(run, 24, 24, (230e9,), 120.0, 5e17, 8.0)
    reqs = [(0, p) for p in points] + [(0, other), (1, _rq(run2, 16, 16, (86e9, 230e9), 70.0))]
    srv = capi.Server(snaps, precision, quantum_jobs=250)
    tickets = [srv.submit(k, r) for k, r in reqs]
    outs = [srv.wait(t) for t in tickets]
    st = srv.stats()
    srv.stop()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    # the four points need 4 launches of 250 of their 768 pixels together, not 4 each
    print(f"\nfp{precision}: {st['launches']} launches")
    assert st["launches"] < 4 * 4
    for (k, r), (res, lat) in zip(reqs, outs):
        alone, _ = snaps[k].render(r, precision)
        assert np.array_equal(res, alone), f"request {r.img_width}x{r.img_height} nf={r.nfreqs}: served != alone"
