"""bench.py's host-side helpers: the ideal queue models its serving report compares with, and the
lower bound on the p99 latency of any schedule."""
import itertools

import numpy as np

from bench import p99_lower_bound, queue_model


def test_fifo_closed_forms():
    # simultaneous arrivals: each waits for all before it
    np.testing.assert_allclose(queue_model(np.zeros(4), np.array([1.0, 2.0, 3.0, 4.0]), "fifo"), [1, 3, 6, 10])
    # arrivals spaced wider than the service: no waiting
    np.testing.assert_allclose(queue_model(np.array([0.0, 5.0, 10.0]), np.array([1.0, 2.0, 3.0]), "fifo"), [1, 2, 3])


def test_processor_sharing_closed_forms():
    # n equal jobs arriving together all finish at n * s
    np.testing.assert_allclose(queue_model(np.zeros(3), np.ones(3), "ps"), [3, 3, 3])
    # two jobs together, sizes 1 and 2: the short one ends at 2, the long one at 3
    np.testing.assert_allclose(queue_model(np.zeros(2), np.array([1.0, 2.0]), "ps"), [2, 3])
    # staggered arrivals: worked by hand (shares 1, then 1/2, then 1/3, then 1/2)
    np.testing.assert_allclose(queue_model(np.array([0.0, 0.1, 0.2]), np.ones(3), "ps"), [2.75, 2.85, 2.8])


def test_srpt_closed_forms():
    # a long job preempted by two short ones arriving during it: they run at once, it ends last
    np.testing.assert_allclose(queue_model(np.array([0.0, 0.1, 0.2]), np.array([1.0, 0.1, 0.1]), "srpt"),
                               [1.2, 0.1, 0.1])
    # simultaneous arrivals: shortest first, so the sum of latencies is the least possible
    np.testing.assert_allclose(queue_model(np.zeros(3), np.array([3.0, 1.0, 2.0]), "srpt"), [6, 1, 3])


def test_work_conservation():
    # every discipline finishes the last request at the same time when the server never idles
    rng = np.random.default_rng(1)
    svc = rng.exponential(1.0, 50)
    arr = np.sort(rng.uniform(0, 5, 50))
    end_f = (queue_model(arr, svc, "fifo") + arr).max()
    end_p = (queue_model(arr, svc, "ps") + arr).max()
    end_s = (queue_model(arr, svc, "srpt") + arr).max()
    assert abs(end_f - end_p) < 1e-9 and abs(end_f - end_s) < 1e-9
    # SRPT minimises the mean latency
    assert queue_model(arr, svc, "srpt").mean() <= min(queue_model(arr, svc, d).mean() for d in ("fifo", "ps")) + 1e-12
    # Little's law style check: FIFO's mean latency is at least the mean service time
    assert queue_model(arr, svc, "fifo").mean() >= svc.mean()


def _order_latencies(arr, svc, order):
    """Latencies when one server takes the requests whole, in the given order, as soon as it is free
    and the request has arrived."""
    t, lat = 0.0, np.zeros(len(arr))
    for i in order:
        t = max(t, arr[i]) + svc[i]
        lat[i] = t - arr[i]
    return lat


def test_p99_lower_bound_by_brute_force():
    # 6 requests: the p99 lies between the two largest latencies, so one request may be sacrificed.
    # Over every order of service, the second-largest latency is never below the bound, and the best
    # order (FIFO of the others, the sacrificed request last) reaches it.
    rng = np.random.default_rng(3)
    for _ in range(5):
        arr = np.sort(rng.uniform(0, 2, 6))
        svc = rng.exponential(0.5, 6)
        lb = p99_lower_bound(arr, svc)
        second = [np.sort(_order_latencies(arr, svc, o))[-2] for o in itertools.permutations(range(6))]
        p99 = [np.percentile(_order_latencies(arr, svc, o), 99) for o in itertools.permutations(range(6))]
        assert abs(min(second) - lb) <= 1e-12 * lb
        assert min(p99) >= lb * (1 - 1e-12)


def test_p99_lower_bound_below_every_discipline():
    # 120 requests (two may be sacrificed): the bound lies below each modelled server's p99
    rng = np.random.default_rng(4)
    arr = np.cumsum(rng.exponential(1.0, 120))
    svc = rng.choice([0.05, 0.2, 0.8, 3.2], 120)
    lb = p99_lower_bound(arr, svc)
    assert lb is not None
    for d in ("fifo", "ps", "srpt"):
        assert lb <= np.percentile(queue_model(arr, svc, d), 99) + 1e-12
