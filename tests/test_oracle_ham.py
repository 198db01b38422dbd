"""The Hamiltonian twin (oracle/ham_twin.cpp): the formulation the mixed-precision march runs
(the null geodesic in Hamiltonian form with p_t and p_phi held exactly, the polarization angle
from the ray's Walker-Penrose constant instead of a transported F1), checked at FP64 on the CPU.

* Hamilton's equations against finite differences of H = g^{mu nu} p_mu p_nu / 2 built from
  the oracle's contravariant metric, and k^mu = g^{mu nu} p_nu;
* the Walker-Penrose constant: zero for f = k, real-linear in f, and constant along a ray
  whose F1 the oracle transports with the reference's Christoffel RK4 (a sign slip in its
  Kerr-Schild form breaks the constancy at the 1e0 level);
* the polarization angle recovered from the constant equal to the projection of the
  transported F1 on an orthonormal screen basis of a moving plasma;
* every golden case marched by the twin against the reference's FP64 build: the acceptance
  gates of the path, total flux within 1e-7, the reference's run statistics to 1e-5.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests import parity
from tests.golden import make_golden as G
from tests.test_oracle_golden import params_from_overrides
from tests.test_oracle_pins import A, HS, metric, near_photon, rand_points, trace


def ham(X, p):
    _, gu = metric(X)
    return 0.5 * p @ gu @ p


def test_hamilton_equations_against_finite_differences(oracle_lib):
    rng = np.random.default_rng(11)
    for X in rand_points(30, 5):
        p = np.array([-1.0, rng.normal() * math.exp(X[1]), rng.normal() * 5, rng.normal() * 3])
        y = np.array([X[1], X[2], X[3], p[1], p[2]])
        d, k = O.ham_rhs(A, HS, p[0], p[3], y)
        _, gu = metric(X)
        np.testing.assert_allclose(k, gu @ p, rtol=1e-12, atol=1e-14 * np.abs(gu @ p).max())
        want = np.zeros(5)
        for i, l in enumerate((1, 2)):  # -dH/dx^l
            h = 1e-6
            Xp, Xm = X.copy(), X.copy()
            Xp[l] += h
            Xm[l] -= h
            want[3 + i] = -(ham(Xp, p) - ham(Xm, p)) / (2 * h)
        for i, j in enumerate((1, 2, 3)):  # dH/dp_j
            h = 1e-6 * max(1.0, abs(p[j]))
            pp, pm = p.copy(), p.copy()
            pp[j] += h
            pm[j] -= h
            want[i] = (ham(X, pp) - ham(X, pm)) / (2 * h)
        np.testing.assert_allclose(d, want, rtol=2e-6, atol=1e-9 * np.abs(want).max())


def transverse_photon(r0):
    """near_photon with its polarization vector made orthogonal to k again (the constant is
    conserved for a parallel-transported f with f.k = 0, as the camera's F1 is)"""
    y0 = near_photon(r0)
    g, _ = metric(y0[:4])
    k, f, e = y0[4:8], y0[8:12], np.array([1.0, 0.0, 0.0, 0.0])
    y0[8:12] = f - (f @ g @ k) / (e @ g @ k) * e
    return y0


def test_walker_penrose_constant_along_a_transported_ray(oracle_lib):
    y0 = transverse_photon(30.0)
    st = trace(y0, -0.02, 3000)
    q0 = O.wp_kappa(A, HS, st[0, :4], st[0, 4:8], st[0, 8:12])
    assert np.linalg.norm(O.wp_kappa(A, HS, st[0, :4], st[0, 4:8], st[0, 4:8])) == 0.0  # kappa(k) = 0
    f2 = np.array([0.3, -1.2, 0.4, 0.05])
    lin = O.wp_kappa(A, HS, st[0, :4], st[0, 4:8], 2.0 * st[0, 8:12] - 3.0 * f2)
    np.testing.assert_allclose(lin, 2.0 * q0 - 3.0 * O.wp_kappa(A, HS, st[0, :4], st[0, 4:8], f2), rtol=1e-12)
    dev = max(np.linalg.norm(O.wp_kappa(A, HS, s[:4], s[4:8], s[8:12]) - q0) for s in st)
    assert dev < 1e-7 * np.linalg.norm(q0)
    assert math.exp(st[:, 1].min()) < 12.0  # the ray passed close to the hole


def screen_basis(g, k, U, b, rng):
    """An orthonormal pair (e1, e2) orthogonal to U and to the spatial direction of k, e2 along b."""
    om = -(k @ g @ U)
    eK = k / om - U
    e2 = b + (b @ g @ U) * U - (b @ g @ eK) * eK
    e2 /= math.sqrt(e2 @ g @ e2)
    v = rng.normal(size=4)
    v = v + (v @ g @ U) * U - (v @ g @ eK) * eK - (v @ g @ e2) * e2
    return v / math.sqrt(v @ g @ v), e2


def test_polarization_angle_from_the_constant(oracle_lib):
    rng = np.random.default_rng(2)
    y0 = transverse_photon(30.0)
    st = trace(y0, -0.02, 3000)
    q0 = O.wp_kappa(A, HS, st[0, :4], st[0, 4:8], st[0, 8:12])
    worst = 0.0
    for s in st[::150]:
        X, k, F1 = s[:4], s[4:8], s[8:12]
        g, gu = metric(X)
        # a plasma moving with a random 3-velocity relative to the normal observer
        alpha = 1.0 / math.sqrt(-gu[0, 0])
        n = -alpha * gu[:, 0]
        v = np.concatenate([[0.0], 0.3 * rng.normal(size=3) / np.sqrt([g[1, 1], g[2, 2], g[3, 3]])])
        v *= 0.5 / max(math.sqrt(v @ g @ v), 0.5)
        U = (n + v) / math.sqrt(1.0 - v @ g @ v)
        b = rng.normal(size=4)
        e1, e2 = screen_basis(g, k, U, b, rng)
        Cp, Sp = e1 @ g @ F1, e2 @ g @ F1
        c2, s2 = (Cp * Cp - Sp * Sp) / (Cp * Cp + Sp * Sp), 2 * Cp * Sp / (Cp * Cp + Sp * Sp)
        k1 = O.wp_kappa(A, HS, X, k, e1)
        k2 = O.wp_kappa(A, HS, X, k, e2)
        f1, f2 = q0[0] * k2[1] - k2[0] * q0[1], k1[0] * q0[1] - q0[0] * k1[1]
        c2w, s2w = (f1 * f1 - f2 * f2) / (f1 * f1 + f2 * f2), 2 * f1 * f2 / (f1 * f1 + f2 * f2)
        worst = max(worst, abs(c2w - c2), abs(s2w - s2))
    assert worst < 1e-6


@pytest.mark.parametrize("case", sorted(G.CASES))
def test_twin_matches_reference_fp64(case, snapdir, oracle_lib):
    from phoenix_amd import snapshot as S

    preset, over = G.CASES[case]
    dump, grid, info = S.ensure_snapshot(preset, snapdir)
    ref, st_ref, meta = G.load(case, 64)
    o = O.Oracle(dump, grid, 64, **params_from_overrides(over))
    res, st = o.render_ham(threads=8)
    o.close()
    figs = parity.compare(res, ref, case=case)
    print(f"\n{case}: {parity.summary(figs)}; stats {st.tolist()} ref {st_ref.tolist()}")
    assert parity.passes(figs), parity.summary(figs)
    # far inside the path's 1e-4 flux gate: the twin carries the polarization by the exact Walker-Penrose
    # constant where the reference integrates F1 with RK4, so Q/U/V differ by that integration error
    # (at most 1.1e-7 of I, on c5_i120_32_2f at 460 GHz)
    for d in figs:
        assert d["flux_dI"] <= 1e-7 and d["flux_dQUV"] <= 3e-7 and d["rel_I_p999"] <= 2e-5
    assert np.all(np.abs(st - st_ref) <= 1e-5 * st_ref + 1)
