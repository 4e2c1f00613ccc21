"""Pins the oracle's pointwise physics by properties the mathematics fixes
(the reference's own unit tests re-expressed: tests/unit/test_riemann.py:35-223,
test_limiters.py:26-110, test_flux_jacobian.py:35-114, test_spalart_allmaras.py:23-148)."""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import dual as M
from oracle import physics as P
from oracle import solver as S

G = 1.4
RNG = np.random.default_rng(0)


def states(n=64, seed=0):
    r = np.random.default_rng(seed)
    rho = r.uniform(0.5, 2.0, n)
    u, v = r.uniform(-0.9, 0.9, n), r.uniform(-0.9, 0.9, n)
    p = r.uniform(0.3, 1.5, n)
    return [rho, u, v, p]


def vectors(n=64, seed=1):
    r = np.random.default_rng(seed)
    return (r.uniform(-2, 2, n), r.uniform(-2, 2, n))


@pytest.mark.parametrize("flux", ["hllc", "roe"])
def test_consistency(flux):
    q, s = states(), vectors()
    F = P.FLUXES[flux](q, q, s, G)
    exact = P.normal_flux(q, s, G)
    for a, b in zip(F, exact):
        assert np.allclose(a, b, rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("flux", ["hllc", "roe"])
def test_conservation(flux):
    qL, qR, s = states(seed=2), states(seed=3), vectors()
    F1 = P.FLUXES[flux](qL, qR, s, G)
    F2 = P.FLUXES[flux](qR, qL, (-s[0], -s[1]), G)
    for a, b in zip(F1, F2):
        assert np.allclose(a, -b, rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("flux", ["hllc", "roe"])
def test_linear_in_face_area(flux):
    qL, qR, s = states(seed=4), states(seed=5), vectors()
    F1 = P.FLUXES[flux](qL, qR, s, G)
    F3 = P.FLUXES[flux](qL, qR, (3.0 * s[0], 3.0 * s[1]), G)
    for a, b in zip(F1, F3):
        assert np.allclose(3.0 * a, b, rtol=1e-12, atol=1e-13)
    Z = P.FLUXES[flux](qL, qR, (0.0 * s[0], 0.0 * s[1]), G)
    assert all(np.all(z == 0.0) for z in Z)


@pytest.mark.parametrize("flux", ["hllc", "roe"])
def test_rotation_invariance(flux):
    qL, qR, s = states(seed=6), states(seed=7), vectors()
    th = 0.7
    c, sn = math.cos(th), math.sin(th)
    rot = lambda x, y: (c * x - sn * y, sn * x + c * y)  # noqa: E731
    F = P.FLUXES[flux](qL, qR, s, G)
    qLr = [qL[0], *rot(qL[1], qL[2]), qL[3]]
    qRr = [qR[0], *rot(qR[1], qR[2]), qR[3]]
    Fr = P.FLUXES[flux](qLr, qRr, rot(*s), G)
    assert np.allclose(Fr[0], F[0], rtol=1e-11, atol=1e-12)
    assert np.allclose(Fr[3], F[3], rtol=1e-11, atol=1e-12)
    mx, my = rot(F[1], F[2])
    assert np.allclose(Fr[1], mx, rtol=1e-11, atol=1e-12)
    assert np.allclose(Fr[2], my, rtol=1e-11, atol=1e-12)


@pytest.mark.parametrize("flux", ["hllc", "roe"])
def test_stationary_contact_is_exact(flux):
    """Equal pressure, zero normal velocity, a density jump and a tangential
    shear: the exact flux is the pressure force alone."""
    n = 16
    s = (np.ones(n), np.zeros(n))
    qL = [np.full(n, 1.0), np.zeros(n), np.full(n, 0.3), np.full(n, 0.7)]
    qR = [np.full(n, 0.25), np.zeros(n), np.full(n, -0.2), np.full(n, 0.7)]
    F = P.FLUXES[flux](qL, qR, s, G)
    assert np.allclose(F[0], 0.0, atol=1e-14)
    assert np.allclose(F[1], 0.7, rtol=1e-13)
    assert np.allclose(F[2], 0.0, atol=1e-14)
    assert np.allclose(F[3], 0.0, atol=1e-14)


def test_roe_matrix_property():
    """Roe's average makes A(qL, qR)(UR - UL) = FR - FL: the Roe flux with
    both speeds of one sign is the upwind analytic flux."""
    n = 32
    r = np.random.default_rng(11)
    qL = [r.uniform(0.8, 1.2, n), r.uniform(2.0, 2.5, n), r.uniform(-0.1, 0.1, n), r.uniform(0.6, 0.8, n)]
    qR = [r.uniform(0.8, 1.2, n), r.uniform(2.0, 2.5, n), r.uniform(-0.1, 0.1, n), r.uniform(0.6, 0.8, n)]
    s = (np.ones(n), np.zeros(n))
    F = P.roe(qL, qR, s, G)
    FL = P.normal_flux(qL, s, G)
    for a, b in zip(F, FL):
        assert np.allclose(a, b, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("name", ["minmod", "van_albada", "van_albada_smooth", "none"])
def test_limiter_identity_and_symmetry(name):
    lim = P.limiter(name)
    a = np.random.default_rng(2).uniform(-3, 3, 200)
    assert np.allclose(lim(a, a), a, rtol=1e-14, atol=1e-14)
    b = np.random.default_rng(3).uniform(-3, 3, 200)
    assert np.allclose(lim(a, b), lim(b, a), rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("name", ["minmod", "van_albada"])
def test_tvd_limiters(name):
    lim = P.limiter(name)
    a = np.random.default_rng(4).uniform(-3, 3, 500)
    b = np.random.default_rng(5).uniform(-3, 3, 500)
    out = lim(a, b)
    same = a * b > 0
    assert np.all(out[~same] == 0.0)
    assert np.all(np.abs(out[same]) <= 2.0 * np.minimum(np.abs(a[same]), np.abs(b[same])) + 1e-15)
    assert np.all(np.sign(out[same]) == np.sign(a[same]))
    if name == "minmod":
        assert np.allclose(np.abs(out[same]), np.minimum(np.abs(a[same]), np.abs(b[same])))


def test_flux_jacobian_homogeneity_and_eigenvalues():
    q, s = states(8), vectors(8)
    A = S.flux_jacobian(np.stack(q), s, G)
    U = np.stack(P.conservative(q, G))
    F = np.stack(P.normal_flux(q, s, G))
    assert np.allclose(np.einsum("ijn,jn->in", A, U), F, rtol=1e-12, atol=1e-13)
    for k in range(8):
        vn = q[1][k] * s[0][k] + q[2][k] * s[1][k]
        a = math.sqrt(G * q[3][k] / q[0][k]) * math.hypot(s[0][k], s[1][k])
        ev = np.sort(np.linalg.eigvals(A[:, :, k]).real)
        assert np.allclose(ev, np.sort([vn - a, vn, vn, vn + a]), rtol=1e-9, atol=1e-10)


def test_dual_tie_rules():
    a, b = M.D(np.array([1.0, 2.0, 3.0]), np.array([1.0, 1.0, 1.0])), M.D(np.array([1.0, 1.0, 4.0]),
                                                                       np.array([3.0, 5.0, 7.0]))
    mx = M.maximum(a, b)
    assert np.allclose(mx.v, [1.0, 2.0, 4.0]) and np.allclose(mx.d, [2.0, 1.0, 7.0])
    mn = M.minimum(a, b)
    assert np.allclose(mn.v, [1.0, 1.0, 3.0]) and np.allclose(mn.d, [2.0, 5.0, 1.0])
    z = M.fabs(M.D(np.array([0.0, -1.0]), np.array([2.0, 2.0])))
    assert np.allclose(z.d, [2.0, -2.0])


# --- Spalart-Allmaras: an independent scalar transcription of the TMR page --------------

def tmr(nu_t, nu, omega, d, rho):
    cb1, sigma, cb2, kappa = 0.1355, 2.0 / 3.0, 0.622, 0.41
    cw2, cw3, cv1, c2, c3, cn1 = 0.3, 2.0, 7.1, 0.7, 0.9, 16.0
    cw1 = cb1 / kappa ** 2 + (1 + cb2) / sigma
    chi = nu_t / nu
    fv1 = chi ** 3 / (chi ** 3 + cv1 ** 3)
    if nu_t >= 0:
        fv2 = 1 - chi / (1 + chi * fv1)
        sb = nu_t * fv2 / (kappa ** 2 * d ** 2)
        st = omega + sb if sb >= -c2 * omega else omega + omega * (c2 ** 2 * omega + c3 * sb) / (
            (c3 - 2 * c2) * omega - sb)
        r = min(nu_t / (st * kappa ** 2 * d ** 2), 10.0)
        g = r + cw2 * (r ** 6 - r)
        fw = g * ((1 + cw3 ** 6) / (g ** 6 + cw3 ** 6)) ** (1 / 6)
        return cb1 * st * nu_t, cw1 * fw * (nu_t / d) ** 2, 1.0, rho * nu_t * fv1
    return cb1 * omega * nu_t, -cw1 * (nu_t / d) ** 2, (cn1 + chi ** 3) / (cn1 - chi ** 3), 0.0


# The reference's sample points (tests/unit/test_spalart_allmaras.py:54-62): log layer,
# sublayer, wake, the S~ clip, r capped, negative nu~.
POINTS = [(4.1e-3, 1e-4, 50.0, 2e-3), (2e-5, 1e-4, 4e3, 1e-5), (3e-4, 1e-4, 0.5, 0.3), (1.5e-4, 1e-4, 0.02, 1e-3),
          (5e-2, 1e-4, 1e-6, 0.05), (-2e-4, 1e-4, 30.0, 1e-3), (-5e-5, 1e-4, 1.0, 0.2)]


@pytest.mark.parametrize("nu_t, nu, omega, d", POINTS)
def test_sa_matches_tmr(nu_t, nu, omega, d):
    Pr, Dr, fnr, mutr = tmr(nu_t, nu, omega, d, 1.3)
    p, dd = P.source(np.array(nu_t), np.array(nu), np.array(omega), np.array(d))
    assert float(p) == pytest.approx(Pr, rel=1e-13)
    assert float(dd) == pytest.approx(Dr, rel=1e-13)
    assert float(P.f_n(np.array(nu_t), np.array(nu))) == pytest.approx(fnr, rel=1e-13)
    assert float(P.eddy_viscosity(np.array(1.3), np.array(nu_t), np.array(nu))) == pytest.approx(mutr, rel=1e-13)


def test_sa_constants():
    assert P.CW1 == pytest.approx(3.2390678, rel=1e-7)
    assert float(P.eddy_viscosity(np.array(1.0), np.array(3.0), np.array(1.0))) == pytest.approx(0.210438, rel=1e-5)
    assert float(P.f_w(np.array(1.0))) == pytest.approx(1.0, rel=1e-15)
