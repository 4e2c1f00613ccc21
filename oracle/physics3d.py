"""Pointwise physics of the oracle in ``d`` dimensions (the 3D path):
primitives, Riemann fluxes, boundary states, viscous stress.  Rows are
Python lists of arrays (or dual numbers): ``q = [rho, v_1 .. v_d, p(, nu_hat)]``.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Each function restates the reference function cited, with the
``d``-component sums in index order as ``src/jaxrans/core/state.py:66-78``
(``dot``; ``magnitude`` is ``sqrt(dot)`` for ``d = 3``).  Scalar laws
(limiters, viscosity, Spalart-Allmaras) are those of ``oracle/physics.py``.
"""

from __future__ import annotations

import numpy as np

from . import dual as M
from .physics import hypot_jnp


def dot(a, b):
    total = a[0] * b[0]
    for k in range(1, len(a)):
        total = total + a[k] * b[k]
    return total


def magnitude_const(a):
    """``magnitude`` of constant (non-dual) vectors."""
    if len(a) == 2:
        return hypot_jnp(a[0], a[1])
    return np.sqrt(dot(a, a))


def primitive(U, gamma, d):
    """(src/jaxrans/core/state.py:88-93)"""
    rho = U[0]
    vel = [U[1 + a] / rho for a in range(d)]
    p = (gamma - 1.0) * (U[d + 1] - 0.5 * rho * dot(vel, vel))
    out = [rho, *vel, p]
    if len(U) > d + 2:
        out.append(U[d + 2] / rho)
    return out


def conservative(q, gamma, d):
    """(src/jaxrans/core/state.py:96-99)"""
    rho, vel, p = q[0], q[1:d + 1], q[d + 1]
    e = p / (gamma - 1.0) + 0.5 * rho * dot(vel, vel)
    out = [rho, *[rho * v for v in vel], e]
    if len(q) > d + 2:
        out.append(rho * q[d + 2])
    return out


def normal_flux(q, n, gamma, d):
    """(src/jaxrans/physics/euler.py:43-52)"""
    rho, vel, p = q[0], q[1:d + 1], q[d + 1]
    vn = dot(vel, n)
    e = p / (gamma - 1.0) + 0.5 * rho * dot(vel, vel)
    return [rho * vn, *[rho * vel[a] * vn + p * n[a] for a in range(d)], (e + p) * vn]


def roe_average(qL, qR, gamma, d):
    """(src/jaxrans/physics/euler.py:74-94)"""
    rl, vl, pl = qL[0], qL[1:d + 1], qL[d + 1]
    rr, vr, pr = qR[0], qR[1:d + 1], qR[d + 1]
    hl = gamma / (gamma - 1.0) * pl / rl + 0.5 * dot(vl, vl)
    hr = gamma / (gamma - 1.0) * pr / rr + 0.5 * dot(vr, vr)
    w = M.sqrt(rr / rl)
    den = 1.0 + w
    rho = M.sqrt(rl * rr)
    vel = [(vl[a] + w * vr[a]) / den for a in range(d)]
    h = (hl + w * hr) / den
    a2 = (gamma - 1.0) * (h - 0.5 * dot(vel, vel))
    return rho, vel, h, M.sqrt(M.maximum(a2, 1e-30))


def unit_normal(s):
    """(src/jaxrans/numerics/riemann.py:47-51)"""
    length = magnitude_const(s)
    safe = np.maximum(length, 1e-300)
    return length, [c / safe for c in s]


def hllc(qL, qR, s, gamma, d):
    """HLLC with Einfeldt speeds (src/jaxrans/numerics/riemann.py:116-198)."""
    length, n = unit_normal(s)
    rl, vl, pl = qL[0], qL[1:d + 1], qL[d + 1]
    rr, vr, pr = qR[0], qR[1:d + 1], qR[d + 1]
    vnl, vnr = dot(vl, n), dot(vr, n)
    al = M.sqrt(gamma * pl / rl)
    ar = M.sqrt(gamma * pr / rr)
    _, vro, _, aro = roe_average(qL, qR, gamma, d)
    vnro = dot(vro, n)
    sl = M.minimum(vnl - al, vnro - aro)
    sr = M.maximum(vnr + ar, vnro + aro)
    ml = rl * (sl - vnl)
    mr = rr * (sr - vnr)
    sm = (pr - pl + ml * vnl - mr * vnr) / M.minimum(ml - mr, -1e-300)
    nf = d + 2

    def star(rho, vel, p, vn, sk, fk, uk, widen):
        pstar = p + rho * (vn - sk) * (vn - sm)
        rstar = rho * (sk - vn) / widen
        estar = ((sk - vn) * uk[d + 1] - p * vn + pstar * sm) / widen
        ustar = [rstar, *[rstar * (vel[a] + (sm - vn) * n[a]) for a in range(d)], estar]
        return [fk[k] + sk * (ustar[k] - uk[k]) for k in range(nf)]

    ucl, ucr = conservative(qL[:nf], gamma, d), conservative(qR[:nf], gamma, d)
    fl, fr = normal_flux(qL, n, gamma, d), normal_flux(qR, n, gamma, d)
    stl = star(rl, vl, pl, vnl, sl, fl, ucl, M.minimum(sl - sm, -1e-300))
    strr = star(rr, vr, pr, vnr, sr, fr, ucr, M.maximum(sr - sm, 1e-300))
    out = []
    for k in range(nf):
        f = M.where(sl >= 0.0, fl[k], M.where(sm >= 0.0, stl[k], M.where(sr >= 0.0, strr[k], fr[k])))
        out.append(length * f)
    return out


def roe(qL, qR, s, gamma, d):
    """Roe flux, no entropy fix (src/jaxrans/numerics/riemann.py:54-113)."""
    length, n = unit_normal(s)
    rho, vel, h, a = roe_average(qL, qR, gamma, d)
    vn = dot(vel, n)
    rl, vl, pl = qL[0], qL[1:d + 1], qL[d + 1]
    rr, vr, pr = qR[0], qR[1:d + 1], qR[d + 1]
    drho, dp = rr - rl, pr - pl
    dv = [vr[c] - vl[c] for c in range(d)]
    dvn = dot(dv, n)
    dvt = [dv[c] - dvn * n[c] for c in range(d)]
    a2 = a * a
    am = (dp - rho * a * dvn) / (2.0 * a2)
    ae = drho - dp / a2
    ap = (dp + rho * a * dvn) / (2.0 * a2)
    lm, lmid, lp = M.fabs(vn - a), M.fabs(vn), M.fabs(vn + a)
    kin = 0.5 * dot(vel, vel)
    one = 1.0
    rm = [one, *[vel[c] - a * n[c] for c in range(d)], h - a * vn]
    re = [one, *vel, kin]
    sh = [0.0, *[rho * dvt[c] for c in range(d)], rho * dot(vel, dvt)]
    rp = [one, *[vel[c] + a * n[c] for c in range(d)], h + a * vn]
    fl, fr = normal_flux(qL, n, gamma, d), normal_flux(qR, n, gamma, d)
    out = []
    for k in range(d + 2):
        diss = lm * am * rm[k] + lmid * ae * re[k] + lmid * sh[k] + lp * ap * rp[k]
        central = 0.5 * (fl[k] + fr[k])
        out.append(length * (central - 0.5 * diss))
    return out


FLUXES = {"hllc": hllc, "roe": roe}


def bc_state(kind: int, q, n, q_inf, gamma, d):
    """Ghost primitives for task ``kind`` from the source cell ``q``, outward
    unit normal ``n`` (``d`` rows) (src/jaxrans/numerics/boundary.py:54-157)."""
    rho, vel, p = q[0], q[1:d + 1], q[d + 1]
    scal = q[d + 2:]
    if kind in (1, 3):                     # slip wall / symmetry: mirror
        vn = dot(vel, n)
        vm = [vel[a] - 2.0 * vn * n[a] for a in range(d)]
        return [rho, *vm, p] + ([-s for s in scal] if kind == 1 else list(scal))
    if kind == 2:                          # no-slip wall
        return [rho, *[-v for v in vel], p] + [-s for s in scal]
    if kind == 4:
        return characteristic(q, n, q_inf, gamma, d)
    if kind == 5:
        return inflow_total(q, q_inf, gamma, d)
    if kind == 6:
        return [rho, *vel, 0.0 * p + q_inf[d + 1]] + list(scal)
    raise ValueError(kind)


def characteristic(q, n, q_inf, gamma, d):
    """(src/jaxrans/numerics/boundary.py:91-122)"""
    rho_i, v_i, p_i = q[0], q[1:d + 1], q[d + 1]
    rho_f, v_f, p_f = q_inf[0], list(q_inf[1:d + 1]), q_inf[d + 1]
    a_i = M.sqrt(gamma * p_i / rho_i)
    a_f = M.sqrt(gamma * p_f / rho_f)
    two_over = 2.0 / (gamma - 1.0)
    vn_i, vn_f = dot(v_i, n), dot(v_f, n)
    r_plus = vn_i + two_over * a_i
    r_minus = vn_f - two_over * a_f
    vn_b = 0.5 * (r_plus + r_minus)
    a_b = 0.25 * (gamma - 1.0) * (r_plus - r_minus)
    out = vn_b > 0.0
    entropy = M.where(out, p_i / rho_i**gamma, p_f / rho_f**gamma + 0.0 * n[0])
    tang = [M.where(out, v_i[a] - vn_i * n[a], v_f[a] - vn_f * n[a]) for a in range(d)]
    rho_b = (a_b * a_b / (gamma * entropy)) ** (1.0 / (gamma - 1.0))
    p_b = rho_b * a_b * a_b / gamma
    res = [rho_b, *[vn_b * n[a] + tang[a] for a in range(d)], p_b]
    for k in range(d + 2, len(q)):
        res.append(M.where(out, q[k], q_inf[k] + 0.0 * n[0]))
    return res


def inflow_total(q, q_inf, gamma, d):
    """(src/jaxrans/numerics/boundary.py:130-147)"""
    g1 = gamma - 1.0
    rho_f, v_f, p_f = q_inf[0], [M.scalar(v) for v in q_inf[1:d + 1]], q_inf[d + 1]
    t_f = gamma * p_f / rho_f
    if M.is_dual(*v_f):   # a freestream carrying a parameter tangent
        speed_f = P.speed_of(v_f[0], v_f[1]) if d == 2 else M.sqrt(dot(v_f, v_f))
    else:
        speed_f = float(magnitude_const(v_f))
    t_tot = t_f + 0.5 * g1 * speed_f * speed_f
    p_tot = p_f * (t_tot / t_f) ** (gamma / g1)
    p_b = q[d + 1]
    mach2 = M.maximum(2.0 / g1 * ((p_tot / p_b) ** (g1 / gamma) - 1.0), 0.0)
    t_b = t_tot / (1.0 + 0.5 * g1 * mach2)
    speed = M.sqrt(mach2 * t_b)
    res = [gamma * p_b / t_b, *[speed * v_f[a] / speed_f for a in range(d)], p_b]
    for k in range(d + 2, len(q)):
        res.append(0.0 * p_b + q_inf[k])
    return res


def temperature(q, gamma, d):
    return gamma * q[d + 1] / q[0]


def viscous_flux(vel, gv, gT, s, mu, k, extra=None):
    """``F_v . S`` rows; ``gv[a][b] = d v_a / d x_b``
    (src/jaxrans/physics/navier_stokes.py:86-116)."""
    d = len(s)
    div = gv[0][0]
    for a in range(1, d):
        div = div + gv[a][a]
    ttd = (2.0 / 3.0) * div

    def tau(a, b):
        if a == b:
            lin = mu * (2.0 * gv[a][a] - ttd)
        else:
            lo, hi = min(a, b), max(a, b)
            lin = mu * (gv[lo][hi] + gv[hi][lo])
        return lin if extra is None else lin + extra[a][b]

    force = [dot([tau(a, b) for b in range(d)], s) for a in range(d)]
    energy = dot(vel, force) + k * dot(gT, s)
    return [0.0 * energy, *force, energy]


def qcr_stress(gv, mu_t, c_cr1):
    """QCR2000 addition (src/jaxrans/physics/spalart_allmaras.py:197-216)."""
    d = len(gv)
    square = None
    for a in range(d):
        for b in range(d):
            term = gv[a][b] * gv[a][b]
            square = term if square is None else square + term
    sq = M.value(square)
    positive = sq > 0.0
    norm = M.where(positive, M.sqrt(M.where(positive, square, 1.0)), 1.0)
    rot = [[(gv[a][b] - gv[b][a]) / norm for b in range(d)] for a in range(d)]
    div = gv[0][0]
    for a in range(1, d):
        div = div + gv[a][a]
    tau = [[mu_t * (gv[a][b] + gv[b][a] - (2.0 / 3.0) * div * (1.0 if a == b else 0.0)) for b in range(d)]
           for a in range(d)]
    prod = [[dot([rot[i][c] for c in range(d)], [tau[j][c] for c in range(d)]) for j in range(d)] for i in range(d)]
    return [[-c_cr1 * (prod[a][b] + prod[b][a]) for b in range(d)] for a in range(d)]


def vorticity_magnitude(g):
    """``|curl V|`` from ``g[b][a] = d v_a / d x_b``; zero tangent at zero
    (src/jaxrans/numerics/turbulence.py:93-103)."""
    if len(g) == 2:
        return M.fabs(g[0][1] - g[1][0])
    curl = [g[1][2] - g[2][1], g[2][0] - g[0][2], g[0][1] - g[1][0]]
    square = dot(curl, curl)
    positive = M.value(square) > 0.0
    return M.where(positive, M.sqrt(M.where(positive, square, 1.0)), 0.0)
