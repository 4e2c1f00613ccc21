"""Pins of the oracle's viscous terms, SA row, boundary states, wave speeds,
chain blocks, Newton control and forces by what the mathematics fixes:
closed forms, exactness on linear fields, orders of accuracy, textbook
identities.  Each check is chosen so that a plausible mistake in the oracle (a
dropped term, a wrong sign or index, a transposed operand) fails it; none of
them re-derives its expected value from the oracle's own formula.

Reference tests restated: ``tests/unit/test_viscous.py:39-115``,
``tests/verification/test_mms.py:50-158``, ``tests/unit/test_rans.py:76-92``,
``tests/unit/test_boundary_conditions.py:48-224``,
``tests/verification/test_implicit.py:56-112``.
"""

from __future__ import annotations

import numpy as np
import pytest

from conformal import conformal_atlas
from conftest import o_atlas, perturbed
from jaxrans_amd.core.config import FlowConditions, ImplicitConfig, NumericsConfig
from jaxrans_amd.mesh.atlas import assign_conditions, build_atlas
from jaxrans_amd.mesh.block import Grid, rectangle_sides
from oracle import physics as P
from oracle import residual as OR
from oracle import solver as OS

GAMMA, MU = 1.4, 0.1
EPS = np.finfo(np.float64).eps


def cartesian_atlas(ni=12, nj=8, dx=0.1, dy=0.05, wall="no_slip_wall", others="far_field"):
    """A uniform Cartesian block ``[0, ni dx] x [0, nj dy]``: the wall on
    ``j_min`` (y = 0), ``others`` on the three remaining sides."""
    x = np.arange(ni + 1)[:, None] * dx + 0.0 * np.arange(nj + 1)[None]
    y = 0.0 * np.arange(ni + 1)[:, None] + np.arange(nj + 1)[None] * dy
    sides = rectangle_sides(ni, nj, j_min="wall", j_max="top", i_min="left", i_max="right")
    mesh = assign_conditions((Grid(x, y, sides),), {"wall": wall, "top": others, "left": others,
                                                    "right": others})
    return build_atlas(mesh)


def interior_faces(at):
    """Masks of the faces of each family that touch an interior cell."""
    I = at.interior
    fi = np.zeros((at.ni + 1, at.nj), bool)
    fi[1:] |= I
    fi[:-1] |= I
    fj = np.zeros((at.ni, at.nj + 1), bool)
    fj[:, 1:] |= I
    fj[:, :-1] |= I
    return fi, fj


def at_centres(at, f):
    """``f(x, y)`` on the one-ghost padded canvas (rows as a list)."""
    return [np.asarray(r, np.float64) for r in f(at.centres[0], at.centres[1])]


def canvas_rows(at, f):
    """``f(x, y)`` at every canvas cell centre (ghost centres included)."""
    return [np.asarray(r, np.float64) for r in f(at.centres[0][1:-1, 1:-1], at.centres[1][1:-1, 1:-1])]


# --- viscous flux, diamond gradients -------------------------------------------------------------


def test_pure_shear_and_heat_flux_closed_forms():
    """``du/dy = 1`` on a face normal to y: shear ``mu``, work ``u mu``; a unit
    temperature gradient along the face normal: heat flux ``k``."""
    k = MU / (P.PR * (GAMMA - 1.0))
    f = P.viscous_flux([0.4, 0.0], [[0.0, 1.0], [0.0, 0.0]], [0.0, 0.0], (0.0, 1.0), MU, k)
    assert np.allclose(f, [0.0, MU, 0.0, 0.4 * MU], rtol=1e-15, atol=0.0)
    f = P.viscous_flux([0.0, 0.0], [[0.0, 0.0], [0.0, 0.0]], [1.0, 0.0], (1.0, 0.0), MU, k)
    assert f[3] == pytest.approx(MU / (0.72 * 0.4), rel=1e-15)
    # normal stress of a pure dilatation du/dx = 1: tau_xx = mu (2 - 2/3), tau_yy = -2/3 mu
    f = P.viscous_flux([0.0, 0.0], [[1.0, 0.0], [0.0, 0.0]], [0.0, 0.0], (1.0, 0.0), MU, k)
    assert f[1] == pytest.approx(MU * 4.0 / 3.0, rel=1e-15) and f[2] == 0.0
    f = P.viscous_flux([0.0, 0.0], [[1.0, 0.0], [0.0, 0.0]], [0.0, 0.0], (0.0, 1.0), MU, k)
    assert f[2] == pytest.approx(-MU * 2.0 / 3.0, rel=1e-15) and f[1] == 0.0


@pytest.mark.parametrize("grid", ["o48", "conformal32", "cartesian"])
def test_face_gradients_are_exact_for_linear_fields(grid):
    """On every face of an interior cell -- wall, far field, periodic seam --
    up to the conditioning of a difference quotient (eps |x| / h)."""
    at = {"o48": lambda: o_atlas(), "conformal32": lambda: conformal_atlas(32),
          "cartesian": lambda: cartesian_atlas()}[grid]()
    phi = at_centres(at, lambda x, y: [1.0 + 0.3 * x - 0.7 * y])
    grads = OR.face_gradients(phi, at.centres)
    x = at.centres[0][1:-1, 1:-1]
    bound = 20.0 * EPS * np.abs(x[at.interior]).max() / np.sqrt(at.area[at.interior].min())
    for (gx, gy), mask in zip(grads, interior_faces(at)):
        assert np.abs(gx[0][mask] - 0.3).max() < bound
        assert np.abs(gy[0][mask] + 0.7).max() < bound


def _laminar(at, mu=MU):
    return OR.Setup(at, GAMMA, np.array([1.0, 0.5, 0.0, 1.0 / GAMMA]), mu_inf=mu)


@pytest.mark.parametrize("grid", ["o48", "conformal32"])
def test_uniform_shear_has_no_net_viscous_momentum(grid):
    """A velocity linear in space has a constant stress, whose divergence is
    zero in every interior cell; the heat flux of a uniform temperature is zero."""
    at = o_atlas() if grid == "o48" else conformal_atlas(32)
    p1 = at_centres(at, lambda x, y: [1.0 + 0 * x, 0.3 * x - 0.7 * y, 0.7 * x + 0.2 * y, 1.0 / GAMMA + 0 * x])
    visc, _ = OR.viscous_face_fluxes(p1, _laminar(at))
    scale = MU * max(np.abs(at.face_i).max(), np.abs(at.face_j).max())
    for c in (1, 2):
        net = OR.divergence((visc[0][c], visc[1][c]))
        assert np.abs(net[at.interior]).max() < 1e-12 * scale, c


def test_uniform_shear_energy_row_is_the_stress_work_on_a_cartesian_grid():
    """Constant stress, linear velocity: ``sum_f (v . tau) . S = (tau : grad v)
    V`` exactly when the face velocity is the velocity at the face midpoint
    (a uniform Cartesian grid); a uniform temperature adds no heat flux."""
    at = cartesian_atlas()
    p1 = at_centres(at, lambda x, y: [1.0 + 0 * x, 0.3 * x - 0.7 * y, 0.7 * x + 0.2 * y, 1.0 / GAMMA + 0 * x])
    visc, _ = OR.viscous_face_fluxes(p1, _laminar(at))
    gv = np.array([[0.3, -0.7], [0.7, 0.2]])
    tau = MU * (gv + gv.T - (2.0 / 3.0) * np.trace(gv) * np.eye(2))
    work = float(np.sum(tau * gv))
    net_e = OR.divergence((visc[0][3], visc[1][3]))
    assert np.allclose(net_e[at.interior], work * at.area[at.interior], rtol=1e-11, atol=0.0)


def test_viscous_operator_is_second_order():
    """``u = y^2``: ``div tau = (2 mu, 0)``; on the conformal O-grids the error
    of the net x-momentum per area falls by > 3.5 per level (cells 0.6 to 5
    chords from mid-chord, maximum norm)."""
    errors = []
    for n in (32, 64, 128):
        at = conformal_atlas(n)
        p1 = at_centres(at, lambda x, y: [1.0 + 0 * x, y * y, 0 * x, 1.0 / GAMMA + 0 * x])
        visc, _ = OR.viscous_face_fluxes(p1, _laminar(at))
        net = OR.divergence((visc[0][1], visc[1][1]))
        xc, yc = at.centres[0][1:-1, 1:-1], at.centres[1][1:-1, 1:-1]
        r = np.hypot(xc - 0.5, yc)
        sel = at.interior & (r > 0.6) & (r < 5.0)
        errors.append(np.abs(net[sel] / at.area[sel] - 2.0 * MU).max())
    assert errors[0] / errors[1] > 3.5 and errors[1] / errors[2] > 3.5, errors


def test_qcr_closed_form_in_simple_shear():
    """QCR2000 in simple shear ``du/dy = s``: the rotation tensor is
    ``sign(s)`` off the diagonal, so the correction is the normal-stress pair
    ``-2 c mu_t |s|`` (xx), ``+2 c mu_t |s|`` (yy) and no shear; a pure strain
    (symmetric gradient) has no rotation and no correction."""
    c, mut = 0.3, 0.02
    for s in (1.7, -0.4):
        q = P.qcr_stress([[0.0, s], [0.0, 0.0]], mut, c)
        assert np.allclose(q, [[-2 * c * mut * abs(s), 0.0], [0.0, 2 * c * mut * abs(s)]], rtol=1e-14, atol=1e-18)
    q = P.qcr_stress([[0.3, 0.5], [0.5, -0.1]], mut, c)
    assert np.allclose(q, 0.0, atol=1e-18)
    assert np.all(np.asarray(P.qcr_stress([[0.0, 0.0], [0.0, 0.0]], mut, c)) == 0.0)


# --- method of manufactured solutions (truncation error) -----------------------------------------


def _manufactured(mu):
    """A smooth non-solution and the Cartesian components of ``F - F_v`` (sympy)."""
    import sympy as sp
    x, y = sp.symbols("x y", real=True)
    M, alpha = 0.5, np.deg2rad(1.0)
    g = sp.exp(-((x - 0.5) ** 2 + y ** 2) / sp.Rational(9, 16))
    rho = 1 + sp.Rational(1, 10) * g * (1 + sp.sin(sp.pi * x) / 2)
    u = M * np.cos(alpha) + sp.Rational(1, 10) * g * sp.cos(sp.pi * y)
    v = M * np.sin(alpha) + sp.Rational(1, 10) * g * sp.sin(sp.pi * x)
    p = 1 / sp.Float(GAMMA) + sp.Rational(1, 20) * g * sp.cos(sp.pi * x / 2 + y)
    E = p / (GAMMA - 1) + rho * (u ** 2 + v ** 2) / 2
    T = GAMMA * p / rho
    div = sp.diff(u, x) + sp.diff(v, y)
    txx = mu * (2 * sp.diff(u, x) - sp.Rational(2, 3) * div)
    tyy = mu * (2 * sp.diff(v, y) - sp.Rational(2, 3) * div)
    txy = mu * (sp.diff(u, y) + sp.diff(v, x))
    k = mu / (0.72 * (GAMMA - 1))
    fx = [rho * u, rho * u ** 2 + p - txx, rho * u * v - txy, (E + p) * u - (u * txx + v * txy) - k * sp.diff(T, x)]
    fy = [rho * v, rho * u * v - txy, rho * v ** 2 + p - tyy, (E + p) * v - (u * txy + v * tyy) - k * sp.diff(T, y)]

    def vec(exprs):
        f = sp.lambdify((x, y), exprs, "numpy")
        return lambda a, b: [np.broadcast_to(np.asarray(c, np.float64), np.shape(a)) for c in f(a, b)]

    return vec([rho, u, v, p]), vec(fx), vec(fy)


def _exact_net_flux(block, fx, fy):
    """``sum_faces int (F - F_v) . S`` per block cell, two-point Gauss on each
    straight face (face vectors: i-faces ``(dy, -dx)``, j-faces ``(-dy, dx)``)."""
    X, Y = block.x_node, block.y_node
    w = (0.5 - 0.5 / np.sqrt(3.0), 0.5 + 0.5 / np.sqrt(3.0))

    def face(xa, ya, xb, yb, sx, sy):
        tot = 0.0
        for t in w:
            px, py = xa + t * (xb - xa), ya + t * (yb - ya)
            gx, gy = np.stack(fx(px, py)), np.stack(fy(px, py))
            tot = tot + 0.5 * (gx * sx + gy * sy)
        return tot

    fi = face(X[:, :-1], Y[:, :-1], X[:, 1:], Y[:, 1:], Y[:, 1:] - Y[:, :-1], -(X[:, 1:] - X[:, :-1]))
    fj = face(X[:-1], Y[:-1], X[1:], Y[1:], -(Y[1:] - Y[:-1]), X[1:] - X[:-1])
    return (fi[:, 1:] - fi[:, :-1]) + (fj[:, :, 1:] - fj[:, :, :-1])


@pytest.mark.parametrize("flux", ["hllc", "roe"])
def test_navier_stokes_truncation_error_is_second_order(flux):
    """The whole viscous discretisation at once -- MUSCL faces (unlimited),
    Riemann flux, diamond gradients, stresses, heat flux, assembly on a
    curvilinear grid -- against the exact flux integral of a manufactured
    field (Re 100): the net flux per area converges at second order in every
    row (observed order > 1.8 between the two finest conformal O-grids; cells
    0.6 to 5 chords from mid-chord, three cells or more from the grid line
    leaving the trailing edge, where the map opens the corner and the cell
    widths jump: there the truncation error is first order, and only the
    discrete solution -- which the reference's own test solves for -- keeps
    second order).  Any term dropped, mis-signed or mis-indexed leaves an O(1)
    error and the order collapses."""
    mu = 0.5 / 100.0
    prim, fx, fy = _manufactured(mu)
    errors = []
    for n in (32, 64, 128):
        at = conformal_atlas(n)
        st = OR.Setup(at, GAMMA, np.array([1.0, 0.5, 0.0, 1.0 / GAMMA]), flux=flux, limiter="none", mu_inf=mu)
        q = canvas_rows(at, prim)
        net = np.stack(OR.net_flux(q, st, first_order=False))
        got = at.unpack(net)[0]
        want = _exact_net_flux(at.mesh[0], fx, fy)
        blk = at.mesh[0]
        r = np.hypot(blk.xc - 0.5, blk.yc)
        i = np.arange(n)[:, None] + 0 * blk.yc
        sel = (r > 0.6) & (r < 5.0) & (i >= 3) & (i <= n - 4)
        errors.append(np.abs((got - want)[:, sel] / blk.area[sel]).max(axis=1))
    e = np.array(errors)
    order = np.log2(e[:-1] / e[1:])
    assert np.all(order[-1] > 1.8), (e, order)


# --- Spalart-Allmaras row ----------------------------------------------------------------------------


def test_sa_row_equals_the_analytic_source_on_a_uniform_field():
    """Uniform flow with ``nu~ = 3 nu_inf``: no gradients, so away from the wall
    the mean-flow rows vanish and the SA row is ``-V rho (P - D) / nu_inf``
    with ``P, D`` of the pointwise model at zero vorticity (reference
    ``tests/unit/test_rans.py:76-92``)."""
    at = conformal_atlas(32, wall="no_slip_wall")
    flow = FlowConditions(mach=0.15, alpha_deg=10.0, reynolds=6e6, sutherland_ratio=198.6 / 540.0,
                          turbulence="sa_noft2")
    st = OR.setup_for(at, flow, NumericsConfig(limiter="none"))
    U = OR.U0_of(st)
    R = OR.residual(U, st)
    blk = at.unpack(R)[0]
    off = np.ones(at.mesh[0].shape, bool)
    off[:, :2] = False                                   # within two cells of the wall
    assert np.abs(blk[:4][:, off]).max() < 1e-12
    nu_inf = flow.mu_inf
    d = at.unpack(at.distance)[0]
    prod, dest = P.source(3.0 * nu_inf, nu_inf, 0.0, d)
    want = -at.mesh[0].area * (prod - dest) / nu_inf
    # round-off of the cancelling convective fluxes rho v.S nu^ (large outer faces)
    convective = 3.0 * flow.mach * max(np.abs(at.face_i).max(), np.abs(at.face_j).max())
    err = np.abs(blk[4][off] - want[off])
    assert err.max() < 1e-9 * np.abs(want[off]).max() + 50 * EPS * convective
    near = off & (np.arange(at.mesh[0].nj)[None] < 8)       # source-dominated cells: relative
    assert np.all(np.abs(blk[4][near] - want[near]) < 1e-9 * np.abs(want[near]))


def test_sa_conservative_diffusion_identity_on_a_linear_field():
    """At rest (rho 1, p 1/gamma, v = 0) with ``nu^ = a + g . x`` on a uniform
    Cartesian grid: the diffusion coefficient ``(mu + rho nu~)/sigma`` is
    linear, so ``div(coef grad nu^) = rho nu_inf |g|^2 / sigma`` exactly, and
    with the ``c_b2`` term the SA row is ``-V rho [nu_inf |g|^2 (1 + c_b2) /
    sigma + (P - D)/nu_inf]`` -- the identity that makes the SA diffusion
    conservative.  ``P, D`` at zero vorticity from the pointwise model."""
    at = cartesian_atlas(ni=10, nj=10, dx=0.1, dy=0.1)
    nu_inf = 1e-4
    g = np.array([0.8, 0.5])
    st = OR.Setup(at, GAMMA, np.array([1.0, 0.0, 0.0, 1.0 / GAMMA, 3.0]), flux="hllc", limiter="none",
                  mu_inf=nu_inf, sa=True)
    q = canvas_rows(at, lambda x, y: [1.0 + 0 * x, 0 * x, 0 * x, 1.0 / GAMMA + 0 * x, 3.0 + g[0] * x + g[1] * y])
    R4 = OR.net_flux(q, st, first_order=False)[4]
    nu_hat = q[4]
    prod, dest = P.source(nu_hat * nu_inf, nu_inf, 0.0, at.distance)
    cb2, sigma = 0.622, 2.0 / 3.0                       # the published SA constants
    want = -at.area * (nu_inf * (g @ g) * (1.0 + cb2) / sigma + (prod - dest) / nu_inf)
    I = at.interior
    assert np.allclose(R4[I], want[I], rtol=1e-11, atol=0.0)


# --- boundary states -------------------------------------------------------------------------------


def _wall_pairs(at, g=2):
    """(ghost, interior) canvas cells across the wall faces of the single
    O-block (wall on block j = 0), and those faces' vectors."""
    blk_i = np.arange(at.mesh[0].ni)
    ghost = (g + blk_i) * at.nj + (g - 1)
    inner = (g + blk_i) * at.nj + g
    S = at.face_j[:, g + blk_i, g]
    return ghost, inner, S


def _filled(at, q_rows, q_inf):
    st = OR.Setup(at, GAMMA, np.asarray(q_inf, np.float64))
    return [r.reshape(at.shape) for r in OR.fill_ghosts([r.reshape(-1) for r in q_rows], st)]


def _subsonic():
    return FlowConditions(mach=0.5, alpha_deg=1.25)


def test_slip_wall_passes_no_mass_or_energy_and_keeps_the_scalars():
    """Mirror ghost: the Roe flux between ghost and cell carries only
    pressure (zero mass and energy flux), density and pressure are copied,
    the normal velocity is negated, the tangential kept."""
    from jaxrans_amd.core.state import q_infinity
    at = o_atlas()
    q_inf = q_infinity(_subsonic())
    rng = np.random.default_rng(0)
    q = [np.full(at.shape, v) * (1.0 + 0.1 * rng.standard_normal(at.shape)) for v in q_inf]
    f = _filled(at, q, q_inf)
    ghost, inner, S = _wall_pairs(at)
    qg = [r.reshape(-1)[ghost] for r in f]
    qi = [r.reshape(-1)[inner] for r in f]
    flux = P.roe(qg, qi, (S[0], S[1]), GAMMA)
    scale = np.abs(qi[3] * np.hypot(S[0], S[1])).max()
    assert np.abs(flux[0]).max() < 1e-13 * scale and np.abs(flux[3]).max() < 1e-13 * scale
    assert np.array_equal(qg[0], qi[0]) and np.array_equal(qg[3], qi[3])
    n = S / np.hypot(S[0], S[1])
    vn_g, vn_i = qg[1] * n[0] + qg[2] * n[1], qi[1] * n[0] + qi[2] * n[1]
    assert np.abs(vn_g + vn_i).max() < 1e-14
    t_g, t_i = -qg[1] * n[1] + qg[2] * n[0], -qi[1] * n[1] + qi[2] * n[0]
    assert np.abs(t_g - t_i).max() < 1e-14


def test_no_slip_face_velocity_is_exactly_zero_and_scalar_parity():
    """No-slip: the face average of cell and ghost velocity vanishes bitwise;
    the turbulence scalar is odd at a wall and even at a symmetry plane
    (reference ``tests/unit/test_rans.py:112-128``)."""
    rng = np.random.default_rng(1)
    q5 = np.array([1.0, 0.3, 0.05, 1.0 / GAMMA, 3.0])
    src = [v * (1.0 + 0.1 * rng.random(7)) for v in q5]
    n = (np.full(7, 0.6), np.full(7, 0.8))
    ns = P.bc_state(2, src, n, q5, GAMMA)
    assert np.all(ns[1] + src[1] == 0.0) and np.all(ns[2] + src[2] == 0.0)
    assert np.array_equal(ns[0], src[0]) and np.array_equal(ns[3], src[3]) and np.all(ns[4] == -src[4])
    sym = P.bc_state(3, src, n, q5, GAMMA)
    slip = P.bc_state(1, src, n, q5, GAMMA)
    assert np.array_equal(sym[4], src[4]) and np.array_equal(slip[4], -src[4])
    assert np.array_equal(sym[1], slip[1]) and np.array_equal(sym[2], slip[2])


@pytest.mark.parametrize("mach, alpha", [(0.5, 0.0), (0.8, 1.25), (0.3, -4.0)])
def test_far_field_returns_the_freestream_exactly_and_transmits_disturbances(mach, alpha):
    """Both Riemann invariants carry the same state, so a freestream interior
    gives the freestream ghost at any incidence; a disturbed interior pressure
    is transmitted, not clamped."""
    from jaxrans_amd.core.state import q_infinity
    q_inf = q_infinity(FlowConditions(mach=mach, alpha_deg=alpha))
    ang = np.linspace(0.0, 2.0 * np.pi, 17)
    n = (np.cos(ang), np.sin(ang))
    src = [np.full(17, v) for v in q_inf]
    out = P.bc_state(4, src, n, q_inf, GAMMA)
    for k in range(4):
        assert np.allclose(out[k], q_inf[k], rtol=1e-12, atol=1e-14), k
    src[3] = src[3] * 1.05
    out = P.bc_state(4, src, n, q_inf, GAMMA)
    assert np.any(np.abs(out[3] - q_inf[3]) > 1e-6 * q_inf[3])
    # outgoing at the boundary state (vn_b > 0) keeps the interior entropy, incoming the freestream's
    vn = out[1] * n[0] + out[2] * n[1]
    s_out = out[3] / out[0] ** GAMMA
    assert np.allclose(s_out[vn > 1e-3], src[3][vn > 1e-3] / src[0][vn > 1e-3] ** GAMMA, rtol=1e-12)
    assert np.allclose(s_out[vn < -1e-3], q_inf[3] / q_inf[0] ** GAMMA, rtol=1e-12)


def test_inflow_total_and_outflow_static_closed_forms():
    """Subsonic inflow: freestream total temperature and total pressure and
    the freestream direction, the interior's static pressure; it returns the
    freestream exactly when the interior pressure is the freestream's.
    Static outflow: the freestream pressure, everything else copied."""
    from jaxrans_amd.core.state import q_infinity
    q_inf = np.concatenate([q_infinity(FlowConditions(mach=0.4, alpha_deg=3.0)), [3.0]])
    n = (np.array([-1.0, -1.0, -1.0]), np.array([0.0, 0.0, 0.0]))
    p = np.array([q_inf[3], 0.97 * q_inf[3], 1.01 * q_inf[3]])
    src = [np.full(3, q_inf[0]) * 1.02, np.full(3, 0.1), np.full(3, 0.02), p, np.full(3, 2.0)]
    out = P.bc_state(5, src, n, q_inf, GAMMA)
    assert np.allclose([o[0] for o in out], q_inf, rtol=1e-13)
    g1 = GAMMA - 1.0
    T0 = GAMMA * q_inf[3] / q_inf[0] * (1.0 + 0.5 * g1 * 0.4 ** 2)
    p0 = q_inf[3] * (1.0 + 0.5 * g1 * 0.4 ** 2) ** (GAMMA / g1)
    T = GAMMA * out[3] / out[0]
    speed2 = out[1] ** 2 + out[2] ** 2
    assert np.allclose(T + 0.5 * g1 * speed2, T0, rtol=1e-13)
    M2 = speed2 / T
    assert np.allclose(out[3] * (1.0 + 0.5 * g1 * M2) ** (GAMMA / g1), p0, rtol=1e-12)
    assert np.allclose(out[2] * q_inf[1], out[1] * q_inf[2], rtol=1e-13)      # freestream direction
    assert np.array_equal(out[3], p) and np.all(out[4] == 3.0)
    o = P.bc_state(6, src, n, q_inf, GAMMA)
    assert np.all(o[3] == q_inf[3]) and all(np.array_equal(o[k], src[k]) for k in (0, 1, 2, 4))


# --- wave speeds and chain blocks --------------------------------------------------------------------


def test_wave_speeds_closed_form_on_a_cartesian_grid():
    """Uniform state (u 0.5, v 0.2, a 1) on dx = 0.1, dy = 0.05 cells: an
    i-face has ``(|u| + a) dy = 0.075``, a j-face ``(|v| + a) dx = 0.12``,
    every interior cell sums 0.39; viscous adds ``2 gamma/Pr mu |S|^2 /
    (rho V)`` with ``V = 0.005``: 0.5 (2.8/0.72) mu on i-faces, 2 (2.8/0.72) mu
    on j-faces."""
    at = cartesian_atlas()
    q = np.array([1.0, 0.5, 0.2, 1.0 / GAMMA])
    U = OR.U0_of(OR.Setup(at, GAMMA, q))
    fi, fj = interior_faces(at)
    for mu in (None, 1e-3):
        st = OR.Setup(at, GAMMA, q, mu_inf=mu)
        li, lj = OR.wave_speeds(U, st)
        add_i = 0.0 if mu is None else 0.5 * 2.8 / 0.72 * mu
        add_j = 0.0 if mu is None else 2.0 * 2.8 / 0.72 * mu
        assert np.allclose(li[fi], 0.075 + add_i, rtol=1e-14)
        assert np.allclose(lj[fj], 0.12 + add_j, rtol=1e-14)
        s = OR.wave_speed_sum((li, lj))
        assert np.allclose(s[at.interior], 0.39 + 2 * add_i + 2 * add_j, rtol=1e-14)
    # face averages: a state linear in i gives the arithmetic mean of the two cells
    U2 = U.copy()
    i = np.arange(at.ni)[:, None] + 0 * np.arange(at.nj)[None]
    U2[0] = 1.0 + 0.01 * i
    lam_i = OR.wave_speeds(U2, OR.Setup(at, GAMMA, q))[0]
    k = at.g + 3
    rho_f = 1.0 + 0.01 * (k - 0.5)
    a = np.sqrt(GAMMA * (1.0 / GAMMA) / rho_f)        # p from E with the mean rho: see below
    qf = [0.5 * (np.asarray(P.primitive([U2[c][k - 1, 4] for c in range(4)], GAMMA))[m]
                 + np.asarray(P.primitive([U2[c][k, 4] for c in range(4)], GAMMA))[m]) for m in range(4)]
    want = abs(qf[1] * 0.05) + np.sqrt(GAMMA * qf[3] / qf[0]) * 0.05
    assert lam_i[k, 4] == pytest.approx(want, rel=1e-14) and a > 0


def _fd_jacobian(q, S, h=1e-6):
    """``d(F(U) . S)/dU`` by central differences of the analytic normal flux."""
    U = np.array(P.conservative(list(q[:4]), GAMMA), np.float64)
    A = np.zeros((4, 4))
    for k in range(4):
        e = np.zeros(4)
        e[k] = h * max(1.0, abs(U[k]))
        fp = np.array(P.normal_flux(P.primitive(list(U + e), GAMMA), S, GAMMA))
        fm = np.array(P.normal_flux(P.primitive(list(U - e), GAMMA), S, GAMMA))
        A[:, k] = (fp - fm) / (2 * e[k])
    return A


def test_flow_chain_blocks_are_the_rusanov_jacobian_along_the_chain():
    """The line preconditioner's blocks are the j-line part of ``V/dt + d R_rus
    / dU`` for the first-order Rusanov residual ``R_rus = sum_f (F(U_c) +
    F(U_nb)) . S_f / 2 - lambda_f (U_nb - U_c) / 2`` (outward ``S_f``, frozen
    ``lambda``): off-diagonal ``A(U_nb, S_out)/2 - lambda_f/2``, diagonal
    ``sum lambda (1/CFL + 1/2)`` (closed cells: ``sum_f A(U_c, S_f) = 0``).
    Checked on every link of three chains against central differences."""
    at = o_atlas()
    st = OR.setup_for(at, FlowConditions(mach=0.5, alpha_deg=1.25), NumericsConfig())
    U = perturbed(OR.U0_of(st), 1e-2)
    lams = OR.wave_speeds(U, st)
    cfl = 7.0
    lower, diag, upper = OS.flow_chain_blocks(U, lams, cfl, st)
    q = np.stack(P.primitive(OR.rows(U), GAMMA))[:4].reshape(4, -1)
    lsum = OR.wave_speed_sum(lams).reshape(-1)
    lam_j = lams[1].reshape(-1)
    Sj = at.face_j.reshape(2, -1)
    L = at.lines
    for c in (0, L.n_lines // 3, L.n_lines - 1):
        s0, s1 = L.start[c], L.start[c + 1]
        for k in range(s1 - s0):
            cell = L.cells[s0 + k]
            assert np.allclose(diag[k, c], lsum[cell] * (1.0 / cfl + 0.5) * np.eye(4), rtol=1e-14)
            if k + 1 < s1 - s0:
                f, sg = L.up_face[s0 + k], L.up_sign[s0 + k]
                nb = L.cells[s0 + k + 1]
                S_out = (sg * Sj[0][f], sg * Sj[1][f])
                want = 0.5 * _fd_jacobian(q[:, nb], S_out) - 0.5 * lam_j[f] * np.eye(4)
                assert np.allclose(upper[k, c], want, rtol=1e-6, atol=1e-9 * np.abs(want).max()), (c, k)
                f2, sg2 = L.down_face[s0 + k], L.down_sign[s0 + k]
                S_back = (sg2 * Sj[0][f2], sg2 * Sj[1][f2])
                assert f2 == f and sg2 == -sg
                want = 0.5 * _fd_jacobian(q[:, cell], S_back) - 0.5 * lam_j[f] * np.eye(4)
                assert np.allclose(lower[k + 1, c], want, rtol=1e-6, atol=1e-9 * np.abs(want).max()), (c, k)
            else:
                assert np.all(upper[k, c] == 0.0)
        assert np.all(lower[0, c] == 0.0)


def test_sa_chain_closed_form_on_a_cartesian_cell():
    """Uniform state (rho 1, u 0.5, v 0.2, nu^ 3) on dx = 0.1, dy = 0.05: along
    a j-chain the outward mass fluxes are ``+-rho v dx = +-0.02``, the thin-layer
    diffusion ``kappa |S_j|^2 / V = kappa 0.01 / 0.005 = 2 kappa`` with
    ``kappa = (mu + rho nu~)/sigma``; upper ``-(0 + 2 kappa) / rho``, lower
    ``-(0.02 + 2 kappa) / rho`` (upwind from below), diagonal ``(out + in) /
    rho`` summed over the four faces + destruction ``2 c_w1 f_w nu_inf nu^ /
    d^2 V`` + ``sum lambda / CFL``."""
    at = cartesian_atlas(ni=6, nj=8)
    nu_inf = 1e-3
    q = np.array([1.0, 0.5, 0.2, 1.0 / GAMMA, 3.0])
    st = OR.Setup(at, GAMMA, q, mu_inf=nu_inf, sa=True)
    U = OR.U0_of(st)
    lams = OR.wave_speeds(U, st)
    cfl = 5.0
    lower, diag, upper = OS.sa_chain_blocks(U, lams, cfl, st)
    kappa = (nu_inf + 3.0 * nu_inf) / (2.0 / 3.0)
    d_j, d_i = 2.0 * kappa, kappa * 0.05 ** 2 / 0.005
    m_j, m_i = 0.5 * 0.0 + 0.2 * 0.1, 0.5 * 0.05
    L = at.lines
    c = L.n_lines // 2
    s0, s1 = L.start[c], L.start[c + 1]
    lsum = OR.wave_speed_sum(lams).reshape(-1)
    for k in range(2, s1 - s0 - 2):
        cell = L.cells[s0 + k]
        assert upper[k, c, 0, 0] == pytest.approx(-(0.0 + d_j), rel=1e-13)
        assert lower[k, c, 0, 0] == pytest.approx(-(m_j + d_j), rel=1e-13)
        # out + in over the four faces: i-faces (m_i, d_i), j-faces (m_j, d_j)
        conv = (m_i + d_i) + (0.0 + d_i) + (m_j + d_j) + (0.0 + d_j)
        chi = 3.0
        d = at.distance.reshape(-1)[cell]
        st_t = P.s_tilde(0.0, 3.0 * nu_inf, nu_inf, d)
        fw = P.f_w(P.r_of(3.0 * nu_inf, st_t, d))
        destr = 2.0 * P.CW1 * fw * nu_inf * chi / d ** 2 * 0.005
        assert diag[k, c, 0, 0] == pytest.approx(conv + destr + lsum[cell] / cfl, rel=1e-12)


# --- Newton control -------------------------------------------------------------------------------------


def test_relaxation_closed_form():
    """``omega = min(1, fraction / worst relative change)`` over density,
    pressure and nu~ (relative to max(|nu^|, 1)); pressure change from
    ``dE - u dm - v dn + |v|^2/2 drho`` (implicit.py:54-81)."""
    U = np.array(P.conservative([1.0, 0.5, 0.1, 1.0 / GAMMA, 0.5], GAMMA), np.float64)[:, None, None] * np.ones((1, 2, 3))
    dU = np.zeros_like(U)
    dU[0, 1, 2] = 0.4                           # 40 % density change, energy held
    _, omega = OS.relax(dU, U, GAMMA, 0.2)
    # dp = (gamma - 1)(0 - 0 - 0 + 0.5 * 0.26 * 0.4) = 0.4 * 0.052 = 0.0208; rel 0.0208 * 1.4 = 0.02912
    assert omega == pytest.approx(0.2 / 0.4, rel=1e-14)
    dU[:] = 0.0
    dU[3, 0, 0] = -0.5                           # energy only: dp = 0.4 * -0.5 = -0.2, rel 0.28
    _, omega = OS.relax(dU, U, GAMMA, 0.2)
    assert omega == pytest.approx(0.2 / (0.2 * GAMMA), rel=1e-14)
    dU[:] = 0.0
    dU[4, 1, 1] = 3.0                            # nu^ 0.5 -> change 3 / max(0.5, 1) = 3
    _, omega = OS.relax(dU, U, GAMMA, 0.2)
    assert omega == pytest.approx(0.2 / 3.0, rel=1e-14)
    dU[:] = 1e-6
    dU2, omega = OS.relax(dU, U, GAMMA, 0.2)
    assert omega == 1.0 and np.array_equal(dU2, dU)
    assert OS.relax(dU, U, GAMMA, None)[1] == 1.0


@pytest.mark.parametrize("accept, lin_red, omega, reduction, want", [
    (True, 0.1, 1.0, 2.0, 13.0),      # growth capped at 1.3
    (True, 0.1, 1.0, 1.1, 11.0),      # grows by the residual reduction
    (True, 0.1, 1.0, 0.5, 10.0),      # never shrinks on an accepted step
    (True, 0.7, 1.0, 2.0, 10.0),      # linear solve not good enough: held
    (True, 0.1, 0.5, 2.0, 10.0),      # relaxed step: held
    (False, 0.1, 1.0, 2.0, 2.5),      # rejected: cut back by 4
])
def test_ser_cfl_rule(accept, lin_red, omega, reduction, want):
    """Switched evolution relaxation with the reference defaults (growth 1.3,
    cutback 4, success 0.5; implicit.py:139-166)."""
    cfg = ImplicitConfig()
    assert OS.next_cfl(10.0, accept, lin_red, omega, reduction, cfg) == pytest.approx(want, rel=1e-15)


def test_ser_cfl_is_clipped():
    cfg = ImplicitConfig(cfl_max=12.0, cfl_min=0.5)
    assert OS.next_cfl(10.0, True, 0.1, 1.0, 5.0, cfg) == 12.0
    assert OS.next_cfl(1.0, False, 0.1, 1.0, 5.0, cfg) == 0.5


@pytest.fixture(scope="module")
def implicit_converged():
    at = o_atlas(32, 16)
    flow = FlowConditions(mach=0.5, alpha_deg=2.0)
    st = OR.setup_for(at, flow, NumericsConfig(limiter="none"))
    cfg = ImplicitConfig(cfl=5.0, cfl_growth=1.5, jacobian="exact", krylov_subspace=20, krylov_restarts=3,
                         krylov_tol=1e-3, line_solver="thomas")
    U0 = OR.U0_of(st)
    s = OS.NewtonState(U0, cfg.cfl)
    hist = []
    for _ in range(30):
        s, n, _ = OS.newton_step(s, st, cfg, exact=True)
        hist.append(n)
    return st, U0, s, np.asarray(hist)


def test_implicit_fixed_point_is_the_explicit_fixed_point(implicit_converged):
    """Both paths drive the same residual, so they share fixed points: an
    explicit RK5 step at the state the Newton path reaches moves it in
    proportion to the residual that remains, not by the distance to some
    other fixed point (reference ``test_implicit.py:56-95``)."""
    from oracle import explicit as OX
    st, U0, s, hist = implicit_converged
    drop = hist[-1] / hist.max()
    assert drop < 1e-2, hist

    def drift(U):
        stepped, _ = OX.rk_step(U, st, OX.RK_SCHEMES["rk5_jameson"], 1.0)
        return np.linalg.norm((stepped - U)[:, st.atlas.interior])

    ratio = drift(s.U) / drift(U0)
    assert ratio < 10.0 * hist[-1] / hist[0], (ratio, hist[-1] / hist[0])


def test_both_paths_report_the_same_norm():
    """The Newton step reports the norm an explicit step of the same state
    reports (reference ``test_implicit.py:98-112``)."""
    from oracle import explicit as OX
    at = o_atlas(32, 16)
    st = OR.setup_for(at, FlowConditions(mach=0.8, alpha_deg=1.25), NumericsConfig(limiter="minmod"))
    U = perturbed(OR.U0_of(st), 1e-3)
    _, n_imp, _ = OS.newton_step(OS.NewtonState(U, 10.0), st, ImplicitConfig())
    _, n_exp = OX.rk_step(U, st, OX.RK_SCHEMES["rk5_jameson"], 1.0)
    assert n_imp == pytest.approx(n_exp, rel=1e-13)


# --- forces ----------------------------------------------------------------------------------------------


def test_forces_closed_form_on_a_flat_plate():
    """A plate ``0 <= x <= 1`` at ``y = 0`` (no-slip wall under a Cartesian
    block), alpha 0: a wall-cell pressure ``1/gamma + c x`` loads the plate
    downward by ``-c sum x_i dx`` (lift), with the nose-up moment about the
    quarter chord ``sum (x_i - 1/4) c x_i dx``; a Couette velocity ``u = k y``
    drags it by ``mu k`` per unit length (all per ``q = M^2 / 2``)."""
    ni, nj, dx, dy = 10, 6, 0.1, 0.05
    at = cartesian_atlas(ni, nj, dx, dy)
    M, mu, k, c = 0.5, 2e-3, 0.3, 0.05
    q_inf = np.array([1.0, M, 0.0, 1.0 / GAMMA])
    xs = (np.arange(ni) + 0.5) * dx
    qdyn = 0.5 * M * M
    # pressure only (inviscid): p = 1/gamma + c x, velocity freestream
    prim = canvas_rows(at, lambda x, y: [1.0 + 0 * x, M + 0 * x, 0 * x, 1.0 / GAMMA + c * x])
    U = np.stack(P.conservative(prim, GAMMA))
    f = OS.forces(U, OR.Setup(at, GAMMA, q_inf), 1.0, (0.25, 0.0))
    assert f["cl"] == pytest.approx(-c * np.sum(xs * dx) / qdyn, rel=1e-13)
    assert abs(f["cd"]) < 1e-15 and f["cd_friction"] == 0.0
    assert f["cm"] == pytest.approx(np.sum((xs - 0.25) * c * xs * dx) / qdyn, rel=1e-13)
    # shear only: u = k y, uniform p, no-slip ghosts are the odd reflection (u(-y) = -u(y))
    prim = canvas_rows(at, lambda x, y: [1.0 + 0 * x, k * y, 0 * x, 1.0 / GAMMA + 0 * x])
    U = np.stack(P.conservative(prim, GAMMA))
    f = OS.forces(U, OR.Setup(at, GAMMA, q_inf, mu_inf=mu), 1.0, (0.25, 0.0))
    assert f["cd_friction"] == pytest.approx(mu * k * 1.0 / qdyn, rel=1e-12)
    assert f["cd"] == pytest.approx(f["cd_friction"], rel=1e-14)
    assert abs(f["cl"]) < 1e-14 and abs(f["cm"]) < 1e-14


def test_forces_rotate_with_the_freestream():
    """The same wall load seen at incidence alpha: lift and drag are its
    components normal and along the freestream (a rotation of (fx, fy))."""
    at = cartesian_atlas(10, 6, 0.1, 0.05)
    prim = canvas_rows(at, lambda x, y: [1.0 + 0 * x, 0.5 + 0 * x, 0 * x, 1.0 / GAMMA + 0.05 * x])
    U = np.stack(P.conservative(prim, GAMMA))
    a = np.deg2rad(3.0)
    q_inf = np.array([1.0, 0.5 * np.cos(a), 0.5 * np.sin(a), 1.0 / GAMMA])
    f = OS.forces(U, OR.Setup(at, GAMMA, q_inf), 1.0, (0.25, 0.0))
    qdyn = 0.5 * 0.25
    assert f["cl"] == pytest.approx((-np.sin(a) * f["fx"] + np.cos(a) * f["fy"]) / qdyn, rel=1e-13)
    assert f["cd"] == pytest.approx((np.cos(a) * f["fx"] + np.sin(a) * f["fy"]) / qdyn, rel=1e-13)
