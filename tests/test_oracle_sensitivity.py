"""The oracle's parameter tangent (``oracle/sensitivity.py``) pinned by
central differences: the freestream's tangent against differences of
``q_infinity``, the residual's parameter JVP against Richardson differences
of the oracle residual at shifted freestreams (Euler, laminar with
Sutherland, SA, the point-vortex far field, total-pressure inflow and
static outflow), the force coefficients' JVP along a state and a parameter
direction at once against differences of the oracle's forces, and the
defect correction against its own linear residual.  A dropped term (the
``U - U0`` rows, ``mu_inf``'s tangent in the viscous or SA terms, the vortex
reference turning with the flow angle, the lift direction turning with it)
fails one of them.  No GPU."""

from __future__ import annotations

from dataclasses import replace

import numpy as np
import pytest

from conftest import c_atlas, o_atlas, perturbed
from grids3d import small_wing
from jaxrans_amd.core.config import FlowConditions, ImplicitConfig, NumericsConfig
from jaxrans_amd.core.state import q_infinity
from jaxrans_amd.mesh.atlas import assign_conditions, build_atlas
from jaxrans_amd.mesh.block import Grid, rectangle_sides
from oracle import residual as OR
from oracle import residual3d as OR3
from oracle import sensitivity as SE
from oracle import solver3d as OS3
from oracle import solver as OS

PARAMETER_NAMES = ("alpha", "mach")
STEP = {"alpha": 1e-4, "mach": 1e-6}   # within the smooth limiter's curvature; no far-field face changes inflow/outflow


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def shifted(flow, name, t):
    key = "alpha_deg" if name == "alpha" else "mach"
    return replace(flow, **{key: getattr(flow, key) + t})


def richardson(f, h):
    d_h = (f(h) - f(-h)) / (2.0 * h)
    d_half = (f(0.5 * h) - f(-0.5 * h)) / h
    return (4.0 * d_half - d_h) / 3.0


def duct():
    ni, nj = 24, 12
    x = np.arange(ni + 1)[:, None] * 0.1 + 0.0 * np.arange(nj + 1)[None]
    y = 0.0 * np.arange(ni + 1)[:, None] + np.arange(nj + 1)[None] * (0.05 + 0.002 * np.arange(ni + 1)[:, None])
    sides = rectangle_sides(ni, nj, j_min="lower", j_max="upper", i_min="inlet", i_max="outlet")
    mesh = assign_conditions((Grid(x, y, sides),), {"lower": "slip_wall", "upper": "slip_wall",
                                                    "inlet": "inflow_total", "outlet": "outflow_static"})
    return build_atlas(mesh)


SUTH = 198.6 / 460.0
CASES = {
    "euler-hllc-vas": (lambda: o_atlas(), FlowConditions(mach=0.5, alpha_deg=1.25),
                       NumericsConfig(flux="hllc", limiter="van_albada_smooth")),
    "laminar-sutherland": (lambda: o_atlas(wall="no_slip_wall"),
                           FlowConditions(mach=0.5, alpha_deg=1.0, reynolds=5e3, sutherland_ratio=SUTH),
                           NumericsConfig(flux="hllc", limiter="van_albada_smooth")),
    "sa-roe": (lambda: c_atlas(), FlowConditions(mach=0.73, alpha_deg=2.5, reynolds=6.5e6, turbulence="sa_noft2"),
               NumericsConfig(flux="roe", limiter="van_albada_smooth")),
    "vortex": (lambda: o_atlas(), FlowConditions(mach=0.6, alpha_deg=3.0),
               NumericsConfig(flux="hllc", limiter="van_albada_smooth", far_field="characteristic_point_vortex")),
    "duct-inflow-outflow": (duct, FlowConditions(mach=0.3, alpha_deg=2.0, reynolds=2e3),
                            NumericsConfig(flux="roe", limiter="none")),
}


def state(atlas, flow, num, amplitude=1e-2):
    st = OR.setup_for(atlas, flow, num)
    return st, perturbed(OR.U0_of(st), amplitude)


@pytest.mark.parametrize("name", PARAMETER_NAMES)
@pytest.mark.parametrize("flow", [FlowConditions(mach=0.5, alpha_deg=1.25),
                                  FlowConditions(mach=0.73, alpha_deg=2.5, reynolds=6.5e6, turbulence="sa_noft2")])
def test_freestream_tangent_is_the_derivative_of_q_infinity(flow, name):
    nvar = 5 if flow.turbulence else 4
    q, mu = SE.freestream_dual(flow, name, nvar)
    dq, dmu = SE.tangents(q, mu)
    assert np.array_equal([float(x.v) for x in q], q_infinity(flow))
    fd = richardson(lambda t: q_infinity(shifted(flow, name, t)), STEP[name])
    assert np.allclose(dq, fd, rtol=1e-9, atol=1e-12)
    if not flow.inviscid:
        assert float(mu.v) == pytest.approx(flow.mu_inf, rel=1e-15)
        fd_mu = richardson(lambda t: np.array([shifted(flow, name, t).mu_inf]), STEP[name])[0]
        assert dmu == pytest.approx(fd_mu, rel=1e-9, abs=1e-20)


@pytest.mark.parametrize("case", sorted(CASES))
def test_residual_parameter_jvp_matches_differences(case):
    make, flow, num = CASES[case]
    atlas = make()
    st, U = state(atlas, flow, num)
    nvar = U.shape[0]
    for name in PARAMETER_NAMES:
        q, mu = SE.freestream_dual(flow, name, nvar)
        got = SE.residual_parameter_jvp(U, st, q, mu)
        fd = richardson(lambda t: OR.residual(U, OR.setup_for(atlas, shifted(flow, name, t), num)), STEP[name])
        err = rel(got, fd)
        print(f"{case} dR/d{name}: |jvp| {np.linalg.norm(got):.3e}  rel to differences {err:.2e}")
        assert err < 1e-7, (name, err)
        # off the interior the tangent is -dU0/dp: the rows hold U - U0
        off = ~atlas.interior
        dU0 = richardson(lambda t: OR.U0_of(OR.setup_for(atlas, shifted(flow, name, t), num)), STEP[name])
        assert np.allclose(got[:, off], -dU0[:, off], rtol=1e-8, atol=1e-13)


@pytest.mark.parametrize("case", ["euler-hllc-vas", "laminar-sutherland", "vortex"])
def test_forces_jvp_matches_differences(case):
    make, flow, num = CASES[case]
    atlas = make()
    st, U = state(atlas, flow, num)
    rng = np.random.default_rng(3)
    dU = rng.standard_normal(U.shape) * np.abs(U).mean(axis=(1, 2), keepdims=True) * 1e-2
    for name in PARAMETER_NAMES:
        q, mu = SE.freestream_dual(flow, name, U.shape[0])
        got = SE.forces_jvp(U, dU, st, q, mu)
        h = STEP[name]

        def coefficients(t):
            f = OS.forces(U + t * dU, OR.setup_for(atlas, shifted(flow, name, t), num))
            return np.array([f[k] for k in ("cl", "cd", "cm", "cd_friction", "fx", "fy")])

        fd = richardson(coefficients, h)
        vec = np.array([got[k] for k in ("cl", "cd", "cm", "cd_friction", "fx", "fy")])
        print(f"{case} dJ/d{name}: {vec}  differences {fd}")
        assert np.allclose(vec, fd, rtol=1e-7, atol=1e-10 * np.abs(fd).max()), (name, vec, fd)


def test_defect_correction_solves_a_linear_system():
    """x <- x + G (b - A x) with G a fixed approximate inverse converges to
    A^-1 b when the spectral radius of I - G A is below one, and stops at
    its tolerance (solvers/sensitivity.py:34-50)."""
    rng = np.random.default_rng(5)
    n = 40
    A = np.eye(n) * 4.0 + rng.standard_normal((n, n)) * 0.3
    G = np.diag(1.0 / np.diag(A))
    b = rng.standard_normal(n)
    assert max(abs(np.linalg.eigvals(np.eye(n) - G @ A))) < 1.0
    x, err, steps = SE.defect_correction(lambda v: A @ v, b, lambda r: G @ r, 1e-12, 1000)
    assert err <= 1e-12 and 0 < steps < 1000
    assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-10, atol=1e-12)
    _, err0, steps0 = SE.defect_correction(lambda v: A @ v, b, lambda r: G @ r, 1e-12, 3)
    assert steps0 == 3 and err0 > 1e-12


def test_newton_inverse_solves_the_pseudo_time_system():
    """G r from ``newton_inverse``: GMRES on V/dt + dR~/dU at cfl_max with
    R~ the first-order residual, line-preconditioned, reduces
    |r - (V/dt + J~) G r| to its tolerance (derivatives.py:94-120).  (With
    the exact Jacobian the same Krylov solve stalls, which is why the
    sensitivity iterates it by defect correction.)"""
    make, flow, num = CASES["euler-hllc-vas"]
    atlas = make()
    st, U = state(atlas, flow, num, 1e-3)
    vdt = OR.wave_speed_sum(OR.wave_speeds(U, st))
    r = np.random.default_rng(9).standard_normal(U.shape) * atlas.interior[None]
    cfg = ImplicitConfig(krylov_subspace=30, krylov_restarts=10, krylov_tol=1e-6, cfl_max=50.0)
    x = SE.newton_inverse(U, st, cfg)(r)
    err = rel(vdt[None] / cfg.cfl_max * x + OR.jvp(U, x, st, first_order=True), r)
    print(f"newton_inverse: |r - A G r| / |r| = {err:.2e}")
    assert err <= 1e-6


WING_FLOWS = {
    "wing-euler": (FlowConditions(mach=0.6, alpha_deg=2.0), "slip_wall"),
    "wing-sa": (FlowConditions(mach=0.6, alpha_deg=2.0, reynolds=1e6, sutherland_ratio=SUTH, turbulence="sa_noft2"),
                "no_slip_wall"),
}


@pytest.mark.parametrize("case", sorted(WING_FLOWS))
def test_3d_parameter_jvps_match_differences(case):
    """3D: the residual's parameter JVP and the force coefficients' JVP (per
    reference area, moment about y) on a small four-block wing against
    Richardson differences at shifted freestreams (velocity M (cos a, 0,
    sin a))."""
    flow, wall = WING_FLOWS[case]
    atlas = small_wing(wall=wall)
    num = NumericsConfig(flux="hllc", limiter="van_albada_smooth")
    st = OR3.setup_for(atlas, flow, num)
    U = perturbed(OR3.U0_of(st), 1e-2)
    rng = np.random.default_rng(5)
    dU = rng.standard_normal(U.shape) * np.abs(U).mean(axis=tuple(range(1, U.ndim)), keepdims=True) * 1e-2
    for name in PARAMETER_NAMES:
        q, mu = SE.freestream_dual(flow, name, U.shape[0], dim=3)
        assert np.allclose([float(x.v) for x in q], st.q_inf, rtol=1e-15, atol=0.0)
        got = SE.residual_parameter_jvp(U, st, q, mu)
        fd = richardson(lambda t: OR3.residual(U, OR3.setup_for(atlas, shifted(flow, name, t), num)), STEP[name])
        err = rel(got, fd)
        print(f"{case} dR/d{name}: rel to differences {err:.2e}")
        assert err < 1e-7, (name, err)
        f = SE.forces_jvp(U, dU, st, q, mu, 1.0, (0.25, 0.0, 0.0), area=2.0)

        def coefficients(t):
            c = OS3.forces(U + t * dU, OR3.setup_for(atlas, shifted(flow, name, t), num), 2.0, 1.0, (0.25, 0.0, 0.0))
            return np.array([c[k] for k in ("cl", "cd", "cm", "cd_friction", "cy")])

        ref = richardson(coefficients, STEP[name])
        vec = np.array([f[k] for k in ("cl", "cd", "cm", "cd_friction", "cy")])
        print(f"{case} dJ/d{name}: {vec}  differences {ref}")
        assert np.allclose(vec, ref, rtol=1e-7, atol=1e-10 * np.abs(ref).max()), (name, vec, ref)
