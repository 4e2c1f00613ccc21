"""Device paths the benchmark cases do not take, against the oracle on the
same seeded inputs: the QCR2000 stress, kappa-MUSCL faces, the subsonic
total-pressure inflow and static-pressure outflow conditions, and the
rejected-step branch of the Newton control (cut-back of the CFL, state kept).
Tolerances as in test_gpu_parity (residual 1e-12, tangent 1e-11 normwise)."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import c_atlas, o_atlas, perturbed
from jaxrans_amd.core.config import FlowConditions, ImplicitConfig, NumericsConfig
from jaxrans_amd.device import context_for
from jaxrans_amd.mesh.atlas import assign_conditions, build_atlas
from jaxrans_amd.mesh.block import Grid, rectangle_sides
from oracle import residual as OR
from oracle import solver as OS

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _check(atlas, flow, num, amplitude=1e-2, seed=7):
    st = OR.setup_for(atlas, flow, num)
    ctx = context_for(atlas, flow, num)
    ctx.set_freestream(st.q_inf, None if flow.inviscid else flow.mu_inf)
    U = perturbed(OR.U0_of(st), amplitude)
    for first_order in (False, True):
        ref = OR.residual(U, st, first_order)
        got = ctx.residual(U, first_order)[0]
        assert rel(got, ref) < 1e-12, (first_order, rel(got, ref))
    rng = np.random.default_rng(seed)
    dU = rng.standard_normal(U.shape) * np.abs(U).mean(axis=(1, 2), keepdims=True) * 1e-3
    for first_order in (True, False):
        ref = OR.jvp(U, dU, st, first_order)
        got = ctx.tangent(U, dU, first_order=first_order)[0]
        assert rel(got, ref) < 1e-11, (first_order, rel(got, ref))
    ctx.close()


def test_qcr_stress_matches_oracle():
    print("[gpu] paths qcr", flush=True)
    flow = FlowConditions(mach=0.73, alpha_deg=2.5, reynolds=6.5e6, turbulence="sa_noft2_qcr")
    _check(c_atlas(), flow, NumericsConfig(flux="roe", limiter="van_albada_smooth"))


@pytest.mark.parametrize("kappa", [1.0 / 3.0, -1.0])
def test_kappa_muscl_matches_oracle(kappa):
    print(f"[gpu] paths kappa {kappa}", flush=True)
    _check(o_atlas(), FlowConditions(mach=0.5, alpha_deg=1.25),
           NumericsConfig(flux="hllc", limiter="none", kappa=kappa))


def _duct():
    """A Cartesian duct: subsonic total-pressure inflow on i_min, static
    outflow on i_max, slip walls below and above."""
    ni, nj = 24, 12
    x = np.arange(ni + 1)[:, None] * 0.1 + 0.0 * np.arange(nj + 1)[None]
    y = 0.0 * np.arange(ni + 1)[:, None] + np.arange(nj + 1)[None] * (0.05 + 0.002 * np.arange(ni + 1)[:, None])
    sides = rectangle_sides(ni, nj, j_min="lower", j_max="upper", i_min="inlet", i_max="outlet")
    mesh = assign_conditions((Grid(x, y, sides),), {"lower": "slip_wall", "upper": "slip_wall",
                                                    "inlet": "inflow_total", "outlet": "outflow_static"})
    return build_atlas(mesh)


@pytest.mark.parametrize("viscous", [False, True])
def test_inflow_and_outflow_conditions_match_oracle(viscous):
    print(f"[gpu] paths duct viscous={viscous}", flush=True)
    flow = FlowConditions(mach=0.4, alpha_deg=0.0, reynolds=2e4 if viscous else None)
    _check(_duct(), flow, NumericsConfig(flux="hllc", limiter="van_albada"), amplitude=2e-2)


def test_rejected_steps_follow_the_oracle():
    """From a 30 % perturbation at CFL 1e4 the oracle's first step is accepted
    and the next three are rejected (non-physical candidates): the state is
    kept and the CFL cut by 4 each time.  The device takes the same decisions,
    reports the same norms and ends on the same state and CFL."""
    print("[gpu] paths reject", flush=True)
    atlas = o_atlas()
    flow = FlowConditions(mach=0.5, alpha_deg=1.25)
    num = NumericsConfig(flux="hllc", limiter="van_albada")
    cfg = ImplicitConfig(krylov_subspace=12, krylov_restarts=2, cfl=1e4, cfl_max=1e5)
    st = OR.setup_for(atlas, flow, num)
    U = perturbed(OR.U0_of(st), 0.3)
    s = OS.NewtonState(U, cfg.cfl)
    norms, accepted, cfls = [], [], []
    for _ in range(4):
        s, n, info = OS.newton_step(s, st, cfg)
        norms.append(n)
        accepted.append(bool(info["accept"]))
        cfls.append(s.cfl)
    assert accepted == [True, False, False, False], accepted
    ctx = context_for(atlas, flow, num, implicit=cfg)
    ctx.set_freestream(st.q_inf, None)
    ctx.set_state(U)
    ctx.newton_init(cfg.cfl)
    got = []
    for k in range(4):
        got.append(ctx.newton_block(1)[0, 0])
        sc = ctx.scalars()[0]
        assert bool(sc[5]) == accepted[k], (k, sc)
        assert sc[0] == pytest.approx(cfls[k], rel=1e-12), (k, sc[0], cfls[k])
    assert np.allclose(got, norms, rtol=1e-9, atol=0.0), (got, norms)
    assert ctx.scalars()[0, 7] == 3
    assert rel(ctx.get_state()[0] - U, s.U - U) < 1e-8
    ctx.close()


@pytest.mark.parametrize("which", ["o_euler", "c_sa"])
def test_point_vortex_far_field_matches_oracle(which):
    """The point-vortex far field (numerics/boundary.py:181-246): residual
    and tangent (through the circulation's dependence on the wall pressure)
    and four Newton steps against the oracle on a lifting state."""
    print(f"[gpu] paths point vortex {which}", flush=True)
    if which == "o_euler":
        atlas, flow = o_atlas(), FlowConditions(mach=0.6, alpha_deg=3.0)
        num = NumericsConfig(flux="hllc", limiter="van_albada", far_field="characteristic_point_vortex")
    else:
        atlas = c_atlas()
        flow = FlowConditions(mach=0.73, alpha_deg=2.5, reynolds=6.5e6, turbulence="sa_noft2")
        num = NumericsConfig(flux="roe", limiter="van_albada_smooth", far_field="characteristic_point_vortex")
    _check(atlas, flow, num)
    cfg = ImplicitConfig(krylov_subspace=12, krylov_restarts=2)
    st = OR.setup_for(atlas, flow, num)
    U = perturbed(OR.U0_of(st), 1e-2)
    s = OS.NewtonState(U, cfg.cfl)
    norms = []
    for _ in range(4):
        s, n, _info = OS.newton_step(s, st, cfg)
        norms.append(n)
    ctx = context_for(atlas, flow, num, implicit=cfg)
    ctx.set_freestream(st.q_inf, None if flow.inviscid else flow.mu_inf)
    ctx.set_state(U)
    ctx.newton_init(cfg.cfl)
    got = ctx.newton_block(4)[:, 0]
    assert np.allclose(got, norms, rtol=1e-9, atol=0.0), (got, norms)
    assert rel(ctx.get_state()[0] - U, s.U - U) < 1e-8
    ctx.close()
