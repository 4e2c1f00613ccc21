"""Sensitivities on the device: ``sensitivity()`` against central differences
of converged ``run()`` results, the end-to-end check of the implicit-function
derivative (``src/jaxrans/derivatives.py:123-147``; the reference's own test
compares with dense Jacobians, ``tests/unit/test_sensitivity.py:62-69``).

The case is the small Euler O-grid (subsonic, M 0.5), and the same grid
laminar at Re 500 (no-slip wall), with unlimited MUSCL
reconstruction (``limiter = "none"``), converged eleven orders, so that the
discrete solution is a smooth function of alpha and M and the converged
forces carry no iteration error.  (With ``van_albada_smooth`` the limiter's
1e-12 regularisation meets the nearly uniform far field, where cell
differences are of its size: there the linearisation at the state and the
differences of converged solutions part by 7% on the outer two cell layers,
in the CPU oracle as on the device, and dCL/dM differs by half.)
"""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from jaxrans_amd.core.config import FlowConditions, NumericsConfig

pytestmark = pytest.mark.gpu

CASE = Path(__file__).resolve().parent / "golden" / "cases" / "euler_o48"
SMOOTH = ("numerical.convective_fluxes.limiter=none", "numerical.convergence.tol=1e-11",
          "numerical.convergence.max_iters=3000")


LAMINAR = ("case.freestream.reynolds=500.0", 'case.boundary_conditions.wall={"type": "no_slip_wall"}')
FLOWS = {"euler": (), "laminar": LAMINAR}   # laminar: M also moves mu_inf = M / Re


def converged(flow: str, alpha: float | None = None, mach: float | None = None):
    from jaxrans_amd import load_case, run
    extra = SMOOTH + FLOWS[flow]
    if alpha is not None:
        extra += (f"case.freestream.alpha_deg={alpha!r}",)
    if mach is not None:
        extra += (f"case.freestream.mach={mach!r}",)
    result = run(load_case(CASE, extra))
    assert result.history.reason == "converged", result.history.summary()
    return result


@pytest.fixture(scope="module", params=sorted(FLOWS))
def base(request):
    return request.param, converged(request.param)


def test_derivatives_match_differences_of_converged_runs(base):
    from jaxrans_amd import sensitivity
    flow, base = base
    s = sensitivity(base, ("cl", "cd", "cm"), ("alpha", "mach"), tol=1e-12, maxiter=3000)
    print(f"[gpu] sensitivity: {s.summary()}  iterations {s.iterations}", flush=True)
    assert max(s.linear_residuals.values()) <= 1e-12
    # each parameter's solve starts from the case's own freestream and state,
    # whatever the parameters before it moved
    alone = sensitivity(base, ("cl",), ("mach",), tol=1e-12, maxiter=3000)
    assert alone.derivatives["cl"]["mach"] == pytest.approx(s.derivatives["cl"]["mach"], rel=1e-9)
    a0, m0 = base.case.flow.alpha_deg, base.case.flow.mach
    for name, step, kw in (("alpha", 0.01, "alpha"), ("mach", 1e-3, "mach")):
        centre = a0 if name == "alpha" else m0

        def difference(h):
            up = converged(flow, **{kw: centre + h}).forces
            down = converged(flow, **{kw: centre - h}).forces
            return {out: (getattr(up, out) - getattr(down, out)) / (2.0 * h) for out in ("cl", "cd", "cm")}

        coarse, fine = difference(step), difference(step / 2)
        for out in ("cl", "cd", "cm"):
            fd = (4.0 * fine[out] - coarse[out]) / 3.0   # Richardson: O(step^4)
            got = s.derivatives[out][name]
            print(f"[gpu]   {flow} d{out}/d{name}: sensitivity {got:+.10e}  converged differences {fd:+.10e}", flush=True)
            assert got == pytest.approx(fd, rel=1e-4, abs=1e-8), (out, name)


PT_CASES = {
    "euler-hllc-vas": ("o", FlowConditions(mach=0.5, alpha_deg=1.25),
                       NumericsConfig(flux="hllc", limiter="van_albada_smooth")),
    "laminar-sutherland": ("ol", FlowConditions(mach=0.5, alpha_deg=1.0, reynolds=5e3, sutherland_ratio=198.6 / 460.0),
                           NumericsConfig(flux="hllc", limiter="van_albada_smooth")),
    "sa-roe": ("c", FlowConditions(mach=0.73, alpha_deg=2.5, reynolds=6.5e6, turbulence="sa_noft2"),
               NumericsConfig(flux="roe", limiter="van_albada_smooth")),
    "qcr-hllc": ("c", FlowConditions(mach=0.73, alpha_deg=2.5, reynolds=6.5e6, turbulence="sa_noft2_qcr"),
                 NumericsConfig(flux="hllc", limiter="van_albada_smooth")),
    "vortex": ("o", FlowConditions(mach=0.6, alpha_deg=3.0),
               NumericsConfig(flux="hllc", limiter="van_albada_smooth", far_field="characteristic_point_vortex")),
}


@pytest.mark.parametrize("case", sorted(PT_CASES))
def test_parameter_tangents_match_the_oracle(case):
    """jr_tangent_freestream and jr_forces_tangent against the oracle's dual
    freestream (oracle/sensitivity.py) on a perturbed state: the residual's
    parameter JVP, with and without a state direction, to 1e-11 normwise and
    the force coefficients' JVP to 1e-10 relative."""
    from conftest import c_atlas, o_atlas, perturbed
    from jaxrans_amd.derivatives import freestream_tangent
    from jaxrans_amd.device import context_for
    from oracle import residual as OR
    from oracle import sensitivity as SE
    print(f"[gpu] parameter tangent {case}", flush=True)
    kind, flow, num = PT_CASES[case]
    atlas = {"o": lambda: o_atlas(), "ol": lambda: o_atlas(wall="no_slip_wall"), "c": lambda: c_atlas()}[kind]()
    st = OR.setup_for(atlas, flow, num)
    U = perturbed(OR.U0_of(st), 1e-2)
    rng = np.random.default_rng(13)
    dU = rng.standard_normal(U.shape) * np.abs(U).mean(axis=(1, 2), keepdims=True) * 1e-2
    ctx = context_for(atlas, flow, num)
    try:
        ctx.set_freestream(st.q_inf, None if flow.inviscid else flow.mu_inf)
        for name in ("alpha", "mach"):
            q, mu = SE.freestream_dual(flow, name, U.shape[0])
            dq, dmu = freestream_tangent(flow, 2, name)
            assert np.allclose(dq, SE.tangents(q, mu)[0], rtol=1e-14, atol=1e-17)
            ref = SE.residual_parameter_jvp(U, st, q, mu)
            got = ctx.tangent_freestream(U, dq, dmu)[0]
            e0 = np.linalg.norm(got - ref) / np.linalg.norm(ref)
            ref_u = SE.residual_parameter_jvp(U, st, q, mu, dU)
            got_u = ctx.tangent_freestream(U, dq, dmu, dU=dU)[0]
            e1 = np.linalg.norm(got_u - ref_u) / np.linalg.norm(ref_u)
            f_ref = SE.forces_jvp(U, dU, st, q, mu)
            f_got = ctx.forces_tangent(U, dq, dmu, dU=dU)[0]
            keys = ("cl", "cd", "cm", "cd_friction", "fx", "fy")
            fr = np.array([f_ref[k] for k in keys])
            e2 = np.abs(f_got[:6] - fr).max() / np.abs(fr).max()
            print(f"[gpu]   d/d{name}: residual {e0:.2e}, with dU {e1:.2e}, forces {e2:.2e}", flush=True)
            assert e0 < 1e-11 and e1 < 1e-11 and e2 < 1e-10, (name, e0, e1, e2)
        # the state tangent alone through the same kernels is jr_tangent's
        zero = np.zeros(5)
        got = ctx.tangent_freestream(U, zero, 0.0 if not flow.inviscid else None, dU=dU)[0]
        assert np.linalg.norm(got - OR.jvp(U, dU, st, first_order=False)) <= 1e-11 * np.linalg.norm(got)
    finally:
        ctx.close()


@pytest.mark.parametrize("case", ["wing-euler", "wing-sa"])
def test_3d_parameter_tangents_match_the_oracle(case):
    """3D: jr_tangent_freestream and jr_forces3_tangent against the 3D
    oracle's dual freestream on a perturbed state of a small four-block
    wing (slip wall, Euler; no-slip, SA with Sutherland)."""
    from grids3d import small_wing
    from jaxrans_amd.derivatives import freestream_tangent
    from jaxrans_amd.device import context_for
    from oracle import residual3d as OR3
    from oracle import sensitivity as SE
    print(f"[gpu] 3D parameter tangent {case}", flush=True)
    flow, wall = {
        "wing-euler": (FlowConditions(mach=0.6, alpha_deg=2.0), "slip_wall"),
        "wing-sa": (FlowConditions(mach=0.6, alpha_deg=2.0, reynolds=1e6, sutherland_ratio=198.6 / 460.0,
                                   turbulence="sa_noft2"), "no_slip_wall")}[case]
    atlas = small_wing(wall=wall)
    num = NumericsConfig(flux="hllc", limiter="van_albada_smooth")
    st = OR3.setup_for(atlas, flow, num)
    from conftest import perturbed
    U = perturbed(OR3.U0_of(st), 1e-2)
    rng = np.random.default_rng(17)
    dU = rng.standard_normal(U.shape) * np.abs(U).mean(axis=tuple(range(1, U.ndim)), keepdims=True) * 1e-2
    ctx = context_for(atlas, flow, num)
    try:
        ctx.set_freestream(st.q_inf, None if flow.inviscid else flow.mu_inf)
        for name in ("alpha", "mach"):
            q, mu = SE.freestream_dual(flow, name, U.shape[0], dim=3)
            dq, dmu = freestream_tangent(flow, 3, name)
            assert np.allclose(dq, SE.tangents(q, mu)[0], rtol=1e-14, atol=1e-17)
            ref = SE.residual_parameter_jvp(U, st, q, mu, dU)
            got = ctx.tangent_freestream(U, dq, dmu, dU=dU)[0]
            e0 = np.linalg.norm(got - ref) / np.linalg.norm(ref)
            f_ref = SE.forces_jvp(U, dU, st, q, mu, 1.0, (0.25, 0.0, 0.0), area=2.0)
            f_got = ctx.forces3_tangent(U, dq, dmu, dU=dU, area=2.0, length=1.0, centre=(0.25, 0.0, 0.0))[0]
            keys = ("cl", "cd", "cm", "cd_friction", "fx", "fy", "fz", "cy")
            fr = np.array([f_ref[k] for k in keys])
            e1 = np.abs(f_got - fr).max() / np.abs(fr).max()
            print(f"[gpu]   d/d{name}: residual {e0:.2e}, forces {e1:.2e}", flush=True)
            assert e0 < 1e-11 and e1 < 1e-10, (name, e0, e1)
    finally:
        ctx.close()


def test_sensitivity_matches_the_oracle_linearisation():
    """sensitivity() with the case's own smooth limiter against the oracle's
    sensitivity solve at the same converged state: the exact parameter
    tangents, the defect correction with the Newton step's GMRES, and the
    output tangents, to 1e-8 relative."""
    from jaxrans_amd import load_case, run, sensitivity
    from oracle import residual as OR
    from oracle import sensitivity as SE
    case = load_case(CASE, ("numerical.convective_fluxes.limiter=van_albada_smooth", "numerical.convergence.tol=1e-11",
                            "numerical.convergence.max_iters=3000"))
    result = run(case)
    print(f"[gpu] sensitivity vs oracle: {result.history.summary()}", flush=True)
    assert result.history.reason == "converged"
    s = sensitivity(result, ("cl", "cd", "cm"), ("alpha", "mach"), tol=1e-11, maxiter=3000)
    st = OR.setup_for(result.grid, case.flow, case.numerics)
    spec = case.grid_spec
    d, errors, steps = SE.sensitivity(result.U, st, case.flow, case.implicit, ("cl", "cd", "cm"), ("alpha", "mach"),
                                      tol=1e-11, maxiter=3000, length=spec.reference_length,
                                      centre=tuple(spec.moment_centre[:2]))
    print(f"[gpu]   device steps {s.iterations}, oracle steps {steps}", flush=True)
    for out in ("cl", "cd", "cm"):
        for name in ("alpha", "mach"):
            got, ref = s.derivatives[out][name], d[out][name]
            print(f"[gpu]   d{out}/d{name}: device {got:+.12e}  oracle {ref:+.12e}", flush=True)
            assert got == pytest.approx(ref, rel=1e-8, abs=1e-12), (out, name)


def test_state_and_result_are_left_unchanged(base):
    from jaxrans_amd import sensitivity
    base = base[1]
    U = base.U.copy()
    sensitivity(base, ("cl",), ("alpha",), tol=1e-8)
    assert np.array_equal(base.U, U)


def test_linear_solve_is_the_newton_step_solve(base):
    """jr_linear_solve(-R(U)) is the update of one Newton step from U at the
    same CFL (same GMRES, same launches): U_next - U to round-off."""
    from jaxrans_amd.simulation import freestream_of, make_context
    base = base[1]
    case, grid = base.case, base.grid
    ctx = make_context(case, grid)
    try:
        q, mu = freestream_of(case.flow, grid.dim)
        ctx.set_freestream(q, mu)
        rng = np.random.default_rng(7)
        U1 = base.U * (1.0 + 1e-3 * rng.standard_normal(base.U.shape) * grid.interior[None])
        ctx.set_state(U1)
        ctx.newton_init(case.implicit.cfl_max)
        x = ctx.linear_solve(-ctx.residual(U1))[0]
        assert np.array_equal(ctx.get_state()[0], U1)   # the solve leaves the state
        ctx.newton_block(1)
        step = ctx.get_state()[0] - U1
        rel = np.linalg.norm(step - x) / np.linalg.norm(x)
        Ax = ctx.tangent(U1, x, first_order=True, with_vdt=True)[0]
        R1 = ctx.residual(U1)[0]
        lin = np.linalg.norm(Ax + R1) / np.linalg.norm(R1)
        print(f"[gpu] linear solve vs Newton step: {rel:.3e}; |A x + R| / |R| = {lin:.3e}", flush=True)
        assert rel < 1e-12
        assert lin < 0.5
    finally:
        ctx.close()


def test_linear_solve_3d_is_the_newton_step_solve():
    """The 3D context's jr_linear_solve is its Newton step's update too (the
    reference's ONERA M6 grid, SA, relaxation off so the step is x itself)."""
    from jaxrans_amd import load_case
    from jaxrans_amd.core.state import uniform_canvas
    from jaxrans_amd.simulation import freestream_of, make_context
    case = load_case(CASE.parent / "onera_m6_sa", ("numerical.implicit.max_relative_update=null",))
    grid = case.grid()
    ctx = make_context(case, grid)
    try:
        q, mu = freestream_of(case.flow, grid.dim)
        ctx.set_freestream(q, mu)
        rng = np.random.default_rng(11)
        U0 = uniform_canvas(q, grid.shape, case.flow.gamma)
        U1 = U0 * (1.0 + 1e-3 * rng.standard_normal(U0.shape) * grid.interior[None])
        ctx.set_state(U1)
        ctx.newton_init(case.implicit.cfl)
        x = ctx.linear_solve(-ctx.residual(U1))[0]
        assert np.array_equal(ctx.get_state()[0], U1)
        ctx.newton_block(1)
        step = ctx.get_state()[0] - U1
        rel = np.linalg.norm(step - x) / np.linalg.norm(x)
        print(f"[gpu] 3D linear solve vs Newton step: {rel:.3e}", flush=True)
        assert rel < 1e-12
    finally:
        ctx.close()
