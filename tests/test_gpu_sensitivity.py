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
