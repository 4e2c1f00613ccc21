"""Derivatives of a converged result's force coefficients with respect to the
flow parameters, on the HIP library: the reference's ``sensitivity``
(``src/jaxrans/derivatives.py:123-147``) and its tangent solve
(``src/jaxrans/solvers/sensitivity.py:34-64``).

    result = run(load_case("cases/naca0012_o_euler"))
    s = sensitivity(result, outputs=("cl", "cd"), parameters=("alpha", "mach"))
    s.derivatives["cl"]["alpha"]            # dCL/dalpha, per degree

At a converged state ``R(U*(p), p) = 0``, so ``dR/dU dU/dp = -dR/dp`` and
``dJ/dp = dJ/dp|_U + dJ/dU dU/dp``: one linear solve per parameter.  The
solve is the reference's defect correction ``x <- x + G (b - A x)`` from
zero until ``|b - A x| <= tol |b|``, with ``A x`` the exact (second-order)
Jacobian-vector product of the residual (``jr_tangent``) and ``G`` the linear
solve of the case's own Newton step at the end of its CFL ramp
(``jr_linear_solve``: GMRES on ``V/dt + dR~/dU`` at ``cfl_max``,
line-preconditioned).

The parameters reach the residual and the force coefficients only through
the freestream primitives and ``mu_inf = M / Re`` (``flow_inputs``, as
``derivatives.py:65-70``): ``q = (1, M cos a, M sin a, 1/gamma[, nu ratio])``
with ``a`` in radians from degrees, so a parameter direction is the
freestream's tangent ``(dq, dmu)`` in closed form (``freestream_tangent``).
The two parameter products that the reference takes by forward-mode
differentiation -- ``dR/dp`` at fixed ``U`` (``solvers/sensitivity.py:60``)
and the output tangent ``dJ/dp|_U + dJ/dU dU`` (``:63``) -- are exact here
too: the device carries the freestream's tangent through the residual's and
the force integration's dual-number kernels (``jr_tangent_freestream``,
``jr_forces_tangent``, ``jr_forces3_tangent``), in 2D and 3D.  The derivative
exists where the discretisation is differentiable (``minmod`` and
``van_albada`` have kinks: use ``van_albada_smooth`` or no limiter) and is as
accurate as the result is converged.  A swept (2.5D) case runs on its
extruded section's 3D canvas, its parameters turning the freestream of the
section-normal frame, and its coefficients come from the wall loads and
their device tangent (``jr_wall_loads_tangent``) by the complex step.
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass, replace

import numpy as np

from .core.config import FlowConditions
from .simulation import Result, freestream_of, open_device

PARAMETERS: tuple[str, ...] = ("alpha", "mach")
OUTPUTS: tuple[str, ...] = ("cl", "cd", "cm", "cd_friction")


@dataclass(frozen=True)
class Sensitivity:
    """``derivatives[output][parameter]`` (per degree for ``alpha``);
    ``linear_residuals``: each parameter's linear solve, relative residual;
    ``iterations``: its defect-correction steps; ``seconds``: wall clock of
    the solves; ``compile_seconds`` is 0 (the library is built ahead)."""

    derivatives: dict[str, dict[str, float]]
    linear_residuals: dict[str, float]
    seconds: float
    compile_seconds: float = 0.0
    iterations: dict[str, int] | None = None

    def summary(self) -> str:
        lines = [f"d{out.upper()}/d{name} = {value:+.8e}"
                 for out, row in self.derivatives.items() for name, value in row.items()]
        worst = max(self.linear_residuals.values())
        return "\n".join(lines + [f"{len(self.linear_residuals)} linear solve(s), "
                                  f"largest relative residual {worst:.1e}, {self.seconds:.1f} s"])


def flow_inputs(flow: FlowConditions, dim: int, p: dict) -> tuple[np.ndarray, float | None]:
    """Freestream primitives and ``mu_inf`` (``None`` if inviscid) at
    parameters ``p`` (``alpha`` in degrees, ``mach``), the Reynolds number
    held (``derivatives.py:65-70``); a swept flow keeps its sweep (the
    section-normal frame's primitives)."""
    at = replace(flow, alpha_deg=float(p["alpha"]), mach=float(p["mach"]))
    if flow.swept:
        from .simulation import swept_q
        return swept_q(at), (None if at.inviscid else at.mu_inf)
    return freestream_of(at, dim)


def freestream_tangent(flow: FlowConditions, dim: int, name: str) -> tuple[np.ndarray, float | None]:
    """``(dq_inf, dmu_inf)``: the derivative of the freestream primitives and
    of ``mu_inf = M / Re`` along ``name`` (``alpha`` per degree, ``mach``) at
    the flow's parameters (``freestream_vector``, ``core/state.py:176-192``;
    ``None`` for ``dmu_inf`` when inviscid)."""
    a = math.radians(flow.alpha_deg)
    if name == "alpha":
        per = math.pi / 180.0
        du, dw, dm = -flow.mach * math.sin(a) * per, flow.mach * math.cos(a) * per, 0.0
    elif name == "mach":
        du, dw, dm = math.cos(a), math.sin(a), 1.0
    else:
        raise ValueError(f"unknown parameter {name!r}; available: {PARAMETERS}")
    if flow.swept:   # the section-normal frame: M (cos a cos L, sin a, cos a sin L)
        c, s = math.cos(math.radians(flow.sweep_deg)), math.sin(math.radians(flow.sweep_deg))
        velocity = [du * c, dw, du * s]
    else:
        velocity = [du, dw] if dim == 2 else [du, 0.0, dw]
    dq = np.array([0.0, *velocity, 0.0] + ([0.0] if flow.turbulence is not None else []))
    return dq, (None if flow.inviscid else dm / flow.reynolds)


def defect_correction(matvec, b: np.ndarray, approximate_inverse, tol: float, maxiter: int, norm=None):
    """``x <- x + G (b - A x)`` from zero until ``|b - A x| <= tol |b|`` or
    ``maxiter`` steps (``solvers/sensitivity.py:34-50``); returns ``x``, its
    relative residual and the step count.  ``norm``: the vector norm of the
    test (default the 2-norm of the array)."""
    norm = norm or (lambda v: float(np.linalg.norm(v)))
    b_norm = max(norm(b), np.finfo(np.float64).tiny)
    x = np.zeros_like(b)
    r = b
    k = 0
    while k < maxiter and norm(r) > tol * b_norm:
        x = x + approximate_inverse(r)
        r = b - matvec(x)
        k += 1
    return x, norm(r) / b_norm, k


def sensitivity(result: Result, outputs: tuple[str, ...] = ("cl", "cd"), parameters: tuple[str, ...] = ("alpha",),
                tol: float = 1e-10, maxiter: int = 10_000) -> Sensitivity:
    """Derivatives of ``outputs`` (``OUTPUTS``) with respect to ``parameters``
    (``PARAMETERS``) at ``result``'s state, one linear solve per parameter
    (``derivatives.py:123-147``)."""
    for kind, given, known in (("output", outputs, OUTPUTS), ("parameter", parameters, PARAMETERS)):
        unknown = [g for g in given if g not in known]
        if unknown or not given:
            raise ValueError(f"unknown {kind}(s) {unknown or given}; available: {known}")
    case, grid = result.case, result.grid
    flow = case.flow
    started = time.perf_counter()
    dev = open_device(case, grid)   # a swept section runs on its extruded 3D canvas
    ctx, dim = dev.ctx, dev.atlas.dim
    U = dev.to_device(np.asarray(result.U, np.float64), grid)
    try:
        ctx.set_state(U)
        ctx.newton_init(case.implicit.cfl_max)
        spec = case.grid_spec
        rows = {"cl": 0, "cd": 1, "cm": 2, "cd_friction": 3}

        def parameter_rhs(name: str) -> np.ndarray:
            dq, dmu = freestream_tangent(flow, dim, name)
            return ctx.tangent_freestream(U, dq, dmu)[0]

        def output_tangent(name: str, dU: np.ndarray) -> np.ndarray:
            dq, dmu = freestream_tangent(flow, dim, name)
            if flow.swept:   # the swept coefficients from the wall loads and their tangent
                from .mesh.swept import swept_coefficients_tangent
                d = swept_coefficients_tangent(ctx.wall_loads()[0], ctx.wall_loads_tangent(U, dq, dmu, dU=dU)[0],
                                               dev.atlas, dev.q, dq, flow.sweep_deg, spec.reference_length,
                                               tuple(spec.moment_centre), ctx.viscous)
                return np.array([d[out] for out in outputs])
            if dim == 3:
                row = ctx.forces3_tangent(U, dq, dmu, dU=dU, area=spec.reference_area,
                                          length=spec.reference_length, centre=tuple(spec.moment_centre))[0]
            else:
                row = ctx.forces_tangent(U, dq, dmu, dU=dU, length=spec.reference_length,
                                         centre=tuple(spec.moment_centre[:2]))[0]
            return np.array([row[rows[out]] for out in outputs])

        rhs = {name: parameter_rhs(name) for name in parameters}

        def matvec(x: np.ndarray) -> np.ndarray:
            return ctx.tangent(U, x, first_order=False)[0]

        def inverse(r: np.ndarray) -> np.ndarray:
            return ctx.linear_solve(r)[0]

        # the test's norm over the case's own state layout: a swept section's
        # 2D canvas, not its extruded planes and their ghost layers
        norm = None
        if flow.swept:
            def norm(v: np.ndarray) -> float:
                return float(np.linalg.norm(dev.from_device(v, grid)))

        columns, errors, iterations = {}, {}, {}
        for name in parameters:
            dU, errors[name], iterations[name] = defect_correction(matvec, -rhs[name], inverse, tol, maxiter, norm)
            columns[name] = output_tangent(name, dU)
    finally:
        ctx.close()
    derivatives = {out: {name: float(columns[name][i]) for name in parameters} for i, out in enumerate(outputs)}
    return Sensitivity(derivatives, errors, time.perf_counter() - started, 0.0, iterations)
