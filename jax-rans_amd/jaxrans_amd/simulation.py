"""Run a case on the device: the drop-in for the reference's ``run``.

Same call, same results record as ``src/jaxrans/simulation.py:213-284`` and the
driver ``src/jaxrans/solvers/driver.py:123-201``: blocks of ``block_iters``
Newton steps, each block one device call with no host synchronisation inside
it; between blocks the host reads the residual history, integrates the forces
(on the device) and applies the convergence, stall, non-finite and settled
tests.  Every number is computed by the HIP library; there is no CPU path.
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field, replace
from typing import Callable

import numpy as np

from .case.loader import Case, CaseError
from .core.config import FlowConditions
from .core.state import q_infinity, uniform_canvas
from .device import Context, context_for
from .mesh.atlas import Atlas
from .mesh.coarsen import prolong

Progress = Callable[[int, float, float, "Forces"], None]


class ExitReason:
    CONVERGED = "converged"
    SETTLED = "settled"
    MAX_ITERS = "max_iters"
    NON_FINITE = "non_finite"
    STALLED = "stalled"


@dataclass
class ConvergenceHistory:
    """``residual``: ``(n_iters,)`` or ``(n_iters, n_cases)`` (driver.py:42-84)."""

    residual: np.ndarray
    reason: str
    n_iters: int
    seconds: float = float("nan")
    first_block_seconds: float = float("nan")
    case_reasons: tuple = ()

    def case(self, k: int) -> "ConvergenceHistory":
        return ConvergenceHistory(self.residual[:, k], self.case_reasons[k], self.n_iters, self.seconds,
                                  self.first_block_seconds)

    @property
    def converged(self) -> bool:
        return self.reason == ExitReason.CONVERGED

    @property
    def peak(self) -> float:
        finite = self.residual[np.isfinite(self.residual)]
        return float(finite.max()) if finite.size else float("nan")

    @property
    def orders_dropped(self) -> float:
        if self.residual.size == 0:
            return 0.0
        peak, last = self.peak, self.residual[-1]
        if np.ndim(last):
            return float("nan")
        if not (np.isfinite(peak) and np.isfinite(last)) or peak <= 0 or last <= 0:
            return float("nan")
        return float(np.log10(peak / last))

    def summary(self) -> str:
        return (f"{self.reason} after {self.n_iters} iterations, residual peak {self.peak:.6e} -> "
                f"{float(np.ravel(self.residual)[-1]):.6e} ({self.orders_dropped:.2f} orders from peak)")


@dataclass(frozen=True)
class Forces:
    cl: float
    cd: float
    cm: float
    fx: float
    fy: float
    cd_friction: float = 0.0
    bodies: tuple = ()
    fz: float = 0.0
    cy: float = 0.0

    def summary(self) -> str:
        friction = f" (friction {self.cd_friction:+.6f})" if self.cd_friction else ""
        return f"CL = {self.cl:+.6f}   CD = {self.cd:+.6f}{friction}   CM = {self.cm:+.6f}"


def _forces(row: np.ndarray, dim: int = 2) -> Forces:
    if dim == 3:
        return Forces(float(row[0]), float(row[1]), float(row[2]), float(row[4]), float(row[5]), float(row[3]),
                      fz=float(row[6]), cy=float(row[7]))
    return Forces(float(row[0]), float(row[1]), float(row[2]), float(row[4]), float(row[5]), float(row[3]))


def force_rows(ctx: Context, case: Case) -> np.ndarray:
    """The coefficient rows of every case on the device, by the case's force
    reference (``post/forces.py:98-135``): per unit span about ``(x, y)`` in
    2D; per ``reference_area`` with the moment about the ``y`` axis through
    ``moment_centre`` in 3D."""
    spec = case.grid_spec
    if ctx.dim == 3:
        if spec.reference_area is None:
            raise CaseError("a 3D case needs domain.reference_area")
        return ctx.forces3(spec.reference_area, spec.reference_length, tuple(spec.moment_centre))
    return ctx.forces(spec.reference_length, tuple(spec.moment_centre[:2]))


@dataclass(frozen=True)
class Timing:
    """Wall clock of a run (simulation.py:53-101): the first block carries the
    set-up, the later ones give the steady cost."""

    total: float
    first_block: float
    iterations: int
    block_iters: int
    cells: int
    cases: int = 1
    devices: int = 1

    @property
    def per_iteration(self) -> float:
        later = self.iterations - self.block_iters
        return (self.total - self.first_block) / later if later > 0 else float("nan")

    @property
    def compile(self) -> float:
        return self.first_block - self.block_iters * self.per_iteration

    @property
    def case_iterations_per_second(self) -> float:
        return self.cases / self.per_iteration

    def record(self) -> dict:
        return {"total_s": self.total, "compile_s": self.compile, "per_iteration_s": self.per_iteration,
                "case_iterations_per_s": self.case_iterations_per_second,
                "ns_per_cell_iteration": 1e9 / (self.case_iterations_per_second * self.cells),
                "iterations": self.iterations, "cells": self.cells, "cases": self.cases, "devices": self.devices}


SETTLED_CL_SPREAD = 1e-4


def cl_spread(history) -> float:
    tail = np.asarray(history)[3 * len(history) // 4:]
    return float(tail.max() - tail.min()) if tail.size >= 4 else float("nan")


@dataclass(frozen=True)
class Result:
    case: Case
    grid: Atlas
    U: np.ndarray
    forces: Forces
    history: ConvergenceHistory
    cl_history: np.ndarray
    timing: Timing
    levels: tuple = ()
    extra: dict = field(default_factory=dict)

    @property
    def seconds(self) -> float:
        return self.timing.total

    @property
    def cl_spread(self) -> float:
        return cl_spread(self.cl_history)

    @property
    def cells(self) -> int:
        return self.grid.cells

    def surface(self) -> dict:
        """Wall distributions per wall face (``post.surface``): 2D ``x``,
        ``y``, ``body``, ``cp`` [, ``cf``]; 3D ``block``, ``i``, ``k``, ``x``,
        ``y``, ``z``, ``cp`` [, ``cf_x``, ``cf_y``, ``cf_z``].  The loads are
        evaluated on the device at ``U``."""
        from .post.surface import surface, swept_surface
        dev = open_device(self.case, self.grid)
        try:
            dev.ctx.set_state(dev.to_device(np.asarray(self.U), self.grid)[None])
            if dev.swept:
                return swept_surface(dev.ctx, self.grid, self.case.flow.sweep_deg)
            return surface(dev.ctx, 0)
        finally:
            dev.ctx.close()

    def write(self, directory=None, extra: dict | None = None):
        """Metadata, residual and lift histories and fields under the
        reference's file names (``simulation.py:135-160``); returns the directory."""
        from .output import write_result
        return write_result(self, directory, extra)


RK_SCHEMES = {"euler": (1.0,), "rk4_jameson": (0.25, 1.0 / 3.0, 0.5, 1.0),
              "rk5_jameson": (0.25, 1.0 / 6.0, 0.375, 0.5, 1.0)}   # solvers/explicit.py:23-27


def _check_case(case: Case, dim: int | None = None) -> None:
    if case.numerics.solver == "explicit":
        if case.rk_scheme not in RK_SCHEMES:
            raise CaseError(f"unknown RK scheme {case.rk_scheme!r}; available: {sorted(RK_SCHEMES)}")
    if case.flow.swept and case.numerics.far_field != "characteristic":
        raise CaseError("the point-vortex far field with a swept flow is not built by this implementation")


def make_context(case: Case, grid: Atlas, n_cases: int = 1) -> Context:
    return context_for(grid, case.flow, case.numerics, case.implicit, n_cases=n_cases)


def freestream_of(flow: FlowConditions, dim: int = 2) -> tuple[np.ndarray, float | None]:
    return q_infinity(flow, dim), (None if flow.inviscid else flow.mu_inf)


@dataclass
class _Device:
    """What a march runs on the device: the context, the canvas it holds and
    the maps between that canvas and the case's own.  A swept case's canvas
    is its 2D section extruded two periodic cells deep (``mesh.swept``); its
    state goes back to the reference's 2D layout with three velocity
    components."""

    ctx: Context
    atlas: object
    q: np.ndarray
    mu: float | None
    swept: bool
    gamma: float

    def to_device(self, U: np.ndarray, grid) -> np.ndarray:
        if not self.swept:
            return U
        from .mesh.swept import section_to_planes
        return section_to_planes(U, grid, self.atlas, uniform_canvas(self.q, self.atlas.shape, self.gamma))

    def from_device(self, U: np.ndarray, grid) -> np.ndarray:
        if not self.swept:
            return U
        from .mesh.swept import planes_to_section
        return planes_to_section(U, self.atlas, grid)

    def forces(self, case: Case) -> Forces:
        if not self.swept:
            return _forces(force_rows(self.ctx, case)[0], self.atlas.dim)
        from .mesh.swept import swept_coefficients
        spec = case.grid_spec
        c = swept_coefficients(self.ctx.wall_loads()[0], self.atlas, self.q, case.flow.sweep_deg,
                               spec.reference_length, tuple(spec.moment_centre), self.ctx.viscous)
        f = c["force"]
        return Forces(c["cl"], c["cd"], c["cm"], float(f[0]), float(f[1]), c["cd_friction"], fz=float(f[2]),
                      cy=c["cy"])


def swept_q(flow: FlowConditions) -> np.ndarray:
    """A swept case's freestream primitives ``(rho, u, v, w, p[, nu ratio])``."""
    from .mesh.swept import swept_freestream
    q = swept_freestream(flow.mach, flow.alpha_deg, flow.sweep_deg, flow.gamma)
    return np.concatenate([q, [flow.nu_tilde_ratio]]) if flow.turbulence is not None else q


def section_state(q: np.ndarray, grid, gamma: float) -> np.ndarray:
    """A swept case's uniform state on its 2D canvas (three velocity components)."""
    from .core.state import conservative
    u = conservative(np.asarray(q, np.float64), gamma, 3)
    return np.broadcast_to(u.reshape((-1, 1, 1)), (u.size,) + tuple(grid.shape)).copy()


def open_device(case: Case, grid: Atlas) -> _Device:
    """The device context of a march of ``case`` on ``grid``, freestream set."""
    flow = case.flow
    mu = None if flow.inviscid else flow.mu_inf
    if flow.swept:
        from .mesh.swept import lift
        atlas = lift(grid.mesh, case.numerics.n_ghost)
        q = swept_q(flow)
    else:
        atlas, (q, mu) = grid, freestream_of(flow, grid.dim)
    ctx = make_context(case, atlas)
    ctx.set_freestream(q, mu)
    return _Device(ctx, atlas, q, mu, flow.swept, flow.gamma)


def run_to_steady(ctx: Context, *, max_iters: int, tol: float, block_iters: int,
                  on_block: Callable[[int, np.ndarray], bool] | None = None,
                  block: Callable[[int], np.ndarray] | None = None):
    """Blocks of Newton steps until every case converges (``last <= tol *
    running peak``), stalls, goes non-finite, settles or hits ``max_iters``
    (driver.py:123-201, the same tests in the same order)."""
    n_blocks = max(1, -(-max_iters // block_iters))
    blocks = []
    peak = reasons = None
    started = time.perf_counter()
    first = float("nan")
    for b in range(n_blocks):
        res = ctx.newton_block(block_iters) if block is None else block(block_iters)
        if b == 0:
            first = time.perf_counter() - started
        blocks.append(res)
        stop = on_block(b, res) if on_block is not None else False
        cols = res.reshape(res.shape[0], -1)
        if peak is None:
            peak = np.zeros(cols.shape[1])
            reasons = np.full(cols.shape[1], ExitReason.MAX_ITERS, dtype=object)
        finite = np.where(np.isfinite(cols), cols, 0.0)
        peak = np.maximum(peak, finite.max(axis=0))
        last = cols[-1]
        running = reasons == ExitReason.MAX_ITERS
        failed = ~np.all(np.isfinite(cols), axis=0)
        reasons[running & failed] = ExitReason.NON_FINITE
        reasons[running & ~failed & (peak > 0) & (last <= tol * peak)] = ExitReason.CONVERGED
        frozen = (len(cols) > 1) & np.all(cols == cols[0], axis=0)
        reasons[(reasons == ExitReason.MAX_ITERS) & frozen] = ExitReason.STALLED
        if stop:
            reasons[reasons == ExitReason.MAX_ITERS] = ExitReason.SETTLED
        if not np.any(reasons == ExitReason.MAX_ITERS):
            break
    residual = np.concatenate(blocks) if blocks else np.zeros((0, ctx.n_cases))
    failures = [r for r in reasons if r not in (ExitReason.CONVERGED, ExitReason.SETTLED)]
    settled = ExitReason.SETTLED in reasons
    if ctx.n_cases == 1:
        residual = residual[:, 0]
    return ConvergenceHistory(residual=residual,
                              reason=failures[0] if failures else ExitReason.SETTLED if settled else
                              ExitReason.CONVERGED,
                              n_iters=int(residual.shape[0]), seconds=time.perf_counter() - started,
                              first_block_seconds=first,
                              case_reasons=tuple(reasons) if ctx.n_cases > 1 else ())


def run(case: Case, grid: Atlas | None = None, progress: Progress | None = None, on_level=None, devices=None, *,
        initial: np.ndarray | None = None) -> Result:
    """March ``case`` to steady state on the device (simulation.py:213-240)."""
    if devices is not None and len(devices) > 1:
        raise CaseError("this implementation runs one case on one GPU; multi-device slabs are not built")
    _check_case(case)
    grid = case.grid() if grid is None else grid
    if initial is not None:
        return _march(case, grid, progress, initial=initial)
    # grid sequencing (simulation.py:230-240): the coarser grids first, each
    # started from the one before by injection
    levels, start = [], None
    for level in range(case.sequencing, 0, -1):
        coarse = case.grid(coarsen=level)
        if on_level is not None:
            on_level(level, coarse)
        r = _march(case, coarse, progress, start=start)
        levels.append({"cells": coarse.cells, "iterations": r.history.n_iters, "seconds": r.seconds,
                       "exit_reason": r.history.reason, "cl": r.forces.cl, "cd": r.forces.cd})
        start = (r.U, coarse)
    if levels and on_level is not None:
        on_level(0, grid)
    return replace(_march(case, grid, progress, start=start), levels=tuple(levels))


def _march(case: Case, grid: Atlas, progress: Progress | None, *, initial=None, start=None) -> Result:
    flow, numerics = case.flow, case.numerics
    _check_case(case, grid.dim)   # before the device context exists: a refusal leaves nothing open
    dev = open_device(case, grid)
    ctx = dev.ctx
    U0 = section_state(dev.q, grid, flow.gamma) if flow.swept else uniform_canvas(dev.q, grid.shape, flow.gamma)
    if initial is not None:
        initial = np.asarray(initial, np.float64)
        if initial.shape != U0.shape:
            raise ValueError(f"initial field must have shape {U0.shape}, got {initial.shape}")
        U0 = np.where(grid.interior[None], initial, U0)
    elif start is not None:
        U0 = prolong(start[0], start[1], grid, U0)
    ctx.set_state(dev.to_device(U0, grid))
    block = None
    if numerics.solver == "explicit":   # solvers/explicit.py:48-71, CFL numerics.cfl
        ctx.newton_init(numerics.cfl)
        alphas = RK_SCHEMES[case.rk_scheme]
        block = lambda n: ctx.rk_block(n, alphas)   # noqa: E731
    else:
        ctx.newton_init(case.implicit.cfl)
    cl_history, coefficients, peak = [], [], [0.0]

    def on_block(index: int, residuals: np.ndarray) -> bool:
        finite = residuals[np.isfinite(residuals)]
        if finite.size:
            peak[0] = max(peak[0], float(finite.max()))
        f = dev.forces(case)
        cl_history.append(f.cl)
        coefficients.append((f.cl, f.cd, f.cm))
        if progress is not None:
            last = float(residuals[-1, 0])
            orders = float(np.log10(peak[0] / last)) if peak[0] > 0.0 and last > 0.0 else 0.0
            progress((index + 1) * case.block_iters, last, orders, f)
        return numerics.forces_tol is not None and all(
            cl_spread(h) < numerics.forces_tol for h in np.transpose(coefficients))

    history = run_to_steady(ctx, max_iters=numerics.max_iters, tol=numerics.tol, block_iters=case.block_iters,
                            on_block=on_block, block=block)
    U = dev.from_device(ctx.get_state()[0], grid)
    forces = dev.forces(case)
    timing = Timing(history.seconds, history.first_block_seconds, history.n_iters, case.block_iters, grid.cells)
    result = Result(case=case, grid=grid, U=U, forces=forces, history=history, cl_history=np.asarray(cl_history),
                    timing=timing, extra={"scalars": ctx.scalars()[0]})
    ctx.close()
    return result


@dataclass(frozen=True)
class Polar:
    """A batch of freestreams on one grid (``src/jaxrans/polar.py:43-75``):
    one ``Result`` per case, in the order of ``alphas``, and the batch timing."""

    case: Case
    alphas: np.ndarray
    results: list
    timing: Timing

    def table(self) -> np.ndarray:
        """One row per case: alpha, CL, CD, CD friction, CM, orders dropped,
        iterations, CL spread over the last quarter of blocks, settled (1.0 when
        that spread is within the settled threshold)."""
        rows = []
        for alpha, r in zip(self.alphas, self.results):
            spread = r.cl_spread
            rows.append([alpha, r.forces.cl, r.forces.cd, r.forces.cd_friction, r.forces.cm,
                         r.history.orders_dropped, r.history.n_iters, spread, float(spread <= SETTLED_CL_SPREAD)])
        return np.asarray(rows, dtype=np.float64)

    def write(self, directory) -> "Path":
        """``polar.dat`` (the table), ``polar_metadata.json`` (case name and
        set-up, alphas, timing) and each case's result files in
        ``alpha_{alpha:+.2f}/``; returns the directory."""
        import json
        from pathlib import Path

        from .output import _jsonable
        out = Path(directory)
        out.mkdir(parents=True, exist_ok=True)
        np.savetxt(out / "polar.dat", self.table(),
                   header="alpha_deg cl cd cd_friction cm orders_dropped iterations cl_spread settled",
                   fmt=["%.4f", "%.8f", "%.8f", "%.8f", "%.8f", "%.3f", "%d", "%.3e", "%d"])
        meta = {"case": self.case.name, "case_setup": _jsonable(self.case.raw),
                "alphas": [float(a) for a in self.alphas], "timing": self.timing.record(),
                "settled_cl_spread": SETTLED_CL_SPREAD}
        (out / "polar_metadata.json").write_text(json.dumps(meta, indent=2, default=float))
        for alpha, r in zip(self.alphas, self.results):
            r.write(out / f"alpha_{alpha:+.2f}")
        return out


def run_polar(case: Case, alphas, grid: Atlas | None = None, devices=None, progress=None, *,
              machs=None, reynolds=None) -> Polar:
    """``case`` at every angle of ``alphas`` (degrees) in one device context,
    every kernel launch shared by the whole batch (the reference's
    ``run_polar``, ``src/jaxrans/polar.py:80-121``; same positional order).
    ``progress(iteration, CL per case)`` is called after each block.  The
    keyword-only ``machs`` and ``reynolds`` give each case its own Mach
    number and Reynolds number, as the reference's slot scheduler carries
    them per slot (``sweep.py:196-201``); by default every case has the
    case's own.  One GPU: ``devices`` above 1 is refused."""
    if devices is not None and int(devices) > 1:
        raise CaseError("this implementation runs a polar on one GPU; splitting it over devices is not built")
    if case.flow.swept:
        raise CaseError("a swept (2.5D) polar is not built by this implementation; run() each angle")
    grid = case.grid() if grid is None else grid
    _check_case(case, grid.dim)
    alphas = np.atleast_1d(np.asarray(alphas, np.float64))
    n = alphas.size
    machs = np.full(n, case.flow.mach) if machs is None else np.broadcast_to(machs, (n,)).astype(np.float64)
    flows = []
    for a, m in zip(alphas, machs):
        re = case.flow.reynolds if reynolds is None else float(reynolds)
        flows.append(replace(case.flow, alpha_deg=float(a), mach=float(m), reynolds=re))
    ctx = make_context(case, grid, n_cases=n)
    qs = np.stack([q_infinity(f, grid.dim) for f in flows])
    q5 = np.zeros((n, 5 if grid.dim == 2 else 6))
    q5[:, :qs.shape[1]] = qs
    mus = None if case.flow.inviscid else np.array([f.mu_inf for f in flows])
    ctx.set_freestream(q5, mus)
    U0 = np.stack([uniform_canvas(q, grid.shape, f.gamma) for q, f in zip(qs, flows)])
    ctx.set_state(U0)
    numerics = case.numerics
    block = None
    if numerics.solver == "explicit":
        ctx.newton_init(np.full(n, numerics.cfl))
        rk = RK_SCHEMES[case.rk_scheme]
        block = lambda k: ctx.rk_block(k, rk)   # noqa: E731
    else:
        ctx.newton_init(np.full(n, case.implicit.cfl))
    cl_hist = []

    def on_block(index, residuals):
        cl_hist.append(force_rows(ctx, case)[:, 0].copy())
        if progress is not None:
            progress((index + 1) * case.block_iters, cl_hist[-1].copy())
        return False

    history = run_to_steady(ctx, max_iters=numerics.max_iters, tol=numerics.tol, block_iters=case.block_iters,
                            on_block=on_block, block=block)
    U = ctx.get_state()
    F = force_rows(ctx, case)
    cl_hist = np.asarray(cl_hist)
    timing = Timing(history.seconds, history.first_block_seconds, history.n_iters, case.block_iters, grid.cells,
                    cases=n)
    out = []
    for k in range(n):
        h = history.case(k) if n > 1 else history
        sub = replace(case, flow=flows[k], name=f"{case.name}_a{alphas[k]:+.2f}")
        out.append(Result(case=sub, grid=grid, U=U[k], forces=_forces(F[k], grid.dim), history=h,
                          cl_history=cl_hist[:, k] if cl_hist.size else cl_hist, timing=timing))
    ctx.close()
    return Polar(case=case, alphas=alphas, results=out, timing=timing)
