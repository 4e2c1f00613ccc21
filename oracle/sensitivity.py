"""Parameter tangent of a converged state: the sensitivity solve
(``src/jaxrans/solvers/sensitivity.py:34-64``) over the oracle's residual,
linear solve and forces (``src/jaxrans/derivatives.py:65-147``).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``): the checker of the
device's ``jr_tangent_freestream`` / ``jr_forces_tangent`` and of
``sensitivity()``.

The flow parameters reach the residual and the force coefficients only
through the freestream primitives and ``mu_inf = M / Re`` (``flow_inputs``,
``derivatives.py:65-70``, over ``freestream_vector``,
``core/state.py:176-192``), so one parameter direction is the freestream
carried as dual scalars: ``q = (1, M cos a, M sin a, 1/gamma[, nu ratio])``
(3D: velocity ``M (cos a, 0, sin a)``) with ``a`` in radians from degrees.  The residual's parameter JVP is the
oracle residual evaluated on that freestream (ghost states, the ``U - U0``
rows, the viscous and SA terms and the point-vortex reference all see it),
and the output JVP is the oracle's force integration on the state's and the
freestream's tangents together.  The linear solve is the reference's defect
correction with ``G`` = the Newton step's GMRES at ``cfl_max``
(``newton_inverse``, ``derivatives.py:94-120``; the solve is restated for
2D setups, the parameter and output JVPs for 2D and 3D).
"""

from __future__ import annotations

from dataclasses import replace

import numpy as np

from . import dual as M
from .residual import Setup, jvp, residual, wave_speed_sum, wave_speeds
from .solver import LinePreconditioner, forces, gmres_right

PARAMETERS = ("alpha", "mach")


def freestream_dual(flow, name: str, nvar: int, dim: int = 2):
    """The freestream primitives (``nvar`` dual scalars) and ``mu_inf`` (dual,
    or None if inviscid) at the flow's parameters, with tangent one along
    ``name`` (``alpha`` per degree, ``mach``), the Reynolds number held; in
    3D the velocity is ``M (cos a, 0, sin a)``."""
    if name not in PARAMETERS:
        raise ValueError(name)
    a = M.D(np.deg2rad(flow.alpha_deg), np.deg2rad(1.0) if name == "alpha" else 0.0)
    mach = M.D(flow.mach, 1.0 if name == "mach" else 0.0)
    u, w = mach * M.cos(a), mach * M.sin(a)
    vel = [u, w] if dim == 2 else [u, M.D(0.0, 0.0), w]
    q = [M.D(1.0, 0.0), *vel, M.D(1.0 / flow.gamma, 0.0)]
    if nvar > dim + 2:
        q.append(M.D(flow.nu_tilde_ratio, 0.0))
    mu = None if flow.inviscid else mach / flow.reynolds
    return q, mu


def _is_3d(setup) -> bool:
    return len(setup.atlas.shape) == 3


def tangents(q, mu):
    """``(dq_inf, dmu_inf)`` of a dual freestream, in the device's layout."""
    return np.array([float(x.d) for x in q]), (None if mu is None else float(mu.d))


def with_freestream(setup: Setup, q, mu) -> Setup:
    return replace(setup, q_inf=list(q), mu_inf=mu)


def residual_parameter_jvp(U, setup: Setup, q, mu, dU=None) -> np.ndarray:
    """``d/de R(U + e dU; q + e dq, mu + e dmu)``: the exact residual's tangent
    (2D or 3D setup)."""
    dU = np.zeros_like(U) if dU is None else dU
    if _is_3d(setup):
        from .residual3d import residual as residual3
        return residual3(M.D(U, dU), with_freestream(setup, q, mu)).d
    return residual(M.D(U, dU), with_freestream(setup, q, mu)).d


def forces_jvp(U, dU, setup: Setup, q, mu, length=1.0, centre=(0.25, 0.0), area=None) -> dict:
    """Tangent of ``forces`` along the state's ``dU`` and the freestream's
    tangent (3D: ``forces3`` per ``area``, the moment about y through
    ``centre = (x, y, z)``)."""
    dU = np.zeros_like(U) if dU is None else dU
    if _is_3d(setup):
        from .solver3d import forces as forces3
        f = forces3(M.D(U, dU), with_freestream(setup, q, mu), area, length, centre)
    else:
        f = forces(M.D(U, dU), with_freestream(setup, q, mu), length, centre)
    return {k: float(M.tangent(v)) for k, v in f.items()}


def newton_inverse(U, setup: Setup, cfg):
    """``r -> G r``: GMRES on ``V/dt + dR~/dU`` at ``cfg.cfl_max``, right-
    preconditioned by the line solve, with the case's Krylov settings; ``R~``
    first order unless ``cfg.jacobian == "exact"`` (``derivatives.py:94-120``)."""
    lams = wave_speeds(U, setup)
    vdt = wave_speed_sum(lams) / cfg.cfl_max
    first_order = cfg.jacobian != "exact"
    precondition = LinePreconditioner(U, lams, cfg.cfl_max, setup, cfg.line_solver)

    def matvec(v):
        return vdt[None] * v + jvp(U, v, setup, first_order)

    def apply(r):
        x, _ = gmres_right(matvec, r, precondition, cfg.krylov_subspace, cfg.krylov_restarts, cfg.krylov_tol)
        return x

    return apply


def defect_correction(matvec, b, approximate_inverse, tol: float, maxiter: int):
    """``x <- x + G (b - A x)`` from zero until ``|b - A x| <= tol |b|``
    (``solvers/sensitivity.py:34-50``); returns ``x``, the relative residual
    and the step count."""
    b_norm = max(float(np.linalg.norm(b)), np.finfo(np.float64).tiny)
    x, r, k = np.zeros_like(b), b, 0
    while k < maxiter and float(np.linalg.norm(r)) > tol * b_norm:
        x = x + approximate_inverse(r)
        r = b - matvec(x)
        k += 1
    return x, float(np.linalg.norm(r)) / b_norm, k


def sensitivity(U, setup: Setup, flow, cfg, outputs=("cl", "cd"), parameters=("alpha",), tol=1e-10,
                maxiter=10_000, length=1.0, centre=(0.25, 0.0)):
    """``(derivatives[output][parameter], {parameter: relative linear residual},
    {parameter: steps})`` at the state ``U`` (``solvers/sensitivity.py:53-64``)."""
    inverse = newton_inverse(U, setup, cfg)

    def matvec(x):
        return jvp(U, x, setup, first_order=False)

    columns, errors, steps = {}, {}, {}
    for name in parameters:
        q, mu = freestream_dual(flow, name, U.shape[0])
        rhs = residual_parameter_jvp(U, setup, q, mu)
        dU, errors[name], steps[name] = defect_correction(matvec, -rhs, inverse, tol, maxiter)
        columns[name] = forces_jvp(U, dU, setup, q, mu, length, centre)
    return {out: {name: columns[name][out] for name in parameters} for out in outputs}, errors, steps
