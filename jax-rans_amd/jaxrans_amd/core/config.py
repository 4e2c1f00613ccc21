"""Frozen, validated run configuration.

The three records carry the fields, defaults and error messages of the
reference's configs (``src/jaxrans/core/config.py:20-212``), so a case read by
either code base holds the same numbers and is refused with the same words.
Validation is table-driven here: each record lists ``(test, message)`` rules
that ``_enforce`` applies in order; the first rule that fails raises
``ValueError`` with its message formatted from the record's fields.
"""

from __future__ import annotations

import math
from dataclasses import asdict, dataclass
from typing import Callable

import numpy as np

from . import constants

TURBULENCE_MODELS: tuple[str, ...] = ("sa_noft2", "sa_noft2_qcr")
SOLVERS: tuple[str, ...] = ("explicit", "implicit")
JACOBIANS: tuple[str, ...] = ("exact", "first_order")
PRECISIONS: tuple[str, ...] = ("fp64", "mixed")
LINE_SOLVERS: tuple[str, ...] = ("cyclic_reduction", "thomas")

Rule = tuple[Callable[[dict], bool], str]


def _enforce(record, rules: tuple[Rule, ...]) -> None:
    """Raise ``ValueError`` with the message of the first rule ``record``
    breaks; messages are ``str.format``-ed with the record's fields."""
    fields = asdict(record)
    for holds, message in rules:
        if not holds(fields):
            raise ValueError(message.format(**fields))


def _null_or(test: Callable) -> Callable:
    return lambda v: v is None or test(v)


_FLOW_RULES: tuple[Rule, ...] = (
    (lambda f: f["mach"] > 0.0, "mach must be positive, got {mach}"),
    (lambda f: f["sweep_deg"] is None or abs(f["sweep_deg"]) < 90.0, "sweep_deg must lie in (-90, 90), got {sweep_deg}"),
    (lambda f: _in_plane_mach(f["mach"], f["alpha_deg"], f["sweep_deg"]) < 1.0,
     "freestream mach must be subsonic for the characteristic far-field BC, got {mach}"),
    (lambda f: f["gamma"] > 1.0, "gamma must exceed 1, got {gamma}"),
    (lambda f: _null_or(lambda r: r > 0.0)(f["reynolds"]), "reynolds must be positive or None, got {reynolds}"),
    (lambda f: f["sutherland_ratio"] is None or f["reynolds"] is not None, "sutherland_ratio given for an inviscid case"),
    (lambda f: _null_or(lambda s: s > 0.0)(f["sutherland_ratio"]),
     "sutherland_ratio must be positive, got {sutherland_ratio}"),
    (lambda f: f["turbulence"] is None or f["turbulence"] in TURBULENCE_MODELS,
     "turbulence must be one of " + str(TURBULENCE_MODELS).replace("{", "{{").replace("}", "}}") +
     ", got {turbulence!r}"),
    (lambda f: f["turbulence"] is None or f["reynolds"] is not None, "a turbulence model needs a Reynolds number"),
    (lambda f: f["nu_tilde_ratio"] > 0.0, "nu_tilde_ratio must be positive, got {nu_tilde_ratio}"),
    (lambda f: abs(f["alpha_deg"]) <= 90.0, "alpha_deg out of range, got {alpha_deg}"),
)


def _in_plane_mach(mach: float, alpha_deg: float, sweep_deg: float | None) -> float:
    """Mach number of the freestream's component in the section plane: the
    whole freestream unswept; with sweep ``L`` the freestream
    ``M (cos a cos L, sin a, cos a sin L)`` less its spanwise part."""
    if sweep_deg is None:
        return mach
    a, s = math.radians(alpha_deg), math.radians(sweep_deg)
    return mach * math.hypot(math.cos(a) * math.cos(s), math.sin(a))


@dataclass(frozen=True)
class FlowConditions:
    """Freestream and gas, nondimensional.  ``reynolds=None``: inviscid;
    ``sutherland_ratio``: S / T_inf (None: constant viscosity);
    ``sweep_deg``: the 2.5D swept-section angle (None: unswept)."""

    mach: float
    alpha_deg: float = 0.0
    reynolds: float | None = None
    gamma: float = constants.GAMMA_AIR
    sutherland_ratio: float | None = None
    turbulence: str | None = None
    nu_tilde_ratio: float = 3.0
    sweep_deg: float | None = None

    def __post_init__(self) -> None:
        _enforce(self, _FLOW_RULES)

    inviscid = property(lambda self: self.reynolds is None)
    swept = property(lambda self: self.sweep_deg is not None)
    alpha_rad = property(lambda self: math.radians(self.alpha_deg))
    sweep_rad = property(lambda self: None if self.sweep_deg is None else math.radians(self.sweep_deg))
    in_plane_mach = property(lambda self: _in_plane_mach(self.mach, self.alpha_deg, self.sweep_deg))

    @property
    def mu_inf(self) -> float:
        """Freestream viscosity in the nondimensionalisation rho = 1, a = 1,
        chord 1: ``M / Re``."""
        if self.inviscid:
            raise ValueError("mu_inf is undefined for an inviscid case")
        return self.mach / self.reynolds


_NUMERICS_RULES: tuple[Rule, ...] = (
    (lambda n: n["cfl"] > 0.0, "cfl must be positive, got {cfl}"),
    (lambda n: n["max_iters"] >= 1, "max_iters must be >= 1, got {max_iters}"),
    (lambda n: 0.0 < n["tol"] < 1.0, "tol must lie in (0, 1), got {tol}"),
    (lambda n: _null_or(lambda t: t > 0.0)(n["forces_tol"]), "forces_tol must be positive or null, got {forces_tol}"),
    (lambda n: n["n_ghost"] >= 2, "n_ghost must be >= 2 for MUSCL reconstruction, got {n_ghost}"),
    (lambda n: -1.0 <= n["kappa"] <= 1.0, "muscl_kappa must lie in [-1, 1], got {kappa}"),
    (lambda n: n["kappa"] == 0.0 or n["limiter"] == "none",
     "muscl_kappa other than 0 is implemented without a limiter only (limiter 'none'), got limiter {limiter!r}"),
    (lambda n: n["solver"] in SOLVERS,
     "solver must be one of " + str(SOLVERS) + ", got {solver!r}"),
)


@dataclass(frozen=True)
class NumericsConfig:
    """Discretization and convergence settings (``src/jaxrans/core/config.py:112-154``)."""

    flux: str = "hllc"
    limiter: str = "minmod"
    cfl: float = 1.5
    max_iters: int = 40_000
    tol: float = 1.0e-12
    forces_tol: float | None = 1.0e-4
    n_ghost: int = 2
    solver: str = "explicit"
    far_field: str = "characteristic"
    kappa: float = 0.0

    def __post_init__(self) -> None:
        _enforce(self, _NUMERICS_RULES)

    dtype = property(lambda self: np.dtype("float64"))


_IMPLICIT_RULES: tuple[Rule, ...] = (
    (lambda c: c["cfl"] > 0.0, "cfl must be positive, got {cfl}"),
    (lambda c: c["cfl_max"] >= c["cfl"], "cfl_max must be at least the starting cfl {cfl}, got {cfl_max}"),
    (lambda c: 0.0 < c["cfl_min"] <= c["cfl"], "cfl_min must lie in (0, cfl], got {cfl_min}"),
    (lambda c: c["cfl_growth"] >= 1.0, "cfl_growth must be >= 1, got {cfl_growth}"),
    (lambda c: c["cfl_cutback"] > 1.0, "cfl_cutback must exceed 1, got {cfl_cutback}"),
    (lambda c: c["jacobian"] in JACOBIANS, "jacobian must be 'exact' or 'first_order', got {jacobian!r}"),
    (lambda c: 0.0 < c["krylov_tol"] < 1.0, "krylov_tol must lie in (0, 1), got {krylov_tol}"),
    (lambda c: c["krylov_subspace"] >= 1, "krylov_subspace must be >= 1, got {krylov_subspace}"),
    (lambda c: c["krylov_restarts"] >= 1, "krylov_restarts must be >= 1, got {krylov_restarts}"),
    (lambda c: _null_or(lambda r: r > 0.0)(c["max_relative_update"]),
     "max_relative_update must be positive or None, got {max_relative_update}"),
    (lambda c: 0.0 < c["krylov_success"] < 1.0, "krylov_success must lie in (0, 1), got {krylov_success}"),
    (lambda c: c["precision"] in PRECISIONS, "precision must be 'fp64' or 'mixed', got {precision!r}"),
    (lambda c: c["line_solver"] in LINE_SOLVERS,
     "line_solver must be 'cyclic_reduction' or 'thomas', got {line_solver!r}"),
)


@dataclass(frozen=True)
class ImplicitConfig:
    """Pseudo-transient Newton-Krylov settings (``src/jaxrans/core/config.py:157-212``).

    ``precision`` is this implementation's addition: ``"fp64"`` (the drop-in,
    every array double) or ``"mixed"`` (line-preconditioner blocks and their
    factor and solve in float32; residual, tangent, Krylov and update in
    float64).  ``line_solver`` is how each line chain is solved: block Thomas
    (``linalg.py:80-132``, the default: a lane of the wavefront per block entry,
    the fastest exact solve on the device) or block cyclic reduction (the
    reference's, ``src/jaxrans/core/linalg.py:159-200``); the two agree to
    round-off."""

    cfl: float = 10.0
    cfl_max: float = 1.0e3
    cfl_growth: float = 1.3
    cfl_cutback: float = 4.0
    cfl_min: float = 0.5
    preconditioner: str = "line_j"
    jacobian: str = "first_order"
    krylov_tol: float = 1.0e-2
    krylov_subspace: int = 10
    krylov_restarts: int = 2
    krylov_success: float = 0.5
    max_relative_update: float | None = None
    precision: str = "fp64"
    line_solver: str = "thomas"

    def __post_init__(self) -> None:
        _enforce(self, _IMPLICIT_RULES)
