"""Keys of ``case.json`` and ``numerical.json``: names, types, defaults, choices.

The key table mirrors ``src/jaxrans/case/schema.py:67-204`` so a reference case
directory loads unchanged.  Grid sources this implementation does not build
(``file``, ``multi_element``) are accepted by the schema and refused by the
loader with a ``CaseError`` that says so.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any

from ..core import constants
from ..core.config import TURBULENCE_MODELS, ImplicitConfig, NumericsConfig
from ..mesh.block import CONDITION_KINDS
from ..mesh.generate import LEADING_EDGE_SPACING, TRAILING_EDGE_SPACING

REQUIRED = "required"
GENERATED = ("o_grid", "c_grid")
SECTION = ("multi_element",)
MADE = GENERATED + SECTION
GRIDS = ("c_grid", "file", "multi_element", "o_grid")
FLUXES = ("hllc", "roe")
LIMITERS = ("minmod", "none", "van_albada", "van_albada_smooth")
PRECONDITIONERS = ("line_j", "none")
RK_SCHEMES = ("euler", "rk4_jameson", "rk5_jameson")


@dataclass(frozen=True)
class Key:
    name: str
    type: str
    default: Any
    choices: tuple | None = None
    grids: tuple | None = None
    required_for: tuple = ()

    def check(self, path: str, value: Any) -> None:
        if value is None:
            if not self.type.endswith("|null"):
                raise TypeError(f"'{path}' may not be null")
            return
        base = self.type.split("|")[0]
        ok = {"number": lambda v: isinstance(v, (int, float)) and not isinstance(v, bool),
              "integer": lambda v: isinstance(v, int) and not isinstance(v, bool),
              "string": lambda v: isinstance(v, str),
              "boolean": lambda v: isinstance(v, bool),
              "list": lambda v: isinstance(v, list)}[base](value)
        if not ok:
            raise TypeError(f"'{path}' must be {'an' if base[0] in 'aeiou' else 'a'} {base}, got {value!r}")
        if self.choices is not None and value not in self.choices:
            raise ValueError(f"'{path}' = {value!r} is not implemented; choose from {sorted(self.choices)}")


_N, _I = NumericsConfig, ImplicitConfig

CASE: dict[str, tuple[Key, ...]] = {
    "general": (
        Key("case_name", "string", REQUIRED),
        Key("description", "string", None),
        Key("save_path", "string", "results"),
    ),
    "domain": (
        Key("grid", "string", REQUIRED, choices=GRIDS),
        Key("airfoil", "string", None, grids=GENERATED, required_for=GENERATED),
        Key("cells_airfoil", "integer", None, grids=GENERATED, required_for=GENERATED),
        Key("cells_wake", "integer", None, grids=("c_grid",), required_for=("c_grid",)),
        Key("cells_normal", "integer", None, grids=GENERATED, required_for=GENERATED),
        Key("wall_spacing", "number", None, grids=MADE, required_for=MADE),
        Key("trailing_edge", "string", "as_given", choices=("as_given", "close"), grids=GENERATED),
        Key("far_field", "number", 100.0, grids=MADE),
        Key("leading_edge_spacing", "number", LEADING_EDGE_SPACING, grids=MADE),
        Key("trailing_edge_spacing", "number", TRAILING_EDGE_SPACING, grids=MADE),
        Key("elements", "list", [], grids=SECTION, required_for=SECTION),
        Key("cells_per_chord", "integer", None, grids=SECTION, required_for=SECTION),
        Key("layer_cells", "integer", 32, grids=SECTION),
        Key("layer_thickness", "number", 0.02, grids=SECTION),
        Key("gap_cells", "integer", 16, grids=SECTION),
        Key("base_blocks", "boolean", False, grids=("multi_element",)),
        Key("even_cells", "integer", 4, grids=SECTION),
        Key("file", "string", None, grids=("file", "multi_element"), required_for=("file",)),
        Key("format", "string", "auto", choices=("auto", "cgns", "plot3d"), grids=("file",)),
        Key("coordinate_scale", "number", 1.0, grids=("file", "multi_element")),
        Key("patches", "list", [], grids=("file",)),
        Key("reference_length", "number", 1.0),
        Key("reference_area", "number|null", None),
        Key("moment_centre", "list", [0.25, 0.0, 0.0]),
        Key("section", "string", "streamwise", choices=("normal", "streamwise")),
    ),
    "freestream": (
        Key("mach", "number", REQUIRED),
        Key("alpha_deg", "number", 0.0),
        Key("sweep_deg", "number|null", None),
        Key("reynolds", "number|null", None),
        Key("temperature", "number", None),
    ),
    "material_properties": (
        Key("specific_heat_ratio", "number", constants.GAMMA_AIR),
        Key("dynamic_viscosity", "string", "constant", choices=("constant", "sutherland")),
        Key("sutherland_temperature", "number", None),
        Key("prandtl_number", "number", constants.PRANDTL_LAMINAR),
    ),
    "turbulence": (
        Key("model", "string", "none", choices=("none", *TURBULENCE_MODELS)),
        Key("freestream_nu_tilde_ratio", "number", 3.0),
    ),
}

NUMERICAL: dict[str, tuple[Key, ...]] = {
    "convective_fluxes": (
        Key("riemann_solver", "string", _N.flux, choices=FLUXES),
        Key("limiter", "string", _N.limiter, choices=LIMITERS),
        Key("muscl_kappa", "number", _N.kappa),
    ),
    "time_integration": (
        Key("solver", "string", _N.solver, choices=("explicit", "implicit")),
        Key("rk_scheme", "string", "rk5_jameson", choices=RK_SCHEMES),
        Key("cfl", "number", _N.cfl),
    ),
    "implicit": (
        Key("cfl", "number", _I.cfl),
        Key("cfl_max", "number", _I.cfl_max),
        Key("cfl_growth", "number", _I.cfl_growth),
        Key("cfl_cutback", "number", _I.cfl_cutback),
        Key("cfl_min", "number", _I.cfl_min),
        Key("preconditioner", "string", _I.preconditioner, choices=PRECONDITIONERS),
        Key("jacobian", "string", _I.jacobian, choices=("exact", "first_order")),
        Key("krylov_tol", "number", _I.krylov_tol),
        Key("krylov_subspace", "integer", _I.krylov_subspace),
        Key("krylov_restarts", "integer", _I.krylov_restarts),
        Key("krylov_success", "number", _I.krylov_success),
        Key("max_relative_update", "number|null", _I.max_relative_update),
        Key("precision", "string", _I.precision, choices=("fp64", "mixed")),
        Key("line_solver", "string", _I.line_solver, choices=("cyclic_reduction", "thomas")),
    ),
    "convergence": (
        Key("max_iters", "integer", _N.max_iters),
        Key("tol", "number", _N.tol),
        Key("forces_tol", "number|null", _N.forces_tol),
        Key("grid_sequencing", "integer", 0),
        Key("block_iters", "integer", None),
    ),
}

FILES = {"case.json": CASE, "numerical.json": NUMERICAL}
FAR_FIELD_MODELS = {"freestream": "characteristic", "point_vortex": "characteristic_point_vortex"}
BOUNDARY_KEYS = (
    Key("type", "string", REQUIRED, choices=tuple(CONDITION_KINDS)),
    Key("model", "string", "freestream", choices=tuple(FAR_FIELD_MODELS)),
)


def keys(file: str, section: str) -> dict[str, Key]:
    return {k.name: k for k in FILES[file][section]}
