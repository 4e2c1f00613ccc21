"""A case directory (``case.json`` + ``numerical.json``) -> a resolved ``Case``.

Keys, defaults, cross-key rules and error messages follow the reference
loader (``src/jaxrans/case/loader.py:31-373``, ``case/domain.py:404-538``), so
``load_case`` is a drop-in.  Grids are generated here (``o_grid``,
``c_grid``) or read from a 3D Plot3D file (``file``, ``case/domain.py:148-154``:
patches named by ``domain.patches`` box rules); 2D grid files and the
``multi_element`` source are refused.
"""

from __future__ import annotations

import copy
import json
import math
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from ..core import constants
from ..core.config import FlowConditions, ImplicitConfig, NumericsConfig
from ..geometry.airfoil import Airfoil, airfoil_source
from ..mesh.assemble import assemble
from ..mesh.assemble3d import PatchRule, is_3d, load_grid_3d
from ..mesh.atlas import Atlas, assign_conditions, build_atlas, patches
from ..mesh.atlas3d import build_atlas_3d
from ..mesh.block import Grid, folded
from ..mesh.coarsen import coarsen as coarsen_mesh
from ..mesh.generate import c_grid, o_grid
from . import schema

CASE_FILE, NUMERICAL_FILE = "case.json", "numerical.json"
PATCH_RULE_KEYS = ("name", "x_min", "x_max", "y_min", "y_max", "z_min", "z_max")


class CaseError(ValueError):
    """A case that cannot be run as written."""


@dataclass(frozen=True)
class GridSpec:
    grid: str = "c_grid"
    airfoil: str | None = None
    cells_airfoil: int | None = None
    cells_wake: int | None = None
    cells_normal: int | None = None
    wall_spacing: float | None = None
    far_field: float = 100.0
    trailing_edge: str = "as_given"
    leading_edge_spacing: float = 2e-3
    trailing_edge_spacing: float = 2e-3
    file: str | None = None
    format: str = "auto"
    coordinate_scale: float = 1.0
    patches: tuple = ()
    elements: tuple = ()
    cells_per_chord: float | None = None
    layer_cells: int = 32
    layer_thickness: float = 0.02
    gap_cells: int = 16
    even_cells: int = 4
    reference_length: float = 1.0
    reference_area: float | None = None
    moment_centre: tuple = (0.25, 0.0, 0.0)
    section: str = "streamwise"
    base_blocks: bool = False

    def airfoil_shape(self, base: Path) -> Airfoil:
        foil = airfoil_source(self.airfoil, base).unit_chord()
        if self.trailing_edge == "close":
            foil = foil.closed()
        return foil

    def rules(self) -> tuple:
        out = []
        for k, rule in enumerate(self.patches):
            if not isinstance(rule, dict) or "name" not in rule:
                raise CaseError(f"'domain.patches[{k}]' must be an object with a 'name'")
            unknown = sorted(set(rule) - set(PATCH_RULE_KEYS) - {"notes"})
            if unknown:
                raise CaseError(f"unknown key(s) {unknown} in 'domain.patches[{k}]'; allowed: {list(PATCH_RULE_KEYS)}")
            out.append(PatchRule(**{key: v for key, v in rule.items() if key != "notes"}))
        return tuple(out)

    def resolve(self, base: Path) -> Path:
        """The grid file, relative to the case directory or the working directory."""
        for root in (Path(base), Path.cwd()):
            path = Path(self.file) if Path(self.file).is_absolute() else root / self.file
            if path.is_file():
                return path
        raise FileNotFoundError(f"{self.file!r} not found in {base} or {Path.cwd()}")

    def build(self, base: Path):
        if self.grid == "file":
            path = self.resolve(base)
            if not is_3d(path, self.format):
                raise CaseError(f"{path.name}: 2D grid files are not read by this implementation (3D Plot3D only)")
            try:
                return load_grid_3d(path, self.format, self.coordinate_scale, self.rules())
            except ValueError as error:
                raise CaseError(f"{path.name}: {error}") from None
        if self.grid not in ("o_grid", "c_grid"):
            raise CaseError(f"domain.grid {self.grid!r} is not built by this implementation "
                            f"(generated o_grid and c_grid, 3D grid files)")
        foil = self.airfoil_shape(base)
        common = dict(cells_normal=self.cells_normal, wall_spacing=self.wall_spacing, far_field=self.far_field,
                      le=self.leading_edge_spacing, te=self.trailing_edge_spacing)
        if self.grid == "o_grid":
            raw = [o_grid(foil, self.cells_airfoil, **common)]
        else:
            if not foil.sharp:
                raise CaseError(f"a blunt trailing edge (gap {foil.trailing_edge_gap:.3g}) needs the base and wake "
                                f"blocks of a blunt C-grid, not built here; set domain.trailing_edge = \"close\" "
                                f"or use an o_grid")
            raw = [c_grid(foil, self.cells_airfoil, self.cells_wake, **common)]
        return assemble(raw)

    def describe(self) -> str:
        if self.grid == "file":
            return f"file {Path(self.file).name}"
        wake = f" + 2x{self.cells_wake} wake" if self.grid == "c_grid" else ""
        return (f"{self.grid} round {self.airfoil}: {self.cells_airfoil}{wake} x {self.cells_normal} cells, "
                f"wall spacing {self.wall_spacing:g}, far field {self.far_field:g}")


@dataclass(frozen=True)
class Case:
    name: str
    description: str
    output: Path
    grid_spec: GridSpec
    conditions: dict
    flow: FlowConditions
    numerics: NumericsConfig
    implicit: ImplicitConfig
    rk_scheme: str
    block_iters: int
    directory: Path
    raw: dict = field(compare=False, repr=False, default_factory=dict)
    sequencing: int = 0

    def grid(self, coarsen: int = 0) -> Atlas:
        """The case's grid, ``coarsen`` times coarsened by every other node
        (grid sequencing, ``src/jaxrans/mesh/coarsen.py:58-72``)."""
        mesh = self.grid_spec.build(self.directory)
        if self.flow.swept:
            # the 2D grid of an infinite swept wing: a streamwise section is
            # meshed normal to the sweep line (x scaled by cos(sweep),
            # src/jaxrans/case/loader.py:60-64); run() extrudes it for the device
            if mesh[0].dim != 2:
                raise CaseError("'freestream.sweep_deg' needs a 2D grid (an infinite swept wing)")
            if self.grid_spec.section == "streamwise":
                mesh = scale_x(mesh, math.cos(self.flow.sweep_rad))
        kinds = resolve_conditions(self.conditions, patches(mesh), self.flow)
        mesh = assign_conditions(mesh, kinds)
        if coarsen:
            try:
                for _ in range(coarsen):
                    mesh = coarsen_mesh(mesh)
            except ValueError as error:
                raise CaseError(str(error)) from None
        if mesh[0].dim == 3:
            if self.grid_spec.reference_area is None:
                raise CaseError("a 3D case needs domain.reference_area")
            return build_atlas_3d(mesh, self.numerics.n_ghost)
        if any(folded(b) for b in mesh):
            raise CaseError("the grid has folded cells")
        return build_atlas(mesh, self.numerics.n_ghost)


def scale_x(mesh, factor: float) -> tuple:
    """Every 2D block with ``x`` scaled by ``factor`` about ``x = 0`` and
    ``y`` kept (the same thickness on a shorter chord)."""
    return tuple(Grid(np.asarray(b.x_node) * factor, np.asarray(b.y_node), b.sides) for b in mesh)


class _Document:
    """One parsed JSON file of a case: its top-level sections checked against
    the schema once, then read section by section -- the keys given (unknown
    ones refused) or the full typed record with defaults filled in."""

    def __init__(self, file: str, doc: dict):
        self.file, self.doc = file, doc
        allowed = set(schema.FILES[file]) | ({"boundary_conditions"} if file == CASE_FILE else set())
        stray = sorted(set(doc) - allowed)
        if stray:
            raise CaseError(f"{file}: unknown section(s) {stray}; allowed: {sorted(schema.FILES[file])}")

    def given(self, name: str, required: bool = False) -> dict:
        """The keys written in section ``name`` (``notes`` dropped)."""
        if name not in self.doc:
            if required:
                raise CaseError(f"missing section '{name}'")
            return {}
        body = self.doc[name]
        if not isinstance(body, dict):
            raise CaseError(f"'{name}' must be an object, got {type(body).__name__}")
        table = schema.keys(self.file, name)
        stray = sorted(f"{name}.{k}" for k in body if k not in table and k != "notes")
        if stray:
            raise CaseError(f"unknown key(s) {stray}; allowed: {sorted(table)}")
        return {k: v for k, v in body.items() if k != "notes"}

    def record(self, name: str, required: bool = False, given: dict | None = None) -> dict:
        """Every key of section ``name``: the written value, type-checked, or
        the schema default; a required key that is absent is refused."""
        given = self.given(name, required) if given is None else given
        out = {}
        for key in schema.FILES[self.file][name]:
            where = f"{name}.{key.name}"
            if key.name not in given:
                if key.default is schema.REQUIRED:
                    raise CaseError(f"missing '{where}'")
                out[key.name] = key.default
                continue
            try:
                key.check(where, given[key.name])
            except (TypeError, ValueError) as error:
                raise CaseError(str(error)) from None
            out[key.name] = given[key.name]
        return out


def _domain(case: _Document) -> GridSpec:
    """``domain``: the grid source decides which keys apply and which are required."""
    written = case.given("domain", required=True)
    if "grid" not in written:
        raise CaseError(f"missing 'domain.grid'; choose from {sorted(schema.GRIDS)}")
    values = case.record("domain", given=written)
    source = values["grid"]
    table = schema.keys(CASE_FILE, "domain")
    misplaced = sorted(k for k in written if table[k].grids is not None and source not in table[k].grids)
    if misplaced:
        applicable = sorted(k for k, key in table.items() if key.grids and source in key.grids)
        raise CaseError(f"'domain.{misplaced[0]}' does not apply to grid {source!r}; it takes {applicable}")
    for name, key in table.items():
        if source in key.required_for and values[name] is None:
            raise CaseError(f"'domain.{name}' is required for grid {source!r}")
    kept = {k: v for k, v in values.items() if v is not None}
    kept.update({k: tuple(values[k]) for k in ("patches", "elements", "moment_centre")})
    return GridSpec(**kept)


# Viscosity laws of ``material_properties.dynamic_viscosity``: a function of
# (freestream record, material record) giving S / T_inf or None.
def _sutherland_ratio(free: dict, material: dict):
    if free["temperature"] is None or material["sutherland_temperature"] is None:
        raise CaseError("Sutherland's law needs 'freestream.temperature' and "
                        "'material_properties.sutherland_temperature', in one absolute unit")
    return material["sutherland_temperature"] / free["temperature"]


def _constant_ratio(free: dict, material: dict):
    if material["sutherland_temperature"] is not None:
        raise CaseError("'material_properties.sutherland_temperature' given with a constant viscosity")
    return None


VISCOSITY_LAWS = {"sutherland": _sutherland_ratio, "constant": _constant_ratio}
VISCOUS_ONLY = ("dynamic_viscosity", "prandtl_number", "sutherland_temperature")


def _flow_conditions(case: _Document) -> FlowConditions:
    """``freestream``, ``material_properties`` and ``turbulence`` -> the flow."""
    written = {name: case.given(name, required=name == "freestream")
               for name in ("freestream", "material_properties", "turbulence")}
    free, material, turb = (case.record(name, given=written[name]) for name in written)
    viscous = free["reynolds"] is not None
    ratio = None
    if not viscous:
        for key in VISCOUS_ONLY:
            if key in written["material_properties"]:
                raise CaseError(f"'material_properties.{key}' given for an inviscid case "
                                f"(freestream.reynolds is null)")
    else:
        ratio = VISCOSITY_LAWS[material["dynamic_viscosity"]](free, material)
        if material["prandtl_number"] != constants.PRANDTL_LAMINAR:
            raise CaseError(f"'material_properties.prandtl_number' = {material['prandtl_number']} is not "
                            f"implemented; only {constants.PRANDTL_LAMINAR}")
    model = None if turb["model"] == "none" else turb["model"]
    if model is not None and not viscous:
        raise CaseError("'turbulence.model' needs a viscous case (freestream.reynolds)")
    if model is None and "freestream_nu_tilde_ratio" in written["turbulence"]:
        raise CaseError("'turbulence.freestream_nu_tilde_ratio' given without a model")
    try:
        return FlowConditions(mach=free["mach"], alpha_deg=free["alpha_deg"], reynolds=free["reynolds"],
                              gamma=material["specific_heat_ratio"], sutherland_ratio=ratio, turbulence=model,
                              nu_tilde_ratio=turb["freestream_nu_tilde_ratio"], sweep_deg=free["sweep_deg"])
    except ValueError as error:
        raise CaseError(str(error)) from None


def default_conditions(flow: FlowConditions) -> dict[str, str]:
    return {"wall": "slip_wall" if flow.inviscid else "no_slip_wall", "farfield": "far_field",
            "symmetry": "symmetry", "inflow": "inflow_total", "outflow": "outflow_static"}


def _patch_entry(patch: str, entry, flow: FlowConditions) -> str:
    """One ``boundary_conditions.<patch>`` object, checked; its condition type."""
    where = f"boundary_conditions.{patch}"
    if not isinstance(entry, dict):
        raise CaseError(f"'{where}' must be an object with a 'type', got {entry!r}")
    table = {k.name: k for k in schema.BOUNDARY_KEYS}
    stray = sorted(f"{where}.{k}" for k in entry if k not in table and k != "notes")
    if stray:
        raise CaseError(f"unknown key(s) {stray}; allowed: {sorted(table)}")
    if "type" not in entry:
        raise CaseError(f"missing '{where}.type'")
    for name, value in entry.items():
        if name == "notes":
            continue
        try:
            table[name].check(f"{where}.{name}", value)
        except (TypeError, ValueError) as error:
            raise CaseError(str(error)) from None
    kind = entry["type"]
    if "model" in entry and kind != "far_field":
        raise CaseError(f"'{where}.model' applies to a far_field patch only")
    if kind == "no_slip_wall" and flow.inviscid:
        raise CaseError(f"'{where}.type' = 'no_slip_wall' in an inviscid case (freestream.reynolds is null)")
    return kind


def _boundaries(case_doc: dict, flow: FlowConditions) -> tuple[dict, str]:
    """Named patch conditions and the one far-field model of the case."""
    section = case_doc.get("boundary_conditions", {})
    if not isinstance(section, dict):
        raise CaseError("'boundary_conditions' must be an object")
    named = {p: _patch_entry(p, e, flow) for p, e in section.items() if p != "notes"}
    models = sorted({e.get("model", "freestream") for p, e in section.items()
                     if p != "notes" and named[p] == "far_field"})
    if len(models) > 1:
        raise CaseError(f"far-field patches name different models {models}; one model per case")
    return named, schema.FAR_FIELD_MODELS[models[0] if models else "freestream"]


def resolve_conditions(named: dict, grid_patches: list, flow: FlowConditions) -> dict:
    unknown = [p for p in named if p not in grid_patches]
    if unknown:
        raise CaseError(f"'boundary_conditions' names patch(es) {unknown} the grid does not have; "
                        f"its patches are {grid_patches}")
    defaults = default_conditions(flow)
    missing = [p for p in grid_patches if p not in named and p not in defaults]
    if missing:
        raise CaseError(f"no boundary condition for patch(es) {missing}: give each a 'type' in "
                        f"'boundary_conditions' (the grid's patches are {grid_patches})")
    return {p: named.get(p, defaults.get(p)) for p in grid_patches}


def parse_case(case_doc: dict, numerical_doc: dict, directory: Path = Path(".")) -> Case:
    """Both documents of a case -> the resolved ``Case`` (``CaseError`` on the
    first thing that cannot run as written)."""
    case, numerical = _Document(CASE_FILE, case_doc), _Document(NUMERICAL_FILE, numerical_doc)
    general = case.record("general", required=True)
    spec = _domain(case)
    flow = _flow_conditions(case)
    if "section" in case_doc["domain"] and not flow.swept:
        raise CaseError("'domain.section' applies to a swept case only ('freestream.sweep_deg')")
    conditions, far_field = _boundaries(case_doc, flow)
    fluxes = numerical.record("convective_fluxes")
    stepping = numerical.record("time_integration")
    implicit = numerical.record("implicit")
    stop = numerical.record("convergence")
    try:
        implicit_cfg = ImplicitConfig(**implicit)
        numerics = NumericsConfig(flux=fluxes["riemann_solver"], limiter=fluxes["limiter"],
                                  kappa=fluxes["muscl_kappa"], cfl=stepping["cfl"], max_iters=stop["max_iters"],
                                  tol=stop["tol"], forces_tol=stop["forces_tol"], solver=stepping["solver"],
                                  far_field=far_field)
    except ValueError as error:
        raise CaseError(str(error)) from None
    # one host synchronisation per block: 20 Newton steps, or 200 explicit ones
    block_iters = stop["block_iters"] or {"explicit": 200, "implicit": 20}[numerics.solver]
    if block_iters < 1:
        raise CaseError(f"'convergence.block_iters' must be >= 1, got {block_iters}")
    if stop["grid_sequencing"] < 0:
        raise CaseError(f"'convergence.grid_sequencing' must be >= 0, got {stop['grid_sequencing']}")
    directory = Path(directory)
    out = Path(general["save_path"])
    return Case(name=general["case_name"], description=general["description"] or general["case_name"],
                output=out if out.is_absolute() else directory / out, grid_spec=spec, conditions=conditions,
                flow=flow, numerics=numerics, implicit=implicit_cfg, rk_scheme=stepping["rk_scheme"],
                block_iters=block_iters, directory=directory, raw={"case": case_doc, "numerical": numerical_doc},
                sequencing=stop["grid_sequencing"])


def apply_override(docs: dict, assignment: str) -> None:
    """Apply ``file.section.key=value`` to ``docs`` in place; the value is
    read as JSON when it parses, else kept as a bare string."""
    path, sep, text = assignment.partition("=")
    if not sep:
        raise CaseError(f"override {assignment!r} is not of the form file.section.key=value")
    target = path.strip().split(".")
    if len(target) != 3 or target[0] not in docs:
        raise CaseError(f"override path {path!r} must be case|numerical.section.key, "
                        f"e.g. numerical.convective_fluxes.limiter")
    try:
        value = json.loads(text)
    except json.JSONDecodeError:
        value = text
    file, section, key = target
    docs[file].setdefault(section, {})[key] = value


def _read_json(path: Path) -> dict:
    if not path.is_file():
        return {}
    try:
        return json.loads(path.read_text())
    except json.JSONDecodeError as error:
        raise CaseError(f"{path}: {error}") from None


def load_case(directory: str | Path, overrides: tuple[str, ...] = ()) -> Case:
    """The case in ``directory`` (``case.json``, optional ``numerical.json``)
    with ``overrides`` applied (``file.section.key=value`` each)."""
    directory = Path(directory)
    if not (directory / CASE_FILE).is_file():
        raise FileNotFoundError(f"{directory} holds no {CASE_FILE}")
    docs = {"case": _read_json(directory / CASE_FILE), "numerical": _read_json(directory / NUMERICAL_FILE)}
    docs = copy.deepcopy(docs)
    for assignment in overrides:
        apply_override(docs, assignment)
    try:
        return parse_case(docs["case"], docs["numerical"], directory)
    except CaseError as error:
        raise CaseError(f"{directory}: {error}") from None
