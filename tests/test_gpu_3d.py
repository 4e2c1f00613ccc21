"""The 3D device path against the 3D oracle on the same seeded inputs: a
small four-block wing (swept, tapered; interfaces between blocks, wake cuts,
corner ghosts composing three conditions) and an extruded O-grid.

Tolerances as the 2D parity tests: residual 1e-12, tangent 1e-11 normwise
relative; wave speeds 1e-13; the FP64 line solve 1e-10, FP32 2e-5; Newton
steps: norms rtol 1e-9, state 1e-8; forces 1e-11.
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import o_atlas, perturbed
from grids3d import extruded_atlas, small_wing
from jaxrans_amd.core.config import FlowConditions, ImplicitConfig, NumericsConfig
from jaxrans_amd.device import context_for
from oracle import residual3d as O3
from oracle import solver as OS
from oracle import solver3d as S3

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def euler():
    return FlowConditions(mach=0.8, alpha_deg=2.0)


def sa():
    return FlowConditions(mach=0.8, alpha_deg=2.0, reynolds=1e7, turbulence="sa_noft2")


def laminar():
    return FlowConditions(mach=0.5, alpha_deg=1.0, reynolds=5e3, sutherland_ratio=198.6 / 460.0)


_ATLAS = {}


def atlas_of(kind):
    if kind not in _ATLAS:
        _ATLAS[kind] = {"wing": lambda: small_wing(), "wing-slip": lambda: small_wing(wall="slip_wall"),
                        "wing-49": lambda: small_wing(cells_normal=49, cells_span=3, blocks=2),
                        "extruded": lambda: extruded_atlas(o_atlas(wall="no_slip_wall"), 3, 0.1)}[kind]()
    return _ATLAS[kind]


CASES = [
    ("wing-euler-hllc-minmod", "wing-slip", euler, "hllc", "minmod"),
    ("wing49-sa-hllc-minmod", "wing-49", sa, "hllc", "minmod"),
    ("wing-euler-roe-vas", "wing-slip", euler, "roe", "van_albada_smooth"),
    ("wing-sa-roe-vas", "wing", sa, "roe", "van_albada_smooth"),
    ("wing-sa-hllc-vanalbada", "wing", sa, "hllc", "van_albada"),
    ("extruded-laminar-hllc-minmod", "extruded", laminar, "hllc", "minmod"),
]


def pick(*names):
    chosen = [c for c in CASES if c[0] in names]
    assert len(chosen) == len(names)
    return pytest.mark.parametrize("name, kind, flowf, flux, lim", chosen, ids=[c[0] for c in chosen])


def _setup(kind, flowf, flux, lim, implicit=None):
    atlas = atlas_of(kind)
    flow = flowf()
    num = NumericsConfig(flux=flux, limiter=lim, solver="implicit")
    st = O3.setup_for(atlas, flow, num)
    ctx = context_for(atlas, flow, num, implicit=implicit)
    ctx.set_freestream(st.q_inf, None if flow.inviscid else flow.mu_inf)
    return atlas, flow, num, st, ctx


def _state(st, amp=1e-2):
    return np.where(st.atlas.interior[None], perturbed(O3.U0_of(st), amp), O3.U0_of(st))


@pytest.mark.parametrize("name, kind, flowf, flux, lim", CASES, ids=[c[0] for c in CASES])
def test_residual_3d_matches_oracle(name, kind, flowf, flux, lim):
    print(f"[gpu] residual3d {name}", flush=True)
    atlas, flow, num, st, ctx = _setup(kind, flowf, flux, lim)
    U = _state(st)
    for first_order in (False, True):
        ref = O3.residual(U, st, first_order)
        got = ctx.residual(U, first_order)[0]
        assert np.all(np.isfinite(got))
        assert rel(got, ref) < 1e-12, (name, first_order, rel(got, ref))
    ctx.close()


@pytest.mark.parametrize("name, kind, flowf, flux, lim", CASES, ids=[c[0] for c in CASES])
def test_tangent_3d_matches_oracle(name, kind, flowf, flux, lim):
    print(f"[gpu] tangent3d {name}", flush=True)
    atlas, flow, num, st, ctx = _setup(kind, flowf, flux, lim)
    U = _state(st)
    rng = np.random.default_rng(7)
    dU = rng.standard_normal(U.shape) * np.abs(U).mean(axis=(1, 2, 3), keepdims=True) * 1e-3
    for first_order in (True, False):
        ref = O3.jvp(U, dU, st, first_order)
        got = ctx.tangent(U, dU, first_order=first_order)[0]
        assert rel(got, ref) < 1e-11, (name, first_order, rel(got, ref))
    ctx.close()


@pick("wing-euler-hllc-minmod", "wing-sa-roe-vas", "extruded-laminar-hllc-minmod")
def test_wave_speeds_3d_match_oracle(name, kind, flowf, flux, lim):
    print(f"[gpu] speeds3d {name}", flush=True)
    atlas, flow, num, st, ctx = _setup(kind, flowf, flux, lim)
    U = _state(st)
    lams = O3.wave_speeds(U, st)
    lam, lsum = ctx.wave_speeds(U)
    ref = np.concatenate([x.ravel() for x in lams])
    assert rel(lam[0], ref) < 1e-13
    assert rel(lsum[0], O3.wave_speed_sum(lams)) < 1e-13
    ctx.close()


@pytest.mark.parametrize("precision, tol", [("fp64", 1e-10), ("mixed", 2e-5)])
@pick("wing-euler-hllc-minmod", "wing-sa-roe-vas", "wing49-sa-hllc-minmod")
def test_line_preconditioner_3d_matches_oracle(name, kind, flowf, flux, lim, precision, tol):
    print(f"[gpu] precond3d {name} {precision}", flush=True)
    cfg = ImplicitConfig(cfl=25.0, precision=precision)
    atlas, flow, num, st, ctx = _setup(kind, flowf, flux, lim, implicit=cfg)
    U = _state(st)
    ctx.set_cfl(25.0)
    ctx.precond_factor(U)
    r = np.random.default_rng(3).standard_normal(U.shape)
    ref = S3.LinePreconditioner3D(U, O3.wave_speeds(U, st), 25.0, st, "thomas")(r)
    got = ctx.precond_apply(r)[0]
    assert rel(got, ref) < tol, rel(got, ref)
    ctx.close()


@pick("wing-euler-roe-vas", "wing-sa-roe-vas", "wing49-sa-hllc-minmod", "extruded-laminar-hllc-minmod")
def test_newton_steps_3d_match_oracle(name, kind, flowf, flux, lim):
    print(f"[gpu] newton3d {name}", flush=True)
    cfg = ImplicitConfig(krylov_subspace=12, krylov_restarts=2, max_relative_update=0.5)
    atlas, flow, num, st, ctx = _setup(kind, flowf, flux, lim, implicit=cfg)
    U = _state(st)
    ctx.set_state(U)
    ctx.newton_init(cfg.cfl)
    s = OS.NewtonState(U, cfg.cfl)
    norms = []
    for _ in range(3):
        s, n, _ = S3.newton_step(s, st, cfg, "thomas")
        norms.append(n)
    hist = ctx.newton_block(3)[:, 0]
    assert np.allclose(hist, norms, rtol=1e-9, atol=0.0), (hist, norms)
    assert rel(ctx.get_state()[0] - U, s.U - U) < 1e-8
    assert ctx.scalars()[0, 0] == pytest.approx(s.cfl, rel=1e-12)
    ctx.close()


@pick("wing-euler-hllc-minmod", "wing-sa-roe-vas", "extruded-laminar-hllc-minmod")
def test_forces_3d_match_oracle(name, kind, flowf, flux, lim):
    print(f"[gpu] forces3d {name}", flush=True)
    atlas, flow, num, st, ctx = _setup(kind, flowf, flux, lim)
    U = _state(st)
    ctx.set_state(U)
    got = ctx.forces3(1.3, 0.7, (0.25, 0.0, 0.0))[0]
    ref = S3.forces(U, st, 1.3, 0.7, (0.25, 0.0, 0.0))
    for k, key in enumerate(("cl", "cd", "cm", "cd_friction")):
        assert got[k] == pytest.approx(ref[key], rel=1e-11, abs=1e-14), key
    assert got[7] == pytest.approx(ref["cy"], rel=1e-11, abs=1e-14)
    ctx.close()


def test_run_on_the_m6_file_case_matches_oracle_newton_steps():
    """``run`` on the reference's ONERA M6 example read from its Plot3D file
    (a pole, an O seam, Sutherland viscosity, SA): four Newton steps in two
    blocks against the oracle's, and the forces by the case's reference area,
    length and moment centre."""
    from pathlib import Path

    from jaxrans_amd.case.loader import load_case
    from jaxrans_amd.simulation import run

    print("[gpu] run m6 file case", flush=True)
    directory = Path(__file__).resolve().parent / "golden" / "cases" / "onera_m6_sa"
    case = load_case(directory, ("numerical.convergence.max_iters=4", "numerical.convergence.block_iters=2"))
    atlas = case.grid()
    st = O3.setup_for(atlas, case.flow, case.numerics)
    U0 = _state(st)          # perturbed: a freestream start leaves a residual of pure cancellation
    result = run(case, atlas, initial=U0)
    s = OS.NewtonState(U0, case.implicit.cfl)
    norms = []
    with np.errstate(divide="ignore", invalid="ignore"):
        for _ in range(4):
            s, n, _ = S3.newton_step(s, st, case.implicit, "thomas")
            norms.append(n)
    assert result.history.n_iters == 4
    assert np.allclose(result.history.residual, norms, rtol=1e-9, atol=0.0), (result.history.residual, norms)
    assert rel(result.U - U0, s.U - U0) < 1e-8
    spec = case.grid_spec
    ref = S3.forces(s.U, st, spec.reference_area, spec.reference_length, tuple(spec.moment_centre))
    for key in ("cl", "cd", "cm", "cd_friction", "cy"):
        assert getattr(result.forces, key) == pytest.approx(ref[key], rel=1e-8, abs=1e-12), key


@pytest.mark.parametrize("which", ["euler_o48", "onera_m6_sa"])
def test_run_polar_equals_each_run_alone(which):
    """``run_polar`` (the reference's batched polar, ``src/jaxrans/polar.py:80-121``)
    on a 2D generated grid and on the 3D M6 file grid: every case of the batch
    has, bit for bit, the residual history and forces of ``run`` on that
    freestream alone."""
    from dataclasses import replace
    from pathlib import Path

    from jaxrans_amd.case.loader import load_case
    from jaxrans_amd.simulation import run, run_polar

    print(f"[gpu] run_polar {which}", flush=True)
    directory = Path(__file__).resolve().parent / "golden" / "cases" / which
    case = load_case(directory, ("numerical.convergence.max_iters=4", "numerical.convergence.block_iters=2",
                                 "numerical.convergence.forces_tol=null"))
    grid = case.grid()
    alphas, machs = [1.0, 3.5], [0.5, 0.7]
    polar = run_polar(case, alphas, grid, None, None, machs=machs)   # the reference's positional order
    batch = polar.results
    assert len(batch) == 2 and polar.table().shape == (2, 9)
    assert np.array_equal(polar.table()[:, 0], alphas)
    for k, (a, m) in enumerate(zip(alphas, machs)):
        alone = run(replace(case, flow=replace(case.flow, alpha_deg=a, mach=m)), grid)
        assert np.array_equal(alone.history.residual, batch[k].history.residual), (k, alone.history.residual,
                                                                                  batch[k].history.residual)
        for key in ("cl", "cd", "cm", "cd_friction", "cy"):
            assert getattr(alone.forces, key) == getattr(batch[k].forces, key), (k, key)


@pick("wing-euler-hllc-minmod", "wing-sa-roe-vas", "extruded-laminar-hllc-minmod")
def test_rk_steps_3d_match_oracle(name, kind, flowf, flux, lim):
    """Explicit multistage steps on a 3D canvas (jr_rk_block; the oracle's
    ``rk_step`` with the 3D local time step): state after 3 RK5 steps to
    1e-12, residual history to 1e-12."""
    from oracle import explicit as OX
    print(f"[gpu] 3D explicit {name}", flush=True)
    atlas, flow, num, st, ctx = _setup(kind, flowf, flux, lim)
    alphas = OX.RK_SCHEMES["rk5_jameson"]
    cfl = 1.0
    U = _state(st, 1e-3)
    ctx.set_state(U)
    ctx.newton_init(cfl)
    hist = ctx.rk_block(3, alphas)[:, 0]
    ref, norms = U, []
    for _ in range(3):
        ref, n = OX.rk_step(ref, st, alphas, cfl)
        norms.append(n)
    got = ctx.get_state()[0]
    ctx.close()
    assert rel(got, ref) < 1e-12, (name, rel(got, ref))
    assert np.allclose(hist, norms, rtol=1e-12, atol=0.0), (hist, norms)
