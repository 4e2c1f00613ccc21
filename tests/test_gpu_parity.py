"""Device kernels against the oracle on the same seeded inputs.

Tolerances: residual and tangent 1e-12 normwise relative (the FP64 path
restates the oracle's arithmetic; ulp-level differences only), wave speeds
1e-13, the FP64 line solve 1e-10 (the same elimination on both sides --
block Thomas or block cyclic reduction -- in a different summation order),
the FP32 line solve 2e-5.
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import c_atlas, o_atlas, perturbed
from jaxrans_amd.core.config import FlowConditions, ImplicitConfig, NumericsConfig
from jaxrans_amd.device import context_for
from oracle import residual as OR
from oracle import solver as OS

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def euler_flow():
    return FlowConditions(mach=0.5, alpha_deg=1.25)


def sa_flow():
    return FlowConditions(mach=0.73, alpha_deg=2.5, reynolds=6.5e6, turbulence="sa_noft2")


def lam_flow():
    return FlowConditions(mach=0.5, alpha_deg=1.0, reynolds=5e3, sutherland_ratio=198.6 / 460.0)


CASES = [
    ("euler-hllc-minmod", "o", euler_flow, "hllc", "minmod"),
    ("euler-hllc-vanalbada", "o", euler_flow, "hllc", "van_albada"),
    ("euler-roe-vas", "o", euler_flow, "roe", "van_albada_smooth"),
    ("euler-roe-none", "o", euler_flow, "roe", "none"),
    ("laminar-hllc-minmod", "ol", lam_flow, "hllc", "minmod"),
    ("sa-roe-vas", "c", sa_flow, "roe", "van_albada_smooth"),
    ("sa-hllc-vanalbada", "c", sa_flow, "hllc", "van_albada"),
]


def _setup(kind, flowf, flux, lim, implicit=None):
    atlas = {"o": lambda: o_atlas(), "ol": lambda: o_atlas(wall="no_slip_wall"), "c": lambda: c_atlas()}[kind]()
    flow = flowf()
    num = NumericsConfig(flux=flux, limiter=lim, solver="implicit")
    st = OR.setup_for(atlas, flow, num)
    ctx = context_for(atlas, flow, num, implicit=implicit)
    ctx.set_freestream(st.q_inf, None if flow.inviscid else flow.mu_inf)
    return atlas, flow, num, st, ctx


@pytest.mark.parametrize("name, kind, flowf, flux, lim", CASES, ids=[c[0] for c in CASES])
def test_residual_matches_oracle(name, kind, flowf, flux, lim):
    print(f"[gpu] residual {name}", flush=True)
    atlas, flow, num, st, ctx = _setup(kind, flowf, flux, lim)
    U = perturbed(OR.U0_of(st), 1e-2)
    for first_order in (False, True):
        ref = OR.residual(U, st, first_order)
        got = ctx.residual(U, first_order)[0]
        assert rel(got, ref) < 1e-12, (name, first_order, rel(got, ref))
        assert np.all(np.isfinite(got))


@pytest.mark.parametrize("name, kind, flowf, flux, lim", CASES, ids=[c[0] for c in CASES])
def test_tangent_matches_oracle(name, kind, flowf, flux, lim):
    print(f"[gpu] tangent {name}", flush=True)
    atlas, flow, num, st, ctx = _setup(kind, flowf, flux, lim)
    U = perturbed(OR.U0_of(st), 1e-2)
    rng = np.random.default_rng(7)
    dU = rng.standard_normal(U.shape) * np.abs(U).mean(axis=(1, 2), keepdims=True) * 1e-3
    for first_order in (True, False):
        ref = OR.jvp(U, dU, st, first_order)
        got = ctx.tangent(U, dU, first_order=first_order)[0]
        assert rel(got, ref) < 1e-11, (name, first_order, rel(got, ref))


@pytest.mark.parametrize("name, kind, flowf, flux, lim", CASES[::2], ids=[c[0] for c in CASES[::2]])
def test_wave_speeds_match_oracle(name, kind, flowf, flux, lim):
    atlas, flow, num, st, ctx = _setup(kind, flowf, flux, lim)
    U = perturbed(OR.U0_of(st), 1e-2)
    lams = OR.wave_speeds(U, st)
    lam, lsum = ctx.wave_speeds(U)
    ref = np.concatenate([lams[0].ravel(), lams[1].ravel()])
    assert rel(lam[0], ref) < 1e-13
    assert rel(lsum[0], OR.wave_speed_sum(lams)) < 1e-13


@pytest.mark.parametrize("solver, method", [("cyclic_reduction", "cr"), ("thomas", "thomas")])
@pytest.mark.parametrize("precision, tol", [("fp64", 1e-10), ("mixed", 2e-5)])
@pytest.mark.parametrize("name, kind, flowf, flux, lim", [CASES[0], CASES[5]], ids=[CASES[0][0], CASES[5][0]])
def test_line_preconditioner_matches_oracle(name, kind, flowf, flux, lim, precision, tol, solver, method):
    print(f"[gpu] precond {name} {precision} {solver}", flush=True)
    cfg = ImplicitConfig(cfl=25.0, cfl_max=1e3, precision=precision, line_solver=solver)
    atlas, flow, num, st, ctx = _setup(kind, flowf, flux, lim, implicit=cfg)
    U = perturbed(OR.U0_of(st), 1e-2)
    ctx.set_cfl(25.0)
    ctx.precond_factor(U)
    rng = np.random.default_rng(3)
    r = rng.standard_normal(U.shape)
    lams = OR.wave_speeds(U, st)
    ref = OS.LinePreconditioner(U, lams, 25.0, st, method)(r)
    got = ctx.precond_apply(r)[0]
    assert rel(got, ref) < tol, rel(got, ref)


@pytest.mark.parametrize("solver", ["cyclic_reduction", "thomas"])
@pytest.mark.parametrize("name, kind, flowf, flux, lim", [CASES[1], CASES[4], CASES[5]],
                         ids=[CASES[1][0], CASES[4][0], CASES[5][0]])
def test_newton_steps_match_oracle(name, kind, flowf, flux, lim, solver):
    """Four Newton steps from a perturbed state (no exact ties in the
    tangent): reported norms, the state and the CFL ramp follow the oracle;
    GMRES internals differ only in round-off.  Both line solvers: the apply
    also normalises the Arnoldi vector and writes the tangent's input."""
    print(f"[gpu] newton {name} {solver}", flush=True)
    cfg = ImplicitConfig(krylov_subspace=12, krylov_restarts=2, line_solver=solver)
    atlas, flow, num, st, ctx = _setup(kind, flowf, flux, lim, implicit=cfg)
    U = perturbed(OR.U0_of(st), 1e-2)
    ctx.set_state(U)
    ctx.newton_init(cfg.cfl)
    s = OS.NewtonState(U, cfg.cfl)
    norms = []
    for _ in range(4):
        s, n, info = OS.newton_step(s, st, cfg)
        norms.append(n)
    hist = ctx.newton_block(4)[:, 0]
    assert np.allclose(hist, norms, rtol=1e-9, atol=0.0), (hist, norms)
    Ug = ctx.get_state()[0]
    assert rel(Ug - U, s.U - U) < 1e-8
    assert ctx.scalars()[0, 0] == pytest.approx(s.cfl, rel=1e-12)


@pytest.mark.parametrize("name, kind, flowf, flux, lim", [CASES[1], CASES[4], CASES[5]],
                         ids=[CASES[1][0], CASES[4][0], CASES[5][0]])
def test_tile_tangent_equals_the_node_face_and_cell_kernels(name, kind, flowf, flux, lim):
    """The Krylov tangent by LDS tiles (k_tangent_tile) against the separate
    node, face and cell kernels: four Newton steps from the same state give
    the same residual history and state to round-off, Euler, laminar and SA
    on the O- and C-grids (tiles cut by the canvas edges in both directions)."""
    print(f"[gpu] tile tangent {name}", flush=True)
    cfg = ImplicitConfig(krylov_subspace=12, krylov_restarts=2)
    runs = []
    for tile in (True, False):
        atlas, flow, num, st, ctx = _setup(kind, flowf, flux, lim, implicit=cfg)
        U = perturbed(OR.U0_of(st), 1e-2)
        ctx.set_tile_tangent(tile)
        ctx.set_state(U)
        ctx.newton_init(cfg.cfl)
        runs.append((ctx.newton_block(4)[:, 0], ctx.get_state()[0] - U))
        ctx.close()
    (h1, d1), (h0, d0) = runs
    print(f"[gpu]   histories {h1} vs {h0}; state {rel(d1, d0):.2e}", flush=True)
    assert np.allclose(h1, h0, rtol=1e-11, atol=0.0)
    assert rel(d1, d0) < 1e-10


def test_run_converges_to_the_oracle_forces():
    """The drop-in run() on the small Euler case converges (1e-10 drop) to
    the oracle's forces within 1e-8 (tests/golden/euler_o48_converged.json)."""
    import json
    from pathlib import Path

    from jaxrans_amd import load_case, run
    root = Path(__file__).resolve().parent / "golden"
    gold = json.loads((root / "euler_o48_converged.json").read_text())
    result = run(load_case(root / "cases" / "euler_o48"))
    print(f"[gpu] run: {result.history.summary()}  {result.forces.summary()}", flush=True)
    assert result.history.reason == "converged"
    assert abs(result.forces.cl - gold["cl"]) < 1e-8
    assert abs(result.forces.cd - gold["cd"]) < 1e-8
    assert abs(result.forces.cm - gold["cm"]) < 1e-8
    assert abs(result.history.n_iters - gold["iterations"]) <= 20


@pytest.mark.parametrize("name, kind, flowf, flux, lim", [CASES[0], CASES[4], CASES[5]],
                         ids=[CASES[0][0], CASES[4][0], CASES[5][0]])
def test_forces_match_oracle(name, kind, flowf, flux, lim):
    atlas, flow, num, st, ctx = _setup(kind, flowf, flux, lim)
    U = perturbed(OR.U0_of(st), 1e-2)
    ctx.set_state(U)
    got = ctx.forces()[0]
    ref = OS.forces(U, st)
    for k, key in enumerate(("cl", "cd", "cm", "cd_friction")):
        assert got[k] == pytest.approx(ref[key], rel=1e-11, abs=1e-14), key


@pytest.mark.parametrize("name, kind, flowf, flux, lim", [CASES[1], CASES[5]], ids=[CASES[1][0], CASES[5][0]])
def test_graph_replays_match_eager_launches(name, kind, flowf, flux, lim):
    """Blocks of Newton steps replayed from the captured HIP graph equal the
    same blocks launched kernel by kernel, bit for bit, replay after replay,
    with host copies and synchronisations between the replays."""
    print(f"[gpu] graph {name}", flush=True)
    cfg = ImplicitConfig(krylov_subspace=12, krylov_restarts=2)
    atlas, flow, num, st, graph = _setup(kind, flowf, flux, lim, implicit=cfg)
    eager = context_for(atlas, flow, num, implicit=cfg)
    eager.set_freestream(st.q_inf, None if flow.inviscid else flow.mu_inf)
    graph.set_graphs(True)
    eager.set_graphs(False)
    U = perturbed(OR.U0_of(st), 1e-2)
    for ctx in (graph, eager):
        ctx.set_state(U)
        ctx.newton_init(cfg.cfl)
    for block in range(3):
        hg = graph.newton_block(3)
        he = eager.newton_block(3)
        assert np.array_equal(hg, he), (block, hg, he)
        assert np.array_equal(graph.get_state(), eager.get_state()), block
        assert np.array_equal(graph.scalars(), eager.scalars()), block
    eager.close()
    graph.close()


@pytest.mark.parametrize("name, kind, flowf, flux, lim", [CASES[1], CASES[5]], ids=[CASES[1][0], CASES[5][0]])
def test_batched_cases_equal_each_case_alone(name, kind, flowf, flux, lim):
    """Three freestreams (alpha, Mach, Reynolds differ) stepped together in one
    context give, case by case, bit for bit the states, residual histories and
    forces of the same cases stepped alone (the batched polar of
    src/jaxrans/polar.py:93-112, the per-slot freestream of sweep.py:196-201)."""
    from dataclasses import replace

    from jaxrans_amd.core.state import q_infinity
    print(f"[gpu] batch {name}", flush=True)
    cfg = ImplicitConfig(krylov_subspace=12, krylov_restarts=2)
    atlas, flow, num, st, _ = _setup(kind, flowf, flux, lim, implicit=cfg)
    _.close()
    flows = [replace(flow, alpha_deg=a, mach=m, reynolds=None if flow.inviscid else r)
             for a, m, r in [(-2.0, 0.45, 4e6), (1.25, 0.5, 6.5e6), (4.0, 0.6, 9e6)]]
    qs = np.stack([q_infinity(f) for f in flows])
    mus = None if flow.inviscid else np.array([f.mu_inf for f in flows])
    U = np.stack([perturbed(OR.U0_of(OR.setup_for(atlas, f, num)), 1e-2) for f in flows])
    batch = context_for(atlas, flow, num, implicit=cfg, n_cases=3)
    batch.set_freestream(qs, mus)
    batch.set_state(U)
    batch.newton_init(np.full(3, cfg.cfl))
    hb = batch.newton_block(4)
    Ub, Fb = batch.get_state(), batch.forces()
    batch.close()
    for k, f in enumerate(flows):
        alone = context_for(atlas, f, num, implicit=cfg)
        alone.set_freestream(qs[k], None if mus is None else mus[k])
        alone.set_state(U[k])
        alone.newton_init(cfg.cfl)
        ha = alone.newton_block(4)
        assert np.array_equal(ha[:, 0], hb[:, k]), (k, ha[:, 0], hb[:, k])
        assert np.array_equal(alone.get_state()[0], Ub[k]), k
        assert np.array_equal(alone.forces()[0], Fb[k]), k
        alone.close()
