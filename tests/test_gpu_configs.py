"""The device path at the benchmark configurations' real sizes.

* 2D Euler, NACA 0012 law, O-grid 128 x 64 (cases/naca0012_o_euler):
  ``run()`` converges (1e-10 drop) to the oracle's converged forces within
  1e-8 (tests/golden/naca0012_o_euler_converged.json, written by
  tests/golden/make_converged.py).
* 2D SA RANS, Karman-Trefftz 10 deg, C-grid 512 x 128 (cases/kt10_c_sa):
  residual (1e-12), first-order tangent (1e-11), the FP64 line
  preconditioner by block cyclic reduction (1e-10) and two Newton steps (norms 1e-9, state 1e-8)
  against the oracle on the same seeded state.
* The same with the block Thomas line solve (lanes over block entries).
* The batched polar on that grid, 32 cases (alpha U[-4, 12], M U[0.3,
  0.78], Re 6e6): every case of the batch equals the case alone bit for bit
  after three Newton steps (a sample of four cases); mixed precision run
  beside it, its per-case force differences printed.
* The served stream's grid sizes (C-grids of 256 x 64 ... 1024 x 256 cells):
  one request of each, served together and one at a time, the same bit for
  bit.
* 3D SA RANS, the four-block swept tapered wing, 1,091,328 cells
  (cases/wing_sa): residual and first-order tangent against the 3D oracle.
"""

from __future__ import annotations

import json
import shutil
import tempfile
from dataclasses import replace
from pathlib import Path

import numpy as np
import pytest

from conftest import perturbed
from jaxrans_amd.core.config import ImplicitConfig
from jaxrans_amd.core.state import q_infinity, uniform_canvas
from jaxrans_amd.device import context_for
from oracle import residual as OR
from oracle import solver as OS

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_euler_o_grid_128x64_run_converges_to_the_oracle_forces():
    from jaxrans_amd import load_case, run
    golden = json.loads((ROOT / "tests/golden/naca0012_o_euler_converged.json").read_text())
    result = run(load_case(ROOT / "cases/naca0012_o_euler"))
    print(f"[gpu] configs O-grid 128x64: {result.history.summary()}; oracle {golden['iterations']} iterations",
          flush=True)
    assert result.history.reason == "converged" and result.cells == golden["cells"]
    for out in ("cl", "cd", "cm"):
        assert abs(getattr(result.forces, out) - golden[out]) < 1e-8, (out, getattr(result.forces, out), golden[out])


def _kt10():
    from jaxrans_amd import load_case
    case = load_case(ROOT / "cases/kt10_c_sa")
    return case, case.grid()


def test_sa_c_grid_512x128_kernels_and_newton_steps_match_the_oracle():
    print("[gpu] configs C-grid 512x128 kernels", flush=True)
    case, atlas = _kt10()
    cfg = replace(case.implicit, krylov_subspace=10, krylov_restarts=2, line_solver="cyclic_reduction")
    st = OR.setup_for(atlas, case.flow, case.numerics)
    ctx = context_for(atlas, case.flow, case.numerics, cfg)
    ctx.set_freestream(st.q_inf, case.flow.mu_inf)
    U = perturbed(OR.U0_of(st), 1e-3)
    assert rel(ctx.residual(U)[0], OR.residual(U, st)) < 1e-12
    rng = np.random.default_rng(11)
    dU = rng.standard_normal(U.shape) * np.abs(U).mean(axis=(1, 2), keepdims=True) * 1e-3
    assert rel(ctx.tangent(U, dU, first_order=True)[0], OR.jvp(U, dU, st, True)) < 1e-11
    ctx.set_cfl(25.0)
    ctx.precond_factor(U)
    r = rng.standard_normal(U.shape)
    ref = OS.LinePreconditioner(U, OR.wave_speeds(U, st), 25.0, st, "cyclic_reduction")(r)
    assert rel(ctx.precond_apply(r)[0], ref) < 1e-10
    ctx.set_state(U)
    ctx.newton_init(cfg.cfl)
    s, norms = OS.NewtonState(U, cfg.cfl), []
    for _ in range(2):
        s, n, _info = OS.newton_step(s, st, cfg)
        norms.append(n)
    hist = ctx.newton_block(2)[:, 0]
    assert np.allclose(hist, norms, rtol=1e-9, atol=0.0), (hist, norms)
    assert rel(ctx.get_state()[0] - U, s.U - U) < 1e-8
    ctx.close()


def test_sa_c_grid_512x128_block_thomas_lines_and_newton_steps_match_the_oracle():
    """The lane-parallel block Thomas apply on the real grid (chains of 128
    cells and of 256 across the wake: two launch classes)."""
    print("[gpu] configs C-grid 512x128 block Thomas", flush=True)
    case, atlas = _kt10()
    cfg = replace(case.implicit, krylov_subspace=10, krylov_restarts=2, line_solver="thomas")
    st = OR.setup_for(atlas, case.flow, case.numerics)
    ctx = context_for(atlas, case.flow, case.numerics, cfg)
    ctx.set_freestream(st.q_inf, case.flow.mu_inf)
    U = perturbed(OR.U0_of(st), 1e-3)
    ctx.set_cfl(25.0)
    ctx.precond_factor(U)
    r = np.random.default_rng(12).standard_normal(U.shape)
    ref = OS.LinePreconditioner(U, OR.wave_speeds(U, st), 25.0, st, "thomas")(r)
    assert rel(ctx.precond_apply(r)[0], ref) < 1e-10
    ctx.set_state(U)
    ctx.newton_init(cfg.cfl)
    s, norms = OS.NewtonState(U, cfg.cfl), []
    for _ in range(2):
        s, n, _info = OS.newton_step(s, st, cfg)
        norms.append(n)
    hist = ctx.newton_block(2)[:, 0]
    assert np.allclose(hist, norms, rtol=1e-9, atol=0.0), (hist, norms)
    assert rel(ctx.get_state()[0] - U, s.U - U) < 1e-8
    ctx.close()


def _polar_flows(flow, n=32):
    rng = np.random.default_rng(2)
    alphas, machs = rng.uniform(-4.0, 12.0, n), rng.uniform(0.3, 0.78, n)
    return [replace(flow, alpha_deg=float(a), mach=float(m), reynolds=6e6) for a, m in zip(alphas, machs)]


def _batch(atlas, flows, case, cfg):
    ctx = context_for(atlas, flows[0], case.numerics, cfg, n_cases=len(flows))
    qs = np.stack([q_infinity(f) for f in flows])
    ctx.set_freestream(qs, np.array([f.mu_inf for f in flows]))
    ctx.set_state(np.stack([uniform_canvas(q, atlas.shape, f.gamma) for q, f in zip(qs, flows)]))
    ctx.newton_init(np.full(len(flows), cfg.cfl))
    return ctx


def test_polar_of_32_cases_equals_each_case_alone():
    print("[gpu] configs polar 32 x 512x128", flush=True)
    case, atlas = _kt10()
    flows = _polar_flows(case.flow)
    results = {}
    for precision in ("fp64", "mixed"):
        cfg = replace(case.implicit, precision=precision)
        batch = _batch(atlas, flows, case, cfg)
        hb = batch.newton_block(3)
        results[precision] = (hb, batch.get_state(), batch.forces())
        batch.close()
    hb, Ub, Fb = results["fp64"]
    assert np.all(np.isfinite(hb))
    cfg = replace(case.implicit, precision="fp64")
    for k in (0, 9, 20, 31):
        alone = _batch(atlas, [flows[k]], case, cfg)
        ha = alone.newton_block(3)
        assert np.array_equal(ha[:, 0], hb[:, k]), (k, ha[:, 0], hb[:, k])
        assert np.array_equal(alone.get_state()[0], Ub[k]), k
        assert np.array_equal(alone.forces()[0], Fb[k]), k
        alone.close()
    Fm = results["mixed"][2]
    d_cl, d_cd = np.abs(Fm[:, 0] - Fb[:, 0]), np.abs(Fm[:, 1] - Fb[:, 1]) * 1e4
    print(f"[gpu] polar mixed - FP64 after 3 steps from the freestream: max |dCL| {d_cl.max():.3e}, "
          f"max |dCD| {d_cd.max():.3e} counts (transient states)", flush=True)
    assert np.all(np.isfinite(Fm))


def test_stream_grid_sizes_alone_and_in_the_mix():
    from jaxrans_amd import serve
    print("[gpu] configs stream sizes, mix vs alone", flush=True)
    sizes = ((256, 64), (512, 128), (768, 192), (1024, 256))
    reqs = [serve.make_requests(1, 0.0, seed=30 + k, sizes=(sz,))[0] for k, sz in enumerate(sizes)]
    reqs = [replace(r, rid=k) for k, r in enumerate(reqs)]
    cfg = serve.default_config("mixed", block_iters=10, max_iters=20, forces_tol=None, max_in_flight=4,
                               reserve_bytes=0, driver="async")
    atlases = {r.rid: serve.request_grid(r, cfg) for r in reqs}
    assert sorted(a.cells for a in atlases.values()) == [256 * 64, 512 * 128, 768 * 192, 1024 * 256]
    mixed = {o.rid: o for o in serve.serve(reqs, atlases, cfg)}
    alone = {o.rid: o for o in serve.serve_serial(reqs, atlases, cfg)}
    for rid in alone:
        a, m = alone[rid], mixed[rid]
        assert (a.iterations, a.reason, a.cells) == (m.iterations, m.reason, m.cells), rid
        assert (a.cl, a.cd, a.cm) == (m.cl, m.cd, m.cm), (rid, a, m)


def test_wing_1p1m_cells_residual_and_tangent_match_the_3d_oracle():
    from jaxrans_amd import load_case
    from jaxrans_amd.mesh.wing import write_wing_grid
    from oracle import residual3d as O3
    print("[gpu] configs wing 1.09M cells", flush=True)
    with tempfile.TemporaryDirectory() as tmp:
        for name in ("case.json", "numerical.json"):
            shutil.copy(ROOT / "cases/wing_sa" / name, tmp)
        case = load_case(tmp)
        write_wing_grid(Path(tmp) / case.grid_spec.file)
        atlas = case.grid()
    assert atlas.cells == 4 * 192 * 49 * 29
    st = O3.setup_for(atlas, case.flow, case.numerics)
    ctx = context_for(atlas, case.flow, case.numerics, case.implicit)
    ctx.set_freestream(st.q_inf, case.flow.mu_inf)
    U0 = O3.U0_of(st)
    U = np.where(atlas.interior[None], perturbed(U0, 1e-3), U0)
    assert rel(ctx.residual(U)[0], O3.residual(U, st)) < 1e-12
    dU = np.where(atlas.interior[None], np.sin(np.arange(U.size)).reshape(U.shape), 0.0) * 1e-3 * np.abs(U0)
    assert rel(ctx.tangent(U, dU, first_order=True)[0], O3.jvp(U, dU, st, True)) < 1e-11
    ctx.close()
