"""The explicit multistage Runge-Kutta path on the device (jr_rk_block)
against the oracle's restatement (``oracle/explicit.py``;
``src/jaxrans/solvers/explicit.py:48-71`` with the local time step of
``numerics/residual.py:142-166``), and ``run()`` with
``time_integration.solver = "explicit"``.  Tolerance: 1e-12 normwise relative
on the state after the steps, and on the residual history (the same
arithmetic in another summation order)."""

from __future__ import annotations

import numpy as np
import pytest

from test_gpu_parity import CASES, _setup, rel
from conftest import perturbed
from oracle import explicit as OX
from oracle import residual as OR

pytestmark = pytest.mark.gpu

SCHEMES = {"euler-hllc-minmod": "rk5_jameson", "laminar-hllc-minmod": "rk4_jameson", "sa-roe-vas": "rk5_jameson",
           "euler-roe-none": "euler"}


@pytest.mark.parametrize("name, kind, flowf, flux, lim", [c for c in CASES if c[0] in SCHEMES],
                         ids=[c[0] for c in CASES if c[0] in SCHEMES])
def test_rk_steps_match_oracle(name, kind, flowf, flux, lim):
    print(f"[gpu] explicit {name} {SCHEMES[name]}", flush=True)
    atlas, flow, num, st, ctx = _setup(kind, flowf, flux, lim)
    alphas = OX.RK_SCHEMES[SCHEMES[name]]
    cfl = 1.2
    # perturbed on the physical cells only: the gap and ghost rows (R = U - U0
    # there) keep the freestream, as a march from it does
    U0 = OR.U0_of(st)
    U = np.where(atlas.interior[None], perturbed(U0, 1e-3), U0)
    ctx.set_state(U)
    ctx.newton_init(cfl)
    hist = ctx.rk_block(3, alphas)[:, 0]
    ref, norms = U, []
    for _ in range(3):
        ref, n = OX.rk_step(ref, st, alphas, cfl)
        norms.append(n)
    got = ctx.get_state()[0]
    assert rel(got, ref) < 1e-12, (name, rel(got, ref))
    assert np.allclose(hist, norms, rtol=1e-12, atol=0.0), (hist, norms)


def test_explicit_run_marches_with_the_oracle():
    """run() on the small Euler case with the explicit solver: its residual
    history and state after 40 steps equal the oracle's RK5 steps."""
    from pathlib import Path

    from jaxrans_amd import load_case, run
    from oracle import residual as OR2
    case = load_case(Path(__file__).resolve().parent / "golden" / "cases" / "euler_o48",
                     ("numerical.time_integration.solver=explicit", "numerical.convergence.max_iters=40",
                      "numerical.convergence.block_iters=20"))
    result = run(case)
    assert result.history.n_iters == 40
    st = OR2.setup_for(result.grid, case.flow, case.numerics)
    U = OR2.U0_of(st)
    norms = []
    for _ in range(40):
        U, n = OX.rk_step(U, st, OX.RK_SCHEMES[case.rk_scheme], case.numerics.cfl)
        norms.append(n)
    print(f"[gpu] explicit run: {result.history.summary()}", flush=True)
    assert np.allclose(result.history.residual, norms, rtol=1e-10, atol=0.0)
    assert rel(result.U, U) < 1e-11


def test_explicit_march_from_freestream_where_host_and_device_round_differently():
    """M 0.33, alpha 0.75 deg: a freestream whose conservative energy the
    device's fused multiply-adds round one ulp away from NumPy's.  The gap and
    ghost rows (R = U - U0) must not amplify that seed: 20 RK5 steps from the
    host freestream stay finite and equal the oracle's steps."""
    from jaxrans_amd.core.config import FlowConditions, NumericsConfig
    from jaxrans_amd.device import context_for
    from conftest import o_atlas
    print("[gpu] explicit march M0.33 a0.75", flush=True)
    atlas = o_atlas()
    flow = FlowConditions(mach=0.33, alpha_deg=0.75)
    num = NumericsConfig(flux="hllc", limiter="minmod", solver="explicit")
    st = OR.setup_for(atlas, flow, num)
    ctx = context_for(atlas, flow, num)
    ctx.set_freestream(st.q_inf, None)
    alphas = OX.RK_SCHEMES["rk5_jameson"]
    U = OR.U0_of(st)
    ctx.set_state(U)
    ctx.newton_init(1.5)
    hist = ctx.rk_block(20, alphas)[:, 0]
    ref, norms = U, []
    for _ in range(20):
        ref, n = OX.rk_step(ref, st, alphas, 1.5)
        norms.append(n)
    got = ctx.get_state()[0]
    ctx.close()
    assert np.all(np.isfinite(got)) and np.all(np.isfinite(hist))
    assert rel(got, ref) < 1e-11, rel(got, ref)
    assert np.allclose(hist, norms, rtol=1e-9, atol=1e-15), (hist, norms)
