"""The Newton block that allocates nothing (jr_newton_block_noalloc, history
staged in the library's static per-device array) against jr_newton_block,
through the C ABI: two calls in a row, the second on a smaller input (fewer
cases, cells and history entries), each bit for bit the existing entry
point's histories, states and per-case scalars; the context's device memory
is unchanged by the calls; an oversized block is refused before any launch."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import c_atlas, o_atlas, perturbed
from jaxrans_amd.core.config import FlowConditions, ImplicitConfig, NumericsConfig
from jaxrans_amd.core.state import q_infinity, uniform_canvas
from jaxrans_amd.device import NOALLOC_HISTORY_MAX, DeviceError, context_for

pytestmark = pytest.mark.gpu


def _pair(atlas, flows, num, cfg, amplitude):
    """Two identical contexts (reference entry point, no-alloc entry point)
    on ``atlas`` with one case per flow, at a perturbed freestream."""
    qs = np.stack([q_infinity(f) for f in flows])
    mu = None if flows[0].inviscid else np.array([f.mu_inf for f in flows])
    U = np.stack([perturbed(uniform_canvas(q, atlas.shape, f.gamma), amplitude) for q, f in zip(qs, flows)])
    out = []
    for _ in range(2):
        c = context_for(atlas, flows[0], num, implicit=cfg, n_cases=len(flows))
        c.set_freestream(qs, mu)
        c.set_state(U)
        c.newton_init(np.full(len(flows), cfg.cfl))
        out.append(c)
    return out


def _same(ref, na, n_steps):
    before = na.device_bytes
    h_na = na.newton_block_noalloc(n_steps)
    h_ref = ref.newton_block(n_steps)
    assert h_na.shape == (n_steps, ref.n_cases)
    assert np.all(np.isfinite(h_ref))
    assert np.array_equal(h_na, h_ref), (h_na, h_ref)
    assert np.array_equal(na.get_state(), ref.get_state())
    assert np.array_equal(na.scalars(), ref.scalars())
    assert na.device_bytes == before


def test_noalloc_block_equals_newton_block_on_two_calls():
    print("[gpu] no-alloc call 1: SA C-grid, 3 cases x 6 steps", flush=True)
    sa = FlowConditions(mach=0.73, alpha_deg=2.5, reynolds=6.5e6, turbulence="sa_noft2")
    flows = [sa, FlowConditions(mach=0.6, alpha_deg=4.0, reynolds=6.5e6, turbulence="sa_noft2"),
             FlowConditions(mach=0.5, alpha_deg=-1.0, reynolds=6.5e6, turbulence="sa_noft2")]
    num = NumericsConfig(flux="roe", limiter="van_albada_smooth", solver="implicit")
    cfg = ImplicitConfig(krylov_subspace=12, krylov_restarts=2)
    ref, na = _pair(c_atlas(), flows, num, cfg, 1e-2)
    _same(ref, na, 6)
    ref.close()
    na.close()

    print("[gpu] no-alloc call 2: Euler O-grid, 1 case x 2 steps", flush=True)
    flows = [FlowConditions(mach=0.5, alpha_deg=1.25)]
    num = NumericsConfig(flux="hllc", limiter="van_albada", solver="implicit")
    cfg = ImplicitConfig(krylov_subspace=10, krylov_restarts=2)
    ref, na = _pair(o_atlas(32, 16), flows, num, cfg, 1e-3)
    _same(ref, na, 2)
    ref.close()
    na.close()


def test_noalloc_block_refuses_an_oversized_history():
    print("[gpu] no-alloc oversized block", flush=True)
    flows = [FlowConditions(mach=0.5, alpha_deg=1.25)] * 2
    num = NumericsConfig(flux="hllc", limiter="van_albada", solver="implicit")
    cfg = ImplicitConfig(krylov_subspace=10, krylov_restarts=2)
    ref, na = _pair(o_atlas(32, 16), flows, num, cfg, 1e-3)
    state = na.get_state()
    with pytest.raises(DeviceError, match="exceeds"):
        na.newton_block_noalloc(NOALLOC_HISTORY_MAX // 2 + 1)
    assert np.array_equal(na.get_state(), state)
    _same(ref, na, 1)
    ref.close()
    na.close()
