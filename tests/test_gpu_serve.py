"""The served stream on the device: asynchronous blocks equal blocking ones,
and each request's result is the same alone and in the mix."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import c_atlas, perturbed
from jaxrans_amd import serve
from jaxrans_amd.core.config import FlowConditions, ImplicitConfig, NumericsConfig
from jaxrans_amd.device import context_for, mem_info
from oracle import residual as OR

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_async_blocks_equal_blocking_blocks(precision):
    """Blocks enqueued asynchronously (every GMRES cycle launched, cycles no
    case needs exiting on the device) give bit for bit the histories, states,
    scalars and forces of jr_newton_block (which skips such cycles on the
    host)."""
    print(f"[gpu] async blocks {precision}", flush=True)
    atlas = c_atlas()
    flow = FlowConditions(mach=0.73, alpha_deg=2.5, reynolds=6.5e6, turbulence="sa_noft2")
    num = NumericsConfig(flux="roe", limiter="van_albada_smooth", solver="implicit")
    cfg = ImplicitConfig(krylov_subspace=12, krylov_restarts=2, precision=precision)
    st = OR.setup_for(atlas, flow, num)
    U = perturbed(OR.U0_of(st), 1e-2)
    ctxs = []
    for _ in range(2):
        c = context_for(atlas, flow, num, implicit=cfg)
        c.set_freestream(st.q_inf, flow.mu_inf)
        c.set_state(U)
        c.newton_init(cfg.cfl)
        ctxs.append(c)
    blocking, asynchronous = ctxs
    ref = (1.0, (0.25, 0.0))
    for block in range(3):
        hb = blocking.newton_block(4)
        fb = blocking.forces(ref[0], ref[1])
        asynchronous.newton_block_async(4, ref)
        while not asynchronous.poll():
            pass
        ha, fa = asynchronous.wait()
        assert np.array_equal(ha, hb), (block, ha, hb)
        assert np.array_equal(fa, fb), block
        assert np.array_equal(asynchronous.get_state(), blocking.get_state()), block
        assert np.array_equal(asynchronous.scalars(), blocking.scalars()), block
    # a block without forces returns the history alone
    asynchronous.newton_block_async(2)
    assert asynchronous.wait().shape == (2, 1)
    for c in ctxs:
        c.close()


def _small_stream():
    sizes = ((96, 24), (128, 32), (160, 32))
    reqs = serve.make_requests(5, 0.0, seed=11, sizes=sizes)
    cfg = serve.default_config("mixed", wall_spacing=1e-3, far_field=15.0, block_iters=10, max_iters=60,
                               forces_tol=None, max_in_flight=5, reserve_bytes=0)
    return reqs, cfg, {r.rid: serve.request_grid(r, cfg) for r in reqs}


@pytest.mark.parametrize("driver", ["threads", "async", "slots"])
def test_served_results_are_the_same_alone_and_in_the_mix(driver):
    """Five requests of two grid sizes, all arriving at once, served five
    in flight (their blocks overlap on the device, each on its own stream)
    and then one at a time through the plain path: each request ends with the same iterations,
    reason and forces, bit for bit."""
    from dataclasses import replace
    print(f"[gpu] served mix vs alone, {driver}", flush=True)
    reqs, cfg, atlases = _small_stream()
    cfg = replace(cfg, driver=driver)
    free0, total = mem_info()
    assert 0 < free0 <= total
    mixed = {o.rid: o for o in serve.serve(reqs, atlases, cfg)}
    alone = {o.rid: o for o in serve.serve_serial(reqs, atlases, cfg)}
    assert sorted(mixed) == sorted(alone) == [r.rid for r in reqs]
    for rid in mixed:
        a, m = alone[rid], mixed[rid]
        assert (a.iterations, a.reason, a.cells) == (m.iterations, m.reason, m.cells), rid
        assert (a.cl, a.cd, a.cm) == (m.cl, m.cd, m.cm), (rid, a, m)
        assert np.isfinite(a.cl) and a.iterations == cfg.max_iters


def test_batched_geometries_equal_each_grid_alone():
    """Three different sections on C-grids of one topology in one context
    (per-case geometry, jr_set_geometry) give, case by case, bit for bit the
    residual histories, states and forces of each grid in its own context."""
    print("[gpu] batched geometries", flush=True)
    reqs = serve.make_requests(3, 0.0, seed=21, sizes=((96, 24),))
    cfg = serve.default_config("fp64", wall_spacing=1e-3, far_field=15.0)
    atlases = [serve.request_grid(r, cfg) for r in reqs]
    flows = [serve.request_flow(r, cfg) for r in reqs]
    from jaxrans_amd.core.state import q_infinity, uniform_canvas
    imp = ImplicitConfig(krylov_subspace=12, krylov_restarts=2, max_relative_update=0.2)
    batch = context_for(atlases[0], flows[0], cfg.numerics, imp, n_cases=3)
    qs = np.zeros((3, 5))
    for k, (a, f) in enumerate(zip(atlases, flows)):
        q = q_infinity(f)
        qs[k, :q.size] = q
        batch.set_geometry(k, a)
    batch.set_freestream(qs, np.array([f.mu_inf for f in flows]))
    U0 = [uniform_canvas(q_infinity(f), a.shape, f.gamma) for a, f in zip(atlases, flows)]
    for k in range(3):
        batch.set_state_case(k, U0[k])
        batch.newton_init_case(k, imp.cfl)
    hb = batch.newton_block(6)
    Fb = batch.forces()
    Ub = batch.get_state()
    other = serve.request_grid(serve.make_requests(1, 0.0, seed=5, sizes=((128, 32),))[0], cfg)
    from jaxrans_amd.device import DeviceError
    with pytest.raises(DeviceError, match="topology"):
        batch.set_geometry(0, other)
    batch.close()
    for k, (a, f) in enumerate(zip(atlases, flows)):
        alone = context_for(a, f, cfg.numerics, imp)
        alone.set_freestream(qs[k], f.mu_inf)
        alone.set_state(U0[k])
        alone.newton_init(imp.cfl)
        ha = alone.newton_block(6)
        assert np.array_equal(ha[:, 0], hb[:, k]), (k, ha[:, 0], hb[:, k])
        assert np.array_equal(alone.get_state()[0], Ub[k]), k
        assert np.array_equal(alone.forces()[0], Fb[k]), k
        alone.close()


def test_parked_slot_keeps_its_state_and_the_others_are_unchanged():
    """A parked case of a batch (jr_park_case) keeps its state bit for bit
    through a block of Newton steps (every pass skips it; its history repeats
    its last norm), and the other cases step exactly as in the same batch
    without parking; released and refilled, it steps again."""
    print("[gpu] parked slot", flush=True)
    from jaxrans_amd.core.state import q_infinity, uniform_canvas
    reqs = serve.make_requests(3, 0.0, seed=21, sizes=((96, 24),))
    cfg = serve.default_config("fp64", wall_spacing=1e-3, far_field=15.0)
    atlases = [serve.request_grid(r, cfg) for r in reqs]
    flows = [serve.request_flow(r, cfg) for r in reqs]
    imp = ImplicitConfig(krylov_subspace=12, krylov_restarts=2, max_relative_update=0.2)
    qs = np.zeros((3, 5))
    for k, f in enumerate(flows):
        q = q_infinity(f)
        qs[k, :q.size] = q
    U0 = [uniform_canvas(q_infinity(f), a.shape, f.gamma) for a, f in zip(atlases, flows)]

    def batch():
        b = context_for(atlases[0], flows[0], cfg.numerics, imp, n_cases=3)
        for k in range(3):
            b.set_geometry(k, atlases[k])
        b.set_freestream(qs, np.array([f.mu_inf for f in flows]))
        for k in range(3):
            b.set_state_case(k, U0[k])
            b.newton_init_case(k, imp.cfl)
        return b

    plain, parked = batch(), batch()
    h_plain = plain.newton_block(8)
    h_first = parked.newton_block(4)
    before = parked.get_state()[1].copy()
    parked.park_case(1, True)
    h_parked = parked.newton_block(4)
    after = parked.get_state()
    assert np.array_equal(after[1], before)
    assert np.all(h_parked[:, 1] == h_first[-1, 1])   # a parked slot's history repeats its last norm
    for k in (0, 2):
        assert np.array_equal(h_parked[:, k], h_plain[4:, k]), k
        assert np.array_equal(after[k], plain.get_state()[k]), k
    parked.set_state_case(1, U0[1])
    parked.newton_init_case(1, imp.cfl)
    parked.park_case(1, False)
    h = parked.newton_block(2)
    assert np.all(np.isfinite(h)) and not np.array_equal(parked.get_state()[1], U0[1])
    plain.close()
    parked.close()
