"""Host logic of the served stream: the request stream and the per-solve
block test (no device calls)."""

from __future__ import annotations

import numpy as np
import pytest

from jaxrans_amd import serve
from jaxrans_amd.simulation import ExitReason, run_to_steady


def test_request_stream_is_seed_fixed_and_in_range():
    a = serve.make_requests(256, 2.0)
    b = serve.make_requests(256, 2.0)
    assert a == b
    assert len(a) == 256 and [r.rid for r in a] == list(range(256))
    t = np.array([r.arrival for r in a])
    assert np.all(np.diff(t) > 0)
    # Poisson arrivals: exponential gaps of mean 1/rate (256 draws: within 20 %)
    assert 0.4 < float(np.mean(np.diff(t))) < 0.6
    for r in a:
        assert -0.12 <= r.centre[0] <= -0.04 and 0.0 <= r.centre[1] <= 0.06
        assert 5.0 <= r.te_angle_deg <= 15.0
        assert 0.3 <= r.mach <= 0.8 and -2.0 <= r.alpha_deg <= 6.0
        assert 1e6 <= r.reynolds <= 2e7
        assert r.cells in serve.GRID_SIZES
    assert {r.cells for r in a} == set(serve.GRID_SIZES)
    assert all(r.reynolds is None for r in serve.make_requests(8, 1.0, viscous=False))
    assert all(r.arrival == 0.0 for r in serve.make_requests(8, 0.0))


class _Blocks:
    """Stands in for a device context: hands out prepared residual blocks and
    the forces after each."""

    n_cases = 1

    def __init__(self, residual, forces, block):
        self.residual, self.forces_rows, self.block, self.k = residual, forces, block, 0

    def newton_block(self, n):
        assert n == self.block
        out = self.residual[self.k * n:(self.k + 1) * n, None]
        self.k += 1
        return out

    def forces(self, *a):
        return self.forces_rows[self.k - 1][None]


def _histories(n_blocks, block):
    n = n_blocks * block
    rng = np.random.default_rng(5)
    decay = 1e-3 * 10.0 ** (-np.arange(n) / 40.0)
    forces_moving = np.cumsum(rng.normal(0, 1e-3, (n_blocks, 8)), axis=0)
    forces_still = np.tile(np.array([0.5, 0.01, -0.05, 0, 0, 0, 0, 0]), (n_blocks, 1))
    nonfinite = decay.copy()
    nonfinite[block * 3 + 2] = np.nan
    stalled = np.full(n, 3e-4)
    stalled[:block] = 1e-3 * (1 + np.arange(block))
    rising = 1e-3 * (1.0 + 0.01 * np.sin(np.arange(n)))
    return {
        "converged": (decay, forces_moving),
        "settled": (rising, forces_still),
        "non_finite": (nonfinite, forces_moving),
        "stalled": (stalled, forces_moving),
        "max_iters": (rising, forces_moving),
    }


@pytest.mark.parametrize("name", ["converged", "settled", "non_finite", "stalled", "max_iters"])
def test_block_test_matches_run_to_steady(name):
    """The served solve's block test ends where run_to_steady (with the
    settle test of simulation.run) ends, for the same reason."""
    block, n_blocks, tol, ftol = 10, 40, 1e-8, 1e-5
    residual, forces = _histories(n_blocks, block)[name]
    ctx = _Blocks(residual, forces, block)
    coefficients = []

    def on_block(index, res):
        coefficients.append(tuple(ctx.forces()[0, :3]))
        from jaxrans_amd.simulation import cl_spread
        return all(cl_spread(h) < ftol for h in np.transpose(coefficients))

    ref = run_to_steady(ctx, max_iters=n_blocks * block, tol=tol, block_iters=block, on_block=on_block)
    cfg = serve.ServeConfig(numerics=None, implicit=None, block_iters=block, max_iters=n_blocks * block, tol=tol,
                            forces_tol=ftol)
    s = serve.Solve(req=None, ctx=None, started=0.0)
    for k in range(n_blocks):
        if s.block_done(residual[k * block:(k + 1) * block, None], forces[k][None], cfg):
            break
    assert s.reason == ref.reason == name
    assert s.iters == ref.n_iters


def test_summary_percentiles():
    outs = [serve.Outcome(k, 0.0, 0.0, float(k + 1), 100, ExitReason.SETTLED, 0.0, 0.0, 0.0, 1000, 3.0)
            for k in range(100)]
    s = serve.summary(outs, 100.0)
    assert s["requests"] == 100 and s["solves_per_min"] == 60.0
    assert s["latency_p50_s"] == 50.5 and s["latency_p99_s"] == pytest.approx(99.01)
    assert s["reasons"] == {ExitReason.SETTLED: 100}
    assert s["newton_steps"] == 10000


def test_summary_counts_only_converged_or_settled_solves():
    reasons = [ExitReason.CONVERGED, ExitReason.SETTLED, ExitReason.MAX_ITERS, ExitReason.STALLED]
    outs = [serve.Outcome(k, 0.0, 0.0, float(k + 1), 100, reasons[k % 4], 0.0, 0.0, 0.0, 1000, 3.0)
            for k in range(100)]
    s = serve.summary(outs, 60.0)
    assert s["requests"] == 100 and s["counted"] == 50
    assert s["solves_per_min"] == 50.0 and s["finished_per_min"] == 100.0
    lat = [k + 1.0 for k in range(100) if k % 4 < 2]
    assert s["latency_p50_s"] == pytest.approx(float(np.percentile(lat, 50)))


def test_slots_per_grid_size():
    cfg = serve.default_config("mixed")
    assert [serve.slots_for(a * n, cfg) for a, n in serve.GRID_SIZES] == [16, 16, 7, 4]
    assert serve.slots_for(10, cfg) == cfg.slots and serve.slots_for(10 ** 9, cfg) == 1
