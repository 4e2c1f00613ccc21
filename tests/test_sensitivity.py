"""Host logic of ``sensitivity()``: the defect correction
(``src/jaxrans/solvers/sensitivity.py:34-50``), the closed-form freestream
tangent against fourth-order central differences, and the refusal of unknown names
(``tests/unit/test_sensitivity.py:72-76``).  No GPU."""

from __future__ import annotations

import numpy as np
import pytest

from jaxrans_amd.derivatives import (OUTPUTS, PARAMETERS, defect_correction, flow_inputs, freestream_tangent,
                                     sensitivity)
from jaxrans_amd.core.config import FlowConditions


def _richardson(f, h: float):
    """(4 D(h/2) - D(h)) / 3 with D(s) = (f(s) - f(-s)) / 2s: f'(0) to O(h^4)."""
    d_h = (f(h) - f(-h)) / (2.0 * h)
    d_half = (f(0.5 * h) - f(-0.5 * h)) / h
    return (4.0 * d_half - d_h) / 3.0


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("name", PARAMETERS)
@pytest.mark.parametrize("flow", [FlowConditions(mach=0.5, alpha_deg=1.25),
                                  FlowConditions(mach=0.3, alpha_deg=-4.0, reynolds=5e3),
                                  FlowConditions(mach=0.73, alpha_deg=2.5, reynolds=6.5e6, turbulence="sa_noft2"),
                                  FlowConditions(mach=0.73, alpha_deg=2.5, reynolds=6.5e6, turbulence="sa_noft2",
                                                 sweep_deg=25.0)])
def test_freestream_tangent_is_the_derivative_of_the_flow_inputs(flow, name, dim):
    """The closed-form (dq_inf, dmu_inf) against Richardson differences of
    flow_inputs (derivatives.py:65-70) in both parameters."""
    if flow.swept and dim == 2:
        pytest.skip("a swept section's freestream has three velocity components")
    dq, dmu = freestream_tangent(flow, dim, name)
    p0 = {"alpha": flow.alpha_deg, "mach": flow.mach}

    def q_at(t):
        q, mu = flow_inputs(flow, dim, {k: v + (t if k == name else 0.0) for k, v in p0.items()})
        return np.concatenate([q, [0.0 if mu is None else mu]])

    fd = _richardson(q_at, 1e-2)   # trigonometric in alpha, linear in M: O(h^4) is round-off here
    assert dq.shape == flow_inputs(flow, dim, p0)[0].shape
    assert np.allclose(dq, fd[:-1], rtol=1e-9, atol=1e-13)
    assert (dmu is None) == flow.inviscid
    if dmu is not None:
        assert dmu == pytest.approx(fd[-1], rel=1e-9)


def test_defect_correction_solves_with_an_approximate_inverse():
    rng = np.random.default_rng(3)
    n = 40
    A = np.eye(n) * 4.0 + rng.standard_normal((n, n)) * 0.3
    G = np.linalg.inv(np.diag(np.diag(A)) + np.tril(A, -1))   # Gauss-Seidel: an approximate inverse
    b = rng.standard_normal(n)
    x, rel, k = defect_correction(lambda v: A @ v, b, lambda r: G @ r, 1e-12, 500)
    assert rel <= 1e-12 and 0 < k < 500
    assert np.allclose(x, np.linalg.solve(A, b), rtol=0, atol=1e-10)


def test_defect_correction_stops_at_maxiter_and_reports_the_residual():
    A = np.array([[2.0, 1.0], [1.0, 3.0]])
    b = np.array([1.0, 2.0])
    x, rel, k = defect_correction(lambda v: A @ v, b, lambda r: 0.1 * r, 1e-14, 3)
    assert k == 3
    assert rel == pytest.approx(np.linalg.norm(b - A @ x) / np.linalg.norm(b), rel=1e-14)


def test_defect_correction_of_a_zero_right_hand_side_takes_no_step():
    x, rel, k = defect_correction(lambda v: v, np.zeros(5), lambda r: r, 1e-10, 10)
    assert k == 0 and rel == 0.0 and not x.any()


def test_flow_inputs_at_the_case_values_are_the_freestream():
    from jaxrans_amd.simulation import freestream_of
    flow = FlowConditions(mach=0.6, alpha_deg=2.5, reynolds=5e3)
    q, mu = flow_inputs(flow, 2, {"alpha": 2.5, "mach": 0.6})
    q0, mu0 = freestream_of(flow, 2)
    assert np.array_equal(q, q0) and mu == mu0
    q1, mu1 = flow_inputs(flow, 2, {"alpha": 2.5, "mach": 0.72})
    assert mu1 == pytest.approx(0.72 / 5e3, rel=1e-15)   # the Reynolds number held
    assert np.hypot(q1[1], q1[2]) == pytest.approx(0.72, rel=1e-15)
    assert flow_inputs(FlowConditions(mach=0.5), 2, {"alpha": 0.0, "mach": 0.5})[1] is None


def test_unknown_names_are_refused():
    with pytest.raises(ValueError, match="unknown parameter"):
        sensitivity(None, parameters=("reynolds",))
    with pytest.raises(ValueError, match="unknown output"):
        sensitivity(None, outputs=("lift",))
    with pytest.raises(ValueError, match="unknown parameter"):
        sensitivity(None, parameters=())
    assert PARAMETERS == ("alpha", "mach") and OUTPUTS == ("cl", "cd", "cm", "cd_friction")
