"""The swept (2.5D) path on the device: the section extruded two periodic
cells deep on the 3D kernels.

* residual and tangent on the extruded canvas against the 3D oracle on the
  same canvas, at 30 deg sweep (1e-12 / 1e-11 normwise);
* ``run()`` at zero sweep converges to the 2D run's forces (the reference's
  ``tests/verification/test_swept.py`` checks the same at the residual);
* the independence principle of an infinite swept wing in inviscid flow: the
  in-plane solution is the 2D one at the in-plane Mach number and incidence,
  the spanwise velocity is the freestream's everywhere, and the energy
  carries its kinetic energy (converged states within 1e-8).
"""

from __future__ import annotations

from dataclasses import replace
from pathlib import Path

import numpy as np
import pytest

from conftest import c_atlas, o_atlas
from jaxrans_amd.core.config import FlowConditions, NumericsConfig
from jaxrans_amd.device import context_for
from jaxrans_amd.mesh.swept import lift, section_to_planes
from jaxrans_amd.simulation import swept_q
from oracle import residual3d as O3

pytestmark = pytest.mark.gpu

CASE = Path(__file__).resolve().parent / "golden" / "cases" / "euler_o48"


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("physics", ["euler", "sa"])
def test_swept_residual_and_tangent_match_the_3d_oracle(physics):
    print(f"[gpu] swept residual {physics}", flush=True)
    if physics == "euler":
        at, flow = o_atlas(), FlowConditions(mach=0.6, alpha_deg=2.5, sweep_deg=30.0)
        num = NumericsConfig(flux="hllc", limiter="van_albada")
    else:
        at = c_atlas()
        flow = FlowConditions(mach=0.5, alpha_deg=2.0, reynolds=1e6, turbulence="sa_noft2", sweep_deg=30.0)
        num = NumericsConfig(flux="roe", limiter="van_albada_smooth")
    a3 = lift(at.mesh)
    q = swept_q(flow)
    st = replace(O3.setup_for(a3, replace(flow, sweep_deg=None), num), q_inf=q)
    U0 = O3.U0_of(st)
    rng = np.random.default_rng(5)
    U25 = np.where(at.interior[None], 1.0 + 0.02 * rng.standard_normal((U0.shape[0],) + at.shape), 1.0)
    from jaxrans_amd.simulation import section_state
    U25 = U25 * section_state(q, at, flow.gamma)
    U = section_to_planes(U25, at, a3, U0)
    ctx = context_for(a3, flow, num)
    ctx.set_freestream(q, None if flow.inviscid else flow.mu_inf)
    for first_order in (False, True):
        ref = O3.residual(U, st, first_order)
        got = ctx.residual(U, first_order)[0]
        assert rel(got, ref) < 1e-12, (first_order, rel(got, ref))
    dU = rng.standard_normal(U.shape) * np.abs(U).mean(axis=(1, 2, 3), keepdims=True) * 1e-3
    for first_order in (True, False):
        ref = O3.jvp(U, dU, st, first_order)
        got = ctx.tangent(U, dU, first_order=first_order)[0]
        assert rel(got, ref) < 1e-11, (first_order, rel(got, ref))
    ctx.close()


def test_unswept_run_converges_to_the_2d_forces():
    from jaxrans_amd import load_case, run
    plain = run(load_case(CASE))
    swept = run(load_case(CASE, ("case.freestream.sweep_deg=0.0",)))
    print(f"[gpu] swept 0 deg: {swept.history.summary()} vs 2D {plain.history.summary()}", flush=True)
    assert swept.history.reason == "converged" and plain.history.reason == "converged"
    assert swept.U.shape == (5,) + plain.grid.shape
    for out in ("cl", "cd", "cm"):
        assert abs(getattr(swept.forces, out) - getattr(plain.forces, out)) < 1e-8, out
    assert abs(swept.forces.cy) < 1e-12


def test_independence_principle_of_the_inviscid_swept_wing():
    from jaxrans_amd import load_case, run
    mach, alpha, sweep = 0.6, 2.0, 30.0
    a, s = np.deg2rad(alpha), np.deg2rad(sweep)
    m_n = float(mach * np.hypot(np.cos(a) * np.cos(s), np.sin(a)))
    a_n = float(np.rad2deg(np.arctan2(np.sin(a), np.cos(a) * np.cos(s))))
    over = ("numerical.convergence.tol=1e-11", "numerical.convergence.max_iters=800")
    swept = run(load_case(CASE, over + (f"case.freestream.mach={mach}", f"case.freestream.alpha_deg={alpha}",
                                         f"case.freestream.sweep_deg={sweep}", 'case.domain.section="normal"')))
    plain = run(load_case(CASE, over + (f"case.freestream.mach={m_n!r}", f"case.freestream.alpha_deg={a_n!r}")))
    print(f"[gpu] independence: swept {swept.history.summary()}; 2D {plain.history.summary()}", flush=True)
    assert swept.history.reason == "converged" and plain.history.reason == "converged"
    inside = plain.grid.interior
    U2, U25 = plain.U[:, inside], swept.U[:, inside]
    w = mach * np.cos(a) * np.sin(s)
    for r in (0, 1, 2):
        assert rel(U25[r], U2[r]) < 1e-8, r
    assert rel(U25[3], U25[0] * w) < 1e-8
    assert rel(U25[4] - 0.5 * U25[0] * w * w, U2[3]) < 1e-8


def test_unswept_sensitivity_equals_the_2d_one():
    """sensitivity() through the swept path at zero sweep -- the section on
    its extruded 3D canvas, the parameter tangents of the 3D kernels, the
    coefficients' tangent from the wall loads' by the complex step -- equals
    the 2D path's at the same converged state (the 2D solution with a zero
    spanwise momentum row)."""
    from dataclasses import replace as _replace

    from jaxrans_amd import load_case, run, sensitivity
    print("[gpu] unswept sensitivity vs 2D", flush=True)
    base = ("numerical.convective_fluxes.limiter=none", "numerical.convergence.tol=1e-11",
            "numerical.convergence.max_iters=3000")
    flat = run(load_case(CASE, base))
    swept = run(load_case(CASE, base + ("case.freestream.sweep_deg=0.0",)))
    assert flat.history.reason == swept.history.reason == "converged"
    U25 = np.concatenate([flat.U[:3], np.zeros_like(flat.U[:1]), flat.U[3:]])
    same = _replace(swept, U=U25)
    # tight solves: the test |b - A x| <= tol |b| is dominated by the rows off
    # the interior (the U - U0 rows' dU0/dM), so tol 1e-11 leaves dCL/dM good
    # to about 1e-7 only, on either path
    s2 = sensitivity(flat, ("cl", "cd", "cm"), ("alpha", "mach"), tol=1e-14, maxiter=5000)
    s3 = sensitivity(same, ("cl", "cd", "cm"), ("alpha", "mach"), tol=1e-14, maxiter=5000)
    print(f"[gpu]   linear residuals 2D {s2.linear_residuals}, swept path {s3.linear_residuals}", flush=True)
    own = sensitivity(swept, ("cl",), ("alpha", "mach"), tol=1e-14, maxiter=5000)
    a = np.array([s2.derivatives[o][n] for o in ("cl", "cd", "cm") for n in ("alpha", "mach")])
    b = np.array([s3.derivatives[o][n] for o in ("cl", "cd", "cm") for n in ("alpha", "mach")])
    for k, (o, n) in enumerate((o, n) for o in ("cl", "cd", "cm") for n in ("alpha", "mach")):
        print(f"[gpu]   d{o}/d{n}: 2D {a[k]:+.10e}  swept path {b[k]:+.10e}", flush=True)
    assert np.linalg.norm(b - a) <= 1e-8 * np.linalg.norm(a)
    assert np.abs(b - a).max() <= 1e-10
    for name in ("alpha", "mach"):   # at the swept path's own converged state: the states' difference
        print(f"[gpu]   dcl/d{name} at the swept run's own state {own.derivatives['cl'][name]:+.10e}", flush=True)
        assert own.derivatives["cl"][name] == pytest.approx(s2.derivatives["cl"][name], rel=1e-6)
