"""The swept (2.5D) path's host side and its equivalence, on the CPU oracles
(the reference's own exact checks, ``tests/verification/test_swept.py``,
restated): the section extruded two periodic cells deep carries the 2D
residual at zero sweep (no spanwise momentum appears), every plane of a
spanwise-uniform state has the same residual bit for bit at any sweep, the
swept freestream, and the swept force coefficients at zero sweep are the 2D
ones."""

from __future__ import annotations

from dataclasses import replace

import numpy as np
import pytest

from conftest import c_atlas, o_atlas, perturbed
from jaxrans_amd.core.config import FlowConditions, NumericsConfig
from jaxrans_amd.mesh.swept import (lift, planes_to_section, section_to_planes, swept_coefficients,
                                    swept_coefficients_tangent, swept_freestream)
from jaxrans_amd.simulation import swept_q
from oracle import residual as OR
from oracle import residual3d as O3
from oracle import solver as OS

PHYSICS = {
    "euler": (FlowConditions(mach=0.6, alpha_deg=2.5), NumericsConfig(flux="hllc", limiter="van_albada")),
    "laminar": (FlowConditions(mach=0.5, alpha_deg=1.0, reynolds=5000.0),
                NumericsConfig(flux="hllc", limiter="van_albada")),
    "sa": (FlowConditions(mach=0.5, alpha_deg=2.0, reynolds=1e6, turbulence="sa_noft2"),
           NumericsConfig(flux="roe", limiter="van_albada_smooth")),
}


def _atlas(physics):
    return o_atlas(wall="no_slip_wall") if physics == "laminar" else c_atlas() if physics == "sa" else o_atlas()


def _setups(physics, sweep):
    flow, num = PHYSICS[physics]
    at = _atlas(physics)
    a3 = lift(at.mesh)
    swept = replace(flow, sweep_deg=sweep)
    st3 = replace(O3.setup_for(a3, replace(flow, sweep_deg=None), num), q_inf=swept_q(swept))
    return at, a3, flow, num, st3


@pytest.mark.parametrize("first_order", [True, False])
@pytest.mark.parametrize("physics", sorted(PHYSICS))
def test_unswept_extrusion_residual_is_the_2d_residual(physics, first_order):
    at, a3, flow, num, st3 = _setups(physics, 0.0)
    st2 = OR.setup_for(at, flow, num)
    U0 = OR.U0_of(st2)
    U2 = np.where(at.interior[None], perturbed(U0, 3e-2), U0)
    U25 = np.concatenate([U2[:3], np.zeros_like(U2[:1]), U2[3:]])
    R3 = O3.residual(section_to_planes(U25, at, a3, O3.U0_of(st3)), st3, first_order)
    R25 = planes_to_section(R3, a3, at)
    R2 = OR.residual(U2, st2, first_order)
    rows = [0, 1, 2] + list(range(4, R25.shape[0]))
    scale = np.abs(R2).max(axis=(1, 2))[:, None]
    worst = (np.abs(R25[rows] - R2)[:, at.interior] / scale).max()
    assert worst < 1e-12, worst
    assert np.abs(R25[3][at.interior]).max() == 0.0


@pytest.mark.parametrize("physics", ["euler", "sa"])
def test_swept_planes_have_the_same_residual_bit_for_bit(physics):
    at, a3, flow, num, st3 = _setups(physics, 30.0)
    U3u = O3.U0_of(st3)
    U25 = planes_to_section(U3u, a3, at)
    rng = np.random.default_rng(7)
    U25 = np.where(at.interior[None], U25 * (1.0 + 0.03 * rng.standard_normal(U25.shape)), U25)
    assert np.abs(U25[3][at.interior]).min() > 0.0
    U3 = section_to_planes(U25, at, a3, U3u)
    assert np.array_equal(planes_to_section(U3, a3, at), U25)
    R3 = O3.residual(U3, st3)
    for blk in a3.unpack(R3):
        assert np.array_equal(blk[..., 0], blk[..., 1])


def test_swept_freestream():
    """``freestream_vector`` with a sweep (``core/state.py:123-139``)."""
    q = swept_freestream(0.85, 2.0, 35.0, 1.4)
    a, s = np.deg2rad(2.0), np.deg2rad(35.0)
    assert np.allclose(q[1:4], 0.85 * np.array([np.cos(a) * np.cos(s), np.sin(a), np.cos(a) * np.sin(s)]),
                       rtol=0, atol=1e-15)
    assert (q[0], q[4]) == (1.0, 1.0 / 1.4)
    flow = FlowConditions(mach=0.85, alpha_deg=2.0, sweep_deg=35.0)
    assert np.isclose(flow.in_plane_mach, np.hypot(q[1], q[2]), rtol=1e-14)


def test_unswept_swept_coefficients_are_the_2d_ones():
    """Inviscid: the wall pressure of the extruded state through the swept
    force transformation at zero sweep gives the 2D oracle's CL, CD, CM."""
    at, a3, flow, num, st3 = _setups("euler", 0.0)
    st2 = OR.setup_for(at, flow, num)
    U2 = np.where(at.interior[None], perturbed(OR.U0_of(st2), 2e-2), OR.U0_of(st2))
    U3 = section_to_planes(np.concatenate([U2[:3], np.zeros_like(U2[:1]), U2[3:]]), at, a3, O3.U0_of(st3))
    from oracle import physics3d as P3
    q = P3.primitive([U3[k] for k in range(5)], 1.4, 3)
    p = q[4].reshape(-1)[a3.walls.cell] - 1.0 / 1.4
    loads = np.stack([p, 0 * p, 0 * p, 0 * p])
    c = swept_coefficients(loads, a3, st3.q_inf, 0.0, 1.0, (0.25, 0.0), False)
    ref = OS.forces(U2, st2)
    for k in ("cl", "cd", "cm"):
        assert abs(c[k] - ref[k]) < 1e-13, (k, c[k], ref[k])
    assert abs(c["cy"]) < 1e-15


@pytest.mark.parametrize("viscous", [False, True])
def test_swept_coefficient_tangent_is_the_derivative(viscous):
    """The complex-step tangent of the swept coefficients (loads and
    freestream moving together) against a fourth-order central difference."""
    a3 = lift(o_atlas().mesh)
    n = a3.walls.n
    rng = np.random.default_rng(2)
    loads = np.stack([0.1 * rng.standard_normal(n), *(1e-3 * rng.standard_normal((3, n)))])
    dloads = rng.standard_normal((4, n)) * np.array([[0.1], [1e-3], [1e-3], [1e-3]])
    q = swept_freestream(0.6, 2.5, 30.0, 1.4)
    dq = np.array([0.0, 0.3, -0.2, 0.1, 0.0])
    args = (a3, 30.0, 1.0, (0.25, 0.0), viscous)
    got = swept_coefficients_tangent(loads, dloads, a3, q, dq, *args[1:])

    def f(t):
        c = swept_coefficients(loads + t * dloads, a3, q + t * dq, *args[1:])
        return np.array([c[k] for k in ("cl", "cd", "cm", "cd_friction", "cy")])

    h = 1e-4
    fd = (4.0 * (f(h / 2) - f(-h / 2)) / h - (f(h) - f(-h)) / (2 * h)) / 3.0
    vec = np.array([got[k] for k in ("cl", "cd", "cm", "cd_friction", "cy")])
    assert np.allclose(vec, fd, rtol=1e-8, atol=1e-12), (vec, fd)
