"""Pins the oracle's explicit Runge-Kutta step and local time step
(``oracle/explicit.py``): the closed form of the local time step on a
Cartesian cell (``numerics/residual.py:142-166``), its linearity in the CFL,
the Euler scheme as one forward step and the stage recursion of
``solvers/explicit.py:48-71``."""

from __future__ import annotations

from dataclasses import replace
from types import SimpleNamespace

import numpy as np
import pytest

from conftest import o_atlas, perturbed
from jaxrans_amd.core.config import FlowConditions, NumericsConfig
from oracle import explicit as OX
from oracle import physics as P
from oracle import residual as OR


def cartesian(setup, nx=4, ny=3, dx=0.3, dy=0.2):
    """A uniform nx x ny grid of dx x dy cells as the time step sees it."""
    face_i = np.zeros((2, nx + 1, ny))
    face_i[0] = dy
    face_j = np.zeros((2, nx, ny + 1))
    face_j[1] = dx
    return replace(setup, atlas=SimpleNamespace(face_i=face_i, face_j=face_j, area=np.full((nx, ny), dx * dy)))


@pytest.mark.parametrize("viscous", [False, True])
def test_local_time_step_closed_form(viscous):
    flow = FlowConditions(mach=0.4, alpha_deg=0.0, reynolds=100.0 if viscous else None)
    st = cartesian(OR.setup_for(o_atlas(), flow, NumericsConfig()))
    rho, u, v, p = 1.1, 0.3, -0.2, 0.8
    g = st.gamma
    U = np.zeros((4, 4, 3))
    U[0], U[1], U[2], U[3] = rho, rho * u, rho * v, p / (g - 1) + 0.5 * rho * (u * u + v * v)
    dx, dy = 0.3, 0.2
    a = np.sqrt(g * p / rho)
    total = 2 * (abs(u) * dy + a * dy) + 2 * (abs(v) * dx + a * dx)
    if viscous:
        mu = P.viscosity(g * p / rho, st.mu_inf, st.sutherland)
        total += 4 * g / P.PR * mu / rho * (dx * dx + dy * dy) / (dx * dy)
    dt = OX.local_time_step(U, st, 1.5)
    assert np.allclose(dt, 1.5 * dx * dy / total, rtol=1e-14, atol=0)
    assert np.allclose(OX.local_time_step(U, st, 3.0), 2 * dt, rtol=1e-15, atol=0)


def test_euler_scheme_is_one_forward_step_and_stages_restart_from_u_n():
    st = OR.setup_for(o_atlas(), FlowConditions(mach=0.5, alpha_deg=1.25), NumericsConfig(flux="hllc"))
    U = perturbed(OR.U0_of(st), 1e-3)
    factor = OX.local_time_step(U, st, 1.2) / st.atlas.area
    R = OR.residual(U, st)
    one, n = OX.rk_step(U, st, (1.0,), 1.2)
    assert np.array_equal(one, U - factor[None] * R) and n == OR.norm(R, st)
    # two stages: the second restarts from U^n with the first stage's residual
    two, _ = OX.rk_step(U, st, (0.5, 1.0), 1.2)
    mid = U - 0.5 * factor[None] * R
    assert np.array_equal(two, U - factor[None] * OR.residual(mid, st))


def test_freestream_is_a_fixed_point_off_the_wall():
    st = OR.setup_for(o_atlas(), FlowConditions(mach=0.5, alpha_deg=1.25), NumericsConfig(flux="hllc"))
    U0 = OR.U0_of(st)
    U1, _ = OX.rk_step(U0, st, OX.RK_SCHEMES["rk5_jameson"], 1.5)
    wi, wj = st.atlas.windows[0]
    assert np.abs((U1 - U0)[:, wi, wj][:, :, 6:]).max() < 1e-12
    assert np.abs(U1 - U0).max() > 1e-8   # the wall moves it


@pytest.mark.parametrize("viscous", [False, True])
def test_local_time_step_3d_closed_form(viscous):
    """A box of dx x dy x dz cells at one state: the six faces' spectral
    radii and the diffusive limit in closed form."""
    from oracle import physics3d as P3
    nx, ny, nz, dx, dy, dz = 3, 4, 2, 0.3, 0.2, 0.5
    faces = []
    for a, d in enumerate((dy * dz, dx * dz, dx * dy)):
        shape = [3, nx, ny, nz]
        shape[1 + a] += 1
        f = np.zeros(shape)
        f[a] = d
        faces.append(f)
    flow = FlowConditions(mach=0.4, alpha_deg=0.0, reynolds=100.0 if viscous else None)
    st = SimpleNamespace(atlas=SimpleNamespace(faces=tuple(faces), volume=np.full((nx, ny, nz), dx * dy * dz), dim=3),
                         gamma=1.4, viscous=viscous, mu_inf=flow.mu_inf if viscous else None, sutherland=None, sa=False)
    rho, u, v, w, p = 1.1, 0.3, -0.2, 0.15, 0.8
    g = st.gamma
    U = np.stack([np.full((nx, ny, nz), x) for x in P3.conservative([rho, u, v, w, p], g, 3)])
    a = np.sqrt(g * p / rho)
    total = 2 * (abs(u) * dy * dz + a * dy * dz) + 2 * (abs(v) * dx * dz + a * dx * dz) + 2 * (abs(w) * dx * dy + a * dx * dy)
    if viscous:
        mu = P.viscosity(g * p / rho, st.mu_inf, None)
        total += 4 * g / P.PR * mu / rho * ((dy * dz) ** 2 + (dx * dz) ** 2 + (dx * dy) ** 2) / (dx * dy * dz)
    dt = OX.local_time_step3(U, st, 1.5)
    assert np.allclose(dt, 1.5 * dx * dy * dz / total, rtol=1e-14, atol=0)
