"""The oracle's 3D path, pinned without the reference:

* a 2D grid extruded along the span with symmetry planes at its ends gives,
  cell by cell, ``dy`` times the 2D residual, wave speeds, tangent and
  forces (the reference's ``tests/test_extruded.py`` property), for
  Euler with two flux/limiter pairs, laminar and SA flow;
* the swept, tapered four-block wing preserves the freestream off the wall
  (closed cells, exact interfaces);
* the composed ghost fill (the device's form) equals the staged fill;
* the metrics: closed cells, exact volumes of affine hexahedra;
* the 3D tangent against central differences; the line preconditioner
  inverts its assembled chains.
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import c_atlas, o_atlas, perturbed
from grids3d import extruded_atlas, small_wing
from jaxrans_amd.core.config import FlowConditions, ImplicitConfig, NumericsConfig
from jaxrans_amd.mesh.block3d import Grid3D, box_sides, metrics3d, normal_closure_3d
from oracle import residual as O2
from oracle import residual3d as O3
from oracle import solver as S2
from oracle import solver3d as S3

DY = 0.1
NK = 3

CASES = [
    ("euler-hllc-minmod", lambda: o_atlas(), FlowConditions(mach=0.5, alpha_deg=1.25), "hllc", "minmod"),
    ("euler-roe-vas", lambda: o_atlas(), FlowConditions(mach=0.5, alpha_deg=1.25), "roe", "van_albada_smooth"),
    ("laminar", lambda: o_atlas(wall="no_slip_wall"),
     FlowConditions(mach=0.5, alpha_deg=1.0, reynolds=5e3, sutherland_ratio=198.6 / 460.0), "hllc", "minmod"),
    ("sa-cgrid", lambda: c_atlas(), FlowConditions(mach=0.73, alpha_deg=2.5, reynolds=6.5e6, turbulence="sa_noft2"),
     "roe", "van_albada_smooth"),
]


def _lift_state(U2, at2, at3, s3):
    """The 2D state on the extruded canvas: rows (rho, rho u, 0, rho v, E[, nu])."""
    nv3 = s3.nvar
    U3 = O3.U0_of(s3).copy()
    rows = [0, 1, None, 2, 3] + ([4] if nv3 == 6 else [])
    for w2, w3 in zip(at2.windows, at3.windows):
        for r3, r2 in enumerate(rows):
            blk = np.zeros(U2[(0,) + w2].shape) if r2 is None else U2[(r2,) + w2]
            U3[(r3,) + w3] = np.repeat(blk[:, :, None], w3[2].stop - w3[2].start, axis=2)
    return U3, rows


def _compare(f3, f2, at2, at3, rows, scale):
    err = 0.0
    for w2, w3 in zip(at2.windows, at3.windows):
        for r3, r2 in enumerate(rows):
            a = f3[(r3,) + w3]
            b = np.zeros(a.shape) if r2 is None else scale * f2[(r2,) + w2][:, :, None]
            ref = max(np.abs(scale * f2[(0 if r2 is None else r2,) + w2]).max(), 1e-300)
            err = max(err, float(np.abs(a - b).max() / ref))
    return err


def _setups(make, flow, flux, lim):
    at2 = make()
    at3 = extruded_atlas(at2, NK, DY)
    num = NumericsConfig(flux=flux, limiter=lim, solver="implicit")
    return at2, at3, O2.setup_for(at2, flow, num), O3.setup_for(at3, flow, num)


@pytest.mark.parametrize("name, make, flow, flux, lim", CASES, ids=[c[0] for c in CASES])
def test_extruded_residual_and_tangent_equal_the_2d_ones(name, make, flow, flux, lim):
    at2, at3, s2, s3 = _setups(make, flow, flux, lim)
    assert at3.volume[at3.interior].min() > 0
    U2 = perturbed(O2.U0_of(s2), 1e-2)
    U3, rows = _lift_state(U2, at2, at3, s3)
    assert _compare(O3.residual(U3, s3), O2.residual(U2, s2), at2, at3, rows, DY) < 1e-12
    dU2 = np.random.default_rng(1).standard_normal(U2.shape) * 1e-3
    dU3, _ = _lift_state(dU2, at2, at3, s3)
    dU3 = np.where(at3.interior[None], dU3, 0.0)
    dU3[2] = 0.0
    dU2 = np.where(at2.interior[None], dU2, 0.0)
    assert _compare(O3.jvp(U3, dU3, s3), O2.jvp(U2, dU2, s2), at2, at3, rows, DY) < 1e-11
    l3, l2 = O3.wave_speeds(U3, s3), O2.wave_speeds(U2, s2)
    for a in range(2):
        for w2, w3 in zip(at2.windows, at3.windows):
            i3 = (slice(w3[0].start, w3[0].stop + (a == 0)), slice(w3[1].start, w3[1].stop + (a == 1)), w3[2])
            i2 = (slice(w2[0].start, w2[0].stop + (a == 0)), slice(w2[1].start, w2[1].stop + (a == 1)))
            assert np.allclose(l3[a][i3], DY * l2[a][i2][:, :, None], rtol=1e-12, atol=0)


def test_wing_newton_step_reduces_the_linear_residual():
    """The 3D step assembles pinned parts (residual, tangent, chains, the
    shared GMRES); the pseudo-time term carries the spanwise faces' wave
    speeds, so it is not the extruded 2D step.  A step from a perturbed
    freestream is accepted with GMRES reducing the linear residual."""
    wing = small_wing()
    s = O3.setup_for(wing, FlowConditions(mach=0.8, alpha_deg=2.0, reynolds=1e7, turbulence="sa_noft2"),
                     NumericsConfig(flux="roe", limiter="van_albada_smooth"))
    cfg = ImplicitConfig(krylov_subspace=10, krylov_restarts=1)
    U = np.where(wing.interior[None], perturbed(O3.U0_of(s), 1e-2), O3.U0_of(s))
    new, nrm, info = S3.newton_step(S2.NewtonState(U, 10.0), s, cfg, "cr")
    assert np.isfinite(nrm) and info["accept"] and info["lin_red"] < 0.5
    assert not np.array_equal(new.U, U)


def test_extruded_forces_equal_the_2d_ones():
    at2, at3, s2, s3 = _setups(CASES[0][1], FlowConditions(mach=0.5, alpha_deg=3.0), "hllc", "minmod")
    U2 = perturbed(O2.U0_of(s2), 1e-2)
    U3, _ = _lift_state(U2, at2, at3, s3)
    f2 = S2.forces(U2, s2)
    f3 = S3.forces(U3, s3, area=NK * DY)
    for k in ("cl", "cd", "cm"):
        assert f3[k] == pytest.approx(f2[k], rel=1e-11, abs=1e-14)


@pytest.fixture(scope="module")
def wing():
    return small_wing()


@pytest.mark.parametrize("flux, lim", [("hllc", "minmod"), ("roe", "van_albada"), ("hllc", "none"),
                                       ("roe", "van_albada_smooth")])
def test_wing_preserves_the_freestream_off_the_wall(wing, flux, lim):
    s = O3.setup_for(wing, FlowConditions(mach=0.8, alpha_deg=2.0), NumericsConfig(flux=flux, limiter=lim))
    U = O3.U0_of(s)
    R = O3.residual(U, s)
    mask = np.zeros(wing.shape, bool)
    for w in wing.windows:
        sub = np.zeros(tuple(x.stop - x.start for x in w), bool)
        sub[:, 3:, :] = True
        mask[w] = sub
    assert np.abs(R[:, mask]).max() < 1e-12


def test_wing_sa_flow_rows_preserve_the_freestream_off_the_wall(wing):
    flow = FlowConditions(mach=0.8, alpha_deg=2.0, reynolds=1e7, turbulence="sa_noft2")
    s = O3.setup_for(wing, flow, NumericsConfig(flux="roe", limiter="van_albada_smooth"))
    R = O3.residual(O3.U0_of(s), s)
    mask = np.zeros(wing.shape, bool)
    for w in wing.windows:
        sub = np.zeros(tuple(x.stop - x.start for x in w), bool)
        sub[:, 4:, :] = True
        mask[w] = sub
    assert np.abs(R[:5, mask]).max() < 1e-12


@pytest.mark.parametrize("sa", [False, True])
def test_composed_fill_equals_the_staged_fill(wing, sa):
    flow = (FlowConditions(mach=0.8, alpha_deg=2.0, reynolds=1e7, turbulence="sa_noft2") if sa
            else FlowConditions(mach=0.8, alpha_deg=2.0))
    s = O3.setup_for(wing, flow, NumericsConfig())
    U = perturbed(O3.U0_of(s), 5e-2)
    q = [r.reshape(-1) for r in O3.P3.primitive([U[k] for k in range(s.nvar)], s.gamma, 3)]
    a, b = O3.fill_ghosts(q, s), O3.fill_composed(q, s)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert np.bincount(wing.chains.n_ops).size >= 3       # corner ghosts compose several conditions


def test_metrics_close_cells_and_give_affine_volumes():
    rng = np.random.default_rng(4)
    A = np.eye(3) + 0.2 * rng.standard_normal((3, 3))
    if np.linalg.det(A) < 0:
        A[:, 0] *= -1
    g = np.stack(np.meshgrid(np.arange(4.0), np.arange(3.0), np.arange(5.0), indexing="ij"))
    nodes = np.einsum("ab,b...->a...", A, g)
    c, v, _ = metrics3d(nodes)
    assert np.allclose(v, np.linalg.det(A), rtol=1e-13)
    warped = nodes + 0.05 * rng.standard_normal(nodes.shape)
    blk = Grid3D(warped, box_sides((3, 2, 4), **dict.fromkeys(("i_min", "i_max", "j_min", "j_max", "k_min", "k_max"), "x")))
    assert np.abs(normal_closure_3d(blk)).max() < 1e-13
    # divergence theorem: the volumes sum to the volume the outer faces enclose
    tot = 0.0
    for a, S in enumerate(blk.faces):
        from jaxrans_amd.mesh.block3d import face_vectors
        mid = face_vectors(warped)[a][1]
        flux = np.sum(S * mid, axis=0)
        tot += np.take(flux, -1, axis=a).sum() - np.take(flux, 0, axis=a).sum()
    assert blk.volume.sum() == pytest.approx(tot / 3.0, rel=1e-12)


def test_wing_tangent_matches_central_differences(wing):
    flow = FlowConditions(mach=0.8, alpha_deg=2.0, reynolds=1e7, turbulence="sa_noft2")
    s = O3.setup_for(wing, flow, NumericsConfig(flux="roe", limiter="van_albada_smooth"))
    U = perturbed(O3.U0_of(s), 1e-2)
    dU = np.where(wing.interior[None], np.random.default_rng(2).standard_normal(U.shape), 0.0) * 1e-3 * np.abs(U)
    J = O3.jvp(U, dU, s, first_order=False)
    h = 1e-6
    fd = (O3.residual(U + h * dU, s) - O3.residual(U - h * dU, s)) / (2 * h)
    assert np.linalg.norm(J - fd) / np.linalg.norm(fd) < 1e-6


@pytest.mark.parametrize("method", ["cr", "thomas"])
def test_wing_line_preconditioner_inverts_its_chains(wing, method):
    flow = FlowConditions(mach=0.8, alpha_deg=2.0, reynolds=1e7, turbulence="sa_noft2")
    s = O3.setup_for(wing, flow, NumericsConfig(flux="roe", limiter="van_albada_smooth"))
    U = perturbed(O3.U0_of(s), 1e-2)
    lams = O3.wave_speeds(U, s)
    pre = S3.LinePreconditioner3D(U, lams, 7.0, s, method)
    r = np.random.default_rng(3).standard_normal(U.shape)
    z = pre(r)
    for blocks, rows in ((S3.flow_chain_blocks(U[:5], lams, 7.0, s), slice(0, 5)),
                         (S3.sa_chain_blocks(U, lams, 7.0, s), slice(5, 6))):
        lower, diag, upper = blocks
        lines = wing.lines
        for c in (0, lines.n_lines // 2, lines.n_lines - 1):
            s0, s1 = lines.start[c], lines.start[c + 1]
            cells = lines.cells[s0:s1]
            m = diag.shape[-1]
            n = s1 - s0
            Adense = np.zeros((n * m, n * m))
            for k in range(n):
                Adense[k * m:(k + 1) * m, k * m:(k + 1) * m] = diag[k, c]
                if k > 0:
                    Adense[k * m:(k + 1) * m, (k - 1) * m:k * m] = lower[k, c]
                if k + 1 < n:
                    Adense[k * m:(k + 1) * m, (k + 1) * m:(k + 2) * m] = upper[k, c]
            zc = z[rows].reshape(m, -1)[:, cells].T.reshape(-1)
            rc = r[rows].reshape(m, -1)[:, cells].T.reshape(-1)
            assert np.allclose(Adense @ zc, rc, rtol=1e-9, atol=1e-9 * np.abs(rc).max())
