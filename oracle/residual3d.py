"""The oracle's residual ``R(U)``, tangent, wave speeds and norm on a 3D canvas.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates for ``d = 3`` what ``oracle/residual.py`` restates for 2D:
``src/jaxrans/numerics/residual.py:54-124`` (block residual, viscous
fluxes), ``:169-203`` (wave speeds), ``:237-240`` (norm), ``:312-327``
(ghost fill, edge pad, interior mask); ``numerics/boundary.py:257-267``
(the staged fill, stages and passes of ``mesh/atlas3d.py``);
``numerics/reconstruction.py:29-80`` (MUSCL along each of the three axes);
``numerics/gradients.py:31-95`` (the diamond: cell line and the two face
diagonals, node values and positions the mean of 8 cells);
``numerics/turbulence.py:56-151`` (SA row, vorticity ``|curl V|``).
Rows are ``(rho, u, v, w, p[, nu_hat])``.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import dual as M
from . import physics as P
from . import physics3d as P3

D3 = 3


@dataclass(frozen=True)
class Setup3D:
    atlas: object
    gamma: float
    q_inf: np.ndarray
    flux: str = "hllc"
    limiter: str = "minmod"
    kappa: float = 0.0
    mu_inf: float | None = None
    sutherland: float | None = None
    sa: bool = False
    c_cr1: float = 0.0

    @property
    def nvar(self) -> int:
        return 6 if self.sa else 5

    @property
    def viscous(self) -> bool:
        return self.mu_inf is not None


def setup_for(atlas, flow, numerics) -> Setup3D:
    from jaxrans_amd.core.state import q_infinity
    q_inf = q_infinity(flow, dim=3)
    mu = None if flow.inviscid else flow.mu_inf
    return Setup3D(atlas, flow.gamma, np.asarray(q_inf, np.float64), numerics.flux, numerics.limiter,
                   numerics.kappa, mu, flow.sutherland_ratio, flow.turbulence is not None,
                   0.3 if flow.turbulence == "sa_noft2_qcr" else 0.0)


def U0_of(setup: Setup3D):
    """The uniform freestream state on the canvas; dual when the freestream
    carries a parameter tangent."""
    rows_ = P3.conservative([M.scalar(x) for x in setup.q_inf], setup.gamma, D3)
    if M.is_dual(*rows_):
        return M.stack([M.full(setup.atlas.shape, x) for x in rows_])
    c = np.array([np.float64(x) for x in rows_])
    return np.broadcast_to(c[:, None, None, None], (c.size,) + setup.atlas.shape).copy()


def _assign(row, dest, vals, n):
    """``row`` with entries ``dest`` replaced by ``vals`` (array or dual)."""
    if isinstance(row, M.D) or isinstance(vals, M.D):
        rv = M.value(row).copy()
        rd = M.tangent(row).copy() if isinstance(row, M.D) else np.zeros(n)
        rv[dest] = M.value(vals)
        rd[dest] = M.tangent(vals) if isinstance(vals, M.D) else 0.0
        return M.D(rv, rd)
    out = row.copy()
    out[dest] = vals
    return out


def fill_ghosts(q, setup: Setup3D):
    """Staged fill on flat primitive rows: every stage gathers from the state
    before it (copies across interfaces, boundary states elsewhere)."""
    n = q[0].shape[0]
    for st in setup.atlas.stages:
        src_rows = [M.take(r, st.src) for r in q]
        vals = [None] * len(q)
        pieces = []
        for kind in np.unique(st.kind):
            sel = np.flatnonzero(st.kind == kind)
            part = [M.take(r, sel) for r in src_rows]
            if kind == 0:
                out = part
            else:
                nrm = [st.normal[c, sel] for c in range(D3)]
                out = P3.bc_state(int(kind), part, nrm, setup.q_inf, setup.gamma, D3)
            pieces.append((sel, out))
        for k in range(len(q)):
            if any(isinstance(o[k], M.D) for _, o in pieces):
                v, d = np.zeros(st.dest.size), np.zeros(st.dest.size)
                for sel, o in pieces:
                    v[sel] = M.value(o[k])
                    d[sel] = M.tangent(o[k]) if isinstance(o[k], M.D) else 0.0
                vals[k] = M.D(v, d)
            else:
                v = np.zeros(st.dest.size)
                for sel, o in pieces:
                    v[sel] = o[k]
                vals[k] = v
        q = [_assign(q[k], st.dest, vals[k], n) for k in range(len(q))]
    return q


def fill_composed(q, setup: Setup3D):
    """The same fill evaluated from the composed chains (root cell, boundary
    tasks in order) -- the form the device evaluates."""
    ch = setup.atlas.chains
    n = q[0].shape[0]
    vals = [M.take(r, ch.root) for r in q]
    for step in range(ch.ops.shape[1]):
        active = np.flatnonzero(ch.n_ops > step)
        if active.size == 0:
            break
        task = ch.ops[active, step]
        for kind in np.unique(ch.kind[task]):
            sel = active[ch.kind[task] == kind]
            t = ch.ops[sel, step]
            part = [M.take(v, sel) for v in vals]
            nrm = [ch.normal[c, t] for c in range(D3)]
            out = P3.bc_state(int(kind), part, nrm, setup.q_inf, setup.gamma, D3)
            vals = [_assign(vals[k], sel, out[k], ch.dest.size) for k in range(len(vals))]
    return [_assign(q[k], ch.dest, vals[k], n) for k in range(len(q))]


def _window(r, g, shape, axis, off):
    NI, NJ, NK = shape
    sl = [slice(g, g + NI), slice(g, g + NJ), slice(g, g + NK)]
    sl[axis] = slice(g - 2 + off, g - 2 + off + shape[axis] + 1)
    return r[tuple(sl)]


def face_states(p, g, shape, axis, lim, kappa, first_order):
    """MUSCL states of one face family from edge-padded rows ``p``."""
    qmm, qm, qp, qpp = ([_window(r, g, shape, axis, off) for r in p] for off in range(4))
    if first_order:
        return qm, qp
    if kappa != 0.0:
        lo, hi = 0.25 * (1.0 - kappa), 0.25 * (1.0 + kappa)
        return ([qm[k] + lo * (qm[k] - qmm[k]) + hi * (qp[k] - qm[k]) for k in range(len(p))],
                [qp[k] - lo * (qpp[k] - qp[k]) - hi * (qp[k] - qm[k]) for k in range(len(p))])
    qL = [qm[k] + 0.5 * lim(qm[k] - qmm[k], qp[k] - qm[k]) for k in range(len(p))]
    qR = [qp[k] - 0.5 * lim(qp[k] - qm[k], qpp[k] - qp[k]) for k in range(len(p))]
    return qL, qR


def corner_average(a):
    """Mean of the 8 cells round each node, as three successive halvings
    (``src/jaxrans/numerics/gradients.py:31-40``)."""
    for axis in range(3):
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[axis], hi[axis] = slice(0, -1), slice(1, None)
        a = 0.5 * (a[tuple(lo)] + a[tuple(hi)])
    return a


def _cross(a, b):
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


def face_gradients(phi, centres):
    """Diamond gradients per family: ``grads[a][b][c]`` = d phi_c / d x_b on
    the faces of family ``a`` (``gradients.py:59-95``, ``_solve3``)."""
    shape = tuple(n - 2 for n in centres.shape[1:])
    xn = [corner_average(centres[c]) for c in range(3)]
    pn = [corner_average(f) for f in phi]
    out = []
    for a in range(3):
        def cell(offset):
            idx = [slice(1, 1 + n) for n in shape]
            idx[a] = slice(offset, offset + shape[a] + 1)
            return tuple(idx)

        def node(shifts):
            idx, it = [], iter(shifts)
            for b in range(3):
                start = 0 if b == a else next(it)
                idx.append(slice(start, start + shape[b] + (b == a)))
            return tuple(idx)

        L, R = cell(0), cell(1)
        d1 = [centres[c][R] - centres[c][L] for c in range(3)]
        diag = [(node((0, 0)), node((1, 1))), (node((1, 0)), node((0, 1)))]
        d2 = [xn[c][diag[0][1]] - xn[c][diag[0][0]] for c in range(3)]
        d3 = [xn[c][diag[1][1]] - xn[c][diag[1][0]] for c in range(3)]
        c23, c31, c12 = _cross(d2, d3), _cross(d3, d1), _cross(d1, d2)
        det = d1[0] * c23[0] + d1[1] * c23[1] + d1[2] * c23[2]
        grads = [[None] * len(phi) for _ in range(3)]
        for k in range(len(phi)):
            dp1 = phi[k][R] - phi[k][L]
            dp2 = pn[k][diag[0][1]] - pn[k][diag[0][0]]
            dp3 = pn[k][diag[1][1]] - pn[k][diag[1][0]]
            for b in range(3):
                grads[b][k] = (dp1 * c23[b] + dp2 * c31[b] + dp3 * c12[b]) / det
        out.append(grads)
    return out


def _lr(p1, a):
    shape = tuple(n - 2 for n in p1[0].shape)
    idx_l = [slice(1, 1 + n) for n in shape]
    idx_r = list(idx_l)
    idx_l[a] = slice(0, shape[a] + 1)
    idx_r[a] = slice(1, shape[a] + 2)
    return [r[tuple(idx_l)] for r in p1], [r[tuple(idx_r)] for r in p1]


def divergence(fl):
    """``sum_a F[hi] - F[lo]`` over the three families, in family order."""
    out = fl[0][1:, :, :] - fl[0][:-1, :, :]
    out = out + (fl[1][:, 1:, :] - fl[1][:, :-1, :])
    return out + (fl[2][:, :, 1:] - fl[2][:, :, :-1])


def green_gauss(phi_faces, setup):
    at = setup.atlas
    return [divergence(tuple(phi_faces[a] * at.faces[a][c] for a in range(3))) / at.volume for c in range(3)]


def net_flux(q, setup: Setup3D, first_order: bool):
    at, g = setup.atlas, setup.atlas.g
    shape = at.shape
    nf = 5
    pad = [M.pad_edge(r, ((g, g), (g, g), (g, g))) for r in q]
    lim = P.limiter(setup.limiter)
    flux = P3.FLUXES[setup.flux]
    S = at.faces
    fluxes = []
    for a in range(3):
        qL, qR = face_states(pad[:nf], g, shape, a, lim, setup.kappa, first_order)
        fluxes.append(flux(qL, qR, [S[a][c] for c in range(3)], setup.gamma, D3))
    sa_row = None
    if setup.viscous:
        p1 = [r[g - 1:r.shape[0] - (g - 1), g - 1:r.shape[1] - (g - 1), g - 1:r.shape[2] - (g - 1)] for r in pad]
        T1 = P3.temperature(p1, setup.gamma, D3)
        phi = [p1[1], p1[2], p1[3], T1] + ([p1[5]] if setup.sa else [])
        grads = face_gradients(phi, at.centres)
        visc = []
        for a in range(3):
            left, right = _lr(p1, a)
            gr = grads[a]
            gv = [[gr[b][c] for b in range(3)] for c in range(3)]      # gv[c][b] = d v_c / d x_b
            gT = [gr[b][3] for b in range(3)]
            vel = [0.5 * (left[1 + c] + right[1 + c]) for c in range(3)]
            T = 0.5 * (P3.temperature(left, setup.gamma, D3) + P3.temperature(right, setup.gamma, D3))
            mu = P.viscosity(T, setup.mu_inf, setup.sutherland)
            k = mu / (P.PR * (setup.gamma - 1.0))
            extra = None
            if setup.sa:
                face = [0.5 * (left[c] + right[c]) for c in range(6)]
                Tf = P3.temperature(face, setup.gamma, D3)
                muf = P.viscosity(Tf, setup.mu_inf, setup.sutherland)
                mut = P.eddy_viscosity(face[0], face[5] * setup.mu_inf, muf / face[0])
                mu = mu + mut
                k = k + mut / (P.PR_T * (setup.gamma - 1.0))
                if setup.c_cr1:
                    extra = P3.qcr_stress(gv, mut, setup.c_cr1)
            visc.append(P3.viscous_flux(vel, gv, gT, [S[a][c] for c in range(3)], mu, k, extra))
        if setup.sa:
            sa_row = _sa_row(p1, [fluxes[a][0] for a in range(3)],
                             [[grads[a][b][4] for b in range(3)] for a in range(3)], setup)
        fluxes = [[fluxes[a][c] - visc[a][c] for c in range(nf)] for a in range(3)]
    out = [divergence(tuple(fluxes[a][c] for a in range(3))) for c in range(nf)]
    if sa_row is not None:
        out.append(sa_row)
    return out


def _sa_row(p1, mass, grad_nu, setup):
    at, gamma, nu_inf = setup.atlas, setup.gamma, setup.mu_inf
    S = at.faces
    fl, faces = [], []
    for a in range(3):
        left, right = _lr(p1, a)
        up = M.where(M.value(mass[a]) > 0.0, left[5], right[5])
        face = [0.5 * (left[c] + right[c]) for c in range(6)]
        faces.append(face)
        mu = P.viscosity(P3.temperature(face, gamma, D3), setup.mu_inf, setup.sutherland)
        coef = P.diffusion_coefficient(face[0], face[5] * nu_inf, mu)
        fl.append(mass[a] * up - coef * P3.dot(grad_nu[a], [S[a][c] for c in range(3)]))
    shape = at.shape
    cell = [r[1:shape[0] + 1, 1:shape[1] + 1, 1:shape[2] + 1] for r in p1]
    rho = cell[0]
    mu = P.viscosity(P3.temperature(cell, gamma, D3), setup.mu_inf, setup.sutherland)
    # g[b][a] = d v_a / d x_b by Green-Gauss of the face-averaged velocity
    gvel = [green_gauss(tuple(faces[f][1 + c] for f in range(3)), setup) for c in range(3)]
    g = [[gvel[a][b] for a in range(3)] for b in range(3)]
    omega = P3.vorticity_magnitude(g)
    prod, dest = P.source(cell[5] * nu_inf, mu / rho, omega, at.distance)
    gn = green_gauss(tuple(faces[f][5] for f in range(3)), setup)
    src = rho * ((prod - dest) / nu_inf + P.CB2 / P.SIGMA * nu_inf * P3.dot(gn, gn))
    return divergence(tuple(fl)) - src * at.volume


def residual(U, setup: Setup3D, first_order: bool = False, composed: bool = False):
    """``R(U)`` on the canvas, ``(nvar, NI, NJ, NK)`` (array or dual)."""
    at = setup.atlas
    Ur = [U[k] for k in range(len(U))]
    q = P3.primitive(Ur, setup.gamma, D3)
    flat = [r.reshape(-1) for r in q]
    filled = (fill_composed if composed else fill_ghosts)(flat, setup)
    qc = [r.reshape(at.shape) for r in filled]
    nf = net_flux(qc, setup, first_order)
    U0 = U0_of(setup)
    return M.stack([M.where(at.interior, nf[k], Ur[k] - U0[k]) for k in range(len(Ur))])


def jvp(U, dU, setup: Setup3D, first_order: bool = True):
    return residual(M.D(U, dU), setup, first_order).d


def norm(R, setup: Setup3D) -> float:
    at = setup.atlas
    w = np.where(at.interior, at.volume, 0.0)
    return float(np.sqrt(np.sum(w * R[0] ** 2) / np.sum(w)))


def face_averages(a, atlas):
    """Neighbour-filled face averages per family (``atlas.py:163-183``)."""
    filled = a.reshape(a.shape[:-3] + (-1,))[..., atlas.neighbours].reshape(a.shape)
    lead = filled.ndim - 3
    out = []
    for axis in range(3):
        w = [(0, 0)] * filled.ndim
        w[lead + axis] = (1, 1)
        f = np.pad(filled, w, mode="edge")
        lo = [slice(None)] * f.ndim
        hi = [slice(None)] * f.ndim
        lo[lead + axis], hi[lead + axis] = slice(0, -1), slice(1, None)
        out.append(0.5 * (f[tuple(lo)] + f[tuple(hi)]))
    return out


def wave_speeds(U, setup: Setup3D):
    """Face ``lambda`` per family (``residual.py:169-192``)."""
    at, gamma = setup.atlas, setup.gamma
    q = np.stack(P3.primitive([U[k] for k in range(len(U))], gamma, D3))
    qf = face_averages(q, at)
    lams = []
    vols = face_averages(at.volume, at) if setup.viscous else None
    for a in range(3):
        f, s = qf[a], at.faces[a]
        mag = np.sqrt(P3.dot([s[c] for c in range(3)], [s[c] for c in range(3)]))
        lam = np.abs(P3.dot([f[1 + c] for c in range(3)], [s[c] for c in range(3)])) + np.sqrt(gamma * f[4] / f[0]) * mag
        if setup.viscous:
            mu = P.viscosity(gamma * f[4] / f[0], setup.mu_inf, setup.sutherland)
            if setup.sa:
                mu = mu + P.eddy_viscosity(f[0], f[5] * setup.mu_inf, mu / f[0])
            factor = 2.0 * gamma / P.PR
            lam = lam + factor * mu * P3.dot([s[c] for c in range(3)], [s[c] for c in range(3)]) / (f[0] * vols[a])
        lams.append(lam)
    return lams


def wave_speed_sum(lams):
    l0, l1, l2 = lams
    return (((l0[:-1] + l0[1:]) + l1[:, :-1]) + l1[:, 1:]) + l2[:, :, :-1] + l2[:, :, 1:]
