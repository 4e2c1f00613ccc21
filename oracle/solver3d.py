"""The oracle's implicit solver on a 3D canvas: line preconditioner (5x5 flow
blocks + scalar SA chains), Newton step and forces; GMRES and the block
solves are those of ``oracle/solver.py`` (dimension-free).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates ``src/jaxrans/solvers/preconditioners.py:209-306`` (chains along
each block's line axis, links into the three face families concatenated),
``solvers/implicit.py:54-166`` (``relax`` with ``d`` velocity rows),
``post/forces.py:48-161`` (3D: lift in the x-z plane, per ``q_inf *
area``, moment about ``y`` from pressure and friction).
"""

from __future__ import annotations

import numpy as np

from . import dual as M
from . import physics as P
from . import physics3d as P3
from .residual import _dsum
from .residual3d import (D3, Setup3D, _lr, face_averages, face_gradients, fill_ghosts, green_gauss, jvp, norm,
                         residual, wave_speed_sum, wave_speeds)
from .solver import (NewtonState, _padded_chains, cr_factor, cr_solve, gmres_right, thomas_factor,
                     thomas_solve)


def _rows(U):
    return [U[k] for k in range(len(U))]


def flux_jacobian(q, s, gamma):
    """``d(F.S)/dU`` columns (``euler.py:55-71``); ``q`` rows (5, n), ``s`` (3, n)."""
    U = P3.conservative(list(q[:5]), gamma, D3)
    cols = []
    for k in range(5):
        Ud = [M.D(U[c], np.full_like(U[c], 1.0 if c == k else 0.0)) for c in range(5)]
        f = P3.normal_flux(P3.primitive(Ud, gamma, D3), [s[0], s[1], s[2]], gamma, D3)
        cols.append(np.stack([fc.d for fc in f]))
    return np.stack(cols, axis=1)          # (5 rows, 5 vars, n)


def _concat(fields):
    """Per-family face fields ``(..., face shape)`` flattened and concatenated."""
    return np.concatenate([f.reshape(f.shape[:f.ndim - 3] + (-1,)) for f in fields], axis=-1)


def flow_chain_blocks(U, lams, cfl, setup: Setup3D):
    """``(lower, diag, upper)`` ``(L, n_lines, 5, 5)`` (preconditioners.py:209-232)."""
    at, gamma = setup.atlas, setup.gamma
    lines = at.lines
    idx = _padded_chains(lines)
    valid = idx >= 0
    slot = np.where(valid, idx, 0)
    q = np.stack(P3.primitive(_rows(U), gamma, D3))[:5].reshape(5, -1)
    diag_cell = (wave_speed_sum(lams) * (1.0 / cfl + 0.5)).reshape(-1)
    d = np.where(valid, diag_cell[lines.cells[slot]], 1.0)
    eye = np.eye(5)
    diag = d.T[:, :, None, None] * eye
    S = _concat(at.faces)                 # (3, n_faces)
    lam = _concat(lams)

    def coupling(face, sign, qnb):
        sgn = np.where(valid, sign[slot], 0)
        f = face[slot]
        s = [sgn.reshape(-1) * S[c][f.reshape(-1)] for c in range(3)]
        A = flux_jacobian(qnb.reshape(5, -1), s, gamma).reshape(5, 5, *sgn.shape)
        blk = 0.5 * A - 0.5 * lam[f][None, None] * eye.reshape(5, 5, 1, 1)
        blk = np.where(sgn[None, None] != 0, blk, 0.0)
        return np.transpose(blk, (3, 2, 0, 1))

    chains = q[:, lines.cells[slot]]
    above = np.concatenate([chains[:, :, 1:], chains[:, :, -1:]], axis=2)
    upper = coupling(lines.up_face, lines.up_sign, above)
    below = coupling(lines.down_face, lines.down_sign, chains)
    lower = np.concatenate([np.zeros_like(below[:1]), below[:-1]], axis=0)
    return lower, diag, upper


def sa_chain_blocks(U, lams, cfl, setup: Setup3D):
    """Scalar SA chains ``(L, n_lines, 1, 1)`` (preconditioners.py:235-287)."""
    at, gamma, nu_inf = setup.atlas, setup.gamma, setup.mu_inf
    lines = at.lines
    idx = _padded_chains(lines)
    valid = idx >= 0
    slot = np.where(valid, idx, 0)
    q = np.stack(P3.primitive(_rows(U), gamma, D3))
    qf = face_averages(q, at)
    vols = face_averages(at.volume, at)
    terms = []
    for a in range(3):
        f, s = qf[a], at.faces[a]
        mu = P.viscosity(gamma * f[4] / f[0], setup.mu_inf, setup.sutherland)
        kap = P.diffusion_coefficient(f[0], f[5] * nu_inf, mu)
        sv = [s[c] for c in range(3)]
        terms.append((f[0] * P3.dot([f[1], f[2], f[3]], sv), kap * P3.dot(sv, sv) / vols[a]))
    rho = q[0]
    diag = None
    for a, (m_f, d_f) in enumerate(terms):
        out_f, in_f = np.maximum(m_f, 0.0) + d_f, np.maximum(-m_f, 0.0) + d_f
        hi = [slice(None)] * 3
        lo = [slice(None)] * 3
        hi[a], lo[a] = slice(1, None), slice(0, -1)
        term = out_f[tuple(hi)] + in_f[tuple(lo)]
        diag = term if diag is None else diag + out_f[tuple(hi)] + in_f[tuple(lo)]
    diag = diag / rho
    nu_hat = q[5]
    gvel = [green_gauss(tuple(qf[f][1 + c] for f in range(3)), setup) for c in range(3)]
    g = [[gvel[a][b] for a in range(3)] for b in range(3)]
    omega = P3.vorticity_magnitude(g)
    nu = P.viscosity(gamma * q[4] / q[0], setup.mu_inf, setup.sutherland) / rho
    dist = at.distance
    st = P.s_tilde(omega, nu_hat * nu_inf, nu, dist)
    fw = P.f_w(P.r_of(nu_hat * nu_inf, st, dist))
    destr = np.where(nu_hat > 0.0, 2.0 * P.CW1 * fw * nu_inf * nu_hat / dist**2, 0.0)
    diag = diag + destr * at.volume + wave_speed_sum(lams) / cfl
    d_chain = np.where(valid, diag.reshape(-1)[lines.cells[slot]], 1.0)
    rho_chain = rho.reshape(-1)[lines.cells[slot]]
    mj = _concat([t[0] for t in terms])
    dj = _concat([t[1] for t in terms])

    def coupling(face, sign, rho_nb):
        sgn = np.where(valid, sign[slot], 0)
        f = face[slot]
        blk = -(np.maximum(-sgn * mj[f], 0.0) + dj[f]) / rho_nb
        return np.where(sgn != 0, blk, 0.0).T[:, :, None, None]

    above = np.concatenate([rho_chain[:, 1:], rho_chain[:, -1:]], axis=1)
    upper = coupling(lines.up_face, lines.up_sign, above)
    below = coupling(lines.down_face, lines.down_sign, rho_chain)
    lower = np.concatenate([np.zeros_like(below[:1]), below[:-1]], axis=0)
    return lower, d_chain.T[:, :, None, None], upper


class LinePreconditioner3D:
    """``r -> M^-1 r`` on 3D canvas fields; ``method`` ``"cr"`` or ``"thomas"``."""

    def __init__(self, U, lams, cfl, setup: Setup3D, method: str = "cr"):
        self.setup, self.method = setup, method
        lines = setup.atlas.lines
        self.idx = _padded_chains(lines)
        self.valid = self.idx >= 0
        self.slot = np.where(self.valid, self.idx, 0)
        fac = thomas_factor if method == "thomas" else cr_factor
        self.flow = fac(*flow_chain_blocks(U[:5], lams, cfl, setup))
        self.sa = fac(*sa_chain_blocks(U, lams, cfl, setup)) if setup.sa else None

    def _solve(self, factor, r):
        lines = self.setup.atlas.lines
        m = r.shape[0]
        flat = r.reshape(m, -1)
        rhs = np.where(self.valid[None], flat[:, lines.cells[self.slot]], 0.0)
        sol = thomas_solve if self.method == "thomas" else cr_solve
        x = np.transpose(sol(factor, np.transpose(rhs, (2, 1, 0))), (2, 1, 0))
        out = flat.copy()
        out[:, lines.cells[self.idx[self.valid]]] = x[:, self.valid]
        return out.reshape(r.shape)

    def __call__(self, r):
        flow = self._solve(self.flow, r[:5])
        if self.sa is None:
            return flow
        return np.concatenate([flow, self._solve(self.sa, r[5:])], axis=0)


def is_physical(U, gamma):
    q = P3.primitive(_rows(U), gamma, D3)
    return (q[0] > P.MIN_DENSITY) & (q[4] > P.MIN_PRESSURE) & np.isfinite(q[0]) & np.isfinite(q[4])


def relax(dU, U, gamma, fraction):
    """(implicit.py:54-81) with three velocity rows."""
    if fraction is None:
        return dU, 1.0
    q = P3.primitive(_rows(U), gamma, D3)
    rho, vel, p = q[0], q[1:4], q[4]
    de = dU[4]
    for k in range(3):
        de = de - vel[k] * dU[1 + k]
    dp = (gamma - 1.0) * (de + 0.5 * P3.dot(vel, vel) * dU[0])
    ratios = [np.abs(dU[0]) / rho, np.abs(dp) / p]
    if U.shape[0] > 5:
        nu_hat = U[5] / rho
        ratios.append(np.abs((dU[5] - nu_hat * dU[0]) / rho) / np.maximum(np.abs(nu_hat), 1.0))
    worst = max(float(np.max(r)) for r in ratios)
    omega = min(1.0, fraction / max(worst, 1e-300))
    return omega * dU, omega


def newton_step(state: NewtonState, setup: Setup3D, cfg, method: str = "cr", exact: bool = False):
    """One pseudo-transient Newton-Krylov step (implicit.py:99-166)."""
    U, cfl = state.U, state.cfl
    R = residual(U, setup)
    nrm = norm(R, setup)
    lams = wave_speeds(U, setup)
    vdt = wave_speed_sum(lams) / cfl

    def matvec(dU):
        return vdt[None] * dU + jvp(U, dU, setup, not exact)

    pre = LinePreconditioner3D(U, lams, cfl, setup, method) if cfg.preconditioner == "line_j" else (lambda r: r)
    dU, action = gmres_right(matvec, -R, pre, cfg.krylov_subspace, cfg.krylov_restarts, cfg.krylov_tol)
    dU, omega = relax(dU, U, setup.gamma, cfg.max_relative_update)
    lin_red = np.linalg.norm(omega * action + R) / max(np.linalg.norm(R), 1e-300)
    cand = U + dU
    physical = bool(np.all(is_physical(cand, setup.gamma))) and bool(np.all(np.isfinite(dU)))
    accept = physical and lin_red < 1.0
    reduction = state.residual_norm / max(nrm, 1e-300)
    if accept and lin_red < cfg.krylov_success and omega >= 1.0:
        nxt = cfl * min(max(reduction, 1.0), cfg.cfl_growth)
    elif accept:
        nxt = cfl
    else:
        nxt = cfl / cfg.cfl_cutback
    new = NewtonState(cand if accept else U, min(max(nxt, cfg.cfl_min), cfg.cfl_max),
                      nrm if accept else state.residual_norm)
    return new, nrm, {"accept": accept, "lin_red": lin_red, "omega": omega}


def forces(U, setup: Setup3D, area: float, length: float = 1.0, centre=(0.25, 0.0, 0.0)):
    """CL, CD, CM (pressure and friction about ``y``), CD friction, CY (forces.py:48-161).
    Dual when ``U`` or the freestream carries a tangent (the sensitivity's
    output JVP)."""
    at, gamma = setup.atlas, setup.gamma
    W = at.walls
    q = P3.primitive(_rows(U[:5]), gamma, D3)
    p = q[4].reshape(-1)[W.cell] - 1.0 / gamma
    fam = [f.reshape(3, -1) for f in at.faces]

    def pick(fields):
        if M.is_dual(*fields):
            v = pick([M.value(f) for f in fields])
            return M.D(v, pick([M.tangent(f) for f in fields]))
        out = np.zeros((3, W.n))
        for a in range(3):
            sel = W.family == a
            out[:, sel] = fields[a][:, W.face[sel]]
        return out

    S = W.into_fluid * pick(fam)
    shear = np.zeros_like(S)
    if setup.viscous:
        lam_setup = Setup3D(at, gamma, setup.q_inf[:5], setup.flux, setup.limiter, setup.kappa, setup.mu_inf,
                            setup.sutherland, False, 0.0)
        flat = fill_ghosts([r.reshape(-1) for r in q], lam_setup)
        qc = [M.pad_edge(r.reshape(at.shape), 1) for r in flat]
        T1 = P3.temperature(qc, gamma, D3)
        grads = face_gradients([qc[1], qc[2], qc[3], T1], at.centres)
        vis = []
        for a in range(3):
            left, right = _lr(qc, a)
            gr = grads[a]
            vel = [0.5 * (left[1 + c] + right[1 + c]) for c in range(3)]
            T = 0.5 * (P3.temperature(left, gamma, D3) + P3.temperature(right, gamma, D3))
            mu = P.viscosity(T, setup.mu_inf, setup.sutherland)
            k = mu / (P.PR * (gamma - 1.0))
            gv = [[gr[b][c] for b in range(3)] for c in range(3)]
            f = P3.viscous_flux(vel, gv, [gr[b][3] for b in range(3)], [at.faces[a][c] for c in range(3)], mu, k)
            vis.append(M.stack([f[1 + c].reshape(-1) for c in range(3)]))
        shear = W.into_fluid * pick(vis)
    rho, vel = M.scalar(setup.q_inf[0]), [M.scalar(v) for v in setup.q_inf[1:4]]
    speed = M.sqrt(P3.dot(vel, vel)) if M.is_dual(*vel) else float(np.sqrt(P3.dot(list(vel), list(vel))))
    pressure = -p * S
    load = pressure + shear
    drag = [v / speed for v in vel]
    lift = [-drag[2], 0.0 * drag[0], drag[0]]
    q_area = 0.5 * rho * speed**2 * area
    r = W.mid - np.asarray(centre[:3])
    moment = r[:, 2] * load[0] - r[:, 0] * load[2]

    def along(d, v):
        return d[0] * v[0] + d[1] * v[1] + d[2] * v[2]

    tot = [_dsum(along(lift, load) / q_area), _dsum(along(drag, load) / q_area), _dsum(moment / length / q_area)]
    force = [_dsum(load[c]) for c in range(3)]
    shear_sum = [_dsum(shear[c]) for c in range(3)]
    return {"cl": tot[0], "cd": tot[1], "cm": tot[2], "cd_friction": along(drag, shear_sum) / q_area,
            "cy": force[1] / q_area, "fx": force[0], "fy": force[1], "fz": force[2]}
