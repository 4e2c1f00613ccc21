"""The oracle's implicit solver: line preconditioner, GMRES, Newton step,
forces and the block driver.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates ``src/jaxrans/core/linalg.py:49-200`` (Gauss-Jordan inverse, block
Thomas, block cyclic reduction), ``solvers/preconditioners.py:74-306`` (line
preconditioner), ``solvers/krylov.py:34-137`` (right-preconditioned GMRES with
CGS2 and the Givens least squares), ``solvers/implicit.py:54-166`` (Newton
step), ``post/forces.py:48-200`` (forces) and ``solvers/driver.py:123-201``
(driver).
"""

from __future__ import annotations

import numpy as np

from . import dual as M
from . import physics as P
from .residual import (Setup, face_averages, green_gauss, jvp, norm, residual, rows, wave_speed_sum,
                       wave_speeds, fill_ghosts, _dsum, _lr, face_gradients)

# --- small dense blocks ---------------------------------------------------------------


def invert_blocks(A):
    """Unpivoted Gauss-Jordan on ``[A | I]`` per block ``(..., m, m)`` (linalg.py:49-77)."""
    m = A.shape[-1]
    rows_ = [np.concatenate([A[..., k, :], np.broadcast_to(np.eye(m)[k], A.shape[:-2] + (m,))], axis=-1)
             for k in range(m)]
    for k in range(m):
        rows_[k] = rows_[k] / rows_[k][..., k, None]
        for r in range(m):
            if r != k:
                rows_[r] = rows_[r] - rows_[r][..., k, None] * rows_[k]
    return np.stack([r[..., m:] for r in rows_], axis=-2)


def thomas_factor(lower, diag, upper):
    """Block Thomas forward sweep along axis 0 (linalg.py:80-100)."""
    n = diag.shape[0]
    pinv = np.empty_like(diag)
    chat = np.empty_like(diag)
    c_prev = np.zeros_like(diag[0])
    for k in range(n):
        pinv[k] = invert_blocks(diag[k] - lower[k] @ c_prev)
        chat[k] = pinv[k] @ upper[k]
        c_prev = chat[k]
    return lower, pinv, chat


def thomas_solve(factor, rhs):
    lower, pinv, chat = factor
    n = rhs.shape[0]
    d = np.empty_like(rhs)
    prev = np.zeros_like(rhs[0])
    for k in range(n):
        prev = np.einsum("...ij,...j->...i", pinv[k], rhs[k] - np.einsum("...ij,...j->...i", lower[k], prev))
        d[k] = prev
    x = np.empty_like(rhs)
    nxt = np.zeros_like(rhs[0])
    for k in range(n - 1, -1, -1):
        nxt = d[k] - np.einsum("...ij,...j->...i", chat[k], nxt)
        x[k] = nxt
    return x


def _next(x, n):
    return np.concatenate([x[1:], np.zeros_like(x[:1])])[:n]


def cr_factor(lower, diag, upper):
    """Block odd-even cyclic reduction factor (linalg.py:159-182)."""
    a, d, c = lower, diag, upper
    levels = []
    while d.shape[0] > 1:
        odd = d.shape[0] // 2
        inv, low, upp = invert_blocks(d[0::2]), a[0::2], c[0::2]
        alpha, gam = a[1::2] @ inv[:odd], c[1::2] @ _next(inv, odd)
        levels.append((inv, low, upp, alpha, gam))
        a, d, c = (-(alpha @ low[:odd]), d[1::2] - alpha @ upp[:odd] - gam @ _next(low, odd),
                   -(gam @ _next(upp, odd)))
    return levels, invert_blocks(d)


def cr_solve(factor, rhs):
    levels, last = factor
    mv = lambda A, x: np.einsum("...ij,...j->...i", A, x)  # noqa: E731
    b, evens = rhs, []
    for _, _, _, alpha, gam in levels:
        odd = b.shape[0] // 2
        evens.append(b[0::2])
        b = b[1::2] - mv(alpha, b[0::2][:odd]) - mv(gam, _next(b[0::2], odd))
    x = mv(last, b)
    for (inv, low, upp, _, _), be in zip(reversed(levels), reversed(evens)):
        even, odd = be.shape[0], x.shape[0]
        prev = np.concatenate([np.zeros_like(x[:1]), x])[:even]
        first = np.concatenate([x, np.zeros_like(x[:1])])[:even]
        xe = mv(inv, be - mv(low, prev) - mv(upp, first))
        pairs = np.stack([xe[:odd], x], axis=1).reshape((-1,) + x.shape[1:])
        x = np.concatenate([pairs, xe[odd:]])
    return x


# --- line preconditioner --------------------------------------------------------------


def flux_jacobian(q, s, gamma):
    """``d(F.S)/dU`` columns by forward differentiation of the analytic flux of
    the conservative state (euler.py:55-71); ``q`` rows (4, n), ``s`` (2, n)."""
    U = P.conservative(list(q[:4]), gamma)
    cols = []
    for k in range(4):
        Ud = [M.D(U[c], np.full_like(U[c], 1.0 if c == k else 0.0)) for c in range(4)]
        f = P.normal_flux(P.primitive(Ud, gamma), (s[0], s[1]), gamma)
        cols.append(np.stack([fc.d for fc in f]))
    return np.stack(cols, axis=1)          # (4 rows, 4 vars, n)


def chain_gather(field_flat, lines):
    return field_flat[..., lines.cells]


def _padded_chains(lines):
    """Per chain: slot indices padded to the longest chain (-1 beyond)."""
    n, L = lines.n_lines, lines.max_length
    idx = np.full((n, L), -1, np.int64)
    for c in range(n):
        s0, s1 = lines.start[c], lines.start[c + 1]
        idx[c, :s1 - s0] = np.arange(s0, s1)
    return idx


def flow_chain_blocks(U, lams, cfl, setup: Setup):
    """``(lower, diag, upper)`` ``(L, n_lines, 4, 4)`` (preconditioners.py:209-232)."""
    at, gamma = setup.atlas, setup.gamma
    lines = at.lines
    idx = _padded_chains(lines)
    valid = idx >= 0
    slot = np.where(valid, idx, 0)
    q = np.stack(P.primitive(rows(U), gamma))[:4].reshape(4, -1)
    diag_cell = (wave_speed_sum(lams) * (1.0 / cfl + 0.5)).reshape(-1)
    d = np.where(valid, diag_cell[lines.cells[slot]], 1.0)
    eye = np.eye(4)
    diag = d.T[:, :, None, None] * eye
    Sj = at.face_j.reshape(2, -1)
    lamj = lams[1].reshape(-1)

    def coupling(face, sign, qnb):
        sgn = np.where(valid, sign[slot], 0)
        f = face[slot]
        A = flux_jacobian(qnb.reshape(4, -1), (sgn.reshape(-1) * Sj[0][f.reshape(-1)],
                                              sgn.reshape(-1) * Sj[1][f.reshape(-1)]), gamma)
        A = A.reshape(4, 4, *sgn.shape)
        blk = 0.5 * A - 0.5 * lamj[f][None, None] * eye.reshape(4, 4, 1, 1)
        blk = np.where(sgn[None, None] != 0, blk, 0.0)
        return np.transpose(blk, (3, 2, 0, 1))                  # (L, n, 4, 4)

    chains = q[:, lines.cells[slot]]                             # (4, n, L)
    above = np.concatenate([chains[:, :, 1:], chains[:, :, -1:]], axis=2)
    upper = coupling(lines.up_face, lines.up_sign, above)
    below = coupling(lines.down_face, lines.down_sign, chains)
    lower = np.concatenate([np.zeros_like(below[:1]), below[:-1]], axis=0)
    return lower, diag, upper


def sa_chain_blocks(U, lams, cfl, setup: Setup):
    """Scalar SA chains ``(L, n_lines, 1, 1)`` (preconditioners.py:235-287)."""
    at, gamma, nu_inf = setup.atlas, setup.gamma, setup.mu_inf
    lines = at.lines
    idx = _padded_chains(lines)
    valid = idx >= 0
    slot = np.where(valid, idx, 0)
    q = np.stack(P.primitive(rows(U), gamma))
    qf = face_averages(q, at)
    vols = face_averages(at.area, at)
    S = (at.face_i, at.face_j)
    terms = []
    for a in range(2):
        f, s = qf[a], S[a]
        mu = P.viscosity(gamma * f[3] / f[0], setup.mu_inf, setup.sutherland)
        kap = P.diffusion_coefficient(f[0], f[4] * nu_inf, mu)
        terms.append((f[0] * (f[1] * s[0] + f[2] * s[1]), kap * (s[0] * s[0] + s[1] * s[1]) / vols[a]))
    rho = q[0]
    (m0, d0), (m1, d1) = terms
    out0, in0 = np.maximum(m0, 0.0) + d0, np.maximum(-m0, 0.0) + d0
    out1, in1 = np.maximum(m1, 0.0) + d1, np.maximum(-m1, 0.0) + d1
    diag = (out0[1:, :] + in0[:-1, :]) + out1[:, 1:] + in1[:, :-1]
    diag = diag / rho
    nu_hat = q[4]
    gu = green_gauss((qf[0][1], qf[1][1]), setup)
    gv = green_gauss((qf[0][2], qf[1][2]), setup)
    omega = np.abs(gv[0] - gu[1])
    nu = P.viscosity(gamma * q[3] / q[0], setup.mu_inf, setup.sutherland) / rho
    dist = at.distance
    st = P.s_tilde(omega, nu_hat * nu_inf, nu, dist)
    fw = P.f_w(P.r_of(nu_hat * nu_inf, st, dist))
    destr = np.where(nu_hat > 0.0, 2.0 * P.CW1 * fw * nu_inf * nu_hat / dist**2, 0.0)
    diag = diag + destr * at.area + wave_speed_sum(lams) / cfl
    d_chain = np.where(valid, diag.reshape(-1)[lines.cells[slot]], 1.0)
    rho_chain = rho.reshape(-1)[lines.cells[slot]]
    mj, dj = m1.reshape(-1), d1.reshape(-1)

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


class LinePreconditioner:
    """``r -> M^-1 r`` on canvas fields (preconditioners.py:197-206, 290-306);
    ``method`` ``"cr"`` (the reference's) or ``"thomas"``."""

    def __init__(self, U, lams, cfl, setup: Setup, method: str = "thomas"):
        self.setup, self.method = setup, method
        lines = setup.atlas.lines
        self.idx = _padded_chains(lines)
        self.valid = self.idx >= 0
        self.slot = np.where(self.valid, self.idx, 0)
        fac = thomas_factor if method == "thomas" else cr_factor
        self.flow = fac(*flow_chain_blocks(U[:4], lams, cfl, setup))
        self.sa = fac(*sa_chain_blocks(U, lams, cfl, setup)) if setup.sa else None

    def _solve(self, factor, r):
        lines = self.setup.atlas.lines
        m = r.shape[0]
        flat = r.reshape(m, -1)
        rhs = np.where(self.valid[None], flat[:, lines.cells[self.slot]], 0.0)     # (m, n, L)
        sol = thomas_solve if self.method == "thomas" else cr_solve
        x = sol(factor, np.transpose(rhs, (2, 1, 0)))                             # (L, n, m)
        x = np.transpose(x, (2, 1, 0))
        out = flat.copy()
        cells = lines.cells[self.idx[self.valid]]
        out[:, cells] = x[:, self.valid]
        return out.reshape(r.shape)

    def __call__(self, r):
        flow = self._solve(self.flow, r[:4])
        if self.sa is None:
            return flow
        return np.concatenate([flow, self._solve(self.sa, r[4:])], axis=0)


# --- GMRES ---------------------------------------------------------------------------


def hessenberg_lstsq(H, beta, rcond):
    """Givens least squares with pivot drop (krylov.py:100-137)."""
    m = H.shape[1]
    h = [[H[i, j] for j in range(m)] for i in range(m + 1)]
    g = [beta] + [0.0] * m
    cs, sn = [], []
    for j in range(m):
        for i in range(j):
            a, c = h[i][j], h[i + 1][j]
            h[i][j], h[i + 1][j] = cs[i] * a + sn[i] * c, cs[i] * c - sn[i] * a
        a, c = h[j][j], h[j + 1][j]
        r = np.sqrt(a * a + c * c)
        safe = r if r > 0 else 1.0
        cs.append(a / safe if r > 0 else 1.0)
        sn.append(c / safe if r > 0 else 0.0)
        h[j][j], h[j + 1][j] = r, 0.0
        g[j], g[j + 1] = cs[j] * g[j], -sn[j] * g[j]
    piv = [abs(h[j][j]) for j in range(m)]
    largest = max(piv)
    y = [0.0] * m
    for j in reversed(range(m)):
        rest = g[j]
        for k in range(j + 1, m):
            rest = rest - h[j][k] * y[k]
        y[j] = rest / h[j][j] if piv[j] > rcond * largest else 0.0
    return np.array(y)


def gmres_right(matvec, b, precondition, restart, maxiter, tol):
    """Right-preconditioned restarted GMRES with CGS2; returns ``(x, action)``
    with the action accumulated from the Arnoldi relation (krylov.py:34-97)."""
    bn = np.linalg.norm(b)
    tiny = np.finfo(np.float64).tiny
    rcond = float(np.finfo(np.float64).eps) * (restart + 1)
    x = np.zeros_like(b)
    action = np.zeros_like(b)
    r = b.copy()
    k = 0
    while k < maxiter and np.linalg.norm(r) > tol * bn:
        beta = np.linalg.norm(r)
        V = np.zeros((restart + 1,) + b.shape)
        V[0] = r / max(beta, tiny)
        H = np.zeros((restart + 1, restart))
        for j in range(restart):
            w = matvec(precondition(V[j]))
            h = np.array([np.vdot(V[i], w) if i <= j else 0.0 for i in range(restart + 1)])
            w = w - np.tensordot(h, V, axes=1)
            h2 = np.array([np.vdot(V[i], w) if i <= j else 0.0 for i in range(restart + 1)])
            w = w - np.tensordot(h2, V, axes=1)
            h = h + h2
            nw = np.linalg.norm(w)
            h[j + 1] = nw
            H[:, j] = h
            V[j + 1] = w / max(nw, tiny)
        y = hessenberg_lstsq(H, beta, rcond)
        x = x + precondition(np.tensordot(y, V[:restart], axes=1))
        action = action + np.tensordot(H @ y, V, axes=1)
        r = b - action
        k += 1
    return x, action


# --- Newton step ------------------------------------------------------------------------


def is_physical(U, gamma):
    q = P.primitive(rows(U), gamma)
    return (q[0] > P.MIN_DENSITY) & (q[3] > P.MIN_PRESSURE) & np.isfinite(q[0]) & np.isfinite(q[3])


def relax(dU, U, gamma, fraction):
    """(implicit.py:54-81)"""
    if fraction is None:
        return dU, 1.0
    q = P.primitive(rows(U), gamma)
    rho, u, v, p = q[:4]
    de = dU[3] - u * dU[1] - v * dU[2]
    dp = (gamma - 1.0) * (de + 0.5 * (u * u + v * v) * dU[0])
    ratios = [np.abs(dU[0]) / rho, np.abs(dp) / p]
    if U.shape[0] > 4:
        nu_hat = U[4] / rho
        ratios.append(np.abs((dU[4] - nu_hat * dU[0]) / rho) / np.maximum(np.abs(nu_hat), 1.0))
    worst = max(float(np.max(r)) for r in ratios)
    omega = min(1.0, fraction / max(worst, 1e-300))
    return omega * dU, omega


def next_cfl(cfl, accept, lin_red, omega, reduction, cfg):
    """Switched evolution relaxation (implicit.py:152-157): grow by the
    residual reduction, at most ``cfl_growth``, only after an accepted,
    unrelaxed step whose linear solve reached ``krylov_success``; hold after
    other accepted steps; cut back by ``cfl_cutback`` after a rejection;
    clipped to ``[cfl_min, cfl_max]``."""
    if accept and lin_red < cfg.krylov_success and omega >= 1.0:
        nxt = cfl * min(max(reduction, 1.0), cfg.cfl_growth)
    elif accept:
        nxt = cfl
    else:
        nxt = cfl / cfg.cfl_cutback
    return min(max(nxt, cfg.cfl_min), cfg.cfl_max)


class NewtonState:
    def __init__(self, U, cfl, residual_norm=np.inf):
        self.U, self.cfl, self.residual_norm = U, cfl, residual_norm


def newton_step(state: NewtonState, setup: Setup, cfg, method: str | None = None, exact: bool = False):
    """One pseudo-transient Newton-Krylov step (implicit.py:99-166); returns
    ``(new state, |R| before the update, info)``.  ``method`` (``"cr"`` or
    ``"thomas"``) defaults to the line solver ``cfg`` names."""
    if method is None:
        method = {"cyclic_reduction": "cr", "thomas": "thomas"}[getattr(cfg, "line_solver", "thomas")]
    U, cfl = state.U, state.cfl
    R = residual(U, setup)
    nrm = norm(R, setup)
    lams = wave_speeds(U, setup)
    vdt = wave_speed_sum(lams) / cfl
    first_order = not exact

    def matvec(dU):
        return vdt[None] * dU + jvp(U, dU, setup, first_order)

    pre = LinePreconditioner(U, lams, cfl, setup, method) if cfg.preconditioner == "line_j" else (lambda r: r)
    dU, action = gmres_right(matvec, -R, pre, cfg.krylov_subspace, cfg.krylov_restarts, cfg.krylov_tol)
    dU, omega = relax(dU, U, setup.gamma, cfg.max_relative_update)
    lin_red = np.linalg.norm(omega * action + R) / max(np.linalg.norm(R), 1e-300)
    cand = U + dU
    physical = bool(np.all(is_physical(cand, setup.gamma))) and bool(np.all(np.isfinite(dU)))
    accept = physical and lin_red < 1.0
    reduction = state.residual_norm / max(nrm, 1e-300)
    new = NewtonState(cand if accept else U, next_cfl(cfl, accept, lin_red, omega, reduction, cfg),
                      nrm if accept else state.residual_norm)
    return new, nrm, {"accept": accept, "lin_red": lin_red, "omega": omega}


# --- forces -----------------------------------------------------------------------------


def forces(U, setup: Setup, length=1.0, centre=(0.25, 0.0)):
    """CL, CD, CM (pressure-only moment), friction drag (forces.py:48-161).
    Dual when ``U`` or the freestream (``setup.q_inf`` components,
    ``setup.mu_inf``) carries a tangent: the output JVP of the sensitivity
    (``src/jaxrans/solvers/sensitivity.py:63``)."""
    at, gamma = setup.atlas, setup.gamma
    W = at.walls
    NI, NJ = at.shape
    q = P.primitive(rows(U[:4]), gamma)
    p = q[3].reshape(-1)[W.cell] - 1.0 / gamma
    fi, fj = at.face_i.reshape(2, -1), at.face_j.reshape(2, -1)
    S = W.into_fluid * np.where(W.family == 0, fi[:, np.where(W.family == 0, W.face, 0)],
                                fj[:, np.where(W.family == 1, W.face, 0)])
    shear = np.zeros_like(S)
    if setup.viscous:
        lam_setup = Setup(at, gamma, setup.q_inf[:4], setup.flux, setup.limiter, setup.kappa, setup.mu_inf,
                          setup.sutherland, False, 0.0)
        flat = fill_ghosts([r.reshape(-1) for r in q], lam_setup)
        qc = [M.pad_edge(r.reshape(NI, NJ), 1) for r in flat]
        T1 = P.temperature(qc, gamma)
        grads = face_gradients([qc[1], qc[2], T1], at.centres)
        vis = []
        for a in range(2):
            left, right = _lr(qc, a)
            gx, gy = grads[a]
            vel = [0.5 * (left[1] + right[1]), 0.5 * (left[2] + right[2])]
            T = 0.5 * (P.temperature(left, gamma) + P.temperature(right, gamma))
            mu = P.viscosity(T, setup.mu_inf, setup.sutherland)
            k = mu / (P.PR * (gamma - 1.0))
            f = P.viscous_flux(vel, [[gx[0], gy[0]], [gx[1], gy[1]]], [gx[2], gy[2]],
                               (at.face_i, at.face_j)[a], mu, k)
            vis.append(M.stack([f[1].reshape(-1), f[2].reshape(-1)]))
        shear = W.into_fluid * M.where(W.family == 0, vis[0][:, np.where(W.family == 0, W.face, 0)],
                                       vis[1][:, np.where(W.family == 1, W.face, 0)])
    rho, u, v = (M.scalar(x) for x in setup.q_inf[:3])
    speed = P.speed_of(u, v)
    pressure = -p * S
    load = pressure + shear
    dx, dy = u / speed, v / speed           # drag direction
    lx, ly = -dy, dx                        # lift direction
    q_area = 0.5 * rho * speed**2 * length
    r = W.mid - np.asarray(centre[:2])
    moment = r[:, 1] * pressure[0] - r[:, 0] * pressure[1]
    cl = _dsum((lx * load[0] + ly * load[1]) / q_area)
    cd = _dsum((dx * load[0] + dy * load[1]) / q_area)
    cm = _dsum(moment / length / q_area)
    return {"cl": cl, "cd": cd, "cm": cm, "cd_friction": (dx * _dsum(shear[0]) + dy * _dsum(shear[1])) / q_area,
            "fx": _dsum(load[0]), "fy": _dsum(load[1])}


# --- driver ----------------------------------------------------------------------------


def run_oracle(setup: Setup, cfg, *, tol: float, max_iters: int, block_iters: int = 20, U=None, method=None,
               length=1.0, centre=(0.25, 0.0)):
    """Newton steps in blocks until ``last <= tol * peak`` at a block end
    (driver.py:123-201 without the settled test); returns the final state,
    the residual history and the forces."""
    from .residual import U0_of
    s = NewtonState(U0_of(setup) if U is None else U, cfg.cfl)
    hist, peak = [], 0.0
    while len(hist) < max_iters:
        for _ in range(block_iters):
            s, n, _ = newton_step(s, setup, cfg, method)
            hist.append(n)
        block = np.asarray(hist[-block_iters:])
        if not np.all(np.isfinite(block)):
            break
        peak = max(peak, float(block.max()))
        if block[-1] <= tol * peak:
            break
    return s, np.asarray(hist), forces(s.U, setup, length, centre)
