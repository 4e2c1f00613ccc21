// Residual assembly and its tangent on the canvas: conservative -> primitive,
// staged ghost fill, face fluxes (MUSCL + HLLC/Roe, diamond-gradient viscous
// flux, SA face flux), and the per-cell divergence with the SA source.
//
// Restates src/jaxrans/numerics/residual.py:54-124, 312-327,
// numerics/boundary.py:257-267, reconstruction.py:29-80, gradients.py:31-95,
// turbulence.py:56-151.  Every kernel is templated on the scalar type:
// double gives R(U), Dual gives (R, dR/dU . dU) in one pass.
#pragma once
#include "jr_physics.h"
#include "jr_types.h"

namespace jr {

template <class T> struct Io;
template <> struct Io<double> {
  static __device__ __forceinline__ double ld(const double* a, const double*, size_t i) { return a[i]; }
  static __device__ __forceinline__ void st(double* a, double*, size_t i, double x) { a[i] = x; }
};
template <> struct Io<Dual> {
  static __device__ __forceinline__ Dual ld(const double* a, const double* d, size_t i) { return mkd(a[i], d[i]); }
  static __device__ __forceinline__ void st(double*, double* d, size_t i, Dual x) { d[i] = x.d; }
};

// Primitives of U (and their tangent) with the ghost fill, in one launch
// (residual.py:312-322, boundary.py:257-267).  Threads [0, nc) convert the
// cells that no stage overwrites; the next n_task[0] threads are stage-1
// tasks, the last n_task[1] stage-2 tasks.  Each stage is a gather from the
// state before it, so a ghost value is computed from its source cell's
// conservative state directly: prim(U[src]) through the task's condition, and
// for a stage-2 task whose source is itself a stage-1 ghost, through that
// stage-1 task first.  The host maps guarantee that every cell is written by
// exactly one thread (stage-1 tasks that stage 2 overwrites are skipped).
template <class T, int NV>
__device__ __forceinline__ void prim_of(const double* __restrict__ U, const double* __restrict__ dU, size_t base,
                                        int nc, int c, double gamma, T* q) {
  T u[NV];
  for (int k = 0; k < NV; ++k) u[k] = Io<T>::ld(U, dU, base + (size_t)k * nc + c);
  primitive<T, NV>(u, gamma, q);
}

// The circulation of case cs for the point-vortex far field (k_circulation):
// its value, and for the tangent its derivative along the tangent's dU.
template <class T> __device__ __forceinline__ T circ_of(const double* c, int cs);
template <> __device__ __forceinline__ double circ_of<double>(const double* c, int cs) { return c[2 * cs]; }
template <> __device__ __forceinline__ Dual circ_of<Dual>(const double* c, int cs) {
  return mkd(c[2 * cs], c[2 * cs + 1]);
}

template <class T, int NV, class TQ>
__device__ __forceinline__ void ghost_of(const Plan& P, int stage, int t, const T* q, const TQ* q_inf,
                                         double gamma, T* g, int cs) {
  const int kind = P.t_kind[stage][t];
  if (kind == K_COPY) {
    for (int k = 0; k < NV; ++k) g[k] = q[k];
  } else {
    const int n = P.n_task[stage];
    const double nx = P.t_norm[stage][t], ny = P.t_norm[stage][n + t];
    if (kind == K_FARFIELD && P.vortex) {
      // the reference state of this face: the vortex of the current lift
      // (point_vortex_far_field, boundary.py:222-240)
      T qr[NV];
      vortex_state<T, NV>(P.t_mid[stage][t] - P.vx, P.t_mid[stage][n + t] - P.vy, q_inf, gamma,
                          circ_of<T>(P.circ, cs), qr);
      far_field_state<T, T, NV>(q, nx, ny, qr, gamma, g);
    } else {
      bc_state<T, NV>(kind, q, nx, ny, q_inf, gamma, g);
    }
  }
}

//
// PT (with T = Dual): the freestream carries its tangent along a flow
// parameter (P.dfree), so the tangent is that of the fill along (dU, dq_inf).
template <class T, int NV, bool PT = false>
__global__ void __launch_bounds__(BLOCK) k_fill(Plan P, double gamma, const Ctl* __restrict__ ctl,
                                                const double* __restrict__ U, const double* __restrict__ dU,
                                                double* __restrict__ Q, double* __restrict__ dQ, int check_active) {
  bind_case(P, blockIdx.y);
  int t = blockIdx.x * BLOCK + threadIdx.x;
  const int cs = blockIdx.y;
  if (skip_case(ctl, cs, check_active)) return;
  const int nc = P.nc;
  const size_t base = (size_t)cs * NV * nc;
  typename Free<PT>::type qfree[6];
  for (int k = 0; k < 6; ++k) qfree[k] = free_val<PT>(ctl[cs].q_inf[k], P.dfree, cs, k);
  const typename Free<PT>::type* q_inf = qfree;
  T q[NV], g[NV];
  int dst;
  if (t < nc) {
    if (P.ghost[t]) return;
    prim_of<T, NV>(U, dU, base, nc, t, gamma, q);
    for (int k = 0; k < NV; ++k) Io<T>::st(Q, dQ, base + (size_t)k * nc + t, q[k]);
    return;
  }
  t -= nc;
  if (t < P.n_task[0]) {
    if (P.t_skip0[t]) return;
    prim_of<T, NV>(U, dU, base, nc, P.t_src[0][t], gamma, q);
    ghost_of<T, NV>(P, 0, t, q, q_inf, gamma, g, cs);
    dst = P.t_dest[0][t];
  } else {
    t -= P.n_task[0];
    if (t >= P.n_task[1]) return;
    const int p = P.t_prev1[t];
    if (p >= 0) {
      T q0[NV];
      prim_of<T, NV>(U, dU, base, nc, P.t_src[0][p], gamma, q0);
      ghost_of<T, NV>(P, 0, p, q0, q_inf, gamma, q, cs);
    } else {
      prim_of<T, NV>(U, dU, base, nc, P.t_src[1][t], gamma, q);
    }
    ghost_of<T, NV>(P, 1, t, q, q_inf, gamma, g, cs);
    dst = P.t_dest[1][t];
  }
  for (int k = 0; k < NV; ++k) Io<T>::st(Q, dQ, base + (size_t)k * nc + dst, g[k]);
}

// Jacobian of every final ghost value with respect to its root cell's
// primitives, at the state of the residual (Q): column k is the tangent of
// the fill with a unit tangent on primitive k -- the same dual-number fill
// (ghost_of, through the stage-1 task first for a corner), so applying it to
// a root tangent reproduces the fill's tangent.  jac[case][q][NV][NV]
// (row = ghost primitive, column = root primitive).  Once per Newton step.
template <int NV>
__global__ void __launch_bounds__(BLOCK) k_ghost_jac(Plan P, double gamma, const Ctl* __restrict__ ctl,
                                                     const double* __restrict__ Q, double* __restrict__ jac) {
  bind_case(P, blockIdx.y);
  const int q = blockIdx.x * BLOCK + threadIdx.x;
  const int cs = blockIdx.y;
  if (q >= P.n_final || ctl[cs].parked) return;   // a parked serving slot: nothing to do this step
  const int nc = P.nc, n0 = P.n_task[0];
  const size_t base = (size_t)cs * NV * nc;
  const double* q_inf = ctl[cs].q_inf;
  const int task = P.gq_task[q], root = P.gq_root[q];
  double qr[NV];
  for (int k = 0; k < NV; ++k) qr[k] = Q[base + (size_t)k * nc + root];
  double* J = jac + ((size_t)cs * P.n_final + q) * NV * NV;
  for (int col = 0; col < NV; ++col) {
    Dual s[NV], g[NV];
    for (int k = 0; k < NV; ++k) s[k] = mkd(qr[k], k == col ? 1.0 : 0.0);
    if (task < n0) {
      ghost_of<Dual, NV>(P, 0, task, s, q_inf, gamma, g, cs);
    } else {
      const int t = task - n0, p = P.t_prev1[t];
      if (p >= 0) {
        Dual s1[NV];
        ghost_of<Dual, NV>(P, 0, p, s, q_inf, gamma, s1, cs);
        ghost_of<Dual, NV>(P, 1, t, s1, q_inf, gamma, g, cs);
      } else {
        ghost_of<Dual, NV>(P, 1, t, s, q_inf, gamma, g, cs);
      }
    }
    for (int r = 0; r < NV; ++r) J[r * NV + col] = g[r].d;
  }
}

template <class T, int NV>
__device__ __forceinline__ void load_q(const double* Q, const double* dQ, size_t base, int nc, int c, T* q) {
  for (int k = 0; k < NV; ++k) q[k] = Io<T>::ld(Q, dQ, base + (size_t)k * nc + c);
}

// One face's flux, or its tangent (T = Dual): MUSCL states from the cells
// either side (qm, qp) and, second order, their outer neighbours (qmm,
// qpp), the Roe or HLLC flux through S = (sx, sy), and, viscous, minus the
// viscous flux from the diamond gradient (the face's node values pn0, pn1,
// and its diamond dg = d1x, d1y, d2x, d2y) with mu_inf (TM: Dual when it
// carries a parameter tangent); row 4 (SA) the nu_hat face flux.
// (residual.py:54-124, reconstruction.py:29-80, gradients.py:59-95,
// turbulence.py:78-90.)
// diamond(pn0, pn1, dg) fetches the node values and the diamond when the
// viscous part needs them (after the inviscid flux, so that they are not
// live across it).
template <class T, int NV, bool VISC, class TM, class DiamondFn>
__device__ __forceinline__ void face_flux(const Physics& ph, int first_order, const T* qmm, const T* qm, const T* qp,
                                          const T* qpp, double sx, double sy, DiamondFn diamond, TM mu_inf, T* Fl) {
  const double gamma = ph.gamma;
  T qL[4], qR[4];
  if (first_order) {
    for (int k = 0; k < 4; ++k) { qL[k] = qm[k]; qR[k] = qp[k]; }
  } else if (ph.kappa != 0.0) {
    double lo = 0.25 * (1.0 - ph.kappa), hi = 0.25 * (1.0 + ph.kappa);
    for (int k = 0; k < 4; ++k) {
      qL[k] = qm[k] + lo * (qm[k] - qmm[k]) + hi * (qp[k] - qm[k]);
      qR[k] = qp[k] - lo * (qpp[k] - qp[k]) - hi * (qp[k] - qm[k]);
    }
  } else {
    for (int k = 0; k < 4; ++k) {
      qL[k] = qm[k] + 0.5 * limit<T>(ph.limiter, qm[k] - qmm[k], qp[k] - qm[k]);
      qR[k] = qp[k] - 0.5 * limit<T>(ph.limiter, qp[k] - qm[k], qpp[k] - qp[k]);
    }
  }
  if (ph.flux == FLUX_ROE) roe<T>(qL, qR, sx, sy, gamma, Fl);
  else hllc<T>(qL, qR, sx, sy, gamma, Fl);
  if constexpr (NV > 4) Fl[4] = lift<T>(0.0);

  if constexpr (VISC) {
    T pn0[4], pn1[4];
    double dg[4];
    diamond(pn0, pn1, dg);
    const double d1x = dg[0], d1y = dg[1], d2x = dg[2], d2y = dg[3];
    const double det = d1x * d2y - d2x * d1y;
    constexpr int NPHI = NV > 4 ? 4 : 3;
    T pl[4], pr[4];
    pl[0] = qm[1]; pl[1] = qm[2]; pl[2] = temperature(qm, gamma);
    pr[0] = qp[1]; pr[1] = qp[2]; pr[2] = temperature(qp, gamma);
    if (NPHI > 3) { pl[3] = qm[NV - 1]; pr[3] = qp[NV - 1]; }
    T gx[4], gy[4];
    const double idet = 1.0 / det;
    for (int k = 0; k < NPHI; ++k) {
      T dp1 = pr[k] - pl[k];
      T dp2 = pn1[k] - pn0[k];
      gx[k] = (dp1 * d2y - dp2 * d1y) * idet;
      gy[k] = (d1x * dp2 - d2x * dp1) * idet;
    }
    T gv[2][2] = {{gx[0], gy[0]}, {gx[1], gy[1]}};
    T gT[2] = {gx[2], gy[2]};
    T vel[2] = {0.5 * (qm[1] + qp[1]), 0.5 * (qm[2] + qp[2])};
    T Tf = 0.5 * (temperature(qm, gamma) + temperature(qp, gamma));
    T mu = viscosity<T>(Tf, mu_inf, ph.sutherland);
    T kc = mu / (PR * (gamma - 1.0));
    T extra[2][2];
    bool has_extra = false;
    if constexpr (NV > 4) {
      T face[5];
      for (int k = 0; k < 5; ++k) face[k] = 0.5 * (qm[k] + qp[k]);
      T muf = viscosity<T>(temperature(face, gamma), mu_inf, ph.sutherland);
      T mut = eddy_viscosity<T>(face[0], face[4] * mu_inf, muf / face[0]);
      mu = mu + mut;
      kc = kc + mut / (PR_T * (gamma - 1.0));
      if (ph.c_cr1 != 0.0) { qcr_stress<T>(gv, mut, ph.c_cr1, extra); has_extra = true; }
      // SA face flux: upwind convection on the mass flux, diffusion through the face gradient.
      T up = val(Fl[0]) > 0.0 ? qm[4] : qp[4];
      T coef = diffusion_coefficient<T>(face[0], face[4] * mu_inf, muf);
      Fl[4] = Fl[0] * up - coef * (gx[3] * sx + gy[3] * sy);
    }
    T fv[4];
    // two inlined calls keep extra in registers (a selected pointer sent it to scratch)
    if (has_extra) viscous_flux<T>(vel, gv, gT, sx, sy, mu, kc, extra, fv);
    else viscous_flux<T>(vel, gv, gT, sx, sy, mu, kc, nullptr, fv);
    for (int k = 0; k < 4; ++k) Fl[k] = Fl[k] - fv[k];
  }
}

// Face fluxes of both families.  Face index f < nfi: i-face (fi, j) between
// cells (fi-1, j) and (fi, j); else j-face (i, fj) between (i, fj-1) and (i, fj).
// Faces not adjacent to an interior cell are zero.  Rows 0..3 hold the
// inviscid minus the viscous flux, row 4 (SA) the nu_hat face flux.
// Node values phi = (u, v, T[, nu_hat]) of the 2D diamond: node (ni, nj) is
// the mean of canvas cells (ni-1..ni, nj-1..nj), ((p00 + p10) + p01) + p11
// (gradients.py:31-40).  One thread per node of the (NI+1)(NJ+1) node grid;
// nodes on the canvas edge are never read and are skipped.  Computing each
// node once per evaluation, instead of 2 x 4 cell loads in every face
// thread, keeps the face kernel in registers.  Tangent mode writes dNd
// only: the values are those of the residual at the same state.
template <class T, int NV>
__global__ void __launch_bounds__(BLOCK) k_nodes(Plan P, double gamma, const Ctl* __restrict__ ctl,
                                                 const double* __restrict__ Q, const double* __restrict__ dQ,
                                                 double* __restrict__ Nd, double* __restrict__ dNd,
                                                 int check_active) {
  constexpr int NPHI = NV > 4 ? 4 : 3;
  const int n = blockIdx.x * BLOCK + threadIdx.x;
  const int cs = blockIdx.y;
  if (skip_case(ctl, cs, check_active)) return;
  const int NI = P.NI, NJ = P.NJ, nc = P.nc;
  const int nn = (NI + 1) * (NJ + 1);
  if (n >= nn) return;
  const int ni = n / (NJ + 1), nj = n - ni * (NJ + 1);
  if (ni < 1 || nj < 1 || ni > NI - 1 || nj > NJ - 1) return;
  const size_t base = (size_t)cs * NV * nc;
  auto phi_at = [&](int i, int j, T* out) {
    T q[NV];
    load_q<T, NV>(Q, dQ, base, nc, i * NJ + j, q);
    out[0] = q[1]; out[1] = q[2]; out[2] = temperature(q, gamma);
    if constexpr (NPHI > 3) out[3] = q[NV - 1];
  };
  T p00[4], p10[4], p01[4], p11[4];
  phi_at(ni - 1, nj - 1, p00); phi_at(ni, nj - 1, p10); phi_at(ni - 1, nj, p01); phi_at(ni, nj, p11);
  const size_t o = (size_t)cs * NPHI * nn + n;
  for (int k = 0; k < NPHI; ++k)
    Io<T>::st(Nd, dNd, o + (size_t)k * nn, 0.25 * (((p00[k] + p10[k]) + p01[k]) + p11[k]));
}

// The diamond of every face (gradients.py:31-95): d1 = the centre difference
// of the face's two cells, d2 = the difference of its two nodes, each node
// the mean of the four edge-padded centres around it, ((c00 + c10) + c01) +
// c11.  Pure geometry: computed once per grid (per case with per-case
// geometry) into out[4][nfi + nfj], (d1x, d1y, d2x, d2y); zero on the faces
// the face kernel does not evaluate.  The geometry of case cs.
__global__ void __launch_bounds__(BLOCK) k_diamond(Plan P, int cs, double* __restrict__ out) {
  bind_case(P, cs);
  const int f = blockIdx.x * BLOCK + threadIdx.x;
  const int nft = P.nfi + P.nfj;
  if (f >= nft) return;
  const int NJ = P.NJ;
  const bool ifam = f < P.nfi;
  int a, b, ff;
  if (ifam) {
    ff = f;
    a = f / NJ; b = f - a * NJ;
  } else {
    ff = f - P.nfi;
    a = ff / (NJ + 1); b = ff - a * (NJ + 1);
  }
  int cL, cR;
  bool inside;
  if (ifam) {
    inside = a >= 1 && a <= P.NI - 1;
    cR = a * NJ + b; cL = cR - NJ;
  } else {
    inside = b >= 1 && b <= NJ - 1;
    cR = a * NJ + b; cL = cR - 1;
  }
  double d[4] = {0.0, 0.0, 0.0, 0.0};
  if (inside && (P.interior[cL] || P.interior[cR])) {
    const int NJ2 = NJ + 2;
    const int iL = cL / NJ, jL = cL - iL * NJ, iR = cR / NJ, jR = cR - iR * NJ;
    const size_t ycen = (size_t)(P.NI + 2) * NJ2;
    d[0] = P.cen[(iR + 1) * NJ2 + jR + 1] - P.cen[(iL + 1) * NJ2 + jL + 1];
    d[1] = P.cen[ycen + (iR + 1) * NJ2 + jR + 1] - P.cen[ycen + (iL + 1) * NJ2 + jL + 1];
    // nodes n0, n1 of the face: the lower-left corners (node index = canvas
    // cell index) of cell R and of cell R shifted one step along the face
    const int n0i = iR, n0j = jR;
    const int n1i = ifam ? iR : iR + 1, n1j = ifam ? jR + 1 : jR;
    auto node_c = [&](int ni, int nj, int row) -> double {
      const double* c = P.cen + (size_t)row * ycen;
      return 0.25 * (((c[ni * NJ2 + nj] + c[(ni + 1) * NJ2 + nj]) + c[ni * NJ2 + nj + 1]) + c[(ni + 1) * NJ2 + nj + 1]);
    };
    d[2] = node_c(n1i, n1j, 0) - node_c(n0i, n0j, 0);
    d[3] = node_c(n1i, n1j, 1) - node_c(n0i, n0j, 1);
  }
  for (int k = 0; k < 4; ++k) out[(size_t)k * nft + f] = d[k];
}

// PT (with T = Dual): mu_inf carries its tangent along a flow parameter.
template <class T, int NV, bool VISC, bool PT = false>
__global__ void __launch_bounds__(BLOCK, 2) k_faces(Plan P, Physics ph, int first_order, const Ctl* __restrict__ ctl,
                                                 const double* __restrict__ Q, const double* __restrict__ dQ,
                                                 const double* __restrict__ Nd, const double* __restrict__ dNd,
                                                 double* __restrict__ Fi, double* __restrict__ Fj, int check_active) {
  bind_case(P, blockIdx.y);
  int f = blockIdx.x * BLOCK + threadIdx.x;
  int cs = blockIdx.y;
  if (skip_case(ctl, cs, check_active)) return;
  if (f >= P.nfi + P.nfj) return;
  const int NJ = P.NJ, nc = P.nc;
  const bool ifam = f < P.nfi;
  int a, b, dL, ff;   // (a, b): the face's own index pair; dL: canvas step across the face
  double sx, sy;
  if (ifam) {
    ff = f;
    a = f / NJ; b = f - a * NJ;           // face a between cells a-1, a (row), column b
    dL = NJ;
    sx = P.fi[ff]; sy = P.fi[P.nfi + ff];
  } else {
    ff = f - P.nfi;
    a = ff / (NJ + 1); b = ff - a * (NJ + 1);   // row a, face b between columns b-1, b
    dL = 1;
    sx = P.fj[ff]; sy = P.fj[P.nfj + ff];
  }
  double* F = ifam ? Fi : Fj;
  const int nfam = ifam ? P.nfi : P.nfj;
  const size_t fbase = (size_t)cs * NV * nfam + ff;
  // Cells either side.
  int cL, cR;
  bool inside;
  if (ifam) {
    inside = a >= 1 && a <= P.NI - 1;
    cR = a * NJ + b; cL = cR - NJ;
  } else {
    inside = b >= 1 && b <= NJ - 1;
    cR = a * NJ + b; cL = cR - 1;
  }
  if (!inside || !(P.interior[cL] || P.interior[cR])) {
    for (int k = 0; k < NV; ++k) F[fbase + (size_t)k * nfam] = 0.0;
    return;
  }
  const size_t base = (size_t)cs * NV * nc;
  T qm[NV], qp[NV], qmm[NV], qpp[NV];
  load_q<T, NV>(Q, dQ, base, nc, cL, qm);
  load_q<T, NV>(Q, dQ, base, nc, cR, qp);
  if (!first_order) {
    load_q<T, NV>(Q, dQ, base, nc, cL - dL, qmm);
    load_q<T, NV>(Q, dQ, base, nc, cR + dL, qpp);
  }
  T Fl[NV];
  if (VISC) {
    // phi = (u, v, T[, nu_hat]) at the face's two nodes (k_nodes): node
    // (ni, nj) averages canvas cells (ni-1, nj-1), (ni, nj-1), (ni-1, nj), (ni, nj).
    constexpr int NPHI = NV > 4 ? 4 : 3;
    const int iR = cR / NJ, jR = cR - iR * NJ;
    const int n1i = ifam ? iR : iR + 1, n1j = ifam ? jR + 1 : jR;
    const int nn = (P.NI + 1) * (NJ + 1);
    // the diamond: the cells' centre difference d1 and the face's node
    // difference d2, precomputed per grid (k_diamond)
    const int nft = P.nfi + P.nfj;
    auto diamond = [&](T* pn0, T* pn1, double* dg) {
      for (int k = 0; k < 4; ++k) dg[k] = P.dg[(size_t)k * nft + f];
      const size_t o0 = (size_t)cs * NPHI * nn + (size_t)iR * (NJ + 1) + jR;
      const size_t o1 = (size_t)cs * NPHI * nn + (size_t)n1i * (NJ + 1) + n1j;
      for (int k = 0; k < NPHI; ++k) pn0[k] = Io<T>::ld(Nd, dNd, o0 + (size_t)k * nn);
      for (int k = 0; k < NPHI; ++k) pn1[k] = Io<T>::ld(Nd, dNd, o1 + (size_t)k * nn);
    };
    face_flux<T, NV, true>(ph, first_order, qmm, qm, qp, qpp, sx, sy, diamond,
                           free_val<PT>(ctl[cs].mu_inf, P.dfree, cs, 6), Fl);
  } else {
    auto none = [](T*, T*, double*) {};
    face_flux<T, NV, false>(ph, first_order, qmm, qm, qp, qpp, sx, sy, none, 0.0, Fl);
  }
  for (int k = 0; k < NV; ++k) {
    if (sizeof(T) == sizeof(double)) F[fbase + (size_t)k * nfam] = val(Fl[k]);
    else F[fbase + (size_t)k * nfam] = tan_of(Fl[k]);
  }
}

// Per cell: R = div F (- V * SA source) at interior cells, U - U0 elsewhere.
// Tangent mode writes  out = vdt * dU + dR  (vdt may be null: no pseudo-time term).
// Partials of the 2D SA source at the state, once per residual, for the
// tangents of the Krylov steps: d src / d(rho, p, nu_hat), d src / d(dv/dx -
// du/dy) (through |.|, the dual-number abs's sign), d src / d(grad nu_hat).
constexpr int SALIN2 = 6;

__global__ void __launch_bounds__(BLOCK) k_sa_lin(Plan P, Physics ph, const Ctl* __restrict__ ctl,
                                                  const double* __restrict__ Q, double* __restrict__ salin,
                                                  int check_active) {
  bind_case(P, blockIdx.y);
  const int c = blockIdx.x * BLOCK + threadIdx.x;
  const int cs = blockIdx.y;
  if (c >= P.nc || skip_case(ctl, cs, check_active)) return;
  const int NJ = P.NJ, nc = P.nc;
  double* L = salin + (size_t)cs * SALIN2 * nc + c;
  if (!P.interior[c]) {
    for (int k = 0; k < SALIN2; ++k) L[(size_t)k * nc] = 0.0;
    return;
  }
  const size_t base = (size_t)cs * 5 * nc;
  const int i = c / NJ, j = c - i * NJ;
  const int fi0 = i * NJ + j, fi1 = fi0 + NJ;
  const int fj0 = i * (NJ + 1) + j, fj1 = fj0 + 1;
  const double gamma = ph.gamma, nu_inf = ctl[cs].mu_inf;
  double qc[5], qw[5], qe[5], qs[5], qn[5];
  load_q<double, 5>(Q, nullptr, base, nc, c, qc);
  load_q<double, 5>(Q, nullptr, base, nc, c - NJ, qw);
  load_q<double, 5>(Q, nullptr, base, nc, c + NJ, qe);
  load_q<double, 5>(Q, nullptr, base, nc, c - 1, qs);
  load_q<double, 5>(Q, nullptr, base, nc, c + 1, qn);
  const double sxi0 = P.fi[fi0], syi0 = P.fi[P.nfi + fi0], sxi1 = P.fi[fi1], syi1 = P.fi[P.nfi + fi1];
  const double sxj0 = P.fj[fj0], syj0 = P.fj[P.nfj + fj0], sxj1 = P.fj[fj1], syj1 = P.fj[P.nfj + fj1];
  const double V = P.area[c];
  auto gg = [&](int k, double& gxo, double& gyo) {
    const double fw = 0.5 * (qw[k] + qc[k]), fe = 0.5 * (qc[k] + qe[k]);
    const double fs = 0.5 * (qs[k] + qc[k]), fn = 0.5 * (qc[k] + qn[k]);
    gxo = ((fe * sxi1 - fw * sxi0) + (fn * sxj1 - fs * sxj0)) / V;
    gyo = ((fe * syi1 - fw * syi0) + (fn * syj1 - fs * syj0)) / V;
  };
  double gux, guy, gvx, gvy, gnx, gny;
  gg(1, gux, guy);
  gg(2, gvx, gvy);
  gg(4, gnx, gny);
  const double w = gvx - guy;
  const double omega = fabs(w);
  double part[4];
  for (int sd = 0; sd < 4; ++sd) {
    Dual q[5];
    for (int v = 0; v < 5; ++v) q[v] = mkd(qc[v], 0.0);
    if (sd == 0) q[0].d = 1.0;
    if (sd == 1) q[3].d = 1.0;
    if (sd == 2) q[4].d = 1.0;
    const Dual om = mkd(omega, sd == 3 ? 1.0 : 0.0);
    const Dual rho = q[0];
    const Dual mu = viscosity<Dual>(temperature(q, gamma), nu_inf, ph.sutherland);
    Dual prod, destr;
    sa_source<Dual>(q[4] * nu_inf, mu / rho, om, P.dist[c], prod, destr);
    const Dual src = rho * ((prod - destr) / nu_inf + CB2 / SIGMA * nu_inf * (gnx * gnx + gny * gny));
    part[sd] = src.d;
  }
  L[0] = part[0];
  L[(size_t)nc] = part[1];
  L[(size_t)2 * nc] = part[2];
  L[(size_t)3 * nc] = w >= 0.0 ? part[3] : -part[3];     // d|w| / dw
  const double rg = qc[0] * (CB2 / SIGMA * nu_inf);
  L[(size_t)4 * nc] = rg * (2.0 * gnx);
  L[(size_t)5 * nc] = rg * (2.0 * gny);
}

// PT (with T = Dual): the freestream and mu_inf carry their tangent along a
// flow parameter -- U0 moves with it, and the SA source is differentiated in
// full (its partials at the state hold dU only).
template <class T, int NV, bool PT = false>
__global__ void __launch_bounds__(BLOCK, 3) k_cells(Plan P, Physics ph, const Ctl* __restrict__ ctl,
                                                 const double* __restrict__ U, const double* __restrict__ dU,
                                                 const double* __restrict__ Q, const double* __restrict__ dQ,
                                                 const double* __restrict__ Fi, const double* __restrict__ Fj,
                                                 const double* __restrict__ vdt, const double* __restrict__ salin,
                                                 double* __restrict__ out, int check_active) {
  bind_case(P, blockIdx.y);
  int c = blockIdx.x * BLOCK + threadIdx.x;
  int cs = blockIdx.y;
  if (skip_case(ctl, cs, check_active)) return;
  if (c >= P.nc) return;
  const int NJ = P.NJ, nc = P.nc;
  const size_t base = (size_t)cs * NV * nc;
  constexpr bool tangent = sizeof(T) != sizeof(double);
  double vd = vdt ? vdt[(size_t)cs * nc + c] : 0.0;
  if (!P.interior[c]) {
    if (!tangent) {
      double u0[5];
      conservative<double, NV>(ctl[cs].q_inf, ph.gamma, u0);
      for (int k = 0; k < NV; ++k) out[base + (size_t)k * nc + c] = U[base + (size_t)k * nc + c] - u0[k];
    } else if constexpr (PT) {
      Dual qf[6], u0[5];
      for (int k = 0; k < 6; ++k) qf[k] = free_val<true>(ctl[cs].q_inf[k], P.dfree, cs, k);
      conservative<Dual, NV>(qf, ph.gamma, u0);
      for (int k = 0; k < NV; ++k) out[base + (size_t)k * nc + c] = dU[base + (size_t)k * nc + c] - u0[k].d;
    } else {
      for (int k = 0; k < NV; ++k) {
        double d = dU[base + (size_t)k * nc + c];
        out[base + (size_t)k * nc + c] = vdt ? vd * d + d : d;
      }
    }
    return;
  }
  int i = c / NJ, j = c - i * NJ;
  const size_t bi = (size_t)cs * NV * P.nfi, bj = (size_t)cs * NV * P.nfj;
  const int fi0 = i * NJ + j, fi1 = fi0 + NJ;
  const int fj0 = i * (NJ + 1) + j, fj1 = fj0 + 1;
  double r[NV];
  for (int k = 0; k < NV; ++k) {
    r[k] = (Fi[bi + (size_t)k * P.nfi + fi1] - Fi[bi + (size_t)k * P.nfi + fi0]) +
           (Fj[bj + (size_t)k * P.nfj + fj1] - Fj[bj + (size_t)k * P.nfj + fj0]);
  }
  if constexpr (NV > 4 && tangent && !PT) {
    // the source's tangent from its partials at the state (k_sa_lin): the
    // face-average gradients are linear in dq, the rest is a dot product
    const double* dq = dQ + base;
    auto d_at = [&](int cell, int k) { return dq[(size_t)k * nc + cell]; };
    const double sxi0 = P.fi[fi0], syi0 = P.fi[P.nfi + fi0], sxi1 = P.fi[fi1], syi1 = P.fi[P.nfi + fi1];
    const double sxj0 = P.fj[fj0], syj0 = P.fj[P.nfj + fj0], sxj1 = P.fj[fj1], syj1 = P.fj[P.nfj + fj1];
    const double V = P.area[c];
    auto gg = [&](int k, double& gxo, double& gyo) {
      const double qc = d_at(c, k);
      const double fw = 0.5 * (d_at(c - NJ, k) + qc), fe = 0.5 * (qc + d_at(c + NJ, k));
      const double fs = 0.5 * (d_at(c - 1, k) + qc), fn = 0.5 * (qc + d_at(c + 1, k));
      gxo = ((fe * sxi1 - fw * sxi0) + (fn * sxj1 - fs * sxj0)) / V;
      gyo = ((fe * syi1 - fw * syi0) + (fn * syj1 - fs * syj0)) / V;
    };
    double gux, guy, gvx, gvy, gnx, gny;
    gg(1, gux, guy);
    gg(2, gvx, gvy);
    gg(4, gnx, gny);
    const double* L = salin + (size_t)cs * SALIN2 * nc + c;
    double ds = (L[0] * d_at(c, 0) + L[(size_t)nc] * d_at(c, 3)) + L[(size_t)2 * nc] * d_at(c, 4);
    ds += L[(size_t)3 * nc] * (gvx - guy);
    ds += L[(size_t)4 * nc] * gnx + L[(size_t)5 * nc] * gny;
    r[4] = r[4] - ds * V;
  } else if constexpr (NV > 4) {
    // SA source at the cell (turbulence.py:140-151): vorticity and grad nu_hat by
    // Green-Gauss over the face-averaged neighbours.
    const double gamma = ph.gamma;
    const auto nu_inf = free_val<PT>(ctl[cs].mu_inf, P.dfree, cs, 6);
    T qc[NV], qw[NV], qe[NV], qs[NV], qn[NV];
    load_q<T, NV>(Q, dQ, base, nc, c, qc);
    load_q<T, NV>(Q, dQ, base, nc, c - NJ, qw);
    load_q<T, NV>(Q, dQ, base, nc, c + NJ, qe);
    load_q<T, NV>(Q, dQ, base, nc, c - 1, qs);
    load_q<T, NV>(Q, dQ, base, nc, c + 1, qn);
    const double sxi0 = P.fi[fi0], syi0 = P.fi[P.nfi + fi0], sxi1 = P.fi[fi1], syi1 = P.fi[P.nfi + fi1];
    const double sxj0 = P.fj[fj0], syj0 = P.fj[P.nfj + fj0], sxj1 = P.fj[fj1], syj1 = P.fj[P.nfj + fj1];
    const double V = P.area[c];
    auto gg = [&](int k, T& gxo, T& gyo) {
      T fw = 0.5 * (qw[k] + qc[k]), fe = 0.5 * (qc[k] + qe[k]);
      T fs = 0.5 * (qs[k] + qc[k]), fn = 0.5 * (qc[k] + qn[k]);
      gxo = ((fe * sxi1 - fw * sxi0) + (fn * sxj1 - fs * sxj0)) / V;
      gyo = ((fe * syi1 - fw * syi0) + (fn * syj1 - fs * syj0)) / V;
    };
    T gux, guy, gvx, gvy, gnx, gny;
    gg(1, gux, guy);
    gg(2, gvx, gvy);
    gg(4, gnx, gny);
    T omega = jabs(gvx - guy);
    T rho = qc[0];
    T mu = viscosity<T>(temperature(qc, gamma), nu_inf, ph.sutherland);
    T prod, destr;
    sa_source<T>(qc[4] * nu_inf, mu / rho, omega, P.dist[c], prod, destr);
    T src = rho * ((prod - destr) / nu_inf + CB2 / SIGMA * nu_inf * (gnx * gnx + gny * gny));
    T sv = src * V;
    r[4] = r[4] - (tangent ? tan_of(sv) : val(sv));
  }
  for (int k = 0; k < NV; ++k) {
    if (tangent && vdt) r[k] = vd * dU[base + (size_t)k * nc + c] + r[k];
    out[base + (size_t)k * nc + c] = r[k];
  }
}


// The first-order tangent of the Krylov operator in one pass over tiles of
// TI x TJ cells (out = vdt dU + dR~/dU dU, the node, face and cell kernels
// fused): a workgroup stages its tile's cells with a one-cell halo --
// primitives Q and their tangent dQ, complete on ghost cells (the line
// apply wrote them) -- in LDS, forms the tile's node values from them,
// computes every face of the tile from LDS with face_flux (the arithmetic of
// k_faces), keeps the face fluxes' tangent in LDS and finishes each cell's
// divergence and SA source there (the arithmetic of k_cells).  Each cell's
// primitives are read from HBM once per tile instead of once per face and
// node, and no face flux goes through HBM.  Measured on the 32-case polar
// it takes 460-470 us per launch against 400 us for the node, face and cell
// kernels together (the face arithmetic, not HBM, bounds both), so it is
// off by default (jr_set_tile_tangent).
constexpr int TT_I = 8, TT_J = 32, TT_THREADS = 256;

template <int NV, bool VISC>
__global__ void __launch_bounds__(TT_THREADS, 2) k_tangent_tile(Plan P, Physics ph, const Ctl* __restrict__ ctl,
                                                                const double* __restrict__ dU,
                                                                const double* __restrict__ Q,
                                                                const double* __restrict__ dQ,
                                                                const double* __restrict__ vdt,
                                                                const double* __restrict__ salin,
                                                                double* __restrict__ out, int check_active) {
  constexpr int HI = TT_I + 2, HJ = TT_J + 2;           // cells with the halo
  constexpr int NI1 = TT_I + 1, NJ1 = TT_J + 1;         // nodes
  constexpr int NPHI = NV > 4 ? 4 : 3;
  __shared__ double sq[NV][HI][HJ], sd[NV][HI][HJ];
  __shared__ double snv[VISC ? NPHI : 1][VISC ? NI1 : 1][VISC ? NJ1 : 1];
  __shared__ double snd[VISC ? NPHI : 1][VISC ? NI1 : 1][VISC ? NJ1 : 1];
  __shared__ double sfi[NV][NI1][TT_J], sfj[NV][TT_I][NJ1];
  bind_case(P, blockIdx.y);
  const int cs = blockIdx.y;
  if (skip_case(ctl, cs, check_active)) return;
  const int NI = P.NI, NJ = P.NJ, nc = P.nc;
  const int tiles_j = (NJ + TT_J - 1) / TT_J;
  const int i0 = (int)(blockIdx.x / tiles_j) * TT_I, j0 = (int)(blockIdx.x % tiles_j) * TT_J;
  const size_t base = (size_t)cs * NV * nc;
  const int tid = threadIdx.x;
  // 1. the tile's cells and halo
  for (int e = tid; e < HI * HJ; e += TT_THREADS) {
    const int li = e / HJ, lj = e - li * HJ;
    const int gi = i0 + li - 1, gj = j0 + lj - 1;
    const bool on = gi >= 0 && gi < NI && gj >= 0 && gj < NJ;
    const int c = on ? gi * NJ + gj : 0;
    for (int k = 0; k < NV; ++k) {
      sq[k][li][lj] = on ? Q[base + (size_t)k * nc + c] : 0.0;
      sd[k][li][lj] = on ? dQ[base + (size_t)k * nc + c] : 0.0;
    }
  }
  __syncthreads();
  auto cell = [&](int li, int lj, Dual* q) {
    for (int k = 0; k < NV; ++k) q[k] = mkd(sq[k][li][lj], sd[k][li][lj]);
  };
  // 2. node values phi = (u, v, T[, nu_hat]) and their tangent (k_nodes):
  // node (ni, nj) averages cells (ni-1..ni, nj-1..nj); nodes on the canvas
  // edge are never read
  if constexpr (VISC) {
    const double gamma = ph.gamma;
    for (int e = tid; e < NI1 * NJ1; e += TT_THREADS) {
      const int ln = e / NJ1, mn = e - ln * NJ1;
      const int ni = i0 + ln, nj = j0 + mn;
      if (ni < 1 || nj < 1 || ni > NI - 1 || nj > NJ - 1) continue;
      Dual p[4][4];
      const int ci[4] = {ln, ln + 1, ln, ln + 1}, cj[4] = {mn, mn, mn + 1, mn + 1};
      for (int t = 0; t < 4; ++t) {
        Dual q[NV];
        cell(ci[t], cj[t], q);
        p[t][0] = q[1]; p[t][1] = q[2]; p[t][2] = temperature(q, gamma);
        if constexpr (NPHI > 3) p[t][3] = q[NV - 1];
      }
      for (int k = 0; k < NPHI; ++k) {
        const Dual v = 0.25 * (((p[0][k] + p[1][k]) + p[2][k]) + p[3][k]);
        snv[k][ln][mn] = v.v;
        snd[k][ln][mn] = v.d;
      }
    }
    __syncthreads();
  }
  // 3. the tile's faces: i-faces (a, b) between cells (a-1, b), (a, b) for
  // a in i0..i0+TI; j-faces (a, b) between (a, b-1), (a, b) for b in j0..j0+TJ
  const int nft = P.nfi + P.nfj;
  auto face = [&](bool ifam, int la, int lb) {
    const int a = i0 + la, b = j0 + lb;
    double* F = ifam ? &sfi[0][0][0] : &sfj[0][0][0];
    const int fstride = ifam ? NI1 * TT_J : TT_I * NJ1;
    const int fo = ifam ? la * TT_J + lb : la * NJ1 + lb;
    int f, cR;
    bool inside;
    double sx, sy;
    if (ifam) {
      if (a > NI || b >= NJ) return;
      f = a * NJ + b;
      inside = a >= 1 && a <= NI - 1;
      cR = f;
      sx = P.fi[f]; sy = P.fi[P.nfi + f];
    } else {
      if (a >= NI || b > NJ) return;
      const int ff = a * (NJ + 1) + b;
      f = P.nfi + ff;
      inside = b >= 1 && b <= NJ - 1;
      cR = a * NJ + b;
      sx = P.fj[ff]; sy = P.fj[P.nfj + ff];
    }
    const int cL = cR - (ifam ? NJ : 1);
    if (!inside || !(P.interior[cL] || P.interior[cR])) {
      for (int k = 0; k < NV; ++k) F[k * fstride + fo] = 0.0;
      return;
    }
    // the cells either side in LDS: (a - 1, b) | (a, b - 1) and (a, b)
    Dual qm[NV], qp[NV];
    cell(ifam ? la : la + 1, ifam ? lb + 1 : lb, qm);
    cell(la + 1, lb + 1, qp);
    Dual Fl[NV];
    if constexpr (VISC) {
      // nodes (a, b) and (a, b + 1) | (a + 1, b)
      const int m1 = ifam ? la : la + 1, n1 = ifam ? lb + 1 : lb;
      auto diamond = [&](Dual* pn0, Dual* pn1, double* dg) {
        for (int k = 0; k < 4; ++k) dg[k] = P.dg[(size_t)k * nft + f];
        for (int k = 0; k < NPHI; ++k) {
          pn0[k] = mkd(snv[k][la][lb], snd[k][la][lb]);
          pn1[k] = mkd(snv[k][m1][n1], snd[k][m1][n1]);
        }
      };
      face_flux<Dual, NV, true>(ph, 1, qm, qm, qp, qp, sx, sy, diamond, ctl[cs].mu_inf, Fl);
    } else {
      auto none = [](Dual*, Dual*, double*) {};
      face_flux<Dual, NV, false>(ph, 1, qm, qm, qp, qp, sx, sy, none, 0.0, Fl);
    }
    for (int k = 0; k < NV; ++k) F[k * fstride + fo] = Fl[k].d;
  };
  constexpr int NFI = NI1 * TT_J, NFJ = TT_I * NJ1;
  for (int e = tid; e < NFI + NFJ; e += TT_THREADS) {
    if (e < NFI) face(true, e / TT_J, e % TT_J);
    else face(false, (e - NFI) / NJ1, (e - NFI) % NJ1);
  }
  __syncthreads();
  // 4. the tile's cells (k_cells): divergence of the face tangents, the SA
  // source's tangent from its partials, the pseudo-time term
  for (int e = tid; e < TT_I * TT_J; e += TT_THREADS) {
    const int li = e / TT_J, lj = e - li * TT_J;
    const int i = i0 + li, j = j0 + lj;
    if (i >= NI || j >= NJ) continue;
    const int c = i * NJ + j;
    const double vd = vdt ? vdt[(size_t)cs * nc + c] : 0.0;
    if (!P.interior[c]) {
      for (int k = 0; k < NV; ++k) {
        const double d = dU[base + (size_t)k * nc + c];
        out[base + (size_t)k * nc + c] = vdt ? vd * d + d : d;
      }
      continue;
    }
    double r[NV];
    for (int k = 0; k < NV; ++k) r[k] = (sfi[k][li + 1][lj] - sfi[k][li][lj]) + (sfj[k][li][lj + 1] - sfj[k][li][lj]);
    if constexpr (NV > 4) {
      const int fi0 = i * NJ + j, fi1 = fi0 + NJ;
      const int fj0 = i * (NJ + 1) + j, fj1 = fj0 + 1;
      const int ci = li + 1, cj = lj + 1;
      const double sxi0 = P.fi[fi0], syi0 = P.fi[P.nfi + fi0], sxi1 = P.fi[fi1], syi1 = P.fi[P.nfi + fi1];
      const double sxj0 = P.fj[fj0], syj0 = P.fj[P.nfj + fj0], sxj1 = P.fj[fj1], syj1 = P.fj[P.nfj + fj1];
      const double V = P.area[c];
      auto gg = [&](int k, double& gxo, double& gyo) {
        const double qc = sd[k][ci][cj];
        const double fw = 0.5 * (sd[k][ci - 1][cj] + qc), fe = 0.5 * (qc + sd[k][ci + 1][cj]);
        const double fs = 0.5 * (sd[k][ci][cj - 1] + qc), fn = 0.5 * (qc + sd[k][ci][cj + 1]);
        gxo = ((fe * sxi1 - fw * sxi0) + (fn * sxj1 - fs * sxj0)) / V;
        gyo = ((fe * syi1 - fw * syi0) + (fn * syj1 - fs * syj0)) / V;
      };
      double gux, guy, gvx, gvy, gnx, gny;
      gg(1, gux, guy);
      gg(2, gvx, gvy);
      gg(4, gnx, gny);
      const double* L = salin + (size_t)cs * SALIN2 * nc + c;
      double ds = (L[0] * sd[0][ci][cj] + L[(size_t)nc] * sd[3][ci][cj]) + L[(size_t)2 * nc] * sd[4][ci][cj];
      ds += L[(size_t)3 * nc] * (gvx - guy);
      ds += L[(size_t)4 * nc] * gnx + L[(size_t)5 * nc] * gny;
      r[4] = r[4] - ds * V;
    }
    for (int k = 0; k < NV; ++k) {
      if (vdt) r[k] = vd * dU[base + (size_t)k * nc + c] + r[k];
      out[base + (size_t)k * nc + c] = r[k];
    }
  }
}

}  // namespace jr
