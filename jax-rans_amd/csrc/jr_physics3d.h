// Pointwise physics of the 3D path, templated on the scalar type (double or
// Dual): primitive rows (rho, u, v, w, p[, nu_hat]).  The d-component sums
// run in index order as src/jaxrans/core/state.py:66-78 (dot; magnitude is
// sqrt(dot) in 3D).  Scalar laws (limiters, viscosity, SA) are shared with
// the 2D path (jr_physics.h).
#pragma once
#include "jr_physics.h"

namespace jr {

template <class A, class B>
__device__ __forceinline__ auto dot3(const A* a, const B* b) -> decltype(a[0] * b[0]) {
  return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

// conservative -> primitive (src/jaxrans/core/state.py:88-93)
template <class T, int NV>
__device__ __forceinline__ void primitive3(const T* U, double gamma, T* q) {
  T rho = U[0];
  T u = U[1] / rho, v = U[2] / rho, w = U[3] / rho;
  q[0] = rho; q[1] = u; q[2] = v; q[3] = w;
  q[4] = (gamma - 1.0) * (U[4] - 0.5 * rho * ((u * u + v * v) + w * w));
  if constexpr (NV > 5) q[5] = U[5] / rho;
}

template <class T, int NV>
__device__ __forceinline__ void conservative3(const T* q, double gamma, T* U) {
  T rho = q[0];
  U[0] = rho; U[1] = rho * q[1]; U[2] = rho * q[2]; U[3] = rho * q[3];
  U[4] = q[4] / (gamma - 1.0) + 0.5 * rho * ((q[1] * q[1] + q[2] * q[2]) + q[3] * q[3]);
  if constexpr (NV > 5) U[5] = rho * q[5];
}

// src/jaxrans/physics/euler.py:43-52
template <class T>
__device__ __forceinline__ void normal_flux3(const T* q, const double* n, double gamma, T* f) {
  T rho = q[0], p = q[4];
  T vn = (q[1] * n[0] + q[2] * n[1]) + q[3] * n[2];
  T e = p / (gamma - 1.0) + 0.5 * rho * ((q[1] * q[1] + q[2] * q[2]) + q[3] * q[3]);
  f[0] = rho * vn;
  f[1] = rho * q[1] * vn + p * n[0];
  f[2] = rho * q[2] * vn + p * n[1];
  f[3] = rho * q[3] * vn + p * n[2];
  f[4] = (e + p) * vn;
}

// src/jaxrans/physics/euler.py:74-94
template <class T>
__device__ __forceinline__ void roe_average3(const T* qL, const T* qR, double gamma, T& rho, T* vel, T& h, T& a) {
  T hl = gamma / (gamma - 1.0) * qL[4] / qL[0] + 0.5 * ((qL[1] * qL[1] + qL[2] * qL[2]) + qL[3] * qL[3]);
  T hr = gamma / (gamma - 1.0) * qR[4] / qR[0] + 0.5 * ((qR[1] * qR[1] + qR[2] * qR[2]) + qR[3] * qR[3]);
  T w = jsqrt(qR[0] / qL[0]);
  T den = 1.0 + w;
  rho = jsqrt(qL[0] * qR[0]);
  for (int c = 0; c < 3; ++c) vel[c] = (qL[1 + c] + w * qR[1 + c]) / den;
  h = (hl + w * hr) / den;
  T a2 = (gamma - 1.0) * (h - 0.5 * dot3(vel, vel));
  a = jsqrt(jmax(a2, lift<T>(1e-30)));
}

__device__ __forceinline__ double unit_normal3(const double* s, double* n) {
  double len = sqrt((s[0] * s[0] + s[1] * s[1]) + s[2] * s[2]);
  double safe = len > 1e-300 ? len : 1e-300;
  n[0] = s[0] / safe; n[1] = s[1] / safe; n[2] = s[2] / safe;
  return len;
}

// HLLC with Einfeldt speeds, times |S| (src/jaxrans/numerics/riemann.py:116-198)
template <class T>
__device__ void hllc3(const T* qL, const T* qR, const double* s, double gamma, T* F) {
  double n[3];
  const double len = unit_normal3(s, n);
  T vnl = dot3(qL + 1, n), vnr = dot3(qR + 1, n);
  T al = jsqrt(gamma * qL[4] / qL[0]), ar = jsqrt(gamma * qR[4] / qR[0]);
  T rro, vro[3], hro, aro;
  roe_average3(qL, qR, gamma, rro, vro, hro, aro);
  T vnro = dot3(vro, n);
  T sl = jmin(vnl - al, vnro - aro);
  T sr = jmax(vnr + ar, vnro + aro);
  T ml = qL[0] * (sl - vnl), mr = qR[0] * (sr - vnr);
  T sm = (qR[4] - qL[4] + ml * vnl - mr * vnr) / jmin(ml - mr, lift<T>(-1e-300));
  T f[5];
  if (val(sl) >= 0.0) {
    normal_flux3(qL, n, gamma, f);
  } else if (val(sm) >= 0.0 || val(sr) >= 0.0) {
    const bool left = val(sm) >= 0.0;
    const T* q = left ? qL : qR;
    const T sk = left ? sl : sr;
    const T vn = left ? vnl : vnr;
    T fk[5], uc[5];
    normal_flux3(q, n, gamma, fk);
    conservative3<T, 5>(q, gamma, uc);
    T widen = left ? jmin(sl - sm, lift<T>(-1e-300)) : jmax(sr - sm, lift<T>(1e-300));
    T pstar = q[4] + q[0] * (vn - sk) * (vn - sm);
    T rstar = q[0] * (sk - vn) / widen;
    T estar = ((sk - vn) * uc[4] - q[4] * vn + pstar * sm) / widen;
    T us[5] = {rstar, rstar * (q[1] + (sm - vn) * n[0]), rstar * (q[2] + (sm - vn) * n[1]),
               rstar * (q[3] + (sm - vn) * n[2]), estar};
    for (int k = 0; k < 5; ++k) f[k] = fk[k] + sk * (us[k] - uc[k]);
  } else {
    normal_flux3(qR, n, gamma, f);
  }
  for (int k = 0; k < 5; ++k) F[k] = len * f[k];
}

// Roe flux, no entropy fix, times |S| (src/jaxrans/numerics/riemann.py:54-113)
template <class T>
__device__ void roe3(const T* qL, const T* qR, const double* s, double gamma, T* F) {
  double n[3];
  const double len = unit_normal3(s, n);
  T rho, vel[3], h, a;
  roe_average3(qL, qR, gamma, rho, vel, h, a);
  T vn = dot3(vel, n);
  T drho = qR[0] - qL[0], dp = qR[4] - qL[4];
  T dv[3] = {qR[1] - qL[1], qR[2] - qL[2], qR[3] - qL[3]};
  T dvn = dot3(dv, n);
  T dvt[3] = {dv[0] - dvn * n[0], dv[1] - dvn * n[1], dv[2] - dvn * n[2]};
  T a2 = a * a;
  T am = (dp - rho * a * dvn) / (2.0 * a2);
  T ae = drho - dp / a2;
  T ap = (dp + rho * a * dvn) / (2.0 * a2);
  T lm = jabs(vn - a), lmid = jabs(vn), lp = jabs(vn + a);
  T kin = 0.5 * dot3(vel, vel);
  T one = lift<T>(1.0);
  T rm[5] = {one, vel[0] - a * n[0], vel[1] - a * n[1], vel[2] - a * n[2], h - a * vn};
  T re[5] = {one, vel[0], vel[1], vel[2], kin};
  T sh[5] = {lift<T>(0.0), rho * dvt[0], rho * dvt[1], rho * dvt[2], rho * dot3(vel, dvt)};
  T rp[5] = {one, vel[0] + a * n[0], vel[1] + a * n[1], vel[2] + a * n[2], h + a * vn};
  T fl[5], fr[5];
  normal_flux3(qL, n, gamma, fl);
  normal_flux3(qR, n, gamma, fr);
  for (int k = 0; k < 5; ++k) {
    T diss = lm * am * rm[k] + lmid * ae * re[k] + lmid * sh[k] + lp * ap * rp[k];
    T central = 0.5 * (fl[k] + fr[k]);
    F[k] = len * (central - 0.5 * diss);
  }
}

// Ghost primitives from the source cell (src/jaxrans/numerics/boundary.py:54-157);
// n the outward unit normal, qf the far-field reference primitives.
// qf plain (TQ = double) or carrying a parameter tangent (TQ = T = Dual).
template <class T, int NV, class TQ>
__device__ void bc_state3(int kind, const T* q, const double* n, const TQ* qf, double gamma, T* g) {
  if (kind == K_SLIP || kind == K_SYMMETRY) {
    T vn = dot3(q + 1, n);
    g[0] = q[0];
    for (int c = 0; c < 3; ++c) g[1 + c] = q[1 + c] - 2.0 * vn * n[c];
    g[4] = q[4];
    if constexpr (NV > 5) g[5] = kind == K_SLIP ? -q[5] : q[5];
  } else if (kind == K_NOSLIP) {
    g[0] = q[0]; g[1] = -q[1]; g[2] = -q[2]; g[3] = -q[3]; g[4] = q[4];
    if constexpr (NV > 5) g[5] = -q[5];
  } else if (kind == K_FARFIELD) {
    T a_i = jsqrt(gamma * q[4] / q[0]);
    TQ a_f = jsqrt(gamma * qf[4] / qf[0]);
    double two_over = 2.0 / (gamma - 1.0);
    T vn_i = dot3(q + 1, n);
    TQ vn_f = dot3(qf + 1, n);
    T rp = vn_i + two_over * a_i;
    TQ rm = vn_f - two_over * a_f;
    T vn_b = 0.5 * (rp + rm);
    T a_b = 0.25 * (gamma - 1.0) * (rp - rm);
    bool out = val(vn_b) > 0.0;
    T entropy = out ? q[4] / jpow(q[0], gamma) : to_t<T>(qf[4] / jpow(qf[0], gamma));
    T rho_b = jpow(a_b * a_b / (gamma * entropy), 1.0 / (gamma - 1.0));
    g[0] = rho_b;
    for (int c = 0; c < 3; ++c) {
      T t = out ? q[1 + c] - vn_i * n[c] : to_t<T>(qf[1 + c] - vn_f * n[c]);
      g[1 + c] = vn_b * n[c] + t;
    }
    g[4] = rho_b * a_b * a_b / gamma;
    if constexpr (NV > 5) g[5] = out ? q[5] : to_t<T>(qf[5]);
  } else if (kind == K_INFLOW) {
    double g1 = gamma - 1.0;
    TQ t_f = gamma * qf[4] / qf[0];
    TQ speed_f = jsqrt(dot3(qf + 1, qf + 1));
    TQ t_tot = t_f + 0.5 * g1 * speed_f * speed_f;
    TQ p_tot = qf[4] * jpow(t_tot / t_f, gamma / g1);
    T p_b = q[4];
    T mach2 = jmax(2.0 / g1 * (jpow(p_tot / p_b, g1 / gamma) - 1.0), lift<T>(0.0));
    T t_b = t_tot / (1.0 + 0.5 * g1 * mach2);
    T speed = jsqrt(mach2 * t_b);
    g[0] = gamma * p_b / t_b;
    for (int c = 0; c < 3; ++c) g[1 + c] = speed * qf[1 + c] / speed_f;
    g[4] = p_b;
    if constexpr (NV > 5) g[5] = to_t<T>(qf[5]);
  } else if (kind == K_OUTFLOW) {
    for (int k = 0; k < NV; ++k) g[k] = q[k];
    g[4] = to_t<T>(qf[4]);
  } else {
    for (int k = 0; k < NV; ++k) g[k] = q[k];
  }
}

template <class T>
__device__ __forceinline__ T temperature3(const T* q, double gamma) { return gamma * q[4] / q[0]; }

// F_v . S (src/jaxrans/physics/navier_stokes.py:86-116); gv[a][b] = d v_a / d x_b.
template <class T>
__device__ __forceinline__ void viscous_flux3(const T* vel, T gv[3][3], const T* gT, const double* s, T mu, T k,
                                              const T (*extra)[3], T* f) {
  T div = (gv[0][0] + gv[1][1]) + gv[2][2];
  T ttd = (2.0 / 3.0) * div;
  T tau[3][3];
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      T lin;
      if (a == b) lin = mu * (2.0 * gv[a][a] - ttd);
      else {
        const int lo = a < b ? a : b, hi = a < b ? b : a;
        lin = mu * (gv[lo][hi] + gv[hi][lo]);
      }
      tau[a][b] = extra ? lin + extra[a][b] : lin;
    }
  T force[3];
  for (int a = 0; a < 3; ++a) force[a] = dot3(tau[a], s);
  f[0] = lift<T>(0.0);
  f[1] = force[0]; f[2] = force[1]; f[3] = force[2];
  f[4] = dot3(vel, force) + k * dot3(gT, s);
}

// QCR2000 (src/jaxrans/physics/spalart_allmaras.py:197-216)
template <class T>
__device__ __forceinline__ void qcr_stress3(T gv[3][3], T mu_t, double c_cr1, T out[3][3]) {
  T sq = gv[0][0] * gv[0][0];
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      if (a || b) sq = sq + gv[a][b] * gv[a][b];
  T nrm = val(sq) > 0.0 ? jsqrt(sq) : lift<T>(1.0);
  T div = (gv[0][0] + gv[1][1]) + gv[2][2];
  T rot[3][3], tau[3][3];
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      rot[a][b] = (gv[a][b] - gv[b][a]) / nrm;
      tau[a][b] = mu_t * (gv[a][b] + gv[b][a] - (a == b ? (2.0 / 3.0) * div : lift<T>(0.0)));
    }
  T prod[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) prod[i][j] = dot3(rot[i], tau[j]);
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) out[a][b] = -c_cr1 * (prod[a][b] + prod[b][a]);
}

// |curl V| from g[b][a] = d v_a / d x_b, zero tangent at zero
// (src/jaxrans/numerics/turbulence.py:93-103)
template <class T>
__device__ __forceinline__ T vorticity3(T g[3][3]) {
  T curl[3] = {g[1][2] - g[2][1], g[2][0] - g[0][2], g[0][1] - g[1][0]};
  T sq = dot3(curl, curl);
  if (val(sq) > 0.0) return jsqrt(sq);
  return lift<T>(0.0);
}

}  // namespace jr
