// Pointwise physics on the device, templated on the scalar type (double or
// Dual).  2D: primitive rows (rho, u, v, p[, nu_hat]).  Each function names
// the reference function whose arithmetic it restates (paths relative to the
// reference repository).
#pragma once
#include "jr_scalar.h"

namespace jr {

constexpr double PR = 0.72, PR_T = 0.9;
constexpr double CB1 = 0.1355, SIGMA = 2.0 / 3.0, CB2 = 0.622, KAPPA = 0.41;
constexpr double CW2 = 0.3, CW3 = 2.0, CV1 = 7.1, C2 = 0.7, C3 = 0.9, CN1 = 16.0, RMAX = 10.0;
constexpr double CW1 = CB1 / (KAPPA * KAPPA) + (1.0 + CB2) / SIGMA;

enum Flux { FLUX_HLLC = 0, FLUX_ROE = 1 };
enum Limiter { LIM_MINMOD = 0, LIM_NONE = 1, LIM_VAN_ALBADA = 2, LIM_VAN_ALBADA_SMOOTH = 3 };
enum Kind { K_COPY = 0, K_SLIP = 1, K_NOSLIP = 2, K_SYMMETRY = 3, K_FARFIELD = 4, K_INFLOW = 5, K_OUTFLOW = 6 };

// jnp.hypot's formula on constant data (src/jaxrans/core/state.py:129-131).
__device__ __forceinline__ double hypot_j(double x, double y) {
  x = fabs(x); y = fabs(y);
  double hi = x > y ? x : y, lo = x > y ? y : x;
  if (hi == 0.0) return hi;
  double r = lo / hi;
  return hi * sqrt(1.0 + r * r);
}
// ... and its derivative (x dx + y dy) / hypot for a freestream that carries
// a parameter tangent.
__device__ __forceinline__ Dual hypot_j(Dual x, Dual y) {
  const double h = hypot_j(x.v, y.v);
  return mkd(h, (x.v * x.d + y.v * y.d) / h);
}

// conservative -> primitive (src/jaxrans/core/state.py:141-146, numerics/residual.py:46-51)
template <class T, int NV>
__device__ __forceinline__ void primitive(const T* U, double gamma, T* q) {
  T rho = U[0];
  T u = U[1] / rho, v = U[2] / rho;
  q[0] = rho; q[1] = u; q[2] = v;
  q[3] = (gamma - 1.0) * (U[3] - 0.5 * rho * (u * u + v * v));
  if constexpr (NV > 4) q[4] = U[4] / rho;
}

template <class T, int NV>
__device__ __forceinline__ void conservative(const T* q, double gamma, T* U) {
  T rho = q[0], u = q[1], v = q[2], p = q[3];
  U[0] = rho; U[1] = rho * u; U[2] = rho * v;
  U[3] = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v);
  if constexpr (NV > 4) U[4] = rho * q[4];
}

// Limiters (src/jaxrans/numerics/limiters.py:234-290).
template <class T>
__device__ __forceinline__ T limit(int kind, T a, T b) {
  if (kind == LIM_NONE) return 0.5 * (a + b);
  if (kind == LIM_MINMOD) {
    if (val(a) * val(b) > 0.0) {
      double s = val(a) > 0.0 ? 1.0 : -1.0;
      return s * jmin(jabs(a), jabs(b));
    }
    return lift<T>(0.0);
  }
  if (kind == LIM_VAN_ALBADA) {
    if (val(a) * val(b) > 0.0) return a * b * (a + b) / (a * a + b * b);
    return lift<T>(0.0);
  }
  const double e = 1e-12;
  return ((b * b + e) * a + (a * a + e) * b) / (a * a + b * b + 2.0 * e);
}

// Analytic flux through a unit normal (src/jaxrans/physics/euler.py:43-52).
template <class T>
__device__ __forceinline__ void normal_flux(const T* q, double nx, double ny, double gamma, T* f) {
  T rho = q[0], u = q[1], v = q[2], p = q[3];
  T vn = u * nx + v * ny;
  T e = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v);
  f[0] = rho * vn;
  f[1] = rho * u * vn + p * nx;
  f[2] = rho * v * vn + p * ny;
  f[3] = (e + p) * vn;
}

template <class T>
__device__ __forceinline__ T total_enthalpy(T rho, T u, T v, T p, double gamma) {
  return gamma / (gamma - 1.0) * p / rho + 0.5 * (u * u + v * v);
}

// Roe average (src/jaxrans/physics/euler.py:74-94).
template <class T>
__device__ __forceinline__ void roe_average(const T* qL, const T* qR, double gamma, T& rho, T& u, T& v, T& h, T& a) {
  T hl = total_enthalpy(qL[0], qL[1], qL[2], qL[3], gamma);
  T hr = total_enthalpy(qR[0], qR[1], qR[2], qR[3], gamma);
  T w = jsqrt(qR[0] / qL[0]);
  T den = 1.0 + w;
  rho = jsqrt(qL[0] * qR[0]);
  u = (qL[1] + w * qR[1]) / den;
  v = (qL[2] + w * qR[2]) / den;
  h = (hl + w * hr) / den;
  T a2 = (gamma - 1.0) * (h - 0.5 * (u * u + v * v));
  a = jsqrt(jmax(a2, lift<T>(1e-30)));
}

// HLLC with Einfeldt speeds, times |S| (src/jaxrans/numerics/riemann.py:116-198).
template <class T>
__device__ void hllc(const T* qL, const T* qR, double sx, double sy, double gamma, T* F) {
  double len = hypot_j(sx, sy);
  double safe = len > 1e-300 ? len : 1e-300;
  double nx = sx / safe, ny = sy / safe;
  T rl = qL[0], ul = qL[1], vl = qL[2], pl = qL[3];
  T rr = qR[0], ur = qR[1], vr = qR[2], pr = qR[3];
  T vnl = ul * nx + vl * ny, vnr = ur * nx + vr * ny;
  T al = jsqrt(gamma * pl / rl), ar = jsqrt(gamma * pr / rr);
  T rro, uro, vro, hro, aro;
  roe_average(qL, qR, gamma, rro, uro, vro, hro, aro);
  T vnro = uro * nx + vro * ny;
  T sl = jmin(vnl - al, vnro - aro);
  T sr = jmax(vnr + ar, vnro + aro);
  T ml = rl * (sl - vnl), mr = rr * (sr - vnr);
  T sm = (pr - pl + ml * vnl - mr * vnr) / jmin(ml - mr, lift<T>(-1e-300));
  T f[4];
  if (val(sl) >= 0.0) {
    normal_flux(qL, nx, ny, gamma, f);
  } else if (val(sm) >= 0.0) {
    T fl[4], uc[4];
    normal_flux(qL, nx, ny, gamma, fl);
    conservative<T, 4>(qL, gamma, uc);
    T widen = jmin(sl - sm, lift<T>(-1e-300));
    T pstar = pl + rl * (vnl - sl) * (vnl - sm);
    T rstar = rl * (sl - vnl) / widen;
    T estar = ((sl - vnl) * uc[3] - pl * vnl + pstar * sm) / widen;
    T us[4] = {rstar, rstar * (ul + (sm - vnl) * nx), rstar * (vl + (sm - vnl) * ny), estar};
    for (int k = 0; k < 4; ++k) f[k] = fl[k] + sl * (us[k] - uc[k]);
  } else if (val(sr) >= 0.0) {
    T fr[4], uc[4];
    normal_flux(qR, nx, ny, gamma, fr);
    conservative<T, 4>(qR, gamma, uc);
    T widen = jmax(sr - sm, lift<T>(1e-300));
    T pstar = pr + rr * (vnr - sr) * (vnr - sm);
    T rstar = rr * (sr - vnr) / widen;
    T estar = ((sr - vnr) * uc[3] - pr * vnr + pstar * sm) / widen;
    T us[4] = {rstar, rstar * (ur + (sm - vnr) * nx), rstar * (vr + (sm - vnr) * ny), estar};
    for (int k = 0; k < 4; ++k) f[k] = fr[k] + sr * (us[k] - uc[k]);
  } else {
    normal_flux(qR, nx, ny, gamma, f);
  }
  for (int k = 0; k < 4; ++k) F[k] = len * f[k];
}

// Roe flux, no entropy fix, times |S| (src/jaxrans/numerics/riemann.py:54-113).
template <class T>
__device__ void roe(const T* qL, const T* qR, double sx, double sy, double gamma, T* F) {
  double len = hypot_j(sx, sy);
  double safe = len > 1e-300 ? len : 1e-300;
  double nx = sx / safe, ny = sy / safe;
  T rho, u, v, h, a;
  roe_average(qL, qR, gamma, rho, u, v, h, a);
  T vn = u * nx + v * ny;
  T drho = qR[0] - qL[0], dp = qR[3] - qL[3], du = qR[1] - qL[1], dv = qR[2] - qL[2];
  T dvn = du * nx + dv * ny;
  T dtu = du - dvn * nx, dtv = dv - dvn * ny;
  T a2 = a * a;
  T am = (dp - rho * a * dvn) / (2.0 * a2);
  T ae = drho - dp / a2;
  T ap = (dp + rho * a * dvn) / (2.0 * a2);
  T lm = jabs(vn - a), lmid = jabs(vn), lp = jabs(vn + a);
  T kin = 0.5 * (u * u + v * v);
  T rm[4] = {lift<T>(1.0), u - a * nx, v - a * ny, h - a * vn};
  T re[4] = {lift<T>(1.0), u, v, kin};
  T sh[4] = {lift<T>(0.0), rho * dtu, rho * dtv, rho * (u * dtu + v * dtv)};
  T rp[4] = {lift<T>(1.0), u + a * nx, v + a * ny, h + a * vn};
  T fl[4], fr[4];
  normal_flux(qL, nx, ny, gamma, fl);
  normal_flux(qR, nx, ny, gamma, fr);
  for (int k = 0; k < 4; ++k) {
    T diss = lm * am * rm[k] + lmid * ae * re[k] + lmid * sh[k] + lp * ap * rp[k];
    T central = 0.5 * (fl[k] + fr[k]);
    F[k] = len * (central - 0.5 * diss);
  }
}

// A far-field reference value as the cell's scalar type: a constant
// (freestream) is lifted, a value of the cell's own type kept.
template <class T> __device__ __forceinline__ T to_t(double x) { return lift<T>(x); }
template <class T> __device__ __forceinline__ T to_t(Dual x) { return x; }

// characteristic_state (src/jaxrans/numerics/boundary.py:91-122) against the
// reference primitives qf: one freestream (TR = double) or a per-face state
// that may carry a tangent (TR = T, the point-vortex far field).
template <class T, class TR, int NV>
__device__ void far_field_state(const T* q, double nx, double ny, const TR* qf, double gamma, T* g) {
  T a_i = jsqrt(gamma * q[3] / q[0]);
  TR a_f = jsqrt(gamma * qf[3] / qf[0]);
  double two_over = 2.0 / (gamma - 1.0);
  T vn_i = q[1] * nx + q[2] * ny;
  TR vn_f = qf[1] * nx + qf[2] * ny;
  T rp = vn_i + two_over * a_i;
  TR rm = vn_f - two_over * a_f;
  T vn_b = 0.5 * (rp + rm);
  T a_b = 0.25 * (gamma - 1.0) * (rp - rm);
  bool out = val(vn_b) > 0.0;
  T entropy = out ? q[3] / jpow(q[0], gamma) : to_t<T>(qf[3] / jpow(qf[0], gamma));
  T tu = out ? q[1] - vn_i * nx : to_t<T>(qf[1] - vn_f * nx);
  T tv = out ? q[2] - vn_i * ny : to_t<T>(qf[2] - vn_f * ny);
  T rho_b = jpow(a_b * a_b / (gamma * entropy), 1.0 / (gamma - 1.0));
  g[0] = rho_b;
  g[1] = vn_b * nx + tu;
  g[2] = vn_b * ny + tv;
  g[3] = rho_b * a_b * a_b / gamma;
  if constexpr (NV > 4) g[4] = out ? q[4] : to_t<T>(qf[4]);
}

// The point-vortex far-field reference (point_vortex_state,
// src/jaxrans/numerics/boundary.py:181-214; Thomas and Salas, AIAA J. 24
// (1986) 1074-1080): the freestream qi plus the Prandtl-Glauert-scaled
// compressible vortex of circulation circ (clockwise positive) at offset
// (dx, dy) from its centre, isentropic at the freestream total enthalpy.
// Evaluated without contraction into FMAs, so that no circulation gives the
// freestream bit for bit, as the reference's does.
// qi: the freestream, plain (TQ = double) or carrying a parameter tangent
// (TQ = TR = Dual).
template <class TR, int NV, class TQ>
__device__ void vortex_state(double dx, double dy, const TQ* qi, double gamma, TR circ, TR* qr) {
#pragma clang fp contract(off)
  const TQ rho_f = qi[0], u_f = qi[1], v_f = qi[2], p_f = qi[3];
  const TQ a2_f = gamma * p_f / rho_f;
  const TQ speed2_f = u_f * u_f + v_f * v_f;
  const TQ mach2 = speed2_f / a2_f;
  const TQ alpha = jatan2(v_f, u_f);
  const double r = hypot_j(dx, dy);
  const double theta = atan2(dy, dx);
  const TQ s = jsin(theta - alpha);
  TR strength = circ * jsqrt(1.0 - mach2) / (2.0 * M_PI * r * (1.0 - mach2 * (s * s)));
  TR u = u_f + strength * sin(theta);
  TR v = v_f - strength * cos(theta);
  TR ratio = 1.0 + 0.5 * (gamma - 1.0) * (speed2_f - u * u - v * v) / a2_f;
  qr[0] = rho_f * jpow(ratio, 1.0 / (gamma - 1.0));
  qr[1] = u;
  qr[2] = v;
  qr[3] = p_f * jpow(ratio, gamma / (gamma - 1.0));
  if constexpr (NV > 4) qr[4] = to_t<TR>(qi[4]);
}

// Ghost primitives from the source cell (src/jaxrans/numerics/boundary.py:54-157).
// qf: the far-field reference (freestream) primitives, plain (TQ = double)
// or carrying a parameter tangent (TQ = T = Dual).
template <class T, int NV, class TQ>
__device__ void bc_state(int kind, const T* q, double nx, double ny, const TQ* qf, double gamma, T* g) {
  if (kind == K_SLIP || kind == K_SYMMETRY) {
    T vn = q[1] * nx + q[2] * ny;
    g[0] = q[0];
    g[1] = q[1] - 2.0 * vn * nx;
    g[2] = q[2] - 2.0 * vn * ny;
    g[3] = q[3];
    if constexpr (NV > 4) g[4] = kind == K_SLIP ? -q[4] : q[4];
  } else if (kind == K_NOSLIP) {
    g[0] = q[0]; g[1] = -q[1]; g[2] = -q[2]; g[3] = q[3];
    if constexpr (NV > 4) g[4] = -q[4];
  } else if (kind == K_FARFIELD) {
    far_field_state<T, TQ, NV>(q, nx, ny, qf, gamma, g);
  } else if (kind == K_INFLOW) {
    // total_condition_inflow (boundary.py:130-147)
    double g1 = gamma - 1.0;
    TQ t_f = gamma * qf[3] / qf[0];
    TQ speed_f = hypot_j(qf[1], qf[2]);
    TQ t_tot = t_f + 0.5 * g1 * speed_f * speed_f;
    TQ p_tot = qf[3] * jpow(t_tot / t_f, gamma / g1);
    T p_b = q[3];
    T mach2 = jmax(2.0 / g1 * (jpow(p_tot / p_b, g1 / gamma) - 1.0), lift<T>(0.0));
    T t_b = t_tot / (1.0 + 0.5 * g1 * mach2);
    T speed = jsqrt(mach2 * t_b);
    g[0] = gamma * p_b / t_b;
    g[1] = speed * qf[1] / speed_f;
    g[2] = speed * qf[2] / speed_f;
    g[3] = p_b;
    if constexpr (NV > 4) g[4] = to_t<T>(qf[4]);
  } else if (kind == K_OUTFLOW) {
    g[0] = q[0]; g[1] = q[1]; g[2] = q[2]; g[3] = to_t<T>(qf[3]);
    if constexpr (NV > 4) g[4] = q[4];
  } else {
    for (int k = 0; k < NV; ++k) g[k] = q[k];
  }
}

// Viscosity law (src/jaxrans/physics/navier_stokes.py:154-175).
// mu_inf plain (TM = double) or carrying a parameter tangent (TM = T = Dual).
template <class T, class TM>
__device__ __forceinline__ T viscosity(T Tm, TM mu_inf, double suth) {
  if (suth <= 0.0) return to_t<T>(mu_inf);
  return mu_inf * Tm * jsqrt(Tm) * (1.0 + suth) / (Tm + suth);
}

template <class T>
__device__ __forceinline__ T temperature(const T* q, double gamma) { return gamma * q[3] / q[0]; }

// F_v . S (src/jaxrans/physics/navier_stokes.py:192-222); gv[a][b] = d v_a / d x_b.
template <class T>
__device__ __forceinline__ void viscous_flux(const T* vel, T gv[2][2], const T* gT, double sx, double sy, T mu, T k,
                                             const T (*extra)[2], T* f) {
  T div = gv[0][0] + gv[1][1];
  T ttd = (2.0 / 3.0) * div;
  T t00 = mu * (2.0 * gv[0][0] - ttd);
  T t11 = mu * (2.0 * gv[1][1] - ttd);
  T t01 = mu * (gv[0][1] + gv[1][0]);
  T t10 = t01;
  if (extra) {
    t00 = t00 + extra[0][0]; t01 = t01 + extra[0][1];
    t10 = t10 + extra[1][0]; t11 = t11 + extra[1][1];
  }
  T fx = t00 * sx + t01 * sy;
  T fy = t10 * sx + t11 * sy;
  f[0] = lift<T>(0.0);
  f[1] = fx;
  f[2] = fy;
  f[3] = (vel[0] * fx + vel[1] * fy) + k * (gT[0] * sx + gT[1] * sy);
}

// QCR2000 (src/jaxrans/physics/spalart_allmaras.py:197-216).
template <class T>
__device__ __forceinline__ void qcr_stress(T gv[2][2], T mu_t, double c_cr1, T out[2][2]) {
  T sq = gv[0][0] * gv[0][0] + gv[0][1] * gv[0][1] + gv[1][0] * gv[1][0] + gv[1][1] * gv[1][1];
  T nrm = val(sq) > 0.0 ? jsqrt(sq) : lift<T>(1.0);
  T rot[2][2], tau[2][2];
  T div = gv[0][0] + gv[1][1];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) {
      rot[a][b] = (gv[a][b] - gv[b][a]) / nrm;
      tau[a][b] = mu_t * (gv[a][b] + gv[b][a] - (a == b ? (2.0 / 3.0) * div : lift<T>(0.0)));
    }
  T prod[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) prod[i][j] = rot[i][0] * tau[j][0] + rot[i][1] * tau[j][1];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) out[a][b] = -c_cr1 * (prod[a][b] + prod[b][a]);
}

// Spalart-Allmaras pieces (src/jaxrans/physics/spalart_allmaras.py:119-194).
template <class T> __device__ __forceinline__ T f_v1(T chi) {
  T c3 = ipow3(chi);
  return c3 / (c3 + CV1 * CV1 * CV1);
}
template <class T> __device__ __forceinline__ T f_v2(T chi) { return 1.0 - chi / (1.0 + chi * f_v1(chi)); }

template <class T>
__device__ __forceinline__ T eddy_viscosity(T rho, T nu_t, T nu) {
  if (val(nu_t) > 0.0) return rho * nu_t * f_v1(nu_t / nu);
  return lift<T>(0.0);
}

template <class T>
__device__ __forceinline__ T s_tilde(T omega, T nu_t, T nu, double d) {
  T chi = nu_t / nu;
  T s_bar = nu_t * f_v2(chi) / ((KAPPA * d) * (KAPPA * d));
  if (val(s_bar) >= val(-C2 * omega)) return omega + s_bar;
  T den = (C3 - 2.0 * C2) * omega - s_bar;
  return omega + omega * (C2 * C2 * omega + C3 * s_bar) / den;
}

template <class T> __device__ __forceinline__ T f_w(T r) {
  T g = r + CW2 * (ipow6(r) - r);
  const double c6 = CW3 * CW3 * CW3 * CW3 * CW3 * CW3;
  return g * jpow((1.0 + c6) / (ipow6(g) + c6), 1.0 / 6.0);
}

template <class T>
__device__ __forceinline__ T r_of(T nu_t, T s_t, double d) {
  T den = s_t * ((KAPPA * d) * (KAPPA * d));
  if (val(den) > 0.0) return jmin(nu_t / den, lift<T>(RMAX));
  return lift<T>(RMAX);
}

template <class T>
__device__ __forceinline__ void sa_source(T nu_t, T nu, T omega, double d, T& prod, T& destr) {
  bool pos = val(nu_t) >= 0.0;
  T st = s_tilde(omega, pos ? nu_t : lift<T>(0.0), nu, d);
  T wall = ipow2(nu_t / d);
  prod = (pos ? CB1 * st : CB1 * omega) * nu_t;
  destr = (pos ? CW1 * f_w(r_of(nu_t, st, d)) : lift<T>(-CW1)) * wall;
}

template <class T>
__device__ __forceinline__ T f_n(T nu_t, T nu) {
  if (val(nu_t) >= 0.0) return lift<T>(1.0);
  T chi3 = ipow3(nu_t / nu);
  return (CN1 + chi3) / (CN1 - chi3);
}

template <class T>
__device__ __forceinline__ T diffusion_coefficient(T rho, T nu_t, T mu) {
  T nu = mu / rho;
  return (mu + rho * nu_t * f_n(nu_t, nu)) / SIGMA;
}

}  // namespace jr
