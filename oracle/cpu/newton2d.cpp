// TEST INFRASTRUCTURE ONLY: a C++/OpenMP FP64 restatement of the NumPy
// oracle's 2D Newton-Krylov step (oracle/residual.py, oracle/physics.py,
// oracle/solver.py), for the CPU baseline of the benchmark, which times it
// per component as the reference's scripts/benchmark.py:102-113 does
// (residual, tangent, line-preconditioner set-up, apply, Newton step).  Only
// tests/ and bench.py's cpu_baseline leg load it; the product never does.
//
// It follows the oracle function by function (each names the reference
// function it restates):
//   residual      numerics/residual.py:54-124, 312-327 (ghost fill, edge pad,
//                 MUSCL, HLLC / Roe, diamond viscous flux, SA row, U - U0)
//   jvp           the same code on dual numbers (jax.linearize,
//                 solvers/implicit.py:105-108), first-order or exact
//   wave speeds   numerics/residual.py:169-203
//   line precond  solvers/preconditioners.py:209-306 (Rusanov flow blocks,
//                 scalar SA chains), core/linalg.py:49-132 (Gauss-Jordan
//                 inverse, block Thomas: the same exact solve as the
//                 reference's cyclic reduction)
//   GMRES         solvers/krylov.py:34-137 (right-preconditioned, CGS2,
//                 Givens least squares with pivot drop)
//   Newton step   solvers/implicit.py:54-166 (relaxation, accept, SER CFL)
// Cells, faces and chains are split over OpenMP threads; cases of a batch
// are separate contexts.  Pinned against the NumPy oracle by
// tests/test_oracle_cpu.py.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace oc {

// --- scalar types: double, and a dual number for the tangent ---------------------
struct D {
  double v, d;
};
inline D mk(double v, double d) { return D{v, d}; }
inline D operator+(D a, D b) { return mk(a.v + b.v, a.d + b.d); }
inline D operator+(D a, double b) { return mk(a.v + b, a.d); }
inline D operator+(double a, D b) { return mk(a + b.v, b.d); }
inline D operator-(D a, D b) { return mk(a.v - b.v, a.d - b.d); }
inline D operator-(D a, double b) { return mk(a.v - b, a.d); }
inline D operator-(double a, D b) { return mk(a - b.v, -b.d); }
inline D operator-(D a) { return mk(-a.v, -a.d); }
inline D operator*(D a, D b) { return mk(a.v * b.v, a.d * b.v + a.v * b.d); }
inline D operator*(D a, double b) { return mk(a.v * b, a.d * b); }
inline D operator*(double a, D b) { return mk(a * b.v, a * b.d); }
inline D operator/(D a, D b) {
  const double q = a.v / b.v;
  return mk(q, (a.d - q * b.d) / b.v);
}
inline D operator/(D a, double b) { return mk(a.v / b, a.d / b); }
inline D operator/(double a, D b) {
  const double q = a / b.v;
  return mk(q, -q * b.d / b.v);
}
inline double val(double x) { return x; }
inline double val(D x) { return x.v; }
template <class T> inline T cst(double x);
template <> inline double cst<double>(double x) { return x; }
template <> inline D cst<D>(double x) { return mk(x, 0.0); }
inline double sqrt_(double x) { return std::sqrt(x); }
inline D sqrt_(D x) {
  const double r = std::sqrt(x.v);
  return mk(r, x.d * (0.5 / r));
}
inline double abs_(double x) { return std::fabs(x); }
inline D abs_(D x) { return mk(std::fabs(x.v), x.v >= 0.0 ? x.d : -x.d); }
// jnp.minimum / maximum: the selected argument's tangent, half of each at a tie
inline double min_(double a, double b) { return a < b ? a : (b < a ? b : a); }
inline double max_(double a, double b) { return a > b ? a : (b > a ? b : a); }
inline D min_(D a, D b) { return a.v < b.v ? a : (b.v < a.v ? b : mk(a.v, 0.5 * (a.d + b.d))); }
inline D max_(D a, D b) { return a.v > b.v ? a : (b.v > a.v ? b : mk(a.v, 0.5 * (a.d + b.d))); }
inline D min_(D a, double b) { return min_(a, mk(b, 0.0)); }
inline D max_(D a, double b) { return max_(a, mk(b, 0.0)); }
inline double pow_(double x, double c) { return std::pow(x, c); }
inline D pow_(D x, double c) { return mk(std::pow(x.v, c), c * std::pow(x.v, c - 1.0) * x.d); }
// x ** n for a small integer n, as NumPy's power on an integer-valued exponent
template <class T> inline T ipow(T x, int n) {
  T r = x;
  for (int k = 1; k < n; ++k) r = r * x;
  return r;
}
inline D ipow(D x, int n) { return mk(std::pow(x.v, n), n * std::pow(x.v, n - 1) * x.d); }

// jnp.hypot's formula on constant data: max * sqrt(1 + (min / max)^2)
inline double hypot_j(double x, double y) {
  x = std::fabs(x);
  y = std::fabs(y);
  const double hi = std::max(x, y), lo = std::min(x, y);
  if (hi == 0.0) return hi;
  const double t = lo / hi;
  return hi * std::sqrt(1.0 + t * t);
}

// --- constants (core/constants.py, physics/spalart_allmaras.py:65-89) ----------------
constexpr double PR = 0.72, PR_T = 0.9, MIN_DENSITY = 1e-10, MIN_PRESSURE = 1e-10;
constexpr double CB1 = 0.1355, SIGMA = 2.0 / 3.0, CB2 = 0.622, KAPPA = 0.41;
constexpr double CW2 = 0.3, CW3 = 2.0, CV1 = 7.1, C2 = 0.7, C3 = 0.9, CN1 = 16.0, RMAX = 10.0;
const double CW1 = CB1 / (KAPPA * KAPPA) + (1.0 + CB2) / SIGMA;

}  // namespace oc

extern "C" {

// The atlas (mesh/atlas.py) and the case numerics, host arrays.
typedef struct {
  int32_t NI, NJ;
  const double* area;       // [NI*NJ]
  const double* face_i;     // [2][(NI+1)*NJ]
  const double* face_j;     // [2][NI*(NJ+1)]
  const double* centres;    // [2][(NI+2)*(NJ+2)]
  const double* distance;   // [NI*NJ]
  const uint8_t* interior;  // [NI*NJ]
  const int64_t* neighbours;
  int32_t n_task[2];
  const int64_t* task_dest[2];
  const int64_t* task_src[2];
  const int32_t* task_kind[2];
  const double* task_normal[2];   // [2][n]
  int32_t n_lines, n_slots;
  const int64_t* line_start;
  const int64_t* line_cells;
  const int64_t* line_up_face;
  const int64_t* line_up_sign;
  const int64_t* line_down_face;
  const int64_t* line_down_sign;
} oc_grid;

typedef struct {
  double gamma, kappa;
  int32_t flux;             // 0 hllc, 1 roe
  int32_t limiter;          // 0 minmod, 1 none, 2 van_albada, 3 van_albada_smooth
  int32_t viscous, sa;
  double sutherland;        // <= 0: constant viscosity
  double c_cr1;
  double mu_inf;
  double q_inf[5];
  int32_t krylov_subspace, krylov_restarts;
  double krylov_tol, krylov_success, cfl_growth, cfl_cutback, cfl_min, cfl_max;
  double max_relative_update;   // <= 0: none
  int32_t exact;            // Krylov operator: 1 the exact tangent, 0 the first-order one
} oc_phys;

}  // extern "C"

namespace oc {

struct Ctx {
  oc_grid g;
  oc_phys ph;
  int NI, NJ, nc, nv, nfi, nfj;
  std::vector<double> U0;              // uniform freestream canvas (conservative)
  std::vector<double> lam;             // face wave speeds, i faces then j faces
  std::vector<double> lsum;            // per cell
  // line factor (block Thomas): per slot L (16), P^-1 (16), C (16); SA l, 1/p, c
  std::vector<double> fac, fac_sa;
  std::vector<double> scratch_d;
};

// --- pointwise physics ---------------------------------------------------------------
template <class T> inline void primitive(const T* u, int nv, double g, T* q) {
  const T rho = u[0];
  const T uu = u[1] / rho, vv = u[2] / rho;
  q[0] = rho;
  q[1] = uu;
  q[2] = vv;
  q[3] = (g - 1.0) * (u[3] - 0.5 * rho * (uu * uu + vv * vv));
  if (nv > 4) q[4] = u[4] / rho;
}

template <class T> inline void conservative(const T* q, int nv, double g, T* u) {
  const T e = q[3] / (g - 1.0) + 0.5 * q[0] * (q[1] * q[1] + q[2] * q[2]);
  u[0] = q[0];
  u[1] = q[0] * q[1];
  u[2] = q[0] * q[2];
  u[3] = e;
  if (nv > 4) u[4] = q[0] * q[4];
}

template <class T> inline T sgn0(T a) { return cst<T>(val(a) > 0.0 ? 1.0 : (val(a) < 0.0 ? -1.0 : 0.0)); }

template <class T> inline T limit(int kind, T a, T b) {
  if (kind == 1) return 0.5 * (a + b);
  if (kind == 0) return val(a * b) > 0.0 ? sgn0(a) * min_(abs_(a), abs_(b)) : cst<T>(0.0);
  if (kind == 2) {
    const bool same = val(a * b) > 0.0;
    const T den = same ? a * a + b * b : cst<T>(1.0);
    return same ? a * b * (a + b) / den : cst<T>(0.0);
  }
  const double e = 1e-12;
  return ((b * b + e) * a + (a * a + e) * b) / (a * a + b * b + 2.0 * e);
}

template <class T> inline void normal_flux(const T* q, double nx, double ny, double g, T* f) {
  const T vn = q[1] * nx + q[2] * ny;
  const T e = q[3] / (g - 1.0) + 0.5 * q[0] * (q[1] * q[1] + q[2] * q[2]);
  f[0] = q[0] * vn;
  f[1] = q[0] * q[1] * vn + q[3] * nx;
  f[2] = q[0] * q[2] * vn + q[3] * ny;
  f[3] = (e + q[3]) * vn;
}

template <class T> inline T enthalpy(const T* q, double g) {
  return g / (g - 1.0) * q[3] / q[0] + 0.5 * (q[1] * q[1] + q[2] * q[2]);
}

template <class T> inline void roe_average(const T* l, const T* r, double g, T* o /* rho u v h a */) {
  const T hl = enthalpy(l, g), hr = enthalpy(r, g);
  const T w = sqrt_(r[0] / l[0]);
  const T den = 1.0 + w;
  o[0] = sqrt_(l[0] * r[0]);
  o[1] = (l[1] + w * r[1]) / den;
  o[2] = (l[2] + w * r[2]) / den;
  o[3] = (hl + w * hr) / den;
  const T a2 = (g - 1.0) * (o[3] - 0.5 * (o[1] * o[1] + o[2] * o[2]));
  o[4] = sqrt_(max_(a2, cst<T>(1e-30)));
}

// HLLC with Einfeldt speeds (riemann.py:116-198), times |S|
template <class T> inline void hllc(const T* l, const T* r, double sx, double sy, double g, T* F) {
  const double len = hypot_j(sx, sy), safe = std::max(len, 1e-300), nx = sx / safe, ny = sy / safe;
  const T vnl = l[1] * nx + l[2] * ny, vnr = r[1] * nx + r[2] * ny;
  const T al = sqrt_(g * l[3] / l[0]), ar = sqrt_(g * r[3] / r[0]);
  T ro[5];
  roe_average(l, r, g, ro);
  const T vnro = ro[1] * nx + ro[2] * ny;
  const T sl = min_(vnl - al, vnro - ro[4]);
  const T sr = max_(vnr + ar, vnro + ro[4]);
  const T ml = l[0] * (sl - vnl), mr = r[0] * (sr - vnr);
  const T sm = (r[3] - l[3] + ml * vnl - mr * vnr) / min_(ml - mr, cst<T>(-1e-300));
  T ucl[4], ucr[4], fl[4], fr[4];
  conservative(l, 4, g, ucl);
  conservative(r, 4, g, ucr);
  normal_flux(l, nx, ny, g, fl);
  normal_flux(r, nx, ny, g, fr);
  auto star = [&](const T* q, T vn, T sk, const T* fk, const T* uk, T widen, T* out) {
    const T pstar = q[3] + q[0] * (vn - sk) * (vn - sm);
    const T rstar = q[0] * (sk - vn) / widen;
    const T estar = ((sk - vn) * uk[3] - q[3] * vn + pstar * sm) / widen;
    const T us[4] = {rstar, rstar * (q[1] + (sm - vn) * nx), rstar * (q[2] + (sm - vn) * ny), estar};
    for (int k = 0; k < 4; ++k) out[k] = fk[k] + sk * (us[k] - uk[k]);
  };
  T stl[4], strr[4];
  star(l, vnl, sl, fl, ucl, min_(sl - sm, cst<T>(-1e-300)), stl);
  star(r, vnr, sr, fr, ucr, max_(sr - sm, cst<T>(1e-300)), strr);
  for (int k = 0; k < 4; ++k) {
    const T f = val(sl) >= 0.0 ? fl[k] : (val(sm) >= 0.0 ? stl[k] : (val(sr) >= 0.0 ? strr[k] : fr[k]));
    F[k] = len * f;
  }
}

// Roe, no entropy fix (riemann.py:54-113), times |S|
template <class T> inline void roe(const T* l, const T* r, double sx, double sy, double g, T* F) {
  const double len = hypot_j(sx, sy), safe = std::max(len, 1e-300), nx = sx / safe, ny = sy / safe;
  T ro[5];
  roe_average(l, r, g, ro);
  const T rho = ro[0], u = ro[1], v = ro[2], h = ro[3], a = ro[4];
  const T vn = u * nx + v * ny;
  const T drho = r[0] - l[0], dp = r[3] - l[3], du = r[1] - l[1], dv = r[2] - l[2];
  const T dvn = du * nx + dv * ny;
  const T dvt0 = du - dvn * nx, dvt1 = dv - dvn * ny;
  const T a2 = a * a;
  const T am = (dp - rho * a * dvn) / (2.0 * a2);
  const T ae = drho - dp / a2;
  const T ap = (dp + rho * a * dvn) / (2.0 * a2);
  const T lm = abs_(vn - a), lmid = abs_(vn), lp = abs_(vn + a);
  const T kin = 0.5 * (u * u + v * v);
  const T one = cst<T>(1.0), zero = cst<T>(0.0);
  const T rm[4] = {one, u - a * nx, v - a * ny, h - a * vn};
  const T re[4] = {one, u, v, kin};
  const T sh[4] = {zero, rho * dvt0, rho * dvt1, rho * (u * dvt0 + v * dvt1)};
  const T rp[4] = {one, u + a * nx, v + a * ny, h + a * vn};
  T fl[4], fr[4];
  normal_flux(l, nx, ny, g, fl);
  normal_flux(r, nx, ny, g, fr);
  for (int k = 0; k < 4; ++k) {
    const T diss = lm * am * rm[k] + lmid * ae * re[k] + lmid * sh[k] + lp * ap * rp[k];
    const T central = 0.5 * (fl[k] + fr[k]);
    F[k] = len * (central - 0.5 * diss);
  }
}

template <class T> inline void characteristic(const T* q, int nv, double nx, double ny, const double* qf, double g,
                                              T* o) {
  const T a_i = sqrt_(g * q[3] / q[0]);
  const double a_f = std::sqrt(g * qf[3] / qf[0]);
  const double two = 2.0 / (g - 1.0);
  const T vn_i = q[1] * nx + q[2] * ny;
  const double vn_f = qf[1] * nx + qf[2] * ny;
  const T rp = vn_i + two * a_i;
  const double rm = vn_f - two * a_f;
  const T vn_b = 0.5 * (rp + rm);
  const T a_b = 0.25 * (g - 1.0) * (rp - rm);
  const bool out = val(vn_b) > 0.0;
  const T ent = out ? q[3] / pow_(q[0], g) : cst<T>(qf[3] / std::pow(qf[0], g));
  const T tu = out ? q[1] - vn_i * nx : cst<T>(qf[1] - vn_f * nx);
  const T tv = out ? q[2] - vn_i * ny : cst<T>(qf[2] - vn_f * ny);
  const T rho_b = pow_(a_b * a_b / (g * ent), 1.0 / (g - 1.0));
  o[0] = rho_b;
  o[1] = vn_b * nx + tu;
  o[2] = vn_b * ny + tv;
  o[3] = rho_b * a_b * a_b / g;
  for (int k = 4; k < nv; ++k) o[k] = out ? q[k] : cst<T>(qf[k]);
}

template <class T> inline void bc_state(int kind, const T* q, int nv, double nx, double ny, const double* qf,
                                        double g, T* o) {
  if (kind == 1 || kind == 3) {
    const T vn = q[1] * nx + q[2] * ny;
    o[0] = q[0];
    o[1] = q[1] - 2.0 * vn * nx;
    o[2] = q[2] - 2.0 * vn * ny;
    o[3] = q[3];
    for (int k = 4; k < nv; ++k) o[k] = kind == 1 ? -q[k] : q[k];
  } else if (kind == 2) {
    o[0] = q[0];
    o[1] = -q[1];
    o[2] = -q[2];
    o[3] = q[3];
    for (int k = 4; k < nv; ++k) o[k] = -q[k];
  } else if (kind == 4) {
    characteristic(q, nv, nx, ny, qf, g, o);
  } else if (kind == 5) {
    const double g1 = g - 1.0, t_f = g * qf[3] / qf[0], speed_f = hypot_j(qf[1], qf[2]);
    const double t_tot = t_f + 0.5 * g1 * speed_f * speed_f;
    const double p_tot = qf[3] * std::pow(t_tot / t_f, g / g1);
    const T p_b = q[3];
    const T mach2 = max_(2.0 / g1 * (pow_(p_tot / p_b, g1 / g) - 1.0), cst<T>(0.0));
    const T t_b = t_tot / (1.0 + 0.5 * g1 * mach2);
    const T speed = sqrt_(mach2 * t_b);
    o[0] = g * p_b / t_b;
    o[1] = speed * qf[1] / speed_f;
    o[2] = speed * qf[2] / speed_f;
    o[3] = p_b;
    for (int k = 4; k < nv; ++k) o[k] = cst<T>(qf[k]);
  } else if (kind == 6) {
    o[0] = q[0];
    o[1] = q[1];
    o[2] = q[2];
    o[3] = cst<T>(qf[3]);
    for (int k = 4; k < nv; ++k) o[k] = q[k];
  } else {
    for (int k = 0; k < nv; ++k) o[k] = q[k];
  }
}

template <class T> inline T viscosity(T Tm, double mu_inf, double suth) {
  if (suth <= 0.0) return cst<T>(mu_inf);
  return mu_inf * Tm * sqrt_(Tm) * (1.0 + suth) / (Tm + suth);
}
template <class T> inline T temperature(const T* q, double g) { return g * q[3] / q[0]; }

template <class T> inline T f_v1(T chi) {
  const T chi3 = ipow(chi, 3);
  return chi3 / (chi3 + CV1 * CV1 * CV1);
}
template <class T> inline T f_v2(T chi) { return 1.0 - chi / (1.0 + chi * f_v1(chi)); }
template <class T> inline T eddy_viscosity(T rho, T nu_t, T nu) {
  if (!(val(nu_t) > 0.0)) return cst<T>(0.0);
  const T chi = nu_t / nu;
  return rho * nu_t * f_v1(chi);
}
template <class T> inline T s_tilde(T omega, T nu_t, T nu, double d) {
  const T chi = nu_t / nu;
  const T s_bar = nu_t * f_v2(chi) / ((KAPPA * d) * (KAPPA * d));
  const bool ordinary = val(s_bar) >= val(-C2 * omega);
  const T den = ordinary ? cst<T>(1.0) : (C3 - 2.0 * C2) * omega - s_bar;
  const T clipped = omega + omega * (C2 * C2 * omega + C3 * s_bar) / den;
  return ordinary ? omega + s_bar : clipped;
}
template <class T> inline T f_w(T r) {
  const T g = r + CW2 * (ipow(r, 6) - r);
  const double c6 = std::pow(CW3, 6);
  return g * pow_((1.0 + c6) / (ipow(g, 6) + c6), 1.0 / 6.0);
}
template <class T> inline T r_of(T nu_t, T st, double d) {
  const T den = st * ((KAPPA * d) * (KAPPA * d));
  if (!(val(den) > 0.0)) return cst<T>(RMAX);
  return min_(nu_t / den, cst<T>(RMAX));
}
template <class T> inline void source(T nu_t, T nu, T omega, double d, T* prod, T* destr) {
  const bool pos = val(nu_t) >= 0.0;
  const T st = s_tilde(omega, pos ? nu_t : cst<T>(0.0), nu, d);
  const T wall = ipow(nu_t / d, 2);
  *prod = (pos ? CB1 * st : CB1 * omega) * nu_t;
  *destr = (pos ? CW1 * f_w(r_of(nu_t, st, d)) : cst<T>(-CW1)) * wall;
}
template <class T> inline T f_n(T nu_t, T nu) {
  if (val(nu_t) >= 0.0) return cst<T>(1.0);
  const T chi3 = ipow(nu_t / nu, 3);
  return (CN1 + chi3) / (CN1 - chi3);
}
template <class T> inline T diffusion_coefficient(T rho, T nu_t, T mu) {
  const T nu = mu / rho;
  return (mu + rho * nu_t * f_n(nu_t, nu)) / SIGMA;
}

// --- the residual -----------------------------------------------------------------------
template <class T> struct Work {
  std::vector<T> q;       // primitives on the canvas (ghosts filled) [nv][nc]
  std::vector<T> tmp;     // gather buffer of a stage
  std::vector<T> Fi, Fj;  // face fluxes [nv][faces]
  std::vector<T> nodes;   // viscous node values [nphi][(NI+1)(NJ+1)]
  std::vector<T> fav;     // SA: face averages of u, v, nu_hat per family
};

inline int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }

template <class T> void residual(Ctx& c, const double* U, const double* dU, double* R_out, int first_order) {
  const int NI = c.NI, NJ = c.NJ, nc = c.nc, nv = c.nv;
  const double g = c.ph.gamma;
  Work<T> w;
  w.q.resize((size_t)nv * nc);
  auto ld = [&](size_t i) -> T {
    if constexpr (sizeof(T) == sizeof(double)) return U[i];
    else return mk(U[i], dU ? dU[i] : 0.0);
  };
#pragma omp parallel for schedule(static)
  for (int cell = 0; cell < nc; ++cell) {
    T u[5], q[5];
    for (int k = 0; k < nv; ++k) u[k] = ld((size_t)k * nc + cell);
    primitive(u, nv, g, q);
    for (int k = 0; k < nv; ++k) w.q[(size_t)k * nc + cell] = q[k];
  }
  // staged ghost fill: each stage a gather from the state before it
  for (int s = 0; s < 2; ++s) {
    const int n = c.g.n_task[s];
    w.tmp.assign((size_t)nv * std::max(n, 1), cst<T>(0.0));
#pragma omp parallel for schedule(static)
    for (int t = 0; t < n; ++t) {
      T q[5], o[5];
      const int64_t src = c.g.task_src[s][t];
      for (int k = 0; k < nv; ++k) q[k] = w.q[(size_t)k * nc + src];
      bc_state(c.g.task_kind[s][t], q, nv, c.g.task_normal[s][t], c.g.task_normal[s][n + t], c.ph.q_inf, g, o);
      for (int k = 0; k < nv; ++k) w.tmp[(size_t)k * n + t] = o[k];
    }
#pragma omp parallel for schedule(static)
    for (int t = 0; t < n; ++t)
      for (int k = 0; k < nv; ++k) w.q[(size_t)k * nc + c.g.task_dest[s][t]] = w.tmp[(size_t)k * n + t];
  }
  auto Q = [&](int k, int i, int j) -> T {   // the filled canvas, edge-padded beyond it
    return w.q[(size_t)k * nc + (size_t)clampi(i, 0, NI - 1) * NJ + clampi(j, 0, NJ - 1)];
  };
  const bool visc = c.ph.viscous != 0, sa = c.ph.sa != 0;
  const int nfi = c.nfi, nfj = c.nfj;
  w.Fi.assign((size_t)nv * nfi, cst<T>(0.0));
  w.Fj.assign((size_t)nv * nfj, cst<T>(0.0));
  // viscous node values of (u, v, T[, nu_hat]): the mean of the 4 cells round
  // each node, on the canvas padded by one (edge)
  const int nphi = sa ? 4 : 3;
  if (visc) {
    w.nodes.assign((size_t)nphi * (NI + 1) * (NJ + 1), cst<T>(0.0));
#pragma omp parallel for schedule(static)
    for (int I = 0; I <= NI; ++I)
      for (int J = 0; J <= NJ; ++J) {
        T ph[4][4];
        const int ci[4] = {I - 1, I, I - 1, I}, cj[4] = {J - 1, J - 1, J, J};
        for (int m = 0; m < 4; ++m) {
          T q[5];
          for (int k = 0; k < nv; ++k) q[k] = Q(k, ci[m], cj[m]);
          ph[m][0] = q[1];
          ph[m][1] = q[2];
          ph[m][2] = temperature(q, g);
          if (sa) ph[m][3] = q[4];
        }
        for (int p = 0; p < nphi; ++p)
          w.nodes[(size_t)p * (NI + 1) * (NJ + 1) + (size_t)I * (NJ + 1) + J] =
              0.25 * (((ph[0][p] + ph[1][p]) + ph[2][p]) + ph[3][p]);
      }
  }
  const double* cen = c.g.centres;
  const int CJ = NJ + 2;
  auto cx = [&](int i, int j) { return cen[(size_t)(i + 1) * CJ + (j + 1)]; };                 // canvas cell coords
  auto cy = [&](int i, int j) { return cen[(size_t)(NI + 2) * CJ + (size_t)(i + 1) * CJ + (j + 1)]; };
  auto nx_ = [&](int I, int J) {
    return 0.25 * (((cx(I - 1, J - 1) + cx(I, J - 1)) + cx(I - 1, J)) + cx(I, J));
  };
  auto ny_ = [&](int I, int J) {
    return 0.25 * (((cy(I - 1, J - 1) + cy(I, J - 1)) + cy(I - 1, J)) + cy(I, J));
  };
  auto node = [&](int p, int I, int J) { return w.nodes[(size_t)p * (NI + 1) * (NJ + 1) + (size_t)I * (NJ + 1) + J]; };

  // one face: a = family, (i, j) its index; il, jl / ir, jr its two cells; the
  // diamond's second diagonal from node (I0, J0) to (I1, J1)
  auto face = [&](int a, int f, int il, int jl, int ir, int jr, int mi, int mj, int I0, int J0, int I1, int J1,
                  double sx, double sy, T* F) {
    T s4[4][5];
    const int di = a == 0 ? 1 : 0, dj = a == 0 ? 0 : 1;
    for (int m = 0; m < 4; ++m)
      for (int k = 0; k < nv; ++k) s4[m][k] = Q(k, mi + (m - 2) * di, mj + (m - 2) * dj);
    T qL[5], qR[5];
    for (int k = 0; k < 4; ++k) {
      if (first_order) {
        qL[k] = s4[1][k];
        qR[k] = s4[2][k];
      } else if (c.ph.kappa != 0.0) {
        const double lo = 0.25 * (1.0 - c.ph.kappa), hi = 0.25 * (1.0 + c.ph.kappa);
        qL[k] = s4[1][k] + lo * (s4[1][k] - s4[0][k]) + hi * (s4[2][k] - s4[1][k]);
        qR[k] = s4[2][k] - lo * (s4[3][k] - s4[2][k]) - hi * (s4[2][k] - s4[1][k]);
      } else {
        qL[k] = s4[1][k] + 0.5 * limit(c.ph.limiter, s4[1][k] - s4[0][k], s4[2][k] - s4[1][k]);
        qR[k] = s4[2][k] - 0.5 * limit(c.ph.limiter, s4[2][k] - s4[1][k], s4[3][k] - s4[2][k]);
      }
    }
    T Fl[4];
    if (c.ph.flux == 0) hllc(qL, qR, sx, sy, g, Fl);
    else roe(qL, qR, sx, sy, g, Fl);
    for (int k = 0; k < 4; ++k) F[k] = Fl[k];
    if (nv > 4) F[4] = cst<T>(0.0);
    if (!visc) return;
    T L[5], Rr[5];
    for (int k = 0; k < nv; ++k) {
      L[k] = Q(k, il, jl);
      Rr[k] = Q(k, ir, jr);
    }
    const double d1x = cx(ir, jr) - cx(il, jl), d1y = cy(ir, jr) - cy(il, jl);
    const double d2x = nx_(I1, J1) - nx_(I0, J0), d2y = ny_(I1, J1) - ny_(I0, J0);
    const double det = d1x * d2y - d2x * d1y;
    T phL[4] = {L[1], L[2], temperature(L, g), sa ? L[4] : cst<T>(0.0)};
    T phR[4] = {Rr[1], Rr[2], temperature(Rr, g), sa ? Rr[4] : cst<T>(0.0)};
    T gx[4], gy[4];
    for (int p = 0; p < nphi; ++p) {
      const T dp1 = phR[p] - phL[p];
      const T dp2 = node(p, I1, J1) - node(p, I0, J0);
      gx[p] = (dp1 * d2y - dp2 * d1y) / det;
      gy[p] = (d1x * dp2 - d2x * dp1) / det;
    }
    const T gv[2][2] = {{gx[0], gy[0]}, {gx[1], gy[1]}};
    const T vel0 = 0.5 * (L[1] + Rr[1]), vel1 = 0.5 * (L[2] + Rr[2]);
    const T Tf = 0.5 * (temperature(L, g) + temperature(Rr, g));
    T mu = viscosity(Tf, c.ph.mu_inf, c.ph.sutherland);
    T k_ = mu / (PR * (g - 1.0));
    T extra[2][2] = {{cst<T>(0.0), cst<T>(0.0)}, {cst<T>(0.0), cst<T>(0.0)}};
    bool has_extra = false;
    if (sa) {
      T fc[5];
      for (int k = 0; k < 5; ++k) fc[k] = 0.5 * (L[k] + Rr[k]);
      const T muf = viscosity(temperature(fc, g), c.ph.mu_inf, c.ph.sutherland);
      const T mut = eddy_viscosity(fc[0], fc[4] * c.ph.mu_inf, muf / fc[0]);
      mu = mu + mut;
      k_ = k_ + mut / (PR_T * (g - 1.0));
      if (c.ph.c_cr1 != 0.0) {
        has_extra = true;
        const T sq = gv[0][0] * gv[0][0] + gv[0][1] * gv[0][1] + gv[1][0] * gv[1][0] + gv[1][1] * gv[1][1];
        const bool pos = val(sq) > 0.0;
        const T nrm = pos ? sqrt_(sq) : cst<T>(1.0);
        T rot[2][2], tau[2][2];
        const T dv = gv[0][0] + gv[1][1];
        for (int a_ = 0; a_ < 2; ++a_)
          for (int b = 0; b < 2; ++b) {
            rot[a_][b] = (gv[a_][b] - gv[b][a_]) / nrm;
            tau[a_][b] = mut * (gv[a_][b] + gv[b][a_] - (2.0 / 3.0) * dv * (a_ == b ? 1.0 : 0.0));
          }
        T prod[2][2];
        for (int i2 = 0; i2 < 2; ++i2)
          for (int j2 = 0; j2 < 2; ++j2) prod[i2][j2] = rot[i2][0] * tau[j2][0] + rot[i2][1] * tau[j2][1];
        for (int a_ = 0; a_ < 2; ++a_)
          for (int b = 0; b < 2; ++b) extra[a_][b] = -c.ph.c_cr1 * (prod[a_][b] + prod[b][a_]);
      }
      // SA face flux: upwind nu_hat on the mass flux, minus the diffusion
      const T up = val(Fl[0]) > 0.0 ? L[4] : Rr[4];
      const T coef = diffusion_coefficient(fc[0], fc[4] * c.ph.mu_inf,
                                           viscosity(temperature(fc, g), c.ph.mu_inf, c.ph.sutherland));
      F[4] = Fl[0] * up - coef * (gx[3] * sx + gy[3] * sy);
    }
    const T dv = gv[0][0] + gv[1][1];
    const T ttd = (2.0 / 3.0) * dv;
    auto tau = [&](int a_, int b) {
      T lin = a_ == b ? mu * (2.0 * gv[a_][a_] - ttd) : mu * (gv[std::min(a_, b)][std::max(a_, b)] +
                                                              gv[std::max(a_, b)][std::min(a_, b)]);
      return has_extra ? lin + extra[a_][b] : lin;
    };
    const T f0 = tau(0, 0) * sx + tau(0, 1) * sy, f1 = tau(1, 0) * sx + tau(1, 1) * sy;
    const T energy = (vel0 * f0 + vel1 * f1) + k_ * (gx[2] * sx + gy[2] * sy);
    F[1] = F[1] - f0;
    F[2] = F[2] - f1;
    F[3] = F[3] - energy;
    F[0] = F[0] - 0.0 * energy;
  };
  const double* Si = c.g.face_i;
  const double* Sj = c.g.face_j;
#pragma omp parallel for schedule(static)
  for (int f = 0; f < nfi; ++f) {
    const int i = f / NJ, j = f - i * NJ;
    T F[5];
    face(0, f, i - 1, j, i, j, i, j, i, j, i, j + 1, Si[f], Si[nfi + f], F);
    for (int k = 0; k < nv; ++k) w.Fi[(size_t)k * nfi + f] = F[k];
  }
#pragma omp parallel for schedule(static)
  for (int f = 0; f < nfj; ++f) {
    const int i = f / (NJ + 1), j = f - i * (NJ + 1);
    T F[5];
    face(1, f, i, j - 1, i, j, i, j, i, j, i + 1, j, Sj[f], Sj[nfj + f], F);
    for (int k = 0; k < nv; ++k) w.Fj[(size_t)k * nfj + f] = F[k];
  }
  // cells: divergence, the SA source, U - U0 off the blocks
#pragma omp parallel for schedule(static)
  for (int cell = 0; cell < nc; ++cell) {
    const int i = cell / NJ, j = cell - i * NJ;
    const int a0 = i * NJ + j, a1 = a0 + NJ, b0 = i * (NJ + 1) + j, b1 = b0 + 1;
    if (!c.g.interior[cell]) {
      for (int k = 0; k < nv; ++k) {
        const double u = U[(size_t)k * nc + cell];
        R_out[(size_t)k * nc + cell] =
            sizeof(T) == sizeof(double) ? u - c.U0[(size_t)k * nc + cell] : (dU ? dU[(size_t)k * nc + cell] : 0.0);
      }
      continue;
    }
    T r[5];
    for (int k = 0; k < nv; ++k)
      r[k] = (w.Fi[(size_t)k * nfi + a1] - w.Fi[(size_t)k * nfi + a0]) +
             (w.Fj[(size_t)k * nfj + b1] - w.Fj[(size_t)k * nfj + b0]);
    if (sa) {
      T q[5];
      for (int k = 0; k < nv; ++k) q[k] = w.q[(size_t)k * nc + cell];
      const T rho = q[0];
      const T mu = viscosity(temperature(q, g), c.ph.mu_inf, c.ph.sutherland);
      // face averages of the four faces (canvas padded by one, edge)
      auto fav = [&](int p, int il, int jl, int ir, int jr) { return 0.5 * (Q(p, il, jl) + Q(p, ir, jr)); };
      const double V = c.g.area[cell];
      auto gg = [&](int p, int dir) {
        const T w_ = fav(p, i - 1, j, i, j), e_ = fav(p, i, j, i + 1, j);
        const T s_ = fav(p, i, j - 1, i, j), n_ = fav(p, i, j, i, j + 1);
        return ((e_ * Si[(size_t)dir * nfi + a1] - w_ * Si[(size_t)dir * nfi + a0]) +
                (n_ * Sj[(size_t)dir * nfj + b1] - s_ * Sj[(size_t)dir * nfj + b0])) / V;
      };
      const T gu1 = gg(1, 1), gv0 = gg(2, 0);
      const T omega = abs_(gv0 - gu1);
      T prod, dest;
      source(q[4] * c.ph.mu_inf, mu / rho, omega, c.g.distance[cell], &prod, &dest);
      const T gn0 = gg(4, 0), gn1 = gg(4, 1);
      const T src = rho * ((prod - dest) / c.ph.mu_inf + CB2 / SIGMA * c.ph.mu_inf * (gn0 * gn0 + gn1 * gn1));
      r[4] = r[4] - src * V;
    }
    for (int k = 0; k < nv; ++k) {
      if constexpr (sizeof(T) == sizeof(double)) R_out[(size_t)k * nc + cell] = val(r[k]);
      else R_out[(size_t)k * nc + cell] = r[k].d;
    }
  }
}

// --- wave speeds and the line preconditioner ----------------------------------------------
void prim_cell(const Ctx& c, const double* U, int cell, double* q) {
  double u[5];
  for (int k = 0; k < c.nv; ++k) u[k] = U[(size_t)k * c.nc + cell];
  primitive(u, c.nv, c.ph.gamma, q);
}

// neighbour-filled face averages of the primitives (atlas.py:163-183)
void face_avg(const Ctx& c, const double* U, int a, int f, double* qf, double* vol) {
  const int NI = c.NI, NJ = c.NJ;
  int c1, c2;
  if (a == 0) {
    const int i = f / NJ, j = f - i * NJ;
    c1 = clampi(i - 1, 0, NI - 1) * NJ + j;
    c2 = clampi(i, 0, NI - 1) * NJ + j;
  } else {
    const int i = f / (NJ + 1), j = f - i * (NJ + 1);
    c1 = i * NJ + clampi(j - 1, 0, NJ - 1);
    c2 = i * NJ + clampi(j, 0, NJ - 1);
  }
  c1 = (int)c.g.neighbours[c1];
  c2 = (int)c.g.neighbours[c2];
  double q1[5], q2[5];
  prim_cell(c, U, c1, q1);
  prim_cell(c, U, c2, q2);
  for (int k = 0; k < c.nv; ++k) qf[k] = 0.5 * (q1[k] + q2[k]);
  *vol = 0.5 * (c.g.area[c1] + c.g.area[c2]);
}

void wave_speeds(Ctx& c, const double* U) {
  const double g = c.ph.gamma;
  c.lam.resize((size_t)c.nfi + c.nfj);
#pragma omp parallel for schedule(static)
  for (int t = 0; t < c.nfi + c.nfj; ++t) {
    const int a = t < c.nfi ? 0 : 1, f = a == 0 ? t : t - c.nfi;
    const double* S = a == 0 ? c.g.face_i : c.g.face_j;
    const int nf = a == 0 ? c.nfi : c.nfj;
    const double sx = S[f], sy = S[nf + f];
    double qf[5], vol;
    face_avg(c, U, a, f, qf, &vol);
    double lam = std::fabs(qf[1] * sx + qf[2] * sy) + std::sqrt(g * qf[3] / qf[0]) * hypot_j(sx, sy);
    if (c.ph.viscous) {
      double mu = viscosity(g * qf[3] / qf[0], c.ph.mu_inf, c.ph.sutherland);
      if (c.ph.sa) mu = mu + eddy_viscosity(qf[0], qf[4] * c.ph.mu_inf, mu / qf[0]);
      lam = lam + (2.0 * g / PR) * mu * (sx * sx + sy * sy) / (qf[0] * vol);
    }
    c.lam[t] = lam;
  }
  c.lsum.resize(c.nc);
  const int NJ = c.NJ;
#pragma omp parallel for schedule(static)
  for (int cell = 0; cell < c.nc; ++cell) {
    const int i = cell / NJ, j = cell - i * NJ;
    const double* l0 = c.lam.data();
    const double* l1 = c.lam.data() + c.nfi;
    c.lsum[cell] = ((l0[i * NJ + j] + l0[(i + 1) * NJ + j]) + l1[i * (NJ + 1) + j]) + l1[i * (NJ + 1) + j + 1];
  }
}

// d(F.S)/dU of the analytic flux at q (euler.py:55-71), columns by forward differentiation
void flux_jacobian(const double* q, double sx, double sy, double g, double A[4][4]) {
  double U[4];
  conservative(q, 4, g, U);
  for (int k = 0; k < 4; ++k) {
    D u[4], p[4], f[4];
    for (int r = 0; r < 4; ++r) u[r] = mk(U[r], r == k ? 1.0 : 0.0);
    primitive(u, 4, g, p);
    normal_flux(p, sx, sy, g, f);
    for (int r = 0; r < 4; ++r) A[r][k] = f[r].d;
  }
}

void inv4(const double* A, double* X) {   // unpivoted Gauss-Jordan on [A | I] (linalg.py:49-77)
  double W[4][8];
  for (int r = 0; r < 4; ++r)
    for (int c2 = 0; c2 < 4; ++c2) {
      W[r][c2] = A[r * 4 + c2];
      W[r][4 + c2] = r == c2 ? 1.0 : 0.0;
    }
  for (int k = 0; k < 4; ++k) {
    const double piv = W[k][k];
    for (int c2 = 0; c2 < 8; ++c2) W[k][c2] = W[k][c2] / piv;
    for (int r = 0; r < 4; ++r) {
      if (r == k) continue;
      const double f = W[r][k];
      for (int c2 = 0; c2 < 8; ++c2) W[r][c2] = W[r][c2] - f * W[k][c2];
    }
  }
  for (int r = 0; r < 4; ++r)
    for (int c2 = 0; c2 < 4; ++c2) X[r * 4 + c2] = W[r][4 + c2];
}

void precond_setup(Ctx& c, const double* U, double cfl) {
  wave_speeds(c, U);
  const double g = c.ph.gamma;
  const int nl = c.g.n_lines, NJ = c.NJ;
  c.fac.assign((size_t)c.g.n_slots * 48, 0.0);
  c.fac_sa.assign((size_t)c.g.n_slots * 3, 0.0);
  const double* Sj = c.g.face_j;
  const double* lamj = c.lam.data() + c.nfi;
  // SA chain terms per j face and cell diagonal (preconditioners.py:235-287)
  std::vector<double> mj, dj, sadiag;
  if (c.ph.sa) {
    mj.resize(c.nfj);
    dj.resize(c.nfj);
    std::vector<double> mi(c.nfi), di(c.nfi);
    auto terms = [&](int a, int f, double* m, double* d) {
      const double* S = a == 0 ? c.g.face_i : c.g.face_j;
      const int nf = a == 0 ? c.nfi : c.nfj;
      double qf[5], vol;
      face_avg(c, U, a, f, qf, &vol);
      const double mu = viscosity(g * qf[3] / qf[0], c.ph.mu_inf, c.ph.sutherland);
      const double kap = diffusion_coefficient(qf[0], qf[4] * c.ph.mu_inf, mu);
      const double sx = S[f], sy = S[nf + f];
      *m = qf[0] * (qf[1] * sx + qf[2] * sy);
      *d = kap * (sx * sx + sy * sy) / vol;
    };
#pragma omp parallel for schedule(static)
    for (int f = 0; f < c.nfi; ++f) terms(0, f, &mi[f], &di[f]);
#pragma omp parallel for schedule(static)
    for (int f = 0; f < c.nfj; ++f) terms(1, f, &mj[f], &dj[f]);
    sadiag.resize(c.nc);
#pragma omp parallel for schedule(static)
    for (int cell = 0; cell < c.nc; ++cell) {
      const int i = cell / NJ, j = cell - i * NJ;
      const int a0 = i * NJ + j, a1 = a0 + NJ, b0 = i * (NJ + 1) + j, b1 = b0 + 1;
      double diag = ((std::max(mi[a1], 0.0) + di[a1]) + (std::max(-mi[a0], 0.0) + di[a0])) +
                    (std::max(mj[b1], 0.0) + dj[b1]) + (std::max(-mj[b0], 0.0) + dj[b0]);
      double q[5];
      prim_cell(c, U, cell, q);
      diag = diag / q[0];
      // vorticity by Green-Gauss over the neighbour-filled face averages
      double fa[4][5], vv;
      face_avg(c, U, 0, a0, fa[0], &vv);
      face_avg(c, U, 0, a1, fa[1], &vv);
      face_avg(c, U, 1, b0, fa[2], &vv);
      face_avg(c, U, 1, b1, fa[3], &vv);
      const double* Si = c.g.face_i;
      const int nfi = c.nfi, nfj = c.nfj;
      const double V = c.g.area[cell];
      auto gg = [&](int p, int dir) {
        return ((fa[1][p] * Si[(size_t)dir * nfi + a1] - fa[0][p] * Si[(size_t)dir * nfi + a0]) +
                (fa[3][p] * Sj[(size_t)dir * nfj + b1] - fa[2][p] * Sj[(size_t)dir * nfj + b0])) / V;
      };
      const double omega = std::fabs(gg(2, 0) - gg(1, 1));
      const double nu = viscosity(g * q[3] / q[0], c.ph.mu_inf, c.ph.sutherland) / q[0];
      const double dw = c.g.distance[cell], nh = q[4];
      const double st = s_tilde(omega, nh * c.ph.mu_inf, nu, dw);
      const double fw = f_w(r_of(nh * c.ph.mu_inf, st, dw));
      const double destr = nh > 0.0 ? 2.0 * CW1 * fw * c.ph.mu_inf * nh / (dw * dw) : 0.0;
      sadiag[cell] = diag + destr * V + c.lsum[cell] / cfl;
    }
  }
#pragma omp parallel for schedule(dynamic, 4)
  for (int ch = 0; ch < nl; ++ch) {
    const int s0 = (int)c.g.line_start[ch], s1 = (int)c.g.line_start[ch + 1], n = s1 - s0;
    double Cp[16] = {0};
    double cps = 0.0;
    for (int k = 0; k < n; ++k) {
      const int s = s0 + k;
      const int cell = (int)c.g.line_cells[s];
      double diag = c.lsum[cell] * (1.0 / cfl + 0.5);
      double lower[16] = {0}, upper[16] = {0};
      auto coupling = [&](int nb_cell, int64_t fce, int64_t sg, double* blk) {
        if (sg == 0) return;
        double q[5];
        prim_cell(c, U, nb_cell, q);
        double A[4][4];
        flux_jacobian(q, sg * Sj[fce], sg * Sj[c.nfj + fce], g, A);
        for (int r = 0; r < 4; ++r)
          for (int cc = 0; cc < 4; ++cc) blk[r * 4 + cc] = 0.5 * A[r][cc] - (r == cc ? 0.5 * lamj[fce] : 0.0);
      };
      if (k > 0) coupling((int)c.g.line_cells[s - 1], c.g.line_down_face[s - 1], c.g.line_down_sign[s - 1], lower);
      if (k + 1 < n) coupling((int)c.g.line_cells[s + 1], c.g.line_up_face[s], c.g.line_up_sign[s], upper);
      else coupling(cell, c.g.line_up_face[s], c.g.line_up_sign[s], upper);
      // block Thomas forward factor (linalg.py:80-100)
      double Pm[16], Pi[16], Cn[16];
      for (int r = 0; r < 4; ++r)
        for (int q2 = 0; q2 < 4; ++q2) {
          double acc = 0.0;
          for (int t = 0; t < 4; ++t) acc += lower[r * 4 + t] * Cp[t * 4 + q2];
          Pm[r * 4 + q2] = (r == q2 ? diag : 0.0) - acc;
        }
      inv4(Pm, Pi);
      for (int r = 0; r < 4; ++r)
        for (int q2 = 0; q2 < 4; ++q2) {
          double acc = 0.0;
          for (int t = 0; t < 4; ++t) acc += Pi[r * 4 + t] * upper[t * 4 + q2];
          Cn[r * 4 + q2] = acc;
        }
      double* o = &c.fac[(size_t)s * 48];
      for (int e = 0; e < 16; ++e) {
        o[e] = lower[e];
        o[16 + e] = Pi[e];
        o[32 + e] = Cn[e];
        Cp[e] = Cn[e];
      }
      if (c.ph.sa) {
        auto scoup = [&](int nb_cell, int64_t fce, int64_t sg) {
          if (sg == 0) return 0.0;
          return -(std::max(-sg * mj[fce], 0.0) + dj[fce]) / U[nb_cell];
        };
        const double lo = k > 0 ? scoup((int)c.g.line_cells[s - 1], c.g.line_down_face[s - 1],
                                        c.g.line_down_sign[s - 1]) : 0.0;
        const double up = k + 1 < n ? scoup((int)c.g.line_cells[s + 1], c.g.line_up_face[s], c.g.line_up_sign[s])
                                    : scoup(cell, c.g.line_up_face[s], c.g.line_up_sign[s]);
        const double pinv = 1.0 / (sadiag[cell] - lo * cps);
        const double cn = pinv * up;
        double* os = &c.fac_sa[(size_t)s * 3];
        os[0] = lo;
        os[1] = pinv;
        os[2] = cn;
        cps = cn;
      }
    }
  }
}

// z = M^-1 r: block Thomas along every chain, the identity on cells on no chain
void precond_apply(Ctx& c, const double* r, double* z) {
  const int nc = c.nc, nv = c.nv;
  std::memcpy(z, r, sizeof(double) * (size_t)nv * nc);
  const int nl = c.g.n_lines;
#pragma omp parallel for schedule(dynamic, 4)
  for (int ch = 0; ch < nl; ++ch) {
    const int s0 = (int)c.g.line_start[ch], s1 = (int)c.g.line_start[ch + 1], n = s1 - s0;
    std::vector<double> d((size_t)n * 5);
    double prev[4] = {0, 0, 0, 0}, prevs = 0.0;
    for (int k = 0; k < n; ++k) {
      const int s = s0 + k, cell = (int)c.g.line_cells[s];
      const double* o = &c.fac[(size_t)s * 48];
      double t[4];
      for (int i = 0; i < 4; ++i) {
        double acc = 0.0;
        for (int q = 0; q < 4; ++q) acc += o[i * 4 + q] * prev[q];
        t[i] = r[(size_t)i * nc + cell] - acc;
      }
      for (int i = 0; i < 4; ++i) {
        double acc = 0.0;
        for (int q = 0; q < 4; ++q) acc += o[16 + i * 4 + q] * t[q];
        d[(size_t)k * 5 + i] = acc;
      }
      for (int i = 0; i < 4; ++i) prev[i] = d[(size_t)k * 5 + i];
      if (nv > 4) {
        const double* os = &c.fac_sa[(size_t)s * 3];
        prevs = os[1] * (r[(size_t)4 * nc + cell] - os[0] * prevs);
        d[(size_t)k * 5 + 4] = prevs;
      }
    }
    double nxt[4] = {0, 0, 0, 0}, nxts = 0.0;
    for (int k = n - 1; k >= 0; --k) {
      const int s = s0 + k, cell = (int)c.g.line_cells[s];
      const double* o = &c.fac[(size_t)s * 48];
      double x[4];
      for (int i = 0; i < 4; ++i) {
        double acc = 0.0;
        for (int q = 0; q < 4; ++q) acc += o[32 + i * 4 + q] * nxt[q];
        x[i] = d[(size_t)k * 5 + i] - acc;
      }
      for (int i = 0; i < 4; ++i) {
        nxt[i] = x[i];
        z[(size_t)i * nc + cell] = x[i];
      }
      if (nv > 4) {
        nxts = d[(size_t)k * 5 + 4] - c.fac_sa[(size_t)s * 3 + 2] * nxts;
        z[(size_t)4 * nc + cell] = nxts;
      }
    }
  }
}

// --- GMRES and the Newton step --------------------------------------------------------------
double dot(const double* a, const double* b, size_t n) {
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (size_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

double rnorm(const Ctx& c, const double* R) {   // density row, volume weighted (residual.py:237-240)
  double a = 0.0, b = 0.0;
#pragma omp parallel for reduction(+ : a, b) schedule(static)
  for (int cell = 0; cell < c.nc; ++cell)
    if (c.g.interior[cell]) {
      a += c.g.area[cell] * R[cell] * R[cell];
      b += c.g.area[cell];
    }
  return std::sqrt(a / b);
}

// Givens least squares with pivot drop (krylov.py:100-137); H (m+1) x m row-major
void lstsq(const std::vector<double>& Hin, int m, double beta, double rcond, double* y) {
  std::vector<double> h(Hin);
  std::vector<double> g(m + 1, 0.0), cs(m), sn(m);
  g[0] = beta;
  auto H = [&](int i, int j) -> double& { return h[(size_t)i * m + j]; };
  for (int j = 0; j < m; ++j) {
    for (int i = 0; i < j; ++i) {
      const double a = H(i, j), cc = H(i + 1, j);
      H(i, j) = cs[i] * a + sn[i] * cc;
      H(i + 1, j) = cs[i] * cc - sn[i] * a;
    }
    const double a = H(j, j), cc = H(j + 1, j);
    const double r = std::sqrt(a * a + cc * cc);
    cs[j] = r > 0 ? a / r : 1.0;
    sn[j] = r > 0 ? cc / r : 0.0;
    H(j, j) = r;
    H(j + 1, j) = 0.0;
    const double gj = g[j];
    g[j] = cs[j] * gj;
    g[j + 1] = -sn[j] * gj;
  }
  double largest = 0.0;
  for (int j = 0; j < m; ++j) largest = std::max(largest, std::fabs(H(j, j)));
  for (int j = m - 1; j >= 0; --j) {
    double rest = g[j];
    for (int k = j + 1; k < m; ++k) rest = rest - H(j, k) * y[k];
    y[j] = std::fabs(H(j, j)) > rcond * largest ? rest / H(j, j) : 0.0;
  }
}

}  // namespace oc

using namespace oc;

extern "C" {

void* oc_create(const oc_grid* g, const oc_phys* ph) {
  Ctx* c = new Ctx();
  c->g = *g;
  c->ph = *ph;
  c->NI = g->NI;
  c->NJ = g->NJ;
  c->nc = g->NI * g->NJ;
  c->nv = ph->sa ? 5 : 4;
  c->nfi = (g->NI + 1) * g->NJ;
  c->nfj = g->NI * (g->NJ + 1);
  double u0[5];
  conservative(ph->q_inf, c->nv, ph->gamma, u0);
  c->U0.resize((size_t)c->nv * c->nc);
  for (int k = 0; k < c->nv; ++k)
    for (int i = 0; i < c->nc; ++i) c->U0[(size_t)k * c->nc + i] = u0[k];
  return c;
}

void oc_destroy(void* h) { delete static_cast<Ctx*>(h); }

int oc_threads(void) { return omp_get_max_threads(); }

void oc_residual(void* h, const double* U, double* R, int first_order) {
  residual<double>(*static_cast<Ctx*>(h), U, nullptr, R, first_order);
}

void oc_jvp(void* h, const double* U, const double* dU, double* out, int first_order) {
  residual<D>(*static_cast<Ctx*>(h), U, dU, out, first_order);
}

void oc_wave_speeds(void* h, const double* U, double* lam) {
  Ctx& c = *static_cast<Ctx*>(h);
  wave_speeds(c, U);
  std::memcpy(lam, c.lam.data(), sizeof(double) * c.lam.size());
}

void oc_precond_setup(void* h, const double* U, double cfl) { precond_setup(*static_cast<Ctx*>(h), U, cfl); }

void oc_precond_apply(void* h, const double* r, double* z) { precond_apply(*static_cast<Ctx*>(h), r, z); }

// One pseudo-transient Newton-Krylov step (implicit.py:99-166) on U in place;
// state = {cfl, |R|_prev}; out = {|R(U)|, accepted, linear reduction, omega}.
void oc_newton_step(void* h, double* U, double* state, double* out) {
  Ctx& c = *static_cast<Ctx*>(h);
  const int nc = c.nc, nv = c.nv, m = c.ph.krylov_subspace;
  const size_t N = (size_t)nv * nc;
  const double cfl = state[0], g = c.ph.gamma;
  std::vector<double> R(N), b(N), x(N, 0.0), act(N, 0.0), r(N), w(N), z(N), V((size_t)(m + 1) * N);
  residual<double>(c, U, nullptr, R.data(), 0);
  const double nrm = rnorm(c, R.data());
  precond_setup(c, U, cfl);
  std::vector<double> vdt(nc);
  for (int i = 0; i < nc; ++i) vdt[i] = c.lsum[i] / cfl;
  const int fo = c.ph.exact ? 0 : 1;
  auto matvec = [&](const double* in, double* o) {
    residual<D>(c, U, in, o, fo);
#pragma omp parallel for schedule(static)
    for (size_t e = 0; e < N; ++e) o[e] = vdt[e % nc] * in[e] + o[e];
  };
  for (size_t e = 0; e < N; ++e) b[e] = -R[e];
  const double bn = std::sqrt(dot(b.data(), b.data(), N));
  const double tiny = 2.2250738585072014e-308, rcond = 2.220446049250313e-16 * (m + 1);
  r = b;
  for (int cyc = 0; cyc < c.ph.krylov_restarts && std::sqrt(dot(r.data(), r.data(), N)) > c.ph.krylov_tol * bn; ++cyc) {
    const double beta = std::sqrt(dot(r.data(), r.data(), N));
    std::vector<double> H((size_t)(m + 1) * m, 0.0);
    for (size_t e = 0; e < N; ++e) V[e] = r[e] / std::max(beta, tiny);
    for (int j = 0; j < m; ++j) {
      precond_apply(c, &V[(size_t)j * N], z.data());
      matvec(z.data(), w.data());
      std::vector<double> hh(j + 1), h2(j + 1);
      for (int i = 0; i <= j; ++i) hh[i] = dot(&V[(size_t)i * N], w.data(), N);
#pragma omp parallel for schedule(static)
      for (size_t e = 0; e < N; ++e) {
        double s = 0.0;
        for (int i = 0; i <= j; ++i) s += hh[i] * V[(size_t)i * N + e];
        w[e] -= s;
      }
      for (int i = 0; i <= j; ++i) h2[i] = dot(&V[(size_t)i * N], w.data(), N);
#pragma omp parallel for schedule(static)
      for (size_t e = 0; e < N; ++e) {
        double s = 0.0;
        for (int i = 0; i <= j; ++i) s += h2[i] * V[(size_t)i * N + e];
        w[e] -= s;
      }
      const double nw = std::sqrt(dot(w.data(), w.data(), N));
      for (int i = 0; i <= j; ++i) H[(size_t)i * m + j] = hh[i] + h2[i];
      H[(size_t)(j + 1) * m + j] = nw;
      for (size_t e = 0; e < N; ++e) V[(size_t)(j + 1) * N + e] = w[e] / std::max(nw, tiny);
    }
    std::vector<double> y(m, 0.0), Hy(m + 1, 0.0);
    lstsq(H, m, beta, rcond, y.data());
    for (int i = 0; i <= m; ++i)
      for (int j = 0; j < m; ++j) Hy[i] += H[(size_t)i * m + j] * y[j];
#pragma omp parallel for schedule(static)
    for (size_t e = 0; e < N; ++e) {
      double s = 0.0, a = 0.0;
      for (int j = 0; j < m; ++j) s += y[j] * V[(size_t)j * N + e];
      for (int i = 0; i <= m; ++i) a += Hy[i] * V[(size_t)i * N + e];
      w[e] = s;
      act[e] += a;
    }
    precond_apply(c, w.data(), z.data());
    for (size_t e = 0; e < N; ++e) {
      x[e] += z[e];
      r[e] = b[e] - act[e];
    }
  }
  // relaxation (implicit.py:54-81)
  double omega = 1.0;
  if (c.ph.max_relative_update > 0.0) {
    double worst = 0.0;
#pragma omp parallel for reduction(max : worst) schedule(static)
    for (int cell = 0; cell < nc; ++cell) {
      double q[5];
      prim_cell(c, U, cell, q);
      const double du0 = x[cell], du1 = x[nc + cell], du2 = x[2 * (size_t)nc + cell], du3 = x[3 * (size_t)nc + cell];
      const double de = du3 - q[1] * du1 - q[2] * du2;
      const double dp = (g - 1.0) * (de + 0.5 * (q[1] * q[1] + q[2] * q[2]) * du0);
      double r1 = std::fabs(du0) / q[0], r2 = std::fabs(dp) / q[3];
      double wv = std::max(r1, r2);
      if (nv > 4) {
        const double nh = U[4 * (size_t)nc + cell] / q[0];
        wv = std::max(wv, std::fabs((x[4 * (size_t)nc + cell] - nh * du0) / q[0]) / std::max(std::fabs(nh), 1.0));
      }
      worst = std::max(worst, wv);
    }
    omega = std::min(1.0, c.ph.max_relative_update / std::max(worst, 1e-300));
  }
  double lr = 0.0;
  for (size_t e = 0; e < N; ++e) {
    const double v = omega * act[e] + R[e];
    lr += v * v;
  }
  const double lin_red = std::sqrt(lr) / std::max(std::sqrt(dot(R.data(), R.data(), N)), 1e-300);
  bool physical = true;
  for (int cell = 0; cell < nc && physical; ++cell) {
    double u[5], q[5];
    for (int k = 0; k < nv; ++k) u[k] = U[(size_t)k * nc + cell] + omega * x[(size_t)k * nc + cell];
    primitive(u, nv, g, q);
    physical = q[0] > MIN_DENSITY && q[3] > MIN_PRESSURE && std::isfinite(q[0]) && std::isfinite(q[3]);
  }
  for (size_t e = 0; e < N && physical; ++e) physical = std::isfinite(x[e]);
  const bool accept = physical && lin_red < 1.0;
  const double reduction = state[1] / std::max(nrm, 1e-300);
  double nxt;
  if (accept && lin_red < c.ph.krylov_success && omega >= 1.0) nxt = cfl * std::min(std::max(reduction, 1.0), c.ph.cfl_growth);
  else if (accept) nxt = cfl;
  else nxt = cfl / c.ph.cfl_cutback;
  state[0] = std::min(std::max(nxt, c.ph.cfl_min), c.ph.cfl_max);
  if (accept) {
    for (size_t e = 0; e < N; ++e) U[e] += omega * x[e];
    state[1] = nrm;
  }
  out[0] = nrm;
  out[1] = accept;
  out[2] = lin_red;
  out[3] = omega;
}

}  // extern "C"
