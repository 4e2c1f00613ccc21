// Scalar types of the device code: double, and a forward-mode dual number
// (value, one tangent) so the residual source doubles as its exact JVP.
//
// The non-smooth primitives follow JAX's JVP rules, which define the
// reference's Krylov operator (src/jaxrans/solvers/implicit.py:105-108
// linearizes the residual with jax.linearize):
//   jmin/jmax  derivative of the selected argument, half of each at a tie;
//   jabs       +d where x >= 0, -d where x < 0;
//   sel        tangent of the selected branch only.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace jr {

struct Dual {
  double v, d;
};

__device__ __forceinline__ Dual mkd(double v, double d) { Dual r; r.v = v; r.d = d; return r; }

__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return mkd(a.v + b.v, a.d + b.d); }
__device__ __forceinline__ Dual operator+(Dual a, double b) { return mkd(a.v + b, a.d); }
__device__ __forceinline__ Dual operator+(double a, Dual b) { return mkd(a + b.v, b.d); }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return mkd(a.v - b.v, a.d - b.d); }
__device__ __forceinline__ Dual operator-(Dual a, double b) { return mkd(a.v - b, a.d); }
__device__ __forceinline__ Dual operator-(double a, Dual b) { return mkd(a - b.v, -b.d); }
__device__ __forceinline__ Dual operator-(Dual a) { return mkd(-a.v, -a.d); }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) { return mkd(a.v * b.v, a.d * b.v + a.v * b.d); }
__device__ __forceinline__ Dual operator*(Dual a, double b) { return mkd(a.v * b, a.d * b); }
__device__ __forceinline__ Dual operator*(double a, Dual b) { return mkd(a * b.v, a * b.d); }
// Division through one reciprocal: an FP64 division expands to about ten
// instructions, a multiply to one.  The value differs from a / b by at most
// an ulp, which the tangent kernels only use for their own coefficients.
__device__ __forceinline__ Dual operator/(Dual a, Dual b) {
  const double r = 1.0 / b.v, q = a.v * r;
  return mkd(q, (a.d - q * b.d) * r);
}
__device__ __forceinline__ Dual operator/(Dual a, double b) {
  const double r = 1.0 / b;
  return mkd(a.v * r, a.d * r);
}
__device__ __forceinline__ Dual operator/(double a, Dual b) {
  const double r = 1.0 / b.v, q = a * r;
  return mkd(q, -q * b.d * r);
}

__device__ __forceinline__ double val(double x) { return x; }
__device__ __forceinline__ double val(Dual x) { return x.v; }
__device__ __forceinline__ double tan_of(double) { return 0.0; }
__device__ __forceinline__ double tan_of(Dual x) { return x.d; }

template <class T> __device__ __forceinline__ T lift(double x);
template <> __device__ __forceinline__ double lift<double>(double x) { return x; }
template <> __device__ __forceinline__ Dual lift<Dual>(double x) { return mkd(x, 0.0); }

__device__ __forceinline__ double jsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ Dual jsqrt(Dual x) {
  double r = sqrt(x.v);
  return mkd(r, x.d * (0.5 / r));
}

__device__ __forceinline__ double jabs(double x) { return fabs(x); }
__device__ __forceinline__ Dual jabs(Dual x) { return mkd(fabs(x.v), x.v >= 0.0 ? x.d : -x.d); }

// jnp.minimum / jnp.maximum with JAX's balanced tie rule for the tangent.
__device__ __forceinline__ double jmin(double a, double b) { return a < b ? a : (b < a ? b : a); }
__device__ __forceinline__ double jmax(double a, double b) { return a > b ? a : (b > a ? b : a); }
__device__ __forceinline__ Dual jmin(Dual a, Dual b) {
  if (a.v < b.v) return a;
  if (b.v < a.v) return b;
  return mkd(a.v, 0.5 * (a.d + b.d));
}
__device__ __forceinline__ Dual jmax(Dual a, Dual b) {
  if (a.v > b.v) return a;
  if (b.v > a.v) return b;
  return mkd(a.v, 0.5 * (a.d + b.d));
}
__device__ __forceinline__ Dual jmin(Dual a, double b) { return jmin(a, mkd(b, 0.0)); }
__device__ __forceinline__ Dual jmax(Dual a, double b) { return jmax(a, mkd(b, 0.0)); }

template <class T> __device__ __forceinline__ T sel(bool c, T a, T b) { return c ? a : b; }

// x ** c for a constant real exponent.
__device__ __forceinline__ double jpow(double x, double c) { return pow(x, c); }
__device__ __forceinline__ Dual jpow(Dual x, double c) {
  return mkd(pow(x.v, c), c * pow(x.v, c - 1.0) * x.d);
}
// Trigonometry for the freestream's parameter tangent (the point-vortex
// reference turns with the flow angle): atan2's derivative
// (x dy - y dx) / (x^2 + y^2), sin' = cos, cos' = -sin.
__device__ __forceinline__ double jsin(double x) { return sin(x); }
__device__ __forceinline__ Dual jsin(Dual x) { return mkd(sin(x.v), cos(x.v) * x.d); }
__device__ __forceinline__ double jatan2(double y, double x) { return atan2(y, x); }
__device__ __forceinline__ Dual jatan2(Dual y, Dual x) {
  return mkd(atan2(y.v, x.v), (x.v * y.d - y.v * x.d) / (x.v * x.v + y.v * y.v));
}

// The freestream as the kernels read it: plain values (double), or, in the
// parameter-tangent kernels, values carrying their tangent along a flow
// parameter (Dual).  dfree[case][8] = (dq_inf[0..5], dmu_inf, unused).
template <bool PT> struct Free { using type = double; };
template <> struct Free<true> { using type = Dual; };

template <bool PT>
__device__ __forceinline__ typename Free<PT>::type free_val(double v, const double* dfree, int cs, int k) {
  if constexpr (PT) return mkd(v, dfree[(size_t)cs * 8 + k]);
  else return v;
}

// x ** n for a small constant integer exponent (repeated products).
template <class T> __device__ __forceinline__ T ipow2(T x) { return x * x; }
template <class T> __device__ __forceinline__ T ipow3(T x) { return x * x * x; }
template <class T> __device__ __forceinline__ T ipow6(T x) { T x2 = x * x; return x2 * x2 * x2; }

}  // namespace jr
