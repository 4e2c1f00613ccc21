// Wave speeds, the line-implicit preconditioner (block Thomas per chain), the
// deterministic reductions, GMRES(m) with CGS2 Arnoldi, the Newton update and
// the force integration.
//
// Restates src/jaxrans/numerics/residual.py:169-203 (wave speeds),
// solvers/preconditioners.py:197-306 (line preconditioner),
// core/linalg.py:49-132 (Gauss-Jordan inverse, block Thomas),
// solvers/krylov.py:34-137 (GMRES, Givens least squares),
// solvers/implicit.py:54-166 (Newton step), post/forces.py:48-161 (forces).
#pragma once
#include "jr_residual.h"

namespace jr {

// ---------------------------------------------------------------------------
// Deterministic block reductions with the "last block finishes" pattern: each
// workgroup writes its partial sums, the last one to arrive (atomic counter)
// adds them in a fixed order.  Results depend only on the vector length and
// the launch shape, never on which workgroup finishes last.

__device__ __forceinline__ double wave_sum(double x) {
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}
__device__ __forceinline__ double wave_max(double x) {
  for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o, 64));
  return x;
}

// Sum NV values over the workgroup; result valid in thread 0.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double (*sh)[4]) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int k = 0; k < NV; ++k) {
    double s = wave_sum(v[k]);
    if (lane == 0) sh[k][w] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0)
    for (int k = 0; k < NV; ++k) v[k] = ((sh[k][0] + sh[k][1]) + sh[k][2]) + sh[k][3];
  __syncthreads();
}

// Partial sums -> the last workgroup gets the totals in `tot` (all its threads).
// Returns true in every thread of the last workgroup of this case.
template <int NV>
__device__ bool finish_sum(double (&v)[NV], double* partials, unsigned* counter, double (*sh)[4], double* tot,
                           int nvals = NV) {
  __shared__ int is_last;
  block_sum<NV>(v, sh);
  const int nblk = gridDim.x, blk = blockIdx.x;
  if (threadIdx.x == 0) {
    for (int k = 0; k < nvals; ++k) partials[(size_t)blk * NV + k] = v[k];
    __threadfence();
    unsigned prev = atomicAdd(counter, 1u);
    is_last = prev == (unsigned)(nblk - 1);
  }
  __syncthreads();
  if (!is_last) return false;
  __threadfence();
  double acc[NV];
  for (int k = 0; k < NV; ++k) acc[k] = 0.0;
  for (int b = threadIdx.x; b < nblk; b += BLOCK)
    for (int k = 0; k < nvals; ++k) acc[k] += __hip_atomic_load(&partials[(size_t)b * NV + k], __ATOMIC_RELAXED,
                                                               __HIP_MEMORY_SCOPE_AGENT);
  block_sum<NV>(acc, sh);
  __shared__ double out[NV];
  if (threadIdx.x == 0) {
    for (int k = 0; k < NV; ++k) out[k] = acc[k];
    *counter = 0u;
  }
  __syncthreads();
  for (int k = 0; k < NV; ++k) tot[k] = out[k];
  return true;
}

// ---------------------------------------------------------------------------
// Wave speeds (residual.py:169-203) on every face from the neighbour-filled
// face averages, and for SA the face mass flux and diffusion terms of the SA
// chains (preconditioners.py:252-258).

template <int NV>
__device__ __forceinline__ void prim_at(const double* U, size_t base, int nc, int c, double gamma, double* q) {
  double u[NV];
  for (int k = 0; k < NV; ++k) u[k] = U[base + (size_t)k * nc + c];
  primitive<double, NV>(u, gamma, q);
}

template <int NV>
__global__ void __launch_bounds__(BLOCK) k_face_speeds(Plan P, Physics ph, const Ctl* __restrict__ ctl,
                                                       const double* __restrict__ U, double* __restrict__ lam,
                                                       double* __restrict__ sam, double* __restrict__ sad) {
  int f = blockIdx.x * BLOCK + threadIdx.x;
  int cs = blockIdx.y;
  const int nf = P.nfi + P.nfj;
  if (f >= nf) return;
  const int NI = P.NI, NJ = P.NJ, nc = P.nc;
  int c1, c2;
  double sx, sy;
  if (f < P.nfi) {
    int a = f / NJ, b = f - a * NJ;                // edge padding at the canvas ends
    c1 = (a > 0 ? a - 1 : 0) * NJ + b;
    c2 = (a < NI ? a : NI - 1) * NJ + b;
    sx = P.fi[f]; sy = P.fi[P.nfi + f];
  } else {
    int ff = f - P.nfi;
    int a = ff / (NJ + 1), b = ff - a * (NJ + 1);
    c1 = a * NJ + (b > 0 ? b - 1 : 0);
    c2 = a * NJ + (b < NJ ? b : NJ - 1);
    sx = P.fj[ff]; sy = P.fj[P.nfj + ff];
  }
  c1 = P.nbr[c1];
  c2 = P.nbr[c2];
  const size_t base = (size_t)cs * NV * nc;
  const double gamma = ph.gamma;
  double q1[NV], q2[NV], qf[NV];
  prim_at<NV>(U, base, nc, c1, gamma, q1);
  prim_at<NV>(U, base, nc, c2, gamma, q2);
  for (int k = 0; k < NV; ++k) qf[k] = 0.5 * (q1[k] + q2[k]);
  double l = fabs(qf[1] * sx + qf[2] * sy) + sqrt(gamma * qf[3] / qf[0]) * hypot_j(sx, sy);
  const size_t fo = (size_t)cs * nf + f;
  if (ph.viscous) {
    const double mu_inf = ctl[cs].mu_inf;
    double vol = 0.5 * (P.area[c1] + P.area[c2]);
    double mu = viscosity<double>(gamma * qf[3] / qf[0], mu_inf, ph.sutherland);
    double mul = mu;
    if constexpr (NV > 4) mu = mu + eddy_viscosity<double>(qf[0], qf[4] * mu_inf, mu / qf[0]);
    l = l + 2.0 * gamma / PR * mu * (sx * sx + sy * sy) / (qf[0] * vol);
    if constexpr (NV > 4) {
      double kap = diffusion_coefficient<double>(qf[0], qf[4] * mu_inf, mul);
      sam[fo] = qf[0] * (qf[1] * sx + qf[2] * sy);
      sad[fo] = kap * (sx * sx + sy * sy) / vol;
    }
  }
  lam[fo] = l;
}

// Per cell: sum of face wave speeds, V/dt = sum/cfl, and for SA the diagonal
// of the scalar chains (preconditioners.py:259-275).
template <int NV>
__global__ void __launch_bounds__(BLOCK) k_cell_speeds(Plan P, Physics ph, const Ctl* __restrict__ ctl,
                                                       const double* __restrict__ U, const double* __restrict__ lam,
                                                       const double* __restrict__ sam, const double* __restrict__ sad,
                                                       double* __restrict__ lsum, double* __restrict__ vdt,
                                                       double* __restrict__ sa_diag) {
  int c = blockIdx.x * BLOCK + threadIdx.x;
  int cs = blockIdx.y;
  if (c >= P.nc) return;
  const int NJ = P.NJ, nf = P.nfi + P.nfj;
  int i = c / NJ, j = c - i * NJ;
  const int fi0 = i * NJ + j, fi1 = fi0 + NJ;
  const int fj0 = P.nfi + i * (NJ + 1) + j, fj1 = fj0 + 1;
  const double* L = lam + (size_t)cs * nf;
  double s = ((L[fi0] + L[fi1]) + L[fj0]) + L[fj1];
  const double cfl = ctl[cs].cfl;
  lsum[(size_t)cs * P.nc + c] = s;
  vdt[(size_t)cs * P.nc + c] = s / cfl;
  if constexpr (NV > 4) {
    const double* m = sam + (size_t)cs * nf;
    const double* d = sad + (size_t)cs * nf;
    const double gamma = ph.gamma, mu_inf = ctl[cs].mu_inf;
    // outflow across the upper faces, inflow across the lower ones
    double diag = ((fmax(m[fi1], 0.0) + d[fi1]) + (fmax(-m[fi0], 0.0) + d[fi0])) + (fmax(m[fj1], 0.0) + d[fj1]) +
                  (fmax(-m[fj0], 0.0) + d[fj0]);
    const size_t base = (size_t)cs * NV * P.nc;
    double q[NV];
    prim_at<NV>(U, base, P.nc, c, gamma, q);
    diag = diag / q[0];
    // vorticity by Green-Gauss over the neighbour-filled face averages
    double gu[2] = {0, 0}, gv[2] = {0, 0};
    if (P.interior[c]) {
      int nb[4] = {c - NJ, c + NJ, c - 1, c + 1};
      double qs[4][NV];
      for (int k = 0; k < 4; ++k) prim_at<NV>(U, base, P.nc, P.nbr[nb[k]], gamma, qs[k]);
      double fw[2] = {0.5 * (qs[0][1] + q[1]), 0.5 * (qs[0][2] + q[2])};
      double fe[2] = {0.5 * (q[1] + qs[1][1]), 0.5 * (q[2] + qs[1][2])};
      double fs[2] = {0.5 * (qs[2][1] + q[1]), 0.5 * (qs[2][2] + q[2])};
      double fn[2] = {0.5 * (q[1] + qs[3][1]), 0.5 * (q[2] + qs[3][2])};
      const double V = P.area[c];
      for (int r = 0; r < 2; ++r) {
        const double* Si = P.fi + (size_t)r * P.nfi;
        const double* Sj = P.fj + (size_t)r * P.nfj;
        int gi0 = fi0, gi1 = fi1, gj0 = fj0 - P.nfi, gj1 = fj1 - P.nfi;
        gu[r] = ((fe[0] * Si[gi1] - fw[0] * Si[gi0]) + (fn[0] * Sj[gj1] - fs[0] * Sj[gj0])) / V;
        gv[r] = ((fe[1] * Si[gi1] - fw[1] * Si[gi0]) + (fn[1] * Sj[gj1] - fs[1] * Sj[gj0])) / V;
      }
    }
    double omega = fabs(gv[0] - gu[1]);
    double nu = viscosity<double>(gamma * q[3] / q[0], mu_inf, ph.sutherland) / q[0];
    double dw = P.dist[c];
    double nh = q[4];
    double destr = 0.0;
    if (nh > 0.0) {
      double st = s_tilde<double>(omega, nh * mu_inf, nu, dw);
      double fw_ = f_w<double>(r_of<double>(nh * mu_inf, st, dw));
      destr = 2.0 * CW1 * fw_ * mu_inf * nh / (dw * dw);
    }
    sa_diag[(size_t)cs * P.nc + c] = diag + destr * P.area[c] + s / cfl;
  }
}

// ---------------------------------------------------------------------------
// Line preconditioner: one thread per chain, block Thomas along it.
//   diag  = sum lambda (1/CFL + 1/2) I
//   upper = 1/2 A(q_next, s S_up) - 1/2 lambda_up I,  lower likewise from the
//   previous link (preconditioners.py:209-232).  Factor stores, per slot,
//   lower L, the pivot inverse P^-1 and C = P^-1 U (linalg.py:80-100).
// R is the storage/arithmetic type of the factor: double, or float (mixed).

// d(F.S)/dU of the analytic flux by forward differentiation (euler.py:55-71).
__device__ __forceinline__ void flux_jacobian(const double* q, double sx, double sy, double gamma, double A[4][4]) {
  double U[4];
  conservative<double, 4>(q, gamma, U);
  for (int k = 0; k < 4; ++k) {
    Dual Ud[4], qd[4], f[4];
    for (int c = 0; c < 4; ++c) Ud[c] = mkd(U[c], c == k ? 1.0 : 0.0);
    primitive<Dual, 4>(Ud, gamma, qd);
    normal_flux<Dual>(qd, sx, sy, gamma, f);
    for (int r = 0; r < 4; ++r) A[r][k] = f[r].d;
  }
}

// Unpivoted Gauss-Jordan inverse of a 4x4 block, in the order of invert_blocks.
template <class R>
__device__ __forceinline__ void invert4(R A[4][4], R X[4][4]) {
  R W[4][8];
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) { W[r][c] = A[r][c]; W[r][4 + c] = r == c ? R(1) : R(0); }
  for (int k = 0; k < 4; ++k) {
    R piv = W[k][k];
    for (int c = 0; c < 8; ++c) W[k][c] = W[k][c] / piv;
    for (int r = 0; r < 4; ++r) {
      if (r == k) continue;
      R f = W[r][k];
      for (int c = 0; c < 8; ++c) W[r][c] = W[r][c] - f * W[k][c];
    }
  }
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) X[r][c] = W[r][4 + c];
}

template <class R>
__device__ __forceinline__ void link_block(const double* Ujs, size_t base, int nc, int cell, double sx, double sy,
                                           double lamf, double gamma, R B[4][4]) {
  double q[4], u[4];
  for (int k = 0; k < 4; ++k) u[k] = Ujs[base + (size_t)k * nc + cell];
  primitive<double, 4>(u, gamma, q);
  double A[4][4];
  flux_jacobian(q, sx, sy, gamma, A);
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) B[r][c] = (R)(0.5 * A[r][c] - (r == c ? 0.5 * lamf : 0.0));
}

// Factor storage per slot: [L(16) | Pinv(16) | C(16)] flow, then [l, pinv, c] SA.
template <class R, int NV>
__global__ void __launch_bounds__(64) k_factor(Plan P, Physics ph, const Ctl* __restrict__ ctl,
                                               const double* __restrict__ U, const double* __restrict__ lam,
                                               const double* __restrict__ lsum, const double* __restrict__ sam,
                                               const double* __restrict__ sad, const double* __restrict__ sa_diag,
                                               R* __restrict__ fac, R* __restrict__ fac_sa) {
  int ln = blockIdx.x * 64 + threadIdx.x;
  int cs = blockIdx.y;
  if (ln >= P.n_lines) return;
  const int nc = P.nc, nf = P.nfi + P.nfj;
  const size_t base = (size_t)cs * NV * nc;
  const double gamma = ph.gamma, cfl = ctl[cs].cfl;
  const double* L = lam + (size_t)cs * nf + P.nfi;   // j-face wave speeds
  R* F = fac + (size_t)cs * P.n_slots * 48;
  R Cp[4][4];
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) Cp[r][c] = R(0);
  const int s0 = P.l_start[ln], s1 = P.l_start[ln + 1];
  // SA scalar chain
  R* Fs = fac_sa ? fac_sa + (size_t)cs * P.n_slots * 3 : nullptr;
  const double* mj = sam ? sam + (size_t)cs * nf + P.nfi : nullptr;
  const double* dj = sad ? sad + (size_t)cs * nf + P.nfi : nullptr;
  R cps = R(0);
  for (int s = s0; s < s1; ++s) {
    const int cell = P.l_cell[s];
    R D = (R)(lsum[(size_t)cs * nc + cell] * (1.0 / cfl + 0.5));
    R Lo[4][4], Up[4][4];
    if (s > s0 && P.l_dns[s - 1] != 0) {
      int fce = P.l_dn[s - 1];
      double sg = P.l_dns[s - 1];
      link_block<R>(U, base, nc, P.l_cell[s - 1], sg * P.fj[fce], sg * P.fj[P.nfj + fce], L[fce], gamma, Lo);
    } else {
      for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) Lo[r][c] = R(0);
    }
    if (P.l_ups[s] != 0) {
      int fce = P.l_up[s];
      double sg = P.l_ups[s];
      link_block<R>(U, base, nc, P.l_cell[s + 1], sg * P.fj[fce], sg * P.fj[P.nfj + fce], L[fce], gamma, Up);
    } else {
      for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) Up[r][c] = R(0);
    }
    R Pm[4][4], Pi[4][4], Cn[4][4];
    for (int r = 0; r < 4; ++r)
      for (int c = 0; c < 4; ++c) {
        R acc = R(0);
        for (int k = 0; k < 4; ++k) acc += Lo[r][k] * Cp[k][c];
        Pm[r][c] = (r == c ? D : R(0)) - acc;
      }
    invert4<R>(Pm, Pi);
    for (int r = 0; r < 4; ++r)
      for (int c = 0; c < 4; ++c) {
        R acc = R(0);
        for (int k = 0; k < 4; ++k) acc += Pi[r][k] * Up[k][c];
        Cn[r][c] = acc;
      }
    R* o = F + (size_t)s * 48;
    for (int r = 0; r < 4; ++r)
      for (int c = 0; c < 4; ++c) {
        o[r * 4 + c] = Lo[r][c];
        o[16 + r * 4 + c] = Pi[r][c];
        o[32 + r * 4 + c] = Cn[r][c];
        Cp[r][c] = Cn[r][c];
      }
    if constexpr (NV > 4) {
      // scalar SA chain (preconditioners.py:277-287)
      R ds = (R)sa_diag[(size_t)cs * nc + cell];
      R lo = R(0), up = R(0);
      if (s > s0 && P.l_dns[s - 1] != 0) {
        int fce = P.l_dn[s - 1];
        double sg = P.l_dns[s - 1];
        double rho = U[base + P.l_cell[s - 1]];
        lo = (R)(-(fmax(-sg * mj[fce], 0.0) + dj[fce]) / rho);
      }
      if (P.l_ups[s] != 0) {
        int fce = P.l_up[s];
        double sg = P.l_ups[s];
        double rho = U[base + P.l_cell[s + 1]];
        up = (R)(-(fmax(-sg * mj[fce], 0.0) + dj[fce]) / rho);
      }
      R pv = ds - lo * cps;
      R pinv = R(1) / pv;
      R cn = pinv * up;
      R* os = Fs + (size_t)s * 3;
      os[0] = lo; os[1] = pinv; os[2] = cn;
      cps = cn;
    }
  }
}

// z <- M^-1 r  (accumulate: z += M^-1 r).  Threads beyond the chains copy the
// cells on no chain (identity there, preconditioners.py:110-115).  `scr` holds
// the forward-sweep values, NV per slot per case.
template <class R, int NV>
__global__ void __launch_bounds__(64) k_apply(Plan P, const Ctl* __restrict__ ctl, const R* __restrict__ fac,
                                              const R* __restrict__ fac_sa, R* __restrict__ scr,
                                              const double* __restrict__ r, double* __restrict__ z, int accumulate,
                                              int check_active) {
  int t = blockIdx.x * 64 + threadIdx.x;
  int cs = blockIdx.y;
  if (check_active && !ctl[cs].active) return;
  const int nc = P.nc;
  const size_t base = (size_t)cs * NV * nc;
  if (t >= P.n_lines) {
    int k = t - P.n_lines;
    if (k >= P.n_free) return;
    int cell = P.free_cell[k];
    for (int v = 0; v < NV; ++v) {
      size_t o = base + (size_t)v * nc + cell;
      double x = r[o];
      z[o] = accumulate ? z[o] + x : x;
    }
    return;
  }
  const R* F = fac + (size_t)cs * P.n_slots * 48;
  const R* Fs = NV > 4 ? fac_sa + (size_t)cs * P.n_slots * 3 : nullptr;
  R* S = scr + (size_t)cs * P.n_slots * NV;
  const int s0 = P.l_start[t], s1 = P.l_start[t + 1];
  // forward: d_k = P_k^-1 (r_k - L_k d_{k-1})
  R dp[4] = {R(0), R(0), R(0), R(0)};
  R dps = R(0);
  for (int s = s0; s < s1; ++s) {
    const int cell = P.l_cell[s];
    const R* o = F + (size_t)s * 48;
    R tmp[4];
    for (int i = 0; i < 4; ++i) {
      R acc = (R)r[base + (size_t)i * nc + cell];
      for (int k = 0; k < 4; ++k) acc -= o[i * 4 + k] * dp[k];
      tmp[i] = acc;
    }
    for (int i = 0; i < 4; ++i) {
      R acc = R(0);
      for (int k = 0; k < 4; ++k) acc += o[16 + i * 4 + k] * tmp[k];
      dp[i] = acc;
    }
    R* sv = S + (size_t)s * NV;
    for (int v = 0; v < 4; ++v) sv[v] = dp[v];
    if constexpr (NV > 4) {
      const R* os = Fs + (size_t)s * 3;
      dps = os[1] * ((R)r[base + (size_t)4 * nc + cell] - os[0] * dps);
      sv[4] = dps;
    }
  }
  // backward: x_k = d_k - C_k x_{k+1}
  R xn[4] = {R(0), R(0), R(0), R(0)};
  R xns = R(0);
  for (int s = s1 - 1; s >= s0; --s) {
    const int cell = P.l_cell[s];
    const R* o = F + (size_t)s * 48;
    const R* sv = S + (size_t)s * NV;
    R x[4];
    for (int i = 0; i < 4; ++i) {
      R acc = sv[i];
      for (int k = 0; k < 4; ++k) acc -= o[32 + i * 4 + k] * xn[k];
      x[i] = acc;
    }
    for (int i = 0; i < 4; ++i) {
      xn[i] = x[i];
      size_t oo = base + (size_t)i * nc + cell;
      z[oo] = accumulate ? z[oo] + (double)x[i] : (double)x[i];
    }
    if constexpr (NV > 4) {
      const R* os = Fs + (size_t)s * 3;
      xns = sv[4] - os[2] * xns;
      size_t oo = base + (size_t)4 * nc + cell;
      z[oo] = accumulate ? z[oo] + (double)xns : (double)xns;
    }
  }
}

}  // namespace jr
