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

// Sum NV values over the workgroup; result valid in thread 0.  All loops run
// over the compile-time NV so the values stay in registers.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double (*sh)[4]) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double s = wave_sum(v[k]);
    if (lane == 0) sh[k][w] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = ((sh[k][0] + sh[k][1]) + sh[k][2]) + sh[k][3];
  }
  __syncthreads();
}

// Kutta-Joukowski circulation of each case's wall pressure force for the
// point-vortex far field (point_vortex_far_field,
// src/jaxrans/numerics/boundary.py:222-240): F = -sum (p - p_inf) S over the
// wall faces (S into the fluid), circ = (F_y u - F_x v) / |V| / (rho |V|) of
// the freestream; with dU given, also its derivative along dU.  One workgroup
// per case; circ[case][2] = (value, derivative).
//
// PT (with T = Dual): the freestream carries its tangent along a flow
// parameter (P.dfree), and the derivative is along (dU, dq_inf).
template <class T, int NV, bool PT = false>
__global__ void __launch_bounds__(BLOCK) k_circulation(Plan P, double gamma, const Ctl* __restrict__ ctl,
                                                       const double* __restrict__ U, const double* __restrict__ dU,
                                                       double* __restrict__ circ, int check_active) {
  bind_case(P, blockIdx.x);
  __shared__ double sh[4][4];
  const int cs = blockIdx.x;
  if (skip_case(ctl, cs, check_active)) return;
  const size_t base = (size_t)cs * NV * P.nc;
  typename Free<PT>::type qi[4];
  for (int k = 0; k < 4; ++k) qi[k] = free_val<PT>(ctl[cs].q_inf[k], P.dfree, cs, k);
  double v[4] = {0.0, 0.0, 0.0, 0.0};   // F_x, F_y and their derivatives
  for (int w = threadIdx.x; w < P.n_walls; w += BLOCK) {
    const int fam = P.w_fam[w], ff = P.w_face[w];
    const double into = P.w_into[w];
    const double sx = into * (fam == 0 ? P.fi[ff] : P.fj[ff]);
    const double sy = into * (fam == 0 ? P.fi[P.nfi + ff] : P.fj[P.nfj + ff]);
    T q[NV];
    prim_of<T, NV>(U, dU, base, P.nc, P.w_cell[w], gamma, q);
    const T p = q[3] - qi[3];
    const T fx = -(p * sx), fy = -(p * sy);
    v[0] += val(fx);
    v[1] += val(fy);
    v[2] += tan_of(fx);
    v[3] += tan_of(fy);
  }
  block_sum<4>(v, sh);
  if (threadIdx.x == 0) {
    if constexpr (PT) {
      const Dual fx = mkd(v[0], v[2]), fy = mkd(v[1], v[3]);
      const Dual speed = hypot_j(qi[1], qi[2]);
      const Dual cc = (fy * qi[1] - fx * qi[2]) / speed / (qi[0] * speed);
      circ[2 * cs] = cc.v;
      circ[2 * cs + 1] = cc.d;
    } else {
      const double speed = hypot_j(qi[1], qi[2]);
      circ[2 * cs] = (v[1] * qi[1] - v[0] * qi[2]) / speed / (qi[0] * speed);
      circ[2 * cs + 1] = (v[3] * qi[1] - v[2] * qi[2]) / speed / (qi[0] * speed);
    }
  }
}

// Partial sums -> the last workgroup to finish adds them in block order and
// leaves the NV totals in the shared array `out`.  Returns true in every
// thread of that workgroup.
template <int NV>
__device__ bool finish_sum(double (&v)[NV], double* partials, unsigned* counter, double (*sh)[4], double* out) {
  __shared__ int is_last;
  block_sum<NV>(v, sh);
  const int nblk = gridDim.x, blk = blockIdx.x;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) partials[(size_t)blk * NV + k] = v[k];
    __threadfence();
    unsigned prev = atomicAdd(counter, 1u);
    is_last = prev == (unsigned)(nblk - 1);
  }
  __syncthreads();
  if (!is_last) return false;
  __threadfence();
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = 0.0;
  for (int b = threadIdx.x; b < nblk; b += BLOCK) {
#pragma unroll
    for (int k = 0; k < NV; ++k)
      acc[k] += __hip_atomic_load(&partials[(size_t)b * NV + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  block_sum<NV>(acc, sh);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) out[k] = acc[k];
    *counter = 0u;
  }
  __syncthreads();
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
  bind_case(P, blockIdx.y);
  int f = blockIdx.x * BLOCK + threadIdx.x;
  int cs = blockIdx.y;
  if (ctl[cs].parked) return;   // a parked serving slot: nothing to do this step
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
  bind_case(P, blockIdx.y);
  int c = blockIdx.x * BLOCK + threadIdx.x;
  int cs = blockIdx.y;
  if (ctl[cs].parked) return;   // a parked serving slot: nothing to do this step
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

// --- explicit multistage Runge-Kutta (solvers/explicit.py:48-71) --------------
// Per canvas cell dt / V with dt = CFL V / max(total, 1e-300), total the
// spectral radii |v . S| + a |S| at the cell's own state over its lower and
// upper i-faces, then its lower and upper j-faces, plus, when viscous, the
// diffusive limit 4 (gamma / Pr) (mu / rho) (S_i^2 + S_j^2) / V with S_a the
// mean area of the two faces normal to a (numerics/residual.py:142-166).
// CFL: the case's control value.
template <int NV, bool VISC>
__global__ void __launch_bounds__(BLOCK) k_local_dt(Plan P, Physics ph, const Ctl* __restrict__ ctl,
                                                    const double* __restrict__ U, double* __restrict__ fac) {
  bind_case(P, blockIdx.y);
  const int c = blockIdx.x * BLOCK + threadIdx.x;
  const int cs = blockIdx.y;
  if (c >= P.nc) return;
  const int NJ = P.NJ;
  const int i = c / NJ, j = c - i * NJ;
  const int fi0 = i * NJ + j, fi1 = fi0 + NJ;
  const int fj0 = i * (NJ + 1) + j, fj1 = fj0 + 1;
  const double gamma = ph.gamma;
  double q[NV];
  prim_at<NV>(U, (size_t)cs * NV * P.nc, P.nc, c, gamma, q);
  const double a = sqrt(gamma * q[3] / q[0]);
  const double* Si0 = P.fi;
  const double* Si1 = P.fi + P.nfi;
  const double* Sj0 = P.fj;
  const double* Sj1 = P.fj + P.nfj;
  const double mi0 = hypot_j(Si0[fi0], Si1[fi0]), mi1 = hypot_j(Si0[fi1], Si1[fi1]);
  const double mj0 = hypot_j(Sj0[fj0], Sj1[fj0]), mj1 = hypot_j(Sj0[fj1], Sj1[fj1]);
  const double li0 = fabs(q[1] * Si0[fi0] + q[2] * Si1[fi0]) + a * mi0;
  const double li1 = fabs(q[1] * Si0[fi1] + q[2] * Si1[fi1]) + a * mi1;
  const double lj0 = fabs(q[1] * Sj0[fj0] + q[2] * Sj1[fj0]) + a * mj0;
  const double lj1 = fabs(q[1] * Sj0[fj1] + q[2] * Sj1[fj1]) + a * mj1;
  double total = ((li0 + li1) + lj0) + lj1;
  const double V = P.area[c];
  if constexpr (VISC) {
    const double si = 0.5 * (mi0 + mi1), sj = 0.5 * (mj0 + mj1);
    const double squares = si * si + sj * sj;
    const double mu_inf = ctl[cs].mu_inf;
    double mu = viscosity<double>(gamma * q[3] / q[0], mu_inf, ph.sutherland);
    if constexpr (NV > 4) mu = mu + eddy_viscosity<double>(q[0], q[4] * mu_inf, mu / q[0]);
    const double factor = gamma / PR * mu / q[0];
    total = total + 4.0 * factor * squares / V;
  }
  const double dt = ctl[cs].cfl * V / fmax(total, 1e-300);
  // Gap and ghost rows carry R = U - U0, which the reference's march keeps at
  // exactly zero from a freestream start.  U0 is recomputed here (contracted
  // to FMAs), so it can sit an ulp from the host's U0, and dt/V on those
  // degenerate rows reaches ~80: an ulp seed would grow by ~1e7 per RK5
  // step.  Their update factor is zero, which keeps them at their start
  // value; interior cells never read them (ghosts are filled from sources).
  fac[(size_t)cs * P.nc + c] = P.interior[c] ? dt / V : 0.0;
}

// One stage: out = U0 - alpha (dt / V) R, every row of every cell.  out may
// be U0 itself (each entry is read before it is written).
template <int NV>
__global__ void __launch_bounds__(BLOCK) k_rk_stage(int nc, const double* U0, const double* __restrict__ fac,
                                                    const double* __restrict__ R, double alpha, double* out) {
  const size_t e = (size_t)blockIdx.x * BLOCK + threadIdx.x;
  const int cs = blockIdx.y;
  if (e >= (size_t)NV * nc) return;
  const size_t o = (size_t)cs * NV * nc + e;
  const int c = (int)(e % (size_t)nc);
  out[o] = U0[o] - alpha * fac[(size_t)cs * nc + c] * R[o];
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

// Chain blocks of every slot, in parallel over slots (preconditioners.py:209-287):
// blk[slot] = [d, Lo(16), Up(16), sa_lo, sa_d, sa_up, pad].  A slot couples to
// its predecessor when the predecessor's link is live: the last slot of every
// chain has no link, so the first slot of the next chain never reads across.
// Stored in the factor's precision R: FP32 in the mixed path (each entry
// evaluated in FP64 and rounded once), FP64 otherwise.
constexpr int BLKW = 36;

template <class R, int NV>
__global__ void __launch_bounds__(BLOCK) k_blocks(Plan P, Physics ph, const Ctl* __restrict__ ctl,
                                                  const double* __restrict__ U, const double* __restrict__ lam,
                                                  const double* __restrict__ lsum, const double* __restrict__ sam,
                                                  const double* __restrict__ sad, const double* __restrict__ sa_diag,
                                                  R* __restrict__ blk) {
  bind_case(P, blockIdx.y);
  const int s = blockIdx.x * BLOCK + threadIdx.x;
  const int cs = blockIdx.y;
  if (ctl[cs].parked) return;   // a parked serving slot: nothing to do this step
  if (s >= P.n_slots) return;
  const int nc = P.nc, nf = P.nfi + P.nfj;
  const size_t base = (size_t)cs * NV * nc;
  const double gamma = ph.gamma, cfl = ctl[cs].cfl;
  const double* L = lam + (size_t)cs * nf + P.nfi;
  const int cell = P.l_cell[s];
  R* o = blk + ((size_t)cs * P.n_slots + s) * BLKW;
  o[0] = (R)(lsum[(size_t)cs * nc + cell] * (1.0 / cfl + 0.5));
  double B[4][4];
  const bool has_lo = s > 0 && P.l_dns[s - 1] != 0;
  if (has_lo) {
    int fce = P.l_dn[s - 1];
    double sg = P.l_dns[s - 1];
    link_block<double>(U, base, nc, P.l_cell[s - 1], sg * P.fj[fce], sg * P.fj[P.nfj + fce], L[fce], gamma, B);
  }
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) o[1 + r * 4 + c] = (R)(has_lo ? B[r][c] : 0.0);
  const bool has_up = P.l_ups[s] != 0;
  if (has_up) {
    int fce = P.l_up[s];
    double sg = P.l_ups[s];
    link_block<double>(U, base, nc, P.l_cell[s + 1], sg * P.fj[fce], sg * P.fj[P.nfj + fce], L[fce], gamma, B);
  }
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) o[17 + r * 4 + c] = (R)(has_up ? B[r][c] : 0.0);
  if constexpr (NV > 4) {
    const double* mj = sam + (size_t)cs * nf + P.nfi;
    const double* dj = sad + (size_t)cs * nf + P.nfi;
    double lo = 0.0, up = 0.0;
    if (has_lo) {
      int fce = P.l_dn[s - 1];
      double sg = P.l_dns[s - 1];
      lo = -(fmax(-sg * mj[fce], 0.0) + dj[fce]) / U[base + P.l_cell[s - 1]];
    }
    if (has_up) {
      int fce = P.l_up[s];
      double sg = P.l_ups[s];
      up = -(fmax(-sg * mj[fce], 0.0) + dj[fce]) / U[base + P.l_cell[s + 1]];
    }
    o[33] = (R)lo;
    o[34] = (R)sa_diag[(size_t)cs * nc + cell];
    o[35] = (R)up;
  }
}

// Block Thomas factor (linalg.py:80-100), one lane per chain: a wavefront
// takes 64 chains, order[t * 64 .. t * 64 + 63] (chains by length, longest
// first, so a wavefront's chains end together), and runs the recurrence of
// each in its own lane, the next slot's blocks loaded while the current one is
// factored.  Per slot, in the order the lanes of the line apply read them
// (entry e = 4 r + q):
//   fac[slot][e][2] = (L[r][q], P^-1[q][r]);
//   fac[n_slots*32 + slot*16 + e] = C[q][r] on the even steps counted from
//     the chain's top (slot L - 1, L - 3, ...), C[r][q] on the odd ones;
//   fac_sa[slot][3] = l, 1/p, c  (the scalar SA chain).
template <class R, int NV>
__global__ void __launch_bounds__(64) k_thomas(Plan P, const Ctl* __restrict__ ctl, const int* __restrict__ order,
                                               const R* __restrict__ blk, R* __restrict__ fac, R* __restrict__ fac_sa) {
  const int t = blockIdx.x * 64 + threadIdx.x, cs = blockIdx.y;
  if (t >= P.n_lines || ctl[cs].parked) return;   // a parked serving slot: nothing to do this step
  const int chain = order[t];
  const int s0 = P.l_start[chain], n = P.l_start[chain + 1] - s0;
  const R* B = blk + ((size_t)cs * P.n_slots + s0) * BLKW;
  R* F = fac + (size_t)cs * P.n_slots * 48;
  R* Fc = F + (size_t)P.n_slots * 32;
  R* Fs = fac_sa + (size_t)cs * P.n_slots * 3;
  constexpr int NB = NV > 4 ? BLKW : 33;   // block values read per slot
  R nb[NB];
  if (n > 0)
    for (int e = 0; e < NB; ++e) nb[e] = B[e];
  R Cp[4][4];
  for (int r = 0; r < 4; ++r)
    for (int q = 0; q < 4; ++q) Cp[r][q] = R(0);
  R cps = R(0);
  for (int k = 0; k < n; ++k) {
    R b[NB];
    for (int e = 0; e < NB; ++e) b[e] = nb[e];
    if (k + 1 < n)
      for (int e = 0; e < NB; ++e) nb[e] = B[(size_t)(k + 1) * BLKW + e];
    const R d = (R)b[0];
    R Pm[4][4], Pi[4][4];
    for (int r = 0; r < 4; ++r)
      for (int q = 0; q < 4; ++q) {
        R acc = R(0);
        for (int u = 0; u < 4; ++u) acc += (R)b[1 + r * 4 + u] * Cp[u][q];
        Pm[r][q] = (r == q ? d : R(0)) - acc;
      }
    invert4<R>(Pm, Pi);
    const size_t s = (size_t)s0 + k;
    for (int r = 0; r < 4; ++r)
      for (int q = 0; q < 4; ++q) {
        R acc = R(0);
        for (int u = 0; u < 4; ++u) acc += Pi[r][u] * (R)b[17 + u * 4 + q];
        Cp[r][q] = acc;
      }
    const bool even = (n - 1 - k) % 2 == 0;
    for (int e = 0; e < 16; ++e) {
      const int r = e / 4, q = e % 4;
      F[s * 32 + 2 * e] = (R)b[1 + e];
      F[s * 32 + 2 * e + 1] = Pi[q][r];
      Fc[s * 16 + e] = even ? Cp[q][r] : Cp[r][q];
    }
    if constexpr (NV > 4) {
      const R lo = (R)b[33], ds = (R)b[34], up = (R)b[35];
      const R pinv = R(1) / (ds - lo * cps);
      cps = pinv * up;
      Fs[s * 3] = lo;
      Fs[s * 3 + 1] = pinv;
      Fs[s * 3 + 2] = cps;
    }
  }
}

// Sum of n partials in an order independent of the launch shape: lane l adds
// p[l], p[l + 64], ... in turn, then the xor butterfly leaves the same total in
// every lane.  Called by one whole wave.
__device__ __forceinline__ double wave_sum_parts(const double* __restrict__ p, int n) {
  const int lane = threadIdx.x & 63;
  double s = 0.0;
  for (int b = lane; b < n; b += 64) s += p[b];
  return wave_sum(s);
}

// |w| of Arnoldi column j from its pass-3 partials: the first wave sums them,
// the workgroup reads the result.  Called by every thread of the workgroup.
__device__ __forceinline__ double column_norm(const double* __restrict__ p3, int nblk) {
  __shared__ double s_norm;
  if (threadIdx.x < 64) {
    const double t = wave_sum_parts(p3, nblk);
    if (threadIdx.x == 0) s_norm = sqrt(t);
  }
  __syncthreads();
  return s_norm;
}

// Input of the line apply: r itself, or -- when vout is set -- the Arnoldi
// vector w / max(|w|, tiny) (krylov.py:73-74), which is also written to vout
// (the next basis vector), so normalisation costs no pass of its own.
__device__ __forceinline__ double apply_input(const double* __restrict__ r, double* __restrict__ vout, size_t o,
                                              double den) {
  if (vout) {
    const double x = r[o] / den;
    vout[o] = x;
    return x;
  }
  return r[o];
}

// dq = d(primitive)/dU (U) z at one cell, the same dual-number conversion the
// fill kernel applies, written to dq[at + k nc].
template <int NV>
__device__ __forceinline__ void tangent_primitive(const double* u, const double* z, double gamma, double* out) {
  Dual ud[NV], qd[NV];
  for (int k = 0; k < NV; ++k) ud[k] = mkd(u[k], z[k]);
  primitive<Dual, NV>(ud, gamma, qd);
  for (int k = 0; k < NV; ++k) out[k] = qd[k].d;
}

// Tangent primitives of non-ghost cell `cell` (dq[at + k nc]) and of the
// final ghost cells rooted there (linearised fill: dq_ghost = J dq_root).
template <int NV>
__device__ __forceinline__ void write_tangent_q(const Plan& P, const double* u, const double* z, double gamma,
                                                double* __restrict__ dq, size_t base, int cell,
                                                const double* __restrict__ jac_case) {
  const int nc = P.nc;
  double d[NV];
  tangent_primitive<NV>(u, z, gamma, d);
  for (int k = 0; k < NV; ++k) dq[base + (size_t)k * nc + cell] = d[k];
  for (int e = P.gdep_start[cell]; e < P.gdep_start[cell + 1]; ++e) {
    const int q = P.gdep_q[e];
    const double* J = jac_case + (size_t)q * NV * NV;
    const int dest = P.gq_dest[q];
    for (int r = 0; r < NV; ++r) {
      double acc = 0.0;
      for (int k = 0; k < NV; ++k) acc += J[r * NV + k] * d[k];
      dq[base + (size_t)r * nc + dest] = acc;
    }
  }
}

// |w| of the Arnoldi vector from its per-workgroup partials (p3 of this case,
// nblk values) when the apply normalises; called by every thread.
__device__ __forceinline__ double apply_den(const double* __restrict__ p3, int nblk, const double* vout) {
  return vout ? fmax(column_norm(p3, nblk), 2.2250738585072014e-308) : 1.0;
}

}  // namespace jr
