// Block cyclic reduction of the line chains (src/jaxrans/core/linalg.py:159-200):
// the reference's own line solve.  Odd-even reduction gives ceil(log2 L)
// levels of independent block products, so a chain of L cells is solved in
// about 2 log2 L synchronised steps instead of the 2 L sequential steps of
// block Thomas.
//
// One workgroup per chain, at least L threads (L <= CR_MAX_LEN).  Level l of
// a chain has n_l rows (n_0 = L, n_{l+1} = floor(n_l / 2)); row j of level l
// is the chain's cell (j + 1) 2^l - 1.  The eliminated rows of every level
// are the even rows (j = 2i), the multiplier rows the odd rows (j = 2i + 1);
// counted over all levels, level by level, the chain has E = L - 1 even rows
// (each level eliminates all but its odd rows) and O = sum_{l >= 1} n_l <= L - 1
// odd rows, and one final row.  Odd row g and even row g are
// owned by thread g, which loads all of its factor values before the first
// synchronisation.
//
// Factor storage per chain, structure of arrays so that the threads of a
// wave read consecutive addresses (value k of row g at [k][g]), R values from
// 80 s0 on (80 per chain slot suffice: 32 O + 48 E + 16 <= 80 L):
//   odd rows  [32][O]: alpha (16), gamma (16)           (the multipliers)
//   even rows [48][E]: P^-1 (16), lower (16), upper (16) (the eliminated rows)
//   final row [16]:    its inverse
// The scalar SA chain alike from 5 s0 on: odd [2][O] alpha, gamma; even
// [3][E] 1/p, lower, upper; final 1/p.
#pragma once
#include "jr_solver.h"

namespace jr {

constexpr int CR_MAX_LEVELS = 16;
constexpr int CR_MAX_LEN = 512;    // longest chain the cyclic-reduction kernels take (one thread per row)
constexpr int CR_FAC = 80;         // factor values per chain slot (32 O + 48 E + 16 <= 80 L)
constexpr int CR_FAC_SA = 5;       // SA factor values per chain slot (2 O + 3 E + 1 <= 5 L)
constexpr int CR_DEP_LDS = 160;    // ghost Jacobians of one chain staged in LDS (the rest read in the tail)

struct CrLevels {
  int n[CR_MAX_LEVELS + 1];
  int odd0[CR_MAX_LEVELS + 1];    // first global odd-row index of the level
  int even0[CR_MAX_LEVELS + 1];   // first global even-row index of the level
  int nl;                         // levels with n > 1
};

__device__ __forceinline__ void cr_levels(int L, CrLevels& c) {
  int n = L, o = 0, e = 0, l = 0;
  while (n > 1 && l < CR_MAX_LEVELS) {
    c.n[l] = n;
    c.odd0[l] = o;
    c.even0[l] = e;
    o += n / 2;
    e += (n + 1) / 2;
    n /= 2;
    ++l;
  }
  c.n[l] = n;
  c.odd0[l] = o;
  c.even0[l] = e;
  c.nl = l;
}

__device__ __forceinline__ int cr_row(int l, int j) { return ((j + 1) << l) - 1; }

// A workgroup-uniform value read from the LDS level table, made wave-uniform
// so that the address arithmetic built on it stays in scalar registers.
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }

// Level l and row j within the level (odd: j = 2i + 1, even: j = 2i) of the
// g-th odd (or even) row counted over all levels; l = -1 if there is none.
__device__ __forceinline__ void cr_find(const CrLevels& c, int g, bool odd, int& l, int& j) {
  l = -1;
  j = -1;
  for (int k = 0; k < c.nl; ++k) {
    const int first = odd ? c.odd0[k] : c.even0[k];
    const int count = odd ? c.n[k] / 2 : (c.n[k] + 1) / 2;
    if (g >= first && g < first + count) {
      l = k;
      j = 2 * (g - first) + (odd ? 1 : 0);
      return;
    }
  }
}

template <class R>
__device__ __forceinline__ void mm4(const R* A, const R* B, R* C) {   // C = A B, row-major 4x4
  for (int r = 0; r < 4; ++r)
    for (int q = 0; q < 4; ++q) {
      R acc = R(0);
      for (int t = 0; t < 4; ++t) acc += A[r * 4 + t] * B[t * 4 + q];
      C[r * 4 + q] = acc;
    }
}

template <class R>
__device__ __forceinline__ void inv4(const R* A, R* X) {
  R a[4][4], x[4][4];
  for (int r = 0; r < 4; ++r)
    for (int q = 0; q < 4; ++q) a[r][q] = A[r * 4 + q];
  invert4<R>(a, x);
  for (int r = 0; r < 4; ++r)
    for (int q = 0; q < 4; ++q) X[r * 4 + q] = x[r][q];
}

// Factor.  Level 0 reads the assembled chain blocks (k_blocks); deeper
// levels (diag, lower, upper) live in the workspace, two halves of [48][L]
// (SA [3][L]) at 96 s0 (SA 6 s0), alternating with the level parity.
template <class R, int NV>
__global__ void __launch_bounds__(CR_MAX_LEN) k_cr_factor(Plan P, const Ctl* __restrict__ ctl,
                                                          const R* __restrict__ blk, R* __restrict__ fac,
                                                          R* __restrict__ fac_sa, R* __restrict__ ws,
                                                          R* __restrict__ ws_sa) {
  const int chain = blockIdx.x, cs = blockIdx.y, T = blockDim.x;
  if (ctl[cs].parked) return;   // a parked serving slot (the whole workgroup): nothing to do this step
  const int s0 = P.l_start[chain], s1 = P.l_start[chain + 1];
  const int L = s1 - s0;
  __shared__ CrLevels c;   // level table in LDS: read with runtime level indices
  if (threadIdx.x == 0) cr_levels(L, c);
  __syncthreads();
  const int O = uni(c.odd0[c.nl]), E = uni(c.even0[c.nl]), NL = uni(c.nl);
  const R* B = blk + (size_t)cs * P.n_slots * BLKW + (size_t)s0 * BLKW;
  R* Fo = fac + ((size_t)cs * P.n_slots + s0) * CR_FAC;
  R* Fe = Fo + (size_t)32 * O;
  R* Ff = Fe + (size_t)48 * E;
  R* So = NV > 4 ? fac_sa + ((size_t)cs * P.n_slots + s0) * CR_FAC_SA : nullptr;
  R* Se = NV > 4 ? So + (size_t)2 * O : nullptr;
  R* Sf = NV > 4 ? Se + (size_t)3 * E : nullptr;
  R* W = ws + ((size_t)cs * P.n_slots + s0) * 96;
  R* Ws = NV > 4 ? ws_sa + ((size_t)cs * P.n_slots + s0) * 6 : nullptr;
  // block q (0 diagonal, 1 lower, 2 upper) of row j of level l: level 0 from
  // the assembled chain blocks (diagonal = b[0] I), deeper levels from the
  // workspace half of the level's parity
  auto ld_blk = [&](int l, int j, int q, R* out) {
    if (l == 0) {
      const R* b = B + (size_t)j * BLKW;
      if (q == 0)
        for (int k = 0; k < 16; ++k) out[k] = (k % 5 == 0) ? (R)b[0] : R(0);
      else
        for (int k = 0; k < 16; ++k) out[k] = (R)b[1 + 16 * (q - 1) + k];
    } else {
      const R* w = W + (size_t)(l & 1) * 48 * L + (size_t)(16 * q) * L + j;
      for (int k = 0; k < 16; ++k) out[k] = w[(size_t)k * L];
    }
  };
  // SA row (diagonal, lower, upper)
  auto ld_sa = [&](int l, int j, R& sd, R& sa_, R& su) {
    if constexpr (NV > 4) {
      if (l == 0) {
        const R* b = B + (size_t)j * BLKW;
        sa_ = (R)b[33];
        sd = (R)b[34];
        su = (R)b[35];
      } else {
        const R* w3 = Ws + (size_t)(l & 1) * 3 * L + j;
        sd = w3[0];
        sa_ = w3[L];
        su = w3[2 * (size_t)L];
      }
    }
  };
  for (int l = 0; l <= NL; ++l) {
    const int n = uni(c.n[l]);
    if (l == NL) {   // the final single row: its inverse
      if (threadIdx.x == 0 && n == 1) {
        R d[16], x[16], sd = R(0), sa_ = R(0), su = R(0);
        ld_blk(l, 0, 0, d);
        inv4<R>(d, x);
        for (int k = 0; k < 16; ++k) Ff[k] = x[k];
        ld_sa(l, 0, sd, sa_, su);
        if constexpr (NV > 4) Sf[0] = R(1) / sd;
      }
      break;
    }
    const int e0 = uni(c.even0[l]), o0 = uni(c.odd0[l]);
    // even rows: inverse of the diagonal, lower and upper kept
    for (int i = threadIdx.x; i < (n + 1) / 2; i += T) {
      R* o = Fe + e0 + i;
      {
        R d[16], x[16];
        ld_blk(l, 2 * i, 0, d);
        inv4<R>(d, x);
        for (int k = 0; k < 16; ++k) o[(size_t)k * E] = x[k];
      }
      for (int q = 1; q <= 2; ++q) {
        R a[16];
        ld_blk(l, 2 * i, q, a);
        for (int k = 0; k < 16; ++k) o[(size_t)(16 * q + k) * E] = a[k];
      }
      if constexpr (NV > 4) {
        R sd = R(0), sa_ = R(0), su = R(0);
        ld_sa(l, 2 * i, sd, sa_, su);
        R* os = Se + e0 + i;
        os[0] = R(1) / sd;
        os[E] = sa_;
        os[2 * (size_t)E] = su;
      }
    }
    __threadfence_block();
    __syncthreads();
    // odd rows: multipliers, then the next level's row i.  Operands are loaded
    // one block at a time, just before their product, to bound live registers.
    for (int i = threadIdx.x; i < n / 2; i += T) {
      const int j = 2 * i + 1;
      const bool has1 = j + 1 < n;
      const R* q0 = Fe + e0 + i;                     // even row j - 1
      const R* q1 = Fe + e0 + (has1 ? i + 1 : i);    // even row j + 1
      auto ld16 = [&](const R* q, int off, R* out) {
        for (int k = 0; k < 16; ++k) out[k] = q[(size_t)(off + k) * E];
      };
      R alpha[16], gam[16], op[16], t[16];
      {
        R a[16];
        ld_blk(l, j, 1, a);
        ld16(q0, 0, op);
        mm4<R>(a, op, alpha);                                         // alpha = a_odd P^-1_even(j-1)
      }
      if (has1) {
        R u[16];
        ld_blk(l, j, 2, u);
        ld16(q1, 0, op);
        mm4<R>(u, op, gam);                                           // gamma = c_odd P^-1_even(j+1)
      } else {
        for (int k = 0; k < 16; ++k) gam[k] = R(0);
      }
      R* o = Fo + o0 + i;
      for (int k = 0; k < 16; ++k) {
        o[(size_t)k * O] = alpha[k];
        o[(size_t)(16 + k) * O] = gam[k];
      }
      __builtin_amdgcn_sched_barrier(0);   // stage boundary: keep the later loads from being hoisted
      // next level row i: a' = -alpha low0, d' = d - alpha upp0 - gamma low1, c' = -gamma upp1
      R* w = W + (size_t)((l + 1) & 1) * 48 * L + i;
      ld16(q0, 16, op);
      mm4<R>(alpha, op, t);
      for (int k = 0; k < 16; ++k) w[(size_t)(16 + k) * L] = -t[k];
      __builtin_amdgcn_sched_barrier(0);
      R d[16];
      ld_blk(l, j, 0, d);
      ld16(q0, 32, op);
      mm4<R>(alpha, op, t);
      for (int k = 0; k < 16; ++k) d[k] = d[k] - t[k];
      if (has1) {
        ld16(q1, 16, op);
        mm4<R>(gam, op, t);
        for (int k = 0; k < 16; ++k) d[k] = d[k] - t[k];
        ld16(q1, 32, op);
        mm4<R>(gam, op, t);
        for (int k = 0; k < 16; ++k) w[(size_t)(32 + k) * L] = -t[k];
      } else {
        for (int k = 0; k < 16; ++k) w[(size_t)(32 + k) * L] = R(0);
      }
      for (int k = 0; k < 16; ++k) w[(size_t)k * L] = d[k];
      if constexpr (NV > 4) {
        R sd = R(0), sa_ = R(0), su = R(0);
        ld_sa(l, j, sd, sa_, su);
        const R* es0 = Se + e0 + i;
        const R* es1 = Se + e0 + (has1 ? i + 1 : i);
        const R al = sa_ * es0[0];
        const R ga = has1 ? su * es1[0] : R(0);
        R* os = So + o0 + i;
        os[0] = al;
        os[O] = ga;
        R* w3 = Ws + (size_t)((l + 1) & 1) * 3 * L + i;
        w3[0] = sd - al * es0[2 * (size_t)E] - (has1 ? ga * es1[E] : R(0));
        w3[L] = -(al * es0[E]);
        w3[2 * (size_t)L] = has1 ? -(ga * es1[2 * (size_t)E]) : R(0);
      }
    }
    __threadfence_block();
    __syncthreads();
  }
}

// Dynamic LDS of k_cr_apply: the right-hand sides (max_len x NV values of R)
// and, when the tangent input is written, the chain's tangents and the ghost
// Jacobians rooted on it.
template <class R, int NV>
__host__ __device__ inline size_t cr_apply_lds(int max_len, int max_dep, bool dq) {
  const size_t xb = ((size_t)(max_len > 0 ? max_len : 1) * NV * sizeof(R) + 7) / 8 * 8;
  if (!dq) return xb;
  const size_t k = (size_t)(max_dep < CR_DEP_LDS ? max_dep : CR_DEP_LDS);
  return xb + (size_t)max_len * NV * sizeof(double) + k * (NV * NV * sizeof(double) + 2 * sizeof(int));
}

// z <- M^-1 r (accumulate: z += M^-1 r).  Workgroups past the chains copy
// the cells on no chain (identity there), blockDim.x per workgroup.
// Dynamic LDS: max_len * NV values of R.  With vout set, the input is the
// Arnoldi vector w / |w|, also written to vout (as apply_input).  With dq_out
// set, the tangent of the primitives at (U, z) is written there too on every
// cell, ghosts through the linearised fill (gjac): the input of the
// Jacobian-vector product that follows.  Dynamic LDS then also holds the
// chain's tangents and ghost Jacobians (cr_apply_lds).
template <class R, int NV>
__global__ void __launch_bounds__(CR_MAX_LEN) k_cr_apply(Plan P, const Ctl* __restrict__ ctl,
                                                         const R* __restrict__ fac, const R* __restrict__ fac_sa,
                                                         const double* __restrict__ r, double* __restrict__ z,
                                                         int accumulate, int check_active, double* __restrict__ vout,
                                                         const double* __restrict__ p3, int nblk,
                                                         double* __restrict__ dq_out, const double* __restrict__ Uc,
                                                         double gamma, const double* __restrict__ gjac) {
  extern __shared__ unsigned char dyn[];
  R* xb = reinterpret_cast<R*>(dyn);
  const int chain = blockIdx.x, cs = blockIdx.y, T = blockDim.x, t = threadIdx.x;
  if (skip_case(ctl, cs, check_active)) return;
  const int nc = P.nc;
  const size_t base = (size_t)cs * NV * nc;
  if (chain >= P.n_lines) {   // cells on no chain: the identity
    const int k = (chain - P.n_lines) * T + t;
    const int cell = k < P.n_free ? P.free_cell[k] : -1;
    double x[NV], u[NV];
    if (cell >= 0)
      for (int v = 0; v < NV; ++v) {
        x[v] = r[base + (size_t)v * nc + cell];
        if (dq_out) u[v] = Uc[base + (size_t)v * nc + cell];
      }
    const double den = apply_den(p3 + (size_t)cs * nblk, nblk, vout);   // every thread: a barrier inside
    if (cell < 0) return;
    for (int v = 0; v < NV; ++v) {
      const size_t o = base + (size_t)v * nc + cell;
      x[v] = vout ? x[v] / den : x[v];
      if (vout) vout[o] = x[v];
      z[o] = accumulate ? z[o] + x[v] : x[v];
    }
    if (dq_out && !P.ghost[cell])
      write_tangent_q<NV>(P, u, x, gamma, dq_out, base, cell, gjac + (size_t)cs * P.n_final * NV * NV);
    return;
  }
  const int s0 = P.l_start[chain], s1 = P.l_start[chain + 1];
  const int L = s1 - s0;
  // thread t owns row t of the chain (T >= L): its right-hand side is loaded
  // first, all NV values at once, and stays in flight while the rest is set up
  const int cell = t < L ? P.l_cell[s0 + t] : -1;
  double xr[NV], u[NV];
  bool is_ghost = false;
  if (cell >= 0)
    for (int v = 0; v < NV; ++v) {
      xr[v] = r[base + (size_t)v * nc + cell];
      if (dq_out) u[v] = Uc[base + (size_t)v * nc + cell];
    }
  if (dq_out && cell >= 0) is_ghost = P.ghost[cell] != 0;
  // Linearised fill of the ghosts rooted on this chain: their Jacobians,
  // destinations and root rows go to LDS now, so the tail has no global loads.
  // Dynamic LDS after xb: dqs [max_len][NV] double, Js [max_dep][NV*NV] double,
  // jd, jrow [max_dep] int.
  const size_t xb_bytes = ((size_t)P.max_len * NV * sizeof(R) + 7) / 8 * 8;
  const int kl = P.max_dep < CR_DEP_LDS ? P.max_dep : CR_DEP_LDS;   // staged per chain
  double* dqs = reinterpret_cast<double*>(dyn + xb_bytes);
  double* Js = dqs + (size_t)P.max_len * NV;
  int* jd = reinterpret_cast<int*>(Js + (size_t)kl * NV * NV);
  int* jrow = jd + kl;
  int nd = 0, d0 = 0;
  const double* Jc = gjac + (size_t)cs * P.n_final * NV * NV;
  if (dq_out) {
    d0 = P.cq_start[chain];
    nd = P.cq_start[chain + 1] - d0;
    for (int i = t; i < nd && i < kl; i += T) {
      const int q = P.cq_q[d0 + i];
      for (int k = 0; k < NV * NV; ++k) Js[(size_t)i * NV * NV + k] = Jc[(size_t)q * NV * NV + k];
      jd[i] = P.gq_dest[q];
      jrow[i] = P.gq_row[q];
    }
  }
  __shared__ CrLevels c;   // level table in LDS: read with runtime level indices
  if (t == 0) cr_levels(L, c);
  __syncthreads();
  const int O = uni(c.odd0[c.nl]), E = uni(c.even0[c.nl]), NL = uni(c.nl);
  const R* Fo = fac + ((size_t)cs * P.n_slots + s0) * CR_FAC;
  const R* Fe = Fo + (size_t)32 * O;
  const R* Ff = Fe + (size_t)48 * E;
  const R* So = NV > 4 ? fac_sa + ((size_t)cs * P.n_slots + s0) * CR_FAC_SA : nullptr;
  const R* Se = NV > 4 ? So + (size_t)2 * O : nullptr;
  const R* Sf = NV > 4 ? Se + (size_t)3 * E : nullptr;
  // this thread's odd row (reduction) and even row (back substitution)
  int lo_l, lo_j, le_l, le_j;
  cr_find(c, t, true, lo_l, lo_j);
  cr_find(c, t, false, le_l, le_j);
  R oa[32], ev[48];
  R soa[2] = {R(0), R(0)}, sev[3] = {R(0), R(0), R(0)};
  if (lo_l >= 0) {
#pragma unroll
    for (int k = 0; k < 32; ++k) oa[k] = Fo[(size_t)k * O + t];
    if constexpr (NV > 4) {
      soa[0] = So[t];
      soa[1] = So[O + t];
    }
  }
  if (le_l >= 0) {
#pragma unroll
    for (int k = 0; k < 48; ++k) ev[k] = Fe[(size_t)k * E + t];
    if constexpr (NV > 4) {
      sev[0] = Se[t];
      sev[1] = Se[E + t];
      sev[2] = Se[2 * (size_t)E + t];
    }
  }
  const double den = apply_den(p3 + (size_t)cs * nblk, nblk, vout);   // every thread: a barrier inside
  if (cell >= 0)
    for (int v = 0; v < NV; ++v) {
      const double xv = vout ? xr[v] / den : xr[v];
      if (vout) vout[base + (size_t)v * nc + cell] = xv;
      xb[t * NV + v] = (R)xv;
    }
  __syncthreads();
  // reduction: b_odd -= alpha b_even(j-1) + gamma b_even(j+1)
  for (int l = 0; l < NL; ++l) {
    if (lo_l == l) {
      const int j = lo_j, n = c.n[l];
      R* bo = xb + (size_t)cr_row(l, j) * NV;
      const R* b0 = xb + (size_t)cr_row(l, j - 1) * NV;
      const bool has1 = j + 1 < n;
      const R* b1 = xb + (size_t)cr_row(l, has1 ? j + 1 : j - 1) * NV;
      R out[4];
      for (int i = 0; i < 4; ++i) {
        R acc = bo[i];
        for (int q = 0; q < 4; ++q) acc -= oa[i * 4 + q] * b0[q];
        if (has1)
          for (int q = 0; q < 4; ++q) acc -= oa[16 + i * 4 + q] * b1[q];
        out[i] = acc;
      }
      R os = R(0);
      if constexpr (NV > 4) os = bo[4] - soa[0] * b0[4] - (has1 ? soa[1] * b1[4] : R(0));
      for (int i = 0; i < 4; ++i) bo[i] = out[i];
      if constexpr (NV > 4) bo[4] = os;
    }
    __syncthreads();
  }
  // the final single row
  if (t == 0) {
    R* b = xb + (size_t)cr_row(NL, 0) * NV;
    R x[4];
    for (int i = 0; i < 4; ++i) {
      R acc = R(0);
      for (int q = 0; q < 4; ++q) acc += Ff[i * 4 + q] * b[q];
      x[i] = acc;
    }
    for (int i = 0; i < 4; ++i) b[i] = x[i];
    if constexpr (NV > 4) b[4] = Sf[0] * b[4];
  }
  __syncthreads();
  // back substitution: x_even = P^-1 (b_even - low x(j-1) - upp x(j+1))
  for (int l = NL - 1; l >= 0; --l) {
    if (le_l == l) {
      const int j = le_j, n = c.n[l];
      R* be = xb + (size_t)cr_row(l, j) * NV;
      const bool has0 = j > 0, has1 = j + 1 < n;
      const R* x0 = xb + (size_t)cr_row(l, has0 ? j - 1 : j) * NV;
      const R* x1 = xb + (size_t)cr_row(l, has1 ? j + 1 : j) * NV;
      R tmp[4];
      for (int i = 0; i < 4; ++i) {
        R acc = be[i];
        if (has0)
          for (int q = 0; q < 4; ++q) acc -= ev[16 + i * 4 + q] * x0[q];
        if (has1)
          for (int q = 0; q < 4; ++q) acc -= ev[32 + i * 4 + q] * x1[q];
        tmp[i] = acc;
      }
      R xs = R(0);
      if constexpr (NV > 4)
        xs = sev[0] * (be[4] - (has0 ? sev[1] * x0[4] : R(0)) - (has1 ? sev[2] * x1[4] : R(0)));
      R x[4];
      for (int i = 0; i < 4; ++i) {
        R acc = R(0);
        for (int q = 0; q < 4; ++q) acc += ev[i * 4 + q] * tmp[q];
        x[i] = acc;
      }
      for (int i = 0; i < 4; ++i) be[i] = x[i];
      if constexpr (NV > 4) be[4] = xs;
    }
    __syncthreads();
  }
  if (cell >= 0) {
    for (int v = 0; v < NV; ++v) {
      const size_t o = base + (size_t)v * nc + cell;
      xr[v] = (double)xb[t * NV + v];
      z[o] = accumulate ? z[o] + xr[v] : xr[v];
    }
    if (dq_out) {
      double d[NV];
      tangent_primitive<NV>(u, xr, gamma, d);
      for (int v = 0; v < NV; ++v) dqs[t * NV + v] = d[v];
      if (!is_ghost)
        for (int v = 0; v < NV; ++v) dq_out[base + (size_t)v * nc + cell] = d[v];
    }
  }
  if (nd > 0) {   // the ghosts rooted on this chain: dQ_ghost = J dQ_root
    __syncthreads();
    for (int i = t; i < nd; i += T) {
      const double* J;
      const double* d;
      int dest;
      if (i < kl) {
        J = Js + (size_t)i * NV * NV;
        d = dqs + (size_t)jrow[i] * NV;
        dest = jd[i];
      } else {   // past the staged ones: indices and Jacobian from global memory
        const int q = P.cq_q[d0 + i];
        J = Jc + (size_t)q * NV * NV;
        d = dqs + (size_t)P.gq_row[q] * NV;
        dest = P.gq_dest[q];
      }
      for (int rr = 0; rr < NV; ++rr) {
        double acc = 0.0;
        for (int k = 0; k < NV; ++k) acc += J[rr * NV + k] * d[k];
        dq_out[base + (size_t)rr * nc + dest] = acc;
      }
    }
  }
}

}  // namespace jr
