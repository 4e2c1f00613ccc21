// Block cyclic reduction of the line chains (src/jaxrans/core/linalg.py:159-200):
// the reference's own line solve.  Odd-even reduction gives ceil(log2 L)
// levels of independent block products, so a chain of L cells is solved in
// about 2 log2 L synchronised steps instead of the 2 L sequential steps of
// block Thomas.
//
// One workgroup per chain.  Level l of a chain has n_l rows (n_0 = L,
// n_{l+1} = floor(n_l / 2)); row j of level l is the chain's cell
// (j + 1) 2^l - 1.  Factor storage per chain: 2 L rows of 48 values of R at
// chain row base 2 s0, level l at offset sum_{l' < l} n_l':
//   even row: P^-1 (16), lower (16), upper (16)   (the eliminated rows)
//   odd row:  alpha (16), gamma (16)              (the multipliers)
//   the final single row: its inverse (16).
// The scalar SA chain is stored alike, 5 values per row: inv, low, upp, alpha, gamma.
//
// Thread mapping: the odd rows of all levels, counted level by level, go to
// threads 0, 1, ... in order, and so do the even rows, so with a workgroup of
// at least L threads every thread owns at most one odd and one even row and
// can fetch all of its factor blocks before the first synchronisation.
#pragma once
#include "jr_solver.h"

namespace jr {

constexpr int CR_MAX_LEVELS = 16;

struct CrLevels {
  int n[CR_MAX_LEVELS + 1];
  int off[CR_MAX_LEVELS + 1];
  int odd0[CR_MAX_LEVELS + 1];    // first global odd-row index of the level
  int even0[CR_MAX_LEVELS + 1];   // first global even-row index of the level
  int nl;                         // levels with n > 1
};

__device__ __forceinline__ void cr_levels(int L, CrLevels& c) {
  int n = L, off = 0, o = 0, e = 0, l = 0;
  while (n > 1 && l < CR_MAX_LEVELS) {
    c.n[l] = n;
    c.off[l] = off;
    c.odd0[l] = o;
    c.even0[l] = e;
    off += n;
    o += n / 2;
    e += (n + 1) / 2;
    n /= 2;
    ++l;
  }
  c.n[l] = n;
  c.off[l] = off;
  c.odd0[l] = o;
  c.even0[l] = e;
  c.nl = l;
}

__device__ __forceinline__ int cr_row(int l, int j) { return ((j + 1) << l) - 1; }

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
// levels (diag, lower, upper) live in the workspace ws (R, 48 per row, at the
// chain's 2 s0 rows), double-buffered by level parity.
template <class R, int NV>
__global__ void __launch_bounds__(512) k_cr_factor(Plan P, const double* __restrict__ blk, R* __restrict__ fac,
                                                   R* __restrict__ fac_sa, R* __restrict__ ws,
                                                   R* __restrict__ ws_sa) {
  const int chain = blockIdx.x, cs = blockIdx.y, T = blockDim.x;
  const int s0 = P.l_start[chain], s1 = P.l_start[chain + 1];
  const int L = s1 - s0;
  __shared__ CrLevels c;   // level table in LDS: read with runtime level indices
  if (threadIdx.x == 0) cr_levels(L, c);
  __syncthreads();
  const double* B = blk + (size_t)cs * P.n_slots * BLKW + (size_t)s0 * BLKW;
  R* F = fac + ((size_t)cs * 2 * P.n_slots + 2 * (size_t)s0) * 48;
  R* Fs = NV > 4 ? fac_sa + ((size_t)cs * 2 * P.n_slots + 2 * (size_t)s0) * 5 : nullptr;
  R* W = ws + ((size_t)cs * 2 * P.n_slots + 2 * (size_t)s0) * 48;          // two halves of L rows
  R* Ws = NV > 4 ? ws_sa + ((size_t)cs * 2 * P.n_slots + 2 * (size_t)s0) * 3 : nullptr;
  // row accessors of the current level's (diag, lower, upper)
  auto load_row = [&](int l, int j, R* d, R* a, R* u, R* sd, R* sa_, R* su) {
    if (l == 0) {
      const double* b = B + (size_t)j * BLKW;
      for (int k = 0; k < 16; ++k) {
        d[k] = (k % 5 == 0) ? (R)b[0] : R(0);
        a[k] = (R)b[1 + k];
        u[k] = (R)b[17 + k];
      }
      if constexpr (NV > 4) {
        *sa_ = (R)b[33];
        *sd = (R)b[34];
        *su = (R)b[35];
      }
    } else {
      const R* w = W + ((size_t)(l & 1) * L + j) * 48;
      for (int k = 0; k < 16; ++k) {
        d[k] = w[k];
        a[k] = w[16 + k];
        u[k] = w[32 + k];
      }
      if constexpr (NV > 4) {
        const R* ws3 = Ws + ((size_t)(l & 1) * L + j) * 3;
        *sd = ws3[0];
        *sa_ = ws3[1];
        *su = ws3[2];
      }
    }
  };
  for (int l = 0; l <= c.nl; ++l) {
    const int n = c.n[l];
    R* Fl = F + (size_t)c.off[l] * 48;
    R* Fsl = NV > 4 ? Fs + (size_t)c.off[l] * 5 : nullptr;
    if (l == c.nl) {   // the last single row: its inverse
      if (threadIdx.x == 0 && n == 1) {
        R d[16], a[16], u[16], sd = R(0), sa_ = R(0), su = R(0);
        load_row(l, 0, d, a, u, &sd, &sa_, &su);
        inv4<R>(d, Fl);
        if constexpr (NV > 4) Fsl[0] = R(1) / sd;
      }
      break;
    }
    // even rows: inverse of the diagonal, lower and upper kept
    for (int i = threadIdx.x; i < (n + 1) / 2; i += T) {
      const int j = 2 * i;
      R d[16], a[16], u[16], sd = R(0), sa_ = R(0), su = R(0);
      load_row(l, j, d, a, u, &sd, &sa_, &su);
      R* o = Fl + (size_t)j * 48;
      inv4<R>(d, o);
      for (int k = 0; k < 16; ++k) {
        o[16 + k] = a[k];
        o[32 + k] = u[k];
      }
      if constexpr (NV > 4) {
        R* os = Fsl + (size_t)j * 5;
        os[0] = R(1) / sd;
        os[1] = sa_;
        os[2] = su;
      }
    }
    __threadfence_block();
    __syncthreads();
    // odd rows: multipliers, then the next level's row
    for (int i = threadIdx.x; i < n / 2; i += T) {
      const int j = 2 * i + 1;
      R d[16], a[16], u[16], sd = R(0), sa_ = R(0), su = R(0);
      load_row(l, j, d, a, u, &sd, &sa_, &su);
      const R* e0 = Fl + (size_t)(j - 1) * 48;                        // even row 2i: inv, low, upp
      const bool has1 = j + 1 < n;
      const R* e1 = Fl + (size_t)(j + 1) * 48;                        // even row 2i+2
      R alpha[16], gam[16], t[16];
      mm4<R>(a, e0, alpha);                                           // alpha = a_odd inv_even
      if (has1) mm4<R>(u, e1, gam);
      else
        for (int k = 0; k < 16; ++k) gam[k] = R(0);
      R* o = Fl + (size_t)j * 48;
      for (int k = 0; k < 16; ++k) {
        o[k] = alpha[k];
        o[16 + k] = gam[k];
      }
      // next level row i: a' = -alpha low_e0, d' = d - alpha upp_e0 - gamma low_e1, c' = -gamma upp_e1
      R* w = W + ((size_t)((l + 1) & 1) * L + i) * 48;
      mm4<R>(alpha, e0 + 16, t);
      for (int k = 0; k < 16; ++k) w[16 + k] = -t[k];
      mm4<R>(alpha, e0 + 32, t);
      for (int k = 0; k < 16; ++k) w[k] = d[k] - t[k];
      if (has1) {
        mm4<R>(gam, e1 + 16, t);
        for (int k = 0; k < 16; ++k) w[k] = w[k] - t[k];
        mm4<R>(gam, e1 + 32, t);
        for (int k = 0; k < 16; ++k) w[32 + k] = -t[k];
      } else {
        for (int k = 0; k < 16; ++k) w[32 + k] = R(0);
      }
      if constexpr (NV > 4) {
        const R* es0 = Fsl + (size_t)(j - 1) * 5;
        const R* es1 = Fsl + (size_t)(j + 1) * 5;
        R al = sa_ * es0[0];
        R ga = has1 ? su * es1[0] : R(0);
        R* os = Fsl + (size_t)j * 5;
        os[3] = al;
        os[4] = ga;
        R* ws3 = Ws + ((size_t)((l + 1) & 1) * L + i) * 3;
        ws3[0] = sd - al * es0[2] - (has1 ? ga * es1[1] : R(0));
        ws3[1] = -(al * es0[1]);
        ws3[2] = has1 ? -(ga * es1[2]) : R(0);
      }
    }
    __threadfence_block();
    __syncthreads();
  }
}

// z <- M^-1 r (accumulate: z += M^-1 r).  Workgroups past the chains copy
// the cells on no chain (identity there), blockDim.x per workgroup.
// Dynamic LDS: max_len * NV values of R.
template <class R, int NV>
__global__ void __launch_bounds__(512) k_cr_apply(Plan P, const Ctl* __restrict__ ctl, const R* __restrict__ fac,
                                                  const R* __restrict__ fac_sa, const double* __restrict__ r,
                                                  double* __restrict__ z, int accumulate, int check_active) {
  extern __shared__ unsigned char dyn[];
  R* xb = reinterpret_cast<R*>(dyn);
  const int chain = blockIdx.x, cs = blockIdx.y, T = blockDim.x, t = threadIdx.x;
  if (check_active && !ctl[cs].active) return;
  const int nc = P.nc;
  const size_t base = (size_t)cs * NV * nc;
  if (chain >= P.n_lines) {
    const int k = (chain - P.n_lines) * T + t;
    if (k >= P.n_free) return;
    const int cell = P.free_cell[k];
    for (int v = 0; v < NV; ++v) {
      size_t o = base + (size_t)v * nc + cell;
      z[o] = accumulate ? z[o] + r[o] : r[o];
    }
    return;
  }
  const int s0 = P.l_start[chain], s1 = P.l_start[chain + 1];
  const int L = s1 - s0;
  __shared__ CrLevels c;   // level table in LDS: read with runtime level indices
  if (threadIdx.x == 0) cr_levels(L, c);
  __syncthreads();
  const R* F = fac + ((size_t)cs * 2 * P.n_slots + 2 * (size_t)s0) * 48;
  const R* Fs = NV > 4 ? fac_sa + ((size_t)cs * 2 * P.n_slots + 2 * (size_t)s0) * 5 : nullptr;
  // this thread's odd row (reduction) and even row (back substitution)
  int lo_l, lo_j, le_l, le_j;
  cr_find(c, t, true, lo_l, lo_j);
  cr_find(c, t, false, le_l, le_j);
  R oa[32], ev[48];
  R soa[2] = {R(0), R(0)}, sev[3] = {R(0), R(0), R(0)};
  if (lo_l >= 0) {
    const R* p = F + ((size_t)c.off[lo_l] + lo_j) * 48;
#pragma unroll
    for (int k = 0; k < 32; ++k) oa[k] = p[k];
    if constexpr (NV > 4) {
      const R* ps = Fs + ((size_t)c.off[lo_l] + lo_j) * 5;
      soa[0] = ps[3];
      soa[1] = ps[4];
    }
  }
  if (le_l >= 0) {
    const R* p = F + ((size_t)c.off[le_l] + le_j) * 48;
#pragma unroll
    for (int k = 0; k < 48; ++k) ev[k] = p[k];
    if constexpr (NV > 4) {
      const R* ps = Fs + ((size_t)c.off[le_l] + le_j) * 5;
      sev[0] = ps[0];
      sev[1] = ps[1];
      sev[2] = ps[2];
    }
  }
  for (int k = t; k < L * NV; k += T) {
    const int row = k / NV, v = k - row * NV;
    xb[k] = (R)r[base + (size_t)v * nc + P.l_cell[s0 + row]];
  }
  __syncthreads();
  // reduction: b_odd -= alpha b_even(j-1) + gamma b_even(j+1)
  for (int l = 0; l < c.nl; ++l) {
    const int n = c.n[l];
    auto reduce = [&](int j, const R* A, const R* G, const R* sA) {
      R* bo = xb + (size_t)cr_row(l, j) * NV;
      const R* b0 = xb + (size_t)cr_row(l, j - 1) * NV;
      const bool has1 = j + 1 < n;
      const R* b1 = xb + (size_t)cr_row(l, has1 ? j + 1 : j - 1) * NV;
      R out[4];
      for (int i = 0; i < 4; ++i) {
        R acc = bo[i];
        for (int q = 0; q < 4; ++q) acc -= A[i * 4 + q] * b0[q];
        if (has1)
          for (int q = 0; q < 4; ++q) acc -= G[i * 4 + q] * b1[q];
        out[i] = acc;
      }
      R os = R(0);
      if constexpr (NV > 4) os = bo[4] - sA[0] * b0[4] - (has1 ? sA[1] * b1[4] : R(0));
      for (int i = 0; i < 4; ++i) bo[i] = out[i];
      if constexpr (NV > 4) bo[4] = os;
    };
    if (lo_l == l) reduce(lo_j, oa, oa + 16, soa);
    // rows beyond the workgroup (chains longer than it): factor read in place
    for (int g = t + T; g < c.odd0[c.nl]; g += T) {
      int ll, jj;
      cr_find(c, g, true, ll, jj);
      if (ll != l) continue;
      const R* p = F + ((size_t)c.off[ll] + jj) * 48;   // read in place
      reduce(jj, p, p + 16, NV > 4 ? Fs + ((size_t)c.off[ll] + jj) * 5 + 3 : nullptr);
    }
    __syncthreads();
  }
  // the last single row
  if (t == 0) {
    const int l = c.nl;
    R* b = xb + (size_t)cr_row(l, 0) * NV;
    const R* inv = F + (size_t)c.off[l] * 48;
    R x[4];
    for (int i = 0; i < 4; ++i) {
      R acc = R(0);
      for (int q = 0; q < 4; ++q) acc += inv[i * 4 + q] * b[q];
      x[i] = acc;
    }
    for (int i = 0; i < 4; ++i) b[i] = x[i];
    if constexpr (NV > 4) b[4] = Fs[(size_t)c.off[l] * 5] * b[4];
  }
  __syncthreads();
  // back substitution: x_even = P^-1 (b_even - low x(j-1) - upp x(j+1))
  for (int l = c.nl - 1; l >= 0; --l) {
    const int n = c.n[l];
    auto back = [&](int j, const R* E, const R* sE) {
      R* be = xb + (size_t)cr_row(l, j) * NV;
      const bool has0 = j > 0, has1 = j + 1 < n;
      const R* x0 = xb + (size_t)cr_row(l, has0 ? j - 1 : j) * NV;
      const R* x1 = xb + (size_t)cr_row(l, has1 ? j + 1 : j) * NV;
      R tmp[4];
      for (int i = 0; i < 4; ++i) {
        R acc = be[i];
        if (has0)
          for (int q = 0; q < 4; ++q) acc -= E[16 + i * 4 + q] * x0[q];
        if (has1)
          for (int q = 0; q < 4; ++q) acc -= E[32 + i * 4 + q] * x1[q];
        tmp[i] = acc;
      }
      R xs = R(0);
      if constexpr (NV > 4) xs = sE[0] * (be[4] - (has0 ? sE[1] * x0[4] : R(0)) - (has1 ? sE[2] * x1[4] : R(0)));
      R x[4];
      for (int i = 0; i < 4; ++i) {
        R acc = R(0);
        for (int q = 0; q < 4; ++q) acc += E[i * 4 + q] * tmp[q];
        x[i] = acc;
      }
      for (int i = 0; i < 4; ++i) be[i] = x[i];
      if constexpr (NV > 4) be[4] = xs;
    };
    if (le_l == l) back(le_j, ev, sev);
    for (int g = t + T; g < c.even0[c.nl]; g += T) {
      int ll, jj;
      cr_find(c, g, false, ll, jj);
      if (ll != l) continue;
      const R* p = F + ((size_t)c.off[ll] + jj) * 48;   // read in place
      back(jj, p, NV > 4 ? Fs + ((size_t)c.off[ll] + jj) * 5 : nullptr);
    }
    __syncthreads();
  }
  for (int k = t; k < L * NV; k += T) {
    const int row = k / NV, v = k - row * NV;
    size_t o = base + (size_t)v * nc + P.l_cell[s0 + row];
    double x = (double)xb[k];
    z[o] = accumulate ? z[o] + x : x;
  }
}

}  // namespace jr
