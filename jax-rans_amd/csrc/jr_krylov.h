// GMRES(m) with CGS2 Arnoldi, the Newton update with accept/reject and the
// SER CFL ramp, and the wall force integration -- all device-resident, one
// control block per case, no host synchronisation inside a Newton step.
//
// Restates src/jaxrans/solvers/krylov.py:34-137, solvers/implicit.py:54-166,
// post/forces.py:48-161.
#pragma once
#include "jr_solver.h"

namespace jr {

// Vector kernels run a grid (nblk, n_cases) over the N = NV*nc entries of each
// case, grid-strided.  Reduction scratch per case: partials [nblk][MM+2], one counter.
struct Red {
  double* partials;     // [n_cases][maxblk][MMAX+2]
  unsigned* counters;   // [n_cases][4]
  int stride;           // doubles per case in partials
};

// Incremental Givens QR of the Arnoldi Hessenberg, per case: the rotated
// columns (MMAX+1) x MMAX row-major, then cos[MMAX], sin[MMAX], g[MMAX+1].
constexpr size_t GIV = (size_t)(MMAX + 1) * MMAX + 2 * MMAX + (MMAX + 1);

#define JR_VEC_LOOP(N) for (size_t e = (size_t)blockIdx.x * BLOCK + threadIdx.x; e < (N); e += (size_t)gridDim.x * BLOCK)

// After the residual: the reported norm (density row, volume weighted,
// residual.py:237-240), ||R||, and the GMRES start x = 0, action = 0, r = -R.
template <int NV>
__global__ void __launch_bounds__(BLOCK) k_norms(Plan P, Ctl* __restrict__ ctl, const double* __restrict__ R,
                                                 double* __restrict__ X, double* __restrict__ Act,
                                                 double* __restrict__ Rv, Red red, double* __restrict__ hist,
                                                 int n_cases) {
  bind_case(P, blockIdx.y);
  __shared__ double sh[3][4];
  const int cs = blockIdx.y;
  if (ctl[cs].parked) {   // a parked serving slot keeps its state and control; its history repeats its norm
    if (hist && blockIdx.x == 0 && threadIdx.x == 0) hist[cs] = ctl[cs].norm;
    return;
  }
  const int nc = P.nc;
  const size_t N = (size_t)NV * nc, base = (size_t)cs * N;
  double v[3] = {0.0, 0.0, 0.0};
  const bool gmres_start = X != nullptr;   // null for the explicit solver: norms only
  JR_VEC_LOOP(N) {
    double r = R[base + e];
    if (gmres_start) {
      X[base + e] = 0.0;
      Act[base + e] = 0.0;
      Rv[base + e] = -r;
    }
    v[2] += r * r;
    if (e < (size_t)nc && P.interior[e]) {
      double w = P.area[e];
      v[0] += w * (r * r);
      v[1] += w;
    }
  }
  __shared__ double tot[3];
  if (finish_sum<3>(v, red.partials + (size_t)cs * red.stride, red.counters + cs * 4, sh, tot)) {
    if (threadIdx.x == 0) {
      Ctl& c = ctl[cs];
      c.norm = sqrt(tot[0] / tot[1]);
      c.bnorm = sqrt(tot[2]);
      c.rnorm = c.bnorm;
      c.cycles = 0;
      if (hist) hist[cs] = c.norm;
    }
  }
}

// R <- -b: a linear solve for the right-hand side b runs the Newton step's
// GMRES, whose start is r = -R (k_norms).
__global__ void __launch_bounds__(BLOCK) k_negate(size_t n, const double* __restrict__ b, double* __restrict__ R) {
  const size_t e = (size_t)blockIdx.x * BLOCK + threadIdx.x;
  if (e < n) R[e] = -b[e];
}

// Cycle start: active = cycles < maxiter and |r| > tol |b| (krylov.py:85-87);
// V0 = r / max(beta, tiny).
template <int NV>
__global__ void __launch_bounds__(BLOCK) k_cycle_begin(int nc, Ctl* __restrict__ ctl, const double* __restrict__ Rv,
                                                       double* __restrict__ V0, size_t vstride, int maxiter,
                                                       double tol, double* __restrict__ G) {
  const int cs = blockIdx.y;
  Ctl& c = ctl[cs];
  const bool act = !c.parked && c.cycles < maxiter && c.rnorm > tol * c.bnorm;
  if (blockIdx.x == 0) {
    double* g = G + (size_t)cs * GIV + (size_t)(MMAX + 1) * MMAX + 2 * MMAX;
    for (int k = threadIdx.x; k <= MMAX; k += BLOCK) g[k] = k == 0 ? c.rnorm : 0.0;
    if (threadIdx.x == 0) {
      c.active = act;
      c.beta = c.rnorm;
    }
  }
  if (!act) return;
  const size_t N = (size_t)NV * nc, base = (size_t)cs * N;
  const double den = fmax(c.rnorm, 2.2250738585072014e-308);
  double* v0 = V0 + (size_t)cs * vstride;
  JR_VEC_LOOP(N) v0[e] = Rv[base + e] / den;
}

// --- CGS2 over the basis V_0..V_j ---------------------------------------------
// The kernels of Arnoldi column j are instantiated for NCH = j / KC + 1
// chunks of KC basis vectors: every basis value an element needs (NCH KC of
// them, indices past j clamped to j, masked out of every sum) is loaded in
// one batch before the first is used, so an element costs one memory round
// trip whatever j is.  Chunks are processed in groups of at most GRP so the
// batch fits in registers when m > GRP KC.
constexpr int KC = 8;
constexpr int GRP = 4;

template <int NCH>
struct Cgs {
  static constexpr int NK = NCH * KC;                                  // basis values per element
  static constexpr int NP = NK <= 8 ? 8 : NK <= 16 ? 16 : NK <= 32 ? 32 : 64;   // padded for the reduction
  static constexpr int NG = (NCH + GRP - 1) / GRP;                     // load batches per element
  static constexpr int GS = NCH <= GRP ? NK : GRP * KC;                // values per batch
  static constexpr int NKG = NG * GS;                                  // NK padded to whole batches (<= NP)
};

// v[k] = V_min(k0 + k, j)[e] for k < NB, issued together.
template <int NB>
__device__ __forceinline__ void basis_batch(const double* __restrict__ Vc, size_t vstride, size_t e, int k0, int j,
                                            double (&v)[NB]) {
#pragma unroll
  for (int k = 0; k < NB; ++k) v[k] = Vc[(size_t)min(k0 + k, j) * vstride + e];
  __builtin_amdgcn_sched_barrier(0);   // keep the batch ahead of its first use
}

// PW consecutive entries per thread and loop step (PW = 2: 16-byte loads,
// half the steps of the grid-stride loop; the vectors' case strides and
// lengths must then be even).  v[k][p] = V_min(k0 + k, j)[e + p].
template <int NB, int PW>
__device__ __forceinline__ void basis_batch_pw(const double* __restrict__ Vc, size_t vstride, size_t e, int k0, int j,
                                               double (&v)[NB][PW]) {
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    const double* p = Vc + (size_t)min(k0 + k, j) * vstride + e;
    if constexpr (PW == 2) {
      const double2 t = *reinterpret_cast<const double2*>(p);
      v[k][0] = t.x;
      v[k][1] = t.y;
    } else {
      v[k][0] = *p;
    }
  }
  __builtin_amdgcn_sched_barrier(0);   // keep the batch ahead of its first use
}

template <int PW>
__device__ __forceinline__ void load_pw(const double* __restrict__ a, size_t o, double (&w)[PW]) {
  if constexpr (PW == 2) {
    const double2 t = *reinterpret_cast<const double2*>(a + o);
    w[0] = t.x;
    w[1] = t.y;
  } else {
    w[0] = a[o];
  }
}

template <int PW>
__device__ __forceinline__ void store_pw(double* __restrict__ a, size_t o, const double (&w)[PW]) {
  if constexpr (PW == 2) {
    *reinterpret_cast<double2*>(a + o) = make_double2(w[0], w[1]);
  } else {
    a[o] = w[0];
  }
}

template <int H, int O, int MM>
__device__ __forceinline__ void halve(double (&v)[MM], int lane) {
  if constexpr (H >= 1) {
    const bool up = (lane & O) != 0;
#pragma unroll
    for (int i = 0; i < H; ++i) {
      const double keep = up ? v[i + H] : v[i];
      const double give = up ? v[i] : v[i + H];
      v[i] = keep + __shfl_xor(give, O, 64);
    }
    halve<H / 2, O / 2, MM>(v, lane);
  }
}

template <int MM>
__device__ __forceinline__ double wave_reduce_scatter(double (&v)[MM]) {
  static_assert(MM >= 1 && MM <= 64 && (MM & (MM - 1)) == 0, "MM must be a power of two <= 64");
  const int lane = threadIdx.x & 63;
  halve<MM / 2, 32, MM>(v, lane);
  double s = v[0];
#pragma unroll
  for (int o = 32 / MM; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  return s;
}

// --- Partial sums of the CGS2 passes ---------------------------------------------
// A CGS2 pass ends with each workgroup writing its sums to a partials array
// (no fence, no arrival counter); the kernel that needs the totals adds the
// partials at its start, every workgroup the same partials in the same fixed
// order, so every consumer sees bit-identical totals and runs are
// reproducible.  The kernel boundary makes the partials visible.
constexpr int CG_BLK = 256;   // workgroups per case of the vector kernels (at most)

struct Cg {
  double* p1;   // [case][MMAX][nblk]  h_k partials (pass 1)
  double* p2;   // [case][MMAX][nblk]  h2_k partials (pass 2)
  double* p3;   // [case][nblk]        |w|^2 partials (pass 3)
  int nblk;
};

// Workgroup totals of NP values, written as this workgroup's partials
// parts[k][blockIdx.x] for k < nk.
template <int NP>
__device__ __forceinline__ void put_partials(double (&v)[NP], double* __restrict__ parts, int nblk, int nk) {
  __shared__ double sh[NP][4];
  constexpr int SPAN = 64 / NP;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const double s = wave_reduce_scatter<NP>(v);
  if (lane % SPAN == 0) sh[lane / SPAN][w] = s;
  __syncthreads();
  if (threadIdx.x < nk) {
    const int k = threadIdx.x;
    parts[(size_t)k * nblk + blockIdx.x] = ((sh[k][0] + sh[k][1]) + sh[k][2]) + sh[k][3];
  }
}

// Totals tot[k], k <= j, of the partials parts[k][b] of the previous pass:
// thread t adds workgroups t, t + BLOCK, ..., then the workgroup reduce-
// scatters.  Rows past j are not read.  Called by every thread.
template <int NP>
__device__ __forceinline__ void get_totals(const double* __restrict__ parts, int nblk, int j, double* tot) {
  __shared__ double sh[NP][4];
  constexpr int SPAN = 64 / NP;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double acc[NP];
#pragma unroll
  for (int k = 0; k < NP; ++k) acc[k] = 0.0;
  for (int b = threadIdx.x; b < nblk; b += BLOCK) {
    double v[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) v[k] = parts[(size_t)min(k, j) * nblk + b];
#pragma unroll
    for (int k = 0; k < NP; ++k) acc[k] += k <= j ? v[k] : 0.0;
  }
  const double s = wave_reduce_scatter<NP>(acc);
  if (lane % SPAN == 0) sh[lane / SPAN][w] = s;
  __syncthreads();
  if (threadIdx.x < NP) {
    const int k = threadIdx.x;
    tot[k] = ((sh[k][0] + sh[k][1]) + sh[k][2]) + sh[k][3];
  }
  __syncthreads();
}

// Column j of the Hessenberg (krylov.py:69-76): H[:, j] = h + h2,
// H[j+1, j] = |w|, zero below, then the Givens QR advanced by column j
// (krylov.py:121-131, the same operations in the same order as the
// whole-matrix sweep).  Run by one workgroup after the column's pass 3.
// G holds the rotated columns (MMAX+1) x MMAX, cos[MMAX], sin[MMAX], g[MMAX+1].
__device__ void column_finish(int cs, int j, int m, const double* __restrict__ p3, int nblk,
                              const double* __restrict__ hv, const double* __restrict__ h2v, double* __restrict__ H,
                              double* __restrict__ G) {
  __shared__ double gcol[MMAX + 1];
  __shared__ double gcs[MMAX], gsn[MMAX];
  const double nw = column_norm(p3, nblk);
  double* Hc = H + (size_t)cs * (MMAX + 1) * MMAX;   // H[row][col], row-major (MMAX+1) x MMAX
  double* Gc = G + (size_t)cs * GIV;
  double* Rr = Gc;                                     // rotated columns
  double* cs_ = Gc + (size_t)(MMAX + 1) * MMAX;
  double* sn = cs_ + MMAX;
  double* g = sn + MMAX;
  // the column and the previous rotations staged in LDS by the whole
  // workgroup, so that the serial rotation sweep of thread 0 reads no global
  // memory
  for (int k = threadIdx.x; k <= j; k += blockDim.x) {
    const double h = hv[(size_t)cs * (MMAX + 1) + k] + h2v[(size_t)cs * (MMAX + 1) + k];
    Hc[k * MMAX + j] = h;
    gcol[k] = h;
  }
  for (int i = threadIdx.x; i < j; i += blockDim.x) {
    gcs[i] = cs_[i];
    gsn[i] = sn[i];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    Hc[(j + 1) * MMAX + j] = nw;
    for (int k = j + 2; k <= m; ++k) Hc[k * MMAX + j] = 0.0;
    double* col = gcol;
    col[j + 1] = nw;
    for (int i = 0; i < j; ++i) {
      double a = col[i], c = col[i + 1];
      col[i] = gcs[i] * a + gsn[i] * c;
      col[i + 1] = gcs[i] * c - gsn[i] * a;
    }
    double a = col[j], c = col[j + 1];
    double r = sqrt(a * a + c * c);
    double safe = r > 0 ? r : 1.0;
    cs_[j] = r > 0 ? a / safe : 1.0;
    sn[j] = r > 0 ? c / safe : 0.0;
    col[j] = r;
    col[j + 1] = 0.0;
    const double gj = g[j];
    g[j] = cs_[j] * gj;
    g[j + 1] = -sn[j] * gj;
    for (int k = 0; k <= j + 1; ++k) Rr[k * MMAX + j] = col[k];
  }
  __syncthreads();
}

// CGS2 pass 1: h_k = V_k . w, k <= j.  The last workgroup (the fewest
// grid-stride steps) first finishes column j - 1 (its pass 3 ended the
// previous kernel sequence), with its first step's loads already in flight.
template <int NV, int NCH, int PW>
__global__ void __launch_bounds__(BLOCK) k_dots(int nc, const Ctl* __restrict__ ctl, const double* __restrict__ V,
                                                size_t vstride, size_t vcase, const double* __restrict__ W, int j,
                                                Cg cg, const double* __restrict__ hv,
                                                const double* __restrict__ h2v, double* __restrict__ H,
                                                double* __restrict__ G, int m) {
  constexpr int NP = Cgs<NCH>::NP, GS = Cgs<NCH>::GS;
  const int cs = blockIdx.y;
  if (!ctl[cs].active) return;
  const size_t N = (size_t)NV * nc, base = (size_t)cs * N;
  const double* Vc = V + (size_t)cs * vcase;
  const size_t step = (size_t)gridDim.x * BLOCK * PW;
  size_t e = ((size_t)blockIdx.x * BLOCK + threadIdx.x) * PW;
  double w[PW], v0[GS][PW];
  if (e < N) {
    load_pw<PW>(W, base + e, w);
    basis_batch_pw<GS, PW>(Vc, vstride, e, 0, j, v0);
  }
  if (blockIdx.x == gridDim.x - 1 && j > 0)
    column_finish(cs, j - 1, m, cg.p3 + (size_t)cs * cg.nblk, cg.nblk, hv, h2v, H, G);
  double acc[NP];
#pragma unroll
  for (int k = 0; k < NP; ++k) acc[k] = 0.0;
  while (e < N) {
#pragma unroll
    for (int k = 0; k < GS; ++k)
#pragma unroll
      for (int p = 0; p < PW; ++p) acc[k] += v0[k][p] * w[p];
#pragma unroll
    for (int g = 1; g < Cgs<NCH>::NG; ++g) {
      double v[GS][PW];
      basis_batch_pw<GS, PW>(Vc, vstride, e, g * GS, j, v);
#pragma unroll
      for (int k = 0; k < GS; ++k)
#pragma unroll
        for (int p = 0; p < PW; ++p) acc[g * GS + k] += v[k][p] * w[p];
    }
    e += step;
    __builtin_amdgcn_sched_barrier(0);   // the next step's loads reuse this step's registers
    if (e < N) {
      load_pw<PW>(W, base + e, w);
      basis_batch_pw<GS, PW>(Vc, vstride, e, 0, j, v0);
    }
  }
  put_partials<NP>(acc, cg.p1 + (size_t)cs * MMAX * cg.nblk, cg.nblk, j + 1);
}

// CGS2 pass 2: h = the pass-1 totals; w <- w - sum_{k<=j} h_k V_k, then
// h2_k = V_k . w.  Workgroup 0 records h.  (Loading the first grid-stride
// step ahead of the totals, as the other two passes do, needs a second set
// of NK basis registers here and spills at NK = 32.)  Three waves per SIMD
// where 168 VGPRs hold the batch without scratch (NCH 3, one entry per step).
template <int NV, int NCH, int PW>
__global__ void __launch_bounds__(BLOCK, (NCH == 3 && PW == 1) ? 3 : 1) k_update_dots(int nc, const Ctl* __restrict__ ctl,
                                                       const double* __restrict__ V, size_t vstride, size_t vcase,
                                                       double* __restrict__ W, int j, double* __restrict__ hv,
                                                       Cg cg) {
  constexpr int NK = Cgs<NCH>::NK, NP = Cgs<NCH>::NP;
  // Cases in reverse order: workgroups are dispatched in the order of their
  // index, so pass 1 ends, and pass 3 starts, on the cases whose basis this
  // pass reads first / last, while it is still in the 256 MB Infinity
  // Cache.  Each case's arithmetic is unchanged.
  const int cs = gridDim.y - 1 - blockIdx.y;
  if (!ctl[cs].active) return;
  __shared__ double tot[NP];
  get_totals<NP>(cg.p1 + (size_t)cs * MMAX * cg.nblk, cg.nblk, j, tot);
  if (blockIdx.x == 0 && threadIdx.x <= j) hv[(size_t)cs * (MMAX + 1) + threadIdx.x] = tot[threadIdx.x];
  const size_t N = (size_t)NV * nc, base = (size_t)cs * N;
  const double* Vc = V + (size_t)cs * vcase;
  double h[NK];
#pragma unroll
  for (int k = 0; k < NK; ++k) h[k] = k <= j ? tot[k] : 0.0;
  double acc[NP];
#pragma unroll
  for (int k = 0; k < NP; ++k) acc[k] = 0.0;
  for (size_t e = ((size_t)blockIdx.x * BLOCK + threadIdx.x) * PW; e < N; e += (size_t)gridDim.x * BLOCK * PW) {
    double vk[NK][PW];
    double w[PW];
    load_pw<PW>(W, base + e, w);
    basis_batch_pw<NK, PW>(Vc, vstride, e, 0, j, vk);
#pragma unroll
    for (int p = 0; p < PW; ++p) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < NK; ++k)
        if (k <= j) s += h[k] * vk[k][p];
      w[p] = w[p] - s;
    }
    store_pw<PW>(W, base + e, w);
#pragma unroll
    for (int k = 0; k < NK; ++k)
#pragma unroll
      for (int p = 0; p < PW; ++p) acc[k] += vk[k][p] * w[p];
  }
  put_partials<NP>(acc, cg.p2 + (size_t)cs * MMAX * cg.nblk, cg.nblk, j + 1);
}

// CGS2 pass 3: h2 = the pass-2 totals; w <- w - sum h2_k V_k and |w|^2.
// Workgroup 0 records h2.  The column is finished (H, Givens) by the next
// column's pass 1, or by the cycle's least squares for the last column.
// Each thread's first grid-stride step is loaded before the totals are added
// and every next step during the current one, except for the third and
// fourth chunks of columns at one entry per step: there the step loads its
// own batch, which leaves 136-150 VGPRs and three waves per SIMD (faster
// than two waves with the loads ahead).
template <int NV, int NCH, int PW>
__global__ void __launch_bounds__(BLOCK, ((NCH == 3 || NCH == 4) && PW == 1) ? 3 : 1) k_update_norm(int nc, const Ctl* __restrict__ ctl,
                                                       const double* __restrict__ V, size_t vstride, size_t vcase,
                                                       double* __restrict__ W, int j, double* __restrict__ h2v,
                                                       Cg cg) {
  constexpr int NKG = Cgs<NCH>::NKG, GS = Cgs<NCH>::GS, NP = Cgs<NCH>::NP;
  constexpr bool PRE = !((NCH == 3 || NCH == 4) && PW == 1);   // next step loaded ahead
  __shared__ double sh[1][4];
  const int cs = blockIdx.y;
  if (!ctl[cs].active) return;
  const size_t N = (size_t)NV * nc, base = (size_t)cs * N;
  const double* Vc = V + (size_t)cs * vcase;
  const size_t step = (size_t)gridDim.x * BLOCK * PW;
  size_t e = ((size_t)blockIdx.x * BLOCK + threadIdx.x) * PW;
  double w[PW], v0[GS][PW];
  if (PRE && e < N) {
    load_pw<PW>(W, base + e, w);
    basis_batch_pw<GS, PW>(Vc, vstride, e, 0, j, v0);
  }
  __shared__ double tot[NP];
  get_totals<NP>(cg.p2 + (size_t)cs * MMAX * cg.nblk, cg.nblk, j, tot);
  if (blockIdx.x == 0 && threadIdx.x <= j) h2v[(size_t)cs * (MMAX + 1) + threadIdx.x] = tot[threadIdx.x];
  double h2[NKG];
#pragma unroll
  for (int k = 0; k < NKG; ++k) h2[k] = k <= j ? tot[min(k, NP - 1)] : 0.0;
  double acc[1] = {0.0};
  while (e < N) {
    if constexpr (!PRE) {
      load_pw<PW>(W, base + e, w);
      basis_batch_pw<GS, PW>(Vc, vstride, e, 0, j, v0);
    }
    double s[PW];
#pragma unroll
    for (int p = 0; p < PW; ++p) s[p] = 0.0;
#pragma unroll
    for (int p = 0; p < PW; ++p)
#pragma unroll
      for (int k = 0; k < GS; ++k)
        if (k <= j) s[p] += h2[k] * v0[k][p];
#pragma unroll
    for (int g = 1; g < Cgs<NCH>::NG; ++g) {
      double v[GS][PW];
      basis_batch_pw<GS, PW>(Vc, vstride, e, g * GS, j, v);
#pragma unroll
      for (int p = 0; p < PW; ++p)
#pragma unroll
        for (int k = 0; k < GS; ++k)
          if (g * GS + k <= j) s[p] += h2[g * GS + k] * v[k][p];
    }
#pragma unroll
    for (int p = 0; p < PW; ++p) w[p] = w[p] - s[p];
    store_pw<PW>(W, base + e, w);
#pragma unroll
    for (int p = 0; p < PW; ++p) acc[0] += w[p] * w[p];
    e += step;
    __builtin_amdgcn_sched_barrier(0);   // the next step's loads reuse this step's registers
    if (PRE && e < N) {
      load_pw<PW>(W, base + e, w);
      basis_batch_pw<GS, PW>(Vc, vstride, e, 0, j, v0);
    }
  }
  block_sum<1>(acc, sh);
  if (threadIdx.x == 0) cg.p3[(size_t)cs * cg.nblk + blockIdx.x] = acc[0];
}

// V_{j+1} = w / max(|w|, tiny) (used for the last column of a cycle; the
// other columns are normalised by the next line apply).
template <int NV>
__global__ void __launch_bounds__(BLOCK) k_normalize(int nc, const Ctl* __restrict__ ctl, const double* __restrict__ W,
                                                     double* __restrict__ Vj1, size_t vcase, Cg cg) {
  const int cs = blockIdx.y;
  if (!ctl[cs].active) return;
  const size_t N = (size_t)NV * nc, base = (size_t)cs * N;
  const double den = fmax(column_norm(cg.p3 + (size_t)cs * cg.nblk, cg.nblk), 2.2250738585072014e-308);
  double* v = Vj1 + (size_t)cs * vcase;
  JR_VEC_LOOP(N) v[e] = W[base + e] / den;
}

// Least squares of the cycle (krylov.py:100-137): finishes the last column
// (H, Givens), then the pivot-drop back substitution (a direction whose pivot
// is below eps (m+1) times the largest is dropped) and H y.  One workgroup
// of 64 threads per case.
__global__ void k_lstsq(int n_cases, const Ctl* __restrict__ ctl, double* __restrict__ H, double* __restrict__ G,
                        int m, double* __restrict__ y, double* __restrict__ hy, Cg cg,
                        const double* __restrict__ hv, const double* __restrict__ h2v) {
  __shared__ double ys[MMAX];
  const int cs = blockIdx.x;
  if (!ctl[cs].active) return;
  column_finish(cs, m - 1, m, cg.p3 + (size_t)cs * cg.nblk, cg.nblk, hv, h2v, H, G);
  const double* Hc = H + (size_t)cs * (MMAX + 1) * MMAX;
  const double* Gc = G + (size_t)cs * GIV;
  const double* Rr = Gc;
  const double* g = Gc + (size_t)(MMAX + 1) * MMAX + 2 * MMAX;
  // the triangle and g staged in LDS by the whole workgroup, so the serial
  // back substitution of thread 0 reads no global memory
  __shared__ double Rs[MMAX * MMAX];
  __shared__ double gs[MMAX];
  for (int e = threadIdx.x; e < m * m; e += blockDim.x) {
    const int row = e / m, col = e - row * m;
    Rs[row * MMAX + col] = Rr[row * MMAX + col];
  }
  for (int jj = threadIdx.x; jj < m; jj += blockDim.x) gs[jj] = g[jj];
  __syncthreads();
  if (threadIdx.x == 0) {
    double largest = fabs(Rs[0]);
    for (int jj = 1; jj < m; ++jj) largest = fmax(largest, fabs(Rs[jj * MMAX + jj]));
    const double rcond = 2.220446049250313e-16 * (m + 1);
    for (int jj = m - 1; jj >= 0; --jj) {
      double rest = gs[jj];
      for (int k = jj + 1; k < m; ++k) rest = rest - Rs[jj * MMAX + k] * ys[k];
      const double piv = Rs[jj * MMAX + jj];
      ys[jj] = fabs(piv) > rcond * largest ? rest / piv : 0.0;
    }
  }
  __syncthreads();
  for (int jj = threadIdx.x; jj < m; jj += blockDim.x) y[(size_t)cs * (MMAX + 1) + jj] = ys[jj];
  for (int i = threadIdx.x; i <= m; i += blockDim.x) {
    double s = 0.0;
    for (int jj = 0; jj < m; ++jj) s += Hc[i * MMAX + jj] * ys[jj];
    hy[(size_t)cs * (MMAX + 1) + i] = s;
  }
}

// End of a cycle: T = sum_{k<m} y_k V_k; action += sum_{k<=m} (Hy)_k V_k;
// r = b - action = -R - action; |r|.
template <int NV, int MM>
__global__ void __launch_bounds__(BLOCK) k_combine(int nc, Ctl* __restrict__ ctl, const double* __restrict__ V,
                                                   size_t vstride, size_t vcase, const double* __restrict__ y,
                                                   const double* __restrict__ hy, int m, double* __restrict__ T,
                                                   double* __restrict__ Act, double* __restrict__ Rv,
                                                   const double* __restrict__ R, Red red) {
  __shared__ double sh[1][4];
  const int cs = blockIdx.y;
  if (!ctl[cs].active) return;
  const size_t N = (size_t)NV * nc, base = (size_t)cs * N;
  const double* Vc = V + (size_t)cs * vcase;
  constexpr int NCH = MM / KC + 1;   // k <= m <= MM
  double yc[NCH * KC], hc[NCH * KC];
#pragma unroll
  for (int k = 0; k < NCH * KC; ++k) {
    yc[k] = k < m ? y[(size_t)cs * (MMAX + 1) + k] : 0.0;
    hc[k] = k <= m ? hy[(size_t)cs * (MMAX + 1) + k] : 0.0;
  }
  double acc[1] = {0.0};
  JR_VEC_LOOP(N) {
    double t = 0.0, a = 0.0;
#pragma unroll
    for (int c = 0; c < NCH; ++c)
      if (c * KC <= m) {
        double v[KC];
        basis_batch<KC>(Vc, vstride, e, c * KC, m, v);
#pragma unroll
        for (int i = 0; i < KC; ++i) {
          const int k = c * KC + i;
          if (k < m) t += yc[k] * v[i];
          if (k <= m) a += hc[k] * v[i];
        }
      }
    T[base + e] = t;
    double act = Act[base + e] + a;
    Act[base + e] = act;
    double r = -R[base + e] - act;
    Rv[base + e] = r;
    acc[0] += r * r;
  }
  __shared__ double tot[1];
  if (finish_sum<1>(acc, red.partials + (size_t)cs * red.stride, red.counters + cs * 4, sh, tot)) {
    if (threadIdx.x == 0) {
      ctl[cs].rnorm = sqrt(tot[0]);
      ctl[cs].cycles += 1;
      ctl[cs].cycles_total += 1;
    }
  }
}

// Relaxation factor (implicit.py:54-81): the largest relative change of
// density, linearized pressure and nu_hat; omega = min(1, fraction / worst).
template <int NV>
__global__ void __launch_bounds__(BLOCK) k_relax(Plan P, double gamma, Ctl* __restrict__ ctl,
                                                 const double* __restrict__ U, const double* __restrict__ X,
                                                 Red red, double fraction) {
  __shared__ double sh[1][4];
  __shared__ int is_last;
  const int cs = blockIdx.y;
  if (ctl[cs].parked) return;   // a parked serving slot: nothing to do this step
  const int nc = P.nc;
  const size_t base = (size_t)cs * NV * nc;
  double worst = 0.0;
  for (int c = blockIdx.x * BLOCK + threadIdx.x; c < nc; c += gridDim.x * BLOCK) {
    double u[NV], q[NV], d[NV];
    for (int k = 0; k < NV; ++k) { u[k] = U[base + (size_t)k * nc + c]; d[k] = X[base + (size_t)k * nc + c]; }
    primitive<double, NV>(u, gamma, q);
    double de = d[3] - q[1] * d[1] - q[2] * d[2];
    double dp = (gamma - 1.0) * (de + 0.5 * (q[1] * q[1] + q[2] * q[2]) * d[0]);
    double r1 = fabs(d[0]) / q[0], r2 = fabs(dp) / q[3];
    double w = fmax(r1, r2);
    if constexpr (NV > 4) {
      double nh = u[4] / q[0];
      w = fmax(w, fabs((d[4] - nh * d[0]) / q[0]) / fmax(fabs(nh), 1.0));
    }
    worst = fmax(worst, w);
  }
  // max-reduction: wave, block, then the last block over the partials
  double wv = wave_max(worst);
  if ((threadIdx.x & 63) == 0) sh[0][threadIdx.x >> 6] = wv;
  __syncthreads();
  double* part = red.partials + (size_t)cs * red.stride;
  if (threadIdx.x == 0) {
    part[blockIdx.x] = fmax(fmax(sh[0][0], sh[0][1]), fmax(sh[0][2], sh[0][3]));
    __threadfence();
    unsigned prev = atomicAdd(red.counters + cs * 4, 1u);
    is_last = prev == gridDim.x - 1;
  }
  __syncthreads();
  if (!is_last) return;
  __threadfence();
  double m = 0.0;
  for (int b = threadIdx.x; b < (int)gridDim.x; b += BLOCK)
    m = fmax(m, __hip_atomic_load(&part[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  m = wave_max(m);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[0][threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    m = fmax(fmax(sh[0][0], sh[0][1]), fmax(sh[0][2], sh[0][3]));
    ctl[cs].omega = fmin(1.0, fraction / fmax(m, 1e-300));
    red.counters[cs * 4] = 0u;
  }
}

// Accept/reject and the SER CFL ramp (implicit.py:139-166).
struct Ramp {
  double cfl_growth, cfl_cutback, cfl_min, cfl_max, krylov_success;
  int relaxed;   // max_relative_update given
};

template <int NV>
__global__ void __launch_bounds__(BLOCK) k_check(Plan P, double gamma, Ctl* __restrict__ ctl,
                                                 const double* __restrict__ U, const double* __restrict__ X,
                                                 const double* __restrict__ Act, const double* __restrict__ R,
                                                 Red red, Ramp ramp) {
  __shared__ double sh[2][4];
  const int cs = blockIdx.y;
  if (ctl[cs].parked) return;   // a parked serving slot: nothing to do this step
  const int nc = P.nc;
  const size_t base = (size_t)cs * NV * nc;
  const double omega = ramp.relaxed ? ctl[cs].omega : 1.0;
  double v[2] = {0.0, 0.0};   // bad cells, |omega action + R|^2
  for (int c = blockIdx.x * BLOCK + threadIdx.x; c < nc; c += gridDim.x * BLOCK) {
    double cand[NV];
    bool finite = true;
    for (int k = 0; k < NV; ++k) {
      size_t o = base + (size_t)k * nc + c;
      double dx = omega * X[o];
      finite = finite && isfinite(dx);
      cand[k] = U[o] + dx;
      double lr = omega * Act[o] + R[o];
      v[1] += lr * lr;
    }
    double q[NV];
    primitive<double, NV>(cand, gamma, q);
    bool phys = q[0] > 1e-10 && q[3] > 1e-10 && isfinite(q[0]) && isfinite(q[3]);
    if (!(phys && finite)) v[0] += 1.0;
  }
  __shared__ double tot[2];
  if (finish_sum<2>(v, red.partials + (size_t)cs * red.stride, red.counters + cs * 4, sh, tot)) {
    if (threadIdx.x == 0) {
      Ctl& c = ctl[cs];
      double lin_red = sqrt(tot[1]) / fmax(c.bnorm, 1e-300);
      bool accept = tot[0] == 0.0 && lin_red < 1.0;
      double reduction = c.res_prev / fmax(c.norm, 1e-300);
      double nxt;
      if (accept && lin_red < ramp.krylov_success && omega >= 1.0)
        nxt = c.cfl * fmin(fmax(reduction, 1.0), ramp.cfl_growth);
      else if (accept)
        nxt = c.cfl;
      else
        nxt = c.cfl / ramp.cfl_cutback;
      c.cfl = fmin(fmax(nxt, ramp.cfl_min), ramp.cfl_max);
      if (accept) c.res_prev = c.norm;
      c.lin_red = lin_red;
      c.accept = accept;
      c.steps += 1;
      if (!accept) c.rejected += 1;
    }
  }
}

template <int NV>
__global__ void __launch_bounds__(BLOCK) k_update(int nc, const Ctl* __restrict__ ctl, double* __restrict__ U,
                                                  const double* __restrict__ X, int relaxed) {
  const int cs = blockIdx.y;
  if (!ctl[cs].accept || ctl[cs].parked) return;   // a parked serving slot keeps its state
  const double omega = relaxed ? ctl[cs].omega : 1.0;
  const size_t N = (size_t)NV * nc, base = (size_t)cs * N;
  JR_VEC_LOOP(N) U[base + e] = U[base + e] + omega * X[base + e];
}

// ---------------------------------------------------------------------------
// Forces (forces.py:48-161): gauge wall pressure and laminar wall shear from
// the ghost-filled primitives Q; CL, CD, CM (pressure only, 2D), CD friction,
// Fx, Fy per case.  One workgroup per case, a fixed-order block sum.

// T = Dual: the shear's tangent, from dQ (the fill's tangent) and a mu_inf
// that may carry a parameter tangent (TM = Dual).
template <class T, class TM>
__device__ void wall_shear(const Plan& P, const Physics& ph, const double* Q, const double* dQ, size_t base,
                           TM mu_inf, int fam, int ff, T* fx, T* fy) {
  const int NJ = P.NJ, nc = P.nc, NJ2 = NJ + 2;
  const double gamma = ph.gamma;
  int cL, cR;
  double sx, sy;
  if (fam == 0) {
    int a = ff / NJ, b = ff - a * NJ;
    cR = a * NJ + b; cL = cR - NJ;
    sx = P.fi[ff]; sy = P.fi[P.nfi + ff];
  } else {
    int a = ff / (NJ + 1), b = ff - a * (NJ + 1);
    cR = a * NJ + b; cL = cR - 1;
    sx = P.fj[ff]; sy = P.fj[P.nfj + ff];
  }
  int iL = cL / NJ, jL = cL - iL * NJ, iR = cR / NJ, jR = cR - iR * NJ;
  const size_t ycen = (size_t)(P.NI + 2) * NJ2;
  double d1x = P.cen[(iR + 1) * NJ2 + jR + 1] - P.cen[(iL + 1) * NJ2 + jL + 1];
  double d1y = P.cen[ycen + (iR + 1) * NJ2 + jR + 1] - P.cen[ycen + (iL + 1) * NJ2 + jL + 1];
  int n0i = iR, n0j = jR, n1i = fam == 0 ? iR : iR + 1, n1j = fam == 0 ? jR + 1 : jR;
  auto node_c = [&](int ni, int nj, int row) -> double {
    const double* c = P.cen + (size_t)row * ycen;
    return 0.25 * (((c[ni * NJ2 + nj] + c[(ni + 1) * NJ2 + nj]) + c[ni * NJ2 + nj + 1]) + c[(ni + 1) * NJ2 + nj + 1]);
  };
  double d2x = node_c(n1i, n1j, 0) - node_c(n0i, n0j, 0);
  double d2y = node_c(n1i, n1j, 1) - node_c(n0i, n0j, 1);
  double det = d1x * d2y - d2x * d1y;
  auto phi_at = [&](int i, int j, T* out) {
    T q[4];
    for (int k = 0; k < 4; ++k) q[k] = Io<T>::ld(Q, dQ, base + (size_t)k * nc + i * NJ + j);
    out[0] = q[1]; out[1] = q[2]; out[2] = gamma * q[3] / q[0];
  };
  auto node_phi = [&](int ni, int nj, T* out) {
    T p00[3], p10[3], p01[3], p11[3];
    phi_at(ni - 1, nj - 1, p00); phi_at(ni, nj - 1, p10); phi_at(ni - 1, nj, p01); phi_at(ni, nj, p11);
    for (int k = 0; k < 3; ++k) out[k] = 0.25 * (((p00[k] + p10[k]) + p01[k]) + p11[k]);
  };
  T pn0[3], pn1[3], pl[3], pr[3];
  node_phi(n0i, n0j, pn0);
  node_phi(n1i, n1j, pn1);
  phi_at(iL, jL, pl);
  phi_at(iR, jR, pr);
  T gx[3], gy[3];
  for (int k = 0; k < 3; ++k) {
    T dp1 = pr[k] - pl[k], dp2 = pn1[k] - pn0[k];
    gx[k] = (dp1 * d2y - dp2 * d1y) / det;
    gy[k] = (d1x * dp2 - d2x * dp1) / det;
  }
  T gv[2][2] = {{gx[0], gy[0]}, {gx[1], gy[1]}};
  T gT[2] = {gx[2], gy[2]};
  T vel[2] = {0.5 * (pl[0] + pr[0]), 0.5 * (pl[1] + pr[1])};
  T Tf = 0.5 * (pl[2] + pr[2]);
  T mu = viscosity<T>(Tf, mu_inf, ph.sutherland);
  T kc = mu / (PR * (gamma - 1.0));
  T f[4];
  viscous_flux<T>(vel, gv, gT, sx, sy, mu, kc, nullptr, f);
  *fx = f[1];
  *fy = f[2];
}

template <int NV>
__global__ void __launch_bounds__(BLOCK) k_forces(Plan P, Physics ph, const Ctl* __restrict__ ctl,
                                                  const double* __restrict__ Q, double length, double cx, double cy,
                                                  double* __restrict__ out, double* __restrict__ loads) {
  bind_case(P, blockIdx.x);
  __shared__ double sh[6][4];
  const int cs = blockIdx.x;
  const size_t base = (size_t)cs * NV * P.nc;
  const double gamma = ph.gamma;
  const double* qi = ctl[cs].q_inf;
  const double speed = hypot_j(qi[1], qi[2]);
  const double dx = qi[1] / speed, dy = qi[2] / speed;   // drag direction
  const double lx = -dy, ly = dx;                        // lift direction
  const double q_area = 0.5 * qi[0] * speed * speed * length;
  double v[6] = {0, 0, 0, 0, 0, 0};   // lift, drag, moment, shear x, shear y, (spare)
  double fxs = 0.0, fys = 0.0;
  for (int w = threadIdx.x; w < P.n_walls; w += BLOCK) {
    int cell = P.w_cell[w], fam = P.w_fam[w], ff = P.w_face[w];
    double into = P.w_into[w];
    double sx = into * (fam == 0 ? P.fi[ff] : P.fj[ff]);
    double sy = into * (fam == 0 ? P.fi[P.nfi + ff] : P.fj[P.nfj + ff]);
    double p = Q[base + (size_t)3 * P.nc + cell] - 1.0 / gamma;
    double px = -p * sx, py = -p * sy;
    double tx = 0.0, ty = 0.0;
    if (ph.viscous) {
      wall_shear<double>(P, ph, Q, nullptr, base, ctl[cs].mu_inf, fam, ff, &tx, &ty);
      tx *= into;
      ty *= into;
    }
    if (loads) {   // per wall face: gauge pressure, viscous force on the body (x, y)
      double* lw = loads + (size_t)cs * 3 * P.n_walls + w;
      lw[0] = p;
      lw[P.n_walls] = tx;
      lw[2 * (size_t)P.n_walls] = ty;
    }
    double ldx = px + tx, ldy = py + ty;
    double rx = P.w_mid[w] - cx, ry = P.w_mid[P.n_walls + w] - cy;
    double moment = ry * px - rx * py;
    v[0] += (lx * ldx + ly * ldy) / q_area;
    v[1] += (dx * ldx + dy * ldy) / q_area;
    v[2] += (moment / length) / q_area;
    v[3] += tx;
    v[4] += ty;
    fxs += ldx;
    fys += ldy;
  }
  v[5] = 0.0;
  block_sum<6>(v, sh);
  double f2[6] = {fxs, fys, 0, 0, 0, 0};
  block_sum<6>(f2, sh);
  if (threadIdx.x == 0) {
    double* o = out + (size_t)cs * 8;
    o[0] = v[0];
    o[1] = v[1];
    o[2] = v[2];
    o[3] = (dx * v[3] + dy * v[4]) / q_area;
    o[4] = f2[0];
    o[5] = f2[1];
    o[6] = 0.0;
    o[7] = 0.0;
  }
}

// The tangent of k_forces' coefficients along (dU, dq_inf, dmu_inf): the
// same integration on dual numbers, with the primitives' tangent dQ (the
// parameter-tangent fill) and the freestream's tangent P.dfree, which turns
// the lift and drag directions and scales the dynamic pressure (the
// reference's output JVP, src/jaxrans/solvers/sensitivity.py:63 over
// derivatives.py:82-84).  out[case][8] = d(CL, CD, CM, CD friction, Fx, Fy).
template <int NV>
__global__ void __launch_bounds__(BLOCK) k_forces_tangent(Plan P, Physics ph, const Ctl* __restrict__ ctl,
                                                          const double* __restrict__ Q, const double* __restrict__ dQ,
                                                          double length, double cx, double cy,
                                                          double* __restrict__ out) {
  bind_case(P, blockIdx.x);
  __shared__ double sh[14][4];
  const int cs = blockIdx.x;
  const size_t base = (size_t)cs * NV * P.nc;
  const double gamma = ph.gamma;
  Dual qi[4];
  for (int k = 0; k < 4; ++k) qi[k] = free_val<true>(ctl[cs].q_inf[k], P.dfree, cs, k);
  const Dual mu_inf = free_val<true>(ctl[cs].mu_inf, P.dfree, cs, 6);
  const Dual speed = hypot_j(qi[1], qi[2]);
  const Dual dx = qi[1] / speed, dy = qi[2] / speed;   // drag direction
  const Dual lx = -dy, ly = dx;                        // lift direction
  const Dual q_area = 0.5 * qi[0] * speed * speed * length;
  // lift, drag, moment, shear x, shear y, Fx, Fy: values then tangents
  Dual acc[7];
  for (int k = 0; k < 7; ++k) acc[k] = mkd(0.0, 0.0);
  for (int w = threadIdx.x; w < P.n_walls; w += BLOCK) {
    const int cell = P.w_cell[w], fam = P.w_fam[w], ff = P.w_face[w];
    const double into = P.w_into[w];
    const double sx = into * (fam == 0 ? P.fi[ff] : P.fj[ff]);
    const double sy = into * (fam == 0 ? P.fi[P.nfi + ff] : P.fj[P.nfj + ff]);
    const Dual p = Io<Dual>::ld(Q, dQ, base + (size_t)3 * P.nc + cell) - 1.0 / gamma;
    const Dual px = -p * sx, py = -p * sy;
    Dual tx = mkd(0.0, 0.0), ty = mkd(0.0, 0.0);
    if (ph.viscous) {
      wall_shear<Dual>(P, ph, Q, dQ, base, mu_inf, fam, ff, &tx, &ty);
      tx = tx * into;
      ty = ty * into;
    }
    const Dual ldx = px + tx, ldy = py + ty;
    const double rx = P.w_mid[w] - cx, ry = P.w_mid[P.n_walls + w] - cy;
    const Dual moment = ry * px - rx * py;
    acc[0] = acc[0] + (lx * ldx + ly * ldy) / q_area;
    acc[1] = acc[1] + (dx * ldx + dy * ldy) / q_area;
    acc[2] = acc[2] + (moment / length) / q_area;
    acc[3] = acc[3] + tx;
    acc[4] = acc[4] + ty;
    acc[5] = acc[5] + ldx;
    acc[6] = acc[6] + ldy;
  }
  double v[14];
  for (int k = 0; k < 7; ++k) { v[k] = acc[k].v; v[7 + k] = acc[k].d; }
  block_sum<14>(v, sh);
  if (threadIdx.x == 0) {
    const Dual shear_x = mkd(v[3], v[10]), shear_y = mkd(v[4], v[11]);
    const Dual friction = (dx * shear_x + dy * shear_y) / q_area;
    double* o = out + (size_t)cs * 8;
    o[0] = v[7];
    o[1] = v[8];
    o[2] = v[9];
    o[3] = friction.d;
    o[4] = v[12];
    o[5] = v[13];
    o[6] = 0.0;
    o[7] = 0.0;
  }
}

}  // namespace jr
