// Block Thomas solve of the 2D line chains with the lanes of a wavefront over
// the entries of the 4x4 flow blocks: the forward and backward substitutions
// of src/jaxrans/core/linalg.py:103-132 on the factor k_thomas writes (per
// slot the lower block L, the pivot inverse P^-1 and C = P^-1 U; the scalar SA
// chain l, 1/p, c).
//
// A wavefront serves up to four chains, sixteen lanes each.  Lane (r, q) of a
// chain (r, q = 0..3) holds entry (r, q) of the step's blocks, so a block
// times a vector is one multiply per lane and a two-step DPP butterfly:
//   * over q inside a quad: the result in row form (lane (r, q) holds entry r);
//   * over r across the four quads of the row, with the block read
//     transposed: the result in column form (lane (r, q) holds entry q).
// Both butterflies pair lanes symmetrically, so every lane of a chain ends
// with the same bits.  The forward step d_j = P^-1 (b_j - L_j d_{j-1}) takes
// one of each and keeps d in column form.  The backward step
// x_j = d_j - C_j x_{j+1} takes one per step, alternating the two forms (C
// stored transposed on the even steps counted from the top of the chain).
// The scalar SA chain is a Thomas solve of its own: lane l < 8 holds row
// j0 + l of the current eight-row chunk, and the recurrence hands its value
// to the next lane with a row rotation.
//
// Memory runs LN_RING - 1 chunks ahead of the arithmetic: the chunk buffers
// form a ring in registers, each filled in one burst of loads, so waiting for
// the oldest never waits for a younger one.  The chain's rows -- right-hand
// side, then d, then x, then (with dq_out) the tangent primitives -- and its
// cell indices live in LDS.  The workgroup is one wavefront: lanes exchange
// LDS values in program order, with no barrier.
#pragma once
#include "jr_solver.h"

namespace jr {

constexpr int LN_CH = 8;     // rows per chunk
constexpr int LN_RING = 4;   // chunk buffers in the ring

// DPP controls (GFX9 encodings): quad permutations and row (16-lane) moves.
constexpr int DPP_QUAD_XOR1 = 0xB1;     // quad_perm [1,0,3,2]
constexpr int DPP_QUAD_XOR2 = 0x4E;     // quad_perm [2,3,0,1]
constexpr int DPP_QUAD_REV = 0x1B;      // quad_perm [3,2,1,0]
constexpr int DPP_HALF_MIRROR = 0x141;  // lane i <- 7 - i within each half row
constexpr int DPP_ROR1 = 0x121;         // lane i <- (i - 1) mod 16
constexpr int DPP_ROR8 = 0x128;         // lane i <- (i - 8) mod 16

template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// v moved by CTRL into lanes 12-15 of each row (DPP bank 3); the other lanes
// keep v.  Executed by every lane, so the source lanes are always active.
template <int CTRL>
__device__ __forceinline__ float dpp_bank3(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), CTRL, 0xF, 0x8, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_bank3(double v) {
  const int lo = __builtin_amdgcn_update_dpp(__double2loint(v), __double2loint(v), CTRL, 0xF, 0x8, false);
  const int hi = __builtin_amdgcn_update_dpp(__double2hiint(v), __double2hiint(v), CTRL, 0xF, 0x8, false);
  return __hiloint2double(hi, lo);
}

// c ? a : b on integers by masks, so the compiler keeps it a select (no
// branch that narrows the execution mask).
__device__ __forceinline__ int pick(bool c, int a, int b) {
  const int m = -(int)c;
  return (a & m) | (b & ~m);
}

// a x summed over q of the quad (lanes 4r .. 4r+3): row form.  Products and
// sums are rounded separately (no fused multiply-add), so the symmetric
// butterfly leaves the same bits in every lane of the quad.
template <class R>
__device__ __forceinline__ R dot_in_quad(R a, R x) {
#pragma clang fp contract(off)
  R p = a * x;
  p += dpp_mov<DPP_QUAD_XOR1>(p);
  return p + dpp_mov<DPP_QUAD_XOR2>(p);
}

// a x summed over r of the row's four quads (lanes q, 4+q, 8+q, 12+q): column
// form.  Half-mirror then quad reversal moves quad r ^ 1 onto quad r; the
// rotation by eight moves quad r ^ 2.
template <class R>
__device__ __forceinline__ R dot_over_quads(R a, R x) {
#pragma clang fp contract(off)
  R p = a * x;
  p += dpp_mov<DPP_QUAD_REV>(dpp_mov<DPP_HALF_MIRROR>(p));
  return p + dpp_mov<DPP_ROR8>(p);
}

// Dynamic LDS of one k_line_apply wavefront: `rows` chain rows of NV doubles
// and their cell indices, then a sink of one double per lane that takes the
// per-step stores of lanes with nothing to store (the stores stay
// unconditional: no change of the execution mask inside a step).
template <int NV>
__host__ __device__ inline size_t line_rows_lds(int rows) {
  return (size_t)rows * (NV * sizeof(double) + sizeof(int));
}
template <int NV>
__host__ __device__ inline size_t line_apply_lds(int rows) {
  return line_rows_lds<NV>(rows) + 64 * sizeof(double);
}

// The chains of one k_line_apply launch in length classes, longest first.
// Class c takes the chains order[first[c] .. first[c] + count[c]), at most
// lmax[c] rows each (a multiple of 16), gper[c] per wavefront, in wavefronts
// wave0[c] .. wave0[c + 1] - 1; every class fits gper * lmax <= rows, the LDS
// rows of a wavefront, so long and short chains share one launch.
constexpr int LN_MAX_CLASSES = 8;
struct LineLaunch {
  int n, rows;
  int wave0[LN_MAX_CLASSES + 1];
  int first[LN_MAX_CLASSES], count[LN_MAX_CLASSES], lmax[LN_MAX_CLASSES], gper[LN_MAX_CLASSES];
};

// One lane's factor values of a step: (L, P^-1 transposed) entries, adjacent.
template <class R>
struct alignas(2 * sizeof(R)) Pair {
  R l, p;
};

// z <- M^-1 r (ACC: z += M^-1 r) over the chains of LL, the longest chain of
// each wavefront first (order is by length, longest first).  Workgroups past
// the chains copy the cells on no chain (the identity,
// preconditioners.py:110-115), 64 per workgroup.  With vout set, the input is
// the Arnoldi vector w / |w| (from the pass-3 partials p3), also written to
// vout.  With dq_out set, the tangent of the primitives at (U, z) is written
// on every cell, ghosts through the linearised fill (gjac).
//
// The forward buffer of chunk c holds its factor pairs, SA values and
// right-hand sides; the backward buffer of chunk m its C entries and SA c and
// what finishing chunk m - 1's rows needs (state, ghost flag, z): finishing
// -- z, the tangent primitives, dq_out -- happens at the start of chunk m, so
// its stores overlap that chunk's steps.
template <class R, int NV, bool ACC>
__global__ void __launch_bounds__(64) k_line_apply(Plan P, const Ctl* __restrict__ ctl, const int* __restrict__ order,
                                                   LineLaunch LL, const R* __restrict__ fac,
                                                   const R* __restrict__ fac_sa, const double* __restrict__ r,
                                                   double* __restrict__ z, int check_active,
                                                   double* __restrict__ vout, const double* __restrict__ p3, int nblk,
                                                   double* __restrict__ dq_out, const double* __restrict__ Uc,
                                                   double gamma, const double* __restrict__ gjac) {
  extern __shared__ double xs[];
  const int cs = blockIdx.y, lane = threadIdx.x;
  if (skip_case(ctl, cs, check_active)) return;
  const int nc = P.nc;
  const size_t base = (size_t)cs * NV * nc;
  const double den = apply_den(p3 + (size_t)cs * nblk, nblk, vout);   // every thread: a barrier inside
  const int w = blockIdx.x, waves = LL.wave0[LL.n];
  if (w >= waves) {   // cells on no chain: the identity
    const int k = (w - waves) * 64 + lane;
    if (k >= P.n_free) return;
    const int cell = P.free_cell[k];
    double x[NV];
    for (int v = 0; v < NV; ++v) {
      const size_t o = base + (size_t)v * nc + cell;
      x[v] = vout ? r[o] / den : r[o];
      if (vout) vout[o] = x[v];
      z[o] = ACC ? z[o] + x[v] : x[v];
    }
    if (dq_out && !P.ghost[cell]) {
      double u[NV];
      for (int v = 0; v < NV; ++v) u[v] = Uc[base + (size_t)v * nc + cell];
      write_tangent_q<NV>(P, u, x, gamma, dq_out, base, cell, gjac + (size_t)cs * P.n_final * NV * NV);
    }
    return;
  }
  int cl = 0;
  while (cl + 1 < LL.n && w >= LL.wave0[cl + 1]) ++cl;
  const int lmax = LL.lmax[cl], gper = LL.gper[cl];
  const int g = lane >> 4, l = lane & 15, rr = (lane >> 2) & 3, qq = lane & 3, e = lane & 15;
  const int idx = (w - LL.wave0[cl]) * gper + g;
  const int chain = (g < gper && idx < LL.count[cl]) ? order[LL.first[cl] + idx] : -1;
  const int s0 = chain >= 0 ? P.l_start[chain] : 0;
  const int L = chain >= 0 ? P.l_start[chain + 1] - s0 : 0;
  // chunks, rounded up to whole turns of the ring; lane 0's chain is the
  // wavefront's longest
  const int nch = (uni(L) + LN_CH * LN_RING - 1) / (LN_CH * LN_RING) * LN_RING;
  // this chain's rows and their cells; a lane with no chain reads (never
  // writes) the last group's
  const int gs = g < gper ? g : gper - 1;
  double* X = xs + (size_t)gs * lmax * NV;
  // the lane's sink, as an offset from X
  const int sink = (int)(line_rows_lds<NV>(LL.rows) / sizeof(double)) + lane - gs * lmax * NV;
  int* cells = reinterpret_cast<int*>(xs + (size_t)gper * lmax * NV) + (size_t)gs * lmax;
  const Pair<R>* F = reinterpret_cast<const Pair<R>*>(fac + (size_t)cs * P.n_slots * 48);
  const R* Fc = fac + (size_t)cs * P.n_slots * 48 + (size_t)P.n_slots * 32;
  const R* Fs = fac_sa;
  if constexpr (NV > 4) Fs = fac_sa + (size_t)cs * P.n_slots * 3;
  for (int i = l; i < L; i += 16) cells[i] = P.l_cell[s0 + i];   // the chain's cells, once
  const bool rowl = l < LN_CH;   // lane l holds row j0 + l of a chunk

  // ---- forward: d_j = P^-1_j (b_j - L_j d_{j-1})
  struct Fwd {
    Pair<R> f[LN_CH];   // (L_j (r, q), P^-1_j (q, r)) of the chunk's rows
    R sl, sp;           // SA l, 1/p of row j0 + l
    double x[NV];       // right-hand side of row j0 + l
    int cell;           // its cell (-1: none)
  };
  // Every load of a buffer is issued, rows past the chain reading its last
  // row (or slot 0 and cell 0 for a lane with no chain), so the waits of the
  // ring can count on each load having been issued.  Values of rows past the
  // chain are used as loaded: they only reach rows past the chain, which are
  // never written.
  const int last = L > 0 ? L - 1 : 0;
  auto load_fwd = [&](int c, Fwd& f) {
    const int j0 = c * LN_CH, i = j0 + l;
#pragma unroll
    for (int k = 0; k < LN_CH; ++k) {
      f.f[k] = F[(size_t)(s0 + min(j0 + k, last)) * 16 + e];
    }
    const bool ok = rowl && i < L;
    const int ic = min(i, last);
    const int cv = cells[ic];
    f.cell = pick(ok, cv, -1);
    const int cl0 = pick(ok, cv, 0);
    for (int v = 0; v < NV; ++v) f.x[v] = r[base + (size_t)v * nc + cl0];
    f.sl = R(0);
    f.sp = R(0);
    if constexpr (NV > 4) {
      f.sl = Fs[(size_t)(s0 + ic) * 3];
      f.sp = Fs[(size_t)(s0 + ic) * 3 + 1];
    }
  };
  R dcol = R(0), sd = R(0);
  auto fwd_chunk = [&](int c, const Fwd& cur) {
    const int j0 = c * LN_CH;
    R sb = R(0);
    if (cur.cell >= 0)   // the chunk's right-hand sides to LDS (normalised: also to vout)
      for (int v = 0; v < NV; ++v) {
        const double x = vout ? cur.x[v] / den : cur.x[v];
        if (vout) vout[base + (size_t)v * nc + cur.cell] = x;
        X[(size_t)(j0 + l) * NV + v] = x;
        if (v == 4) sb = (R)x;
      }
    // right-hand sides read two steps ahead of their use, ahead of the step's
    // own LDS store, so their latency overlaps the recurrence
    const int jr = j0 < lmax - LN_CH ? j0 : lmax - LN_CH;   // in-bounds rows past the chain's end
    R bq[LN_CH];
    bq[0] = (R)X[(size_t)jr * NV + rr];
    bq[1] = (R)X[(size_t)(jr + 1) * NV + rr];
    if constexpr (NV > 4) sd = dpp_bank3<DPP_ROR8>(sd);   // lane 15 <- the last row, lane 7
#pragma unroll
    for (int k = 0; k < LN_CH; ++k) {
      const int j = j0 + k;
      if (k + 2 < LN_CH) bq[k + 2] = (R)X[(size_t)(jr + k + 2) * NV + rr];
      const R s = bq[k] - dot_in_quad<R>(cur.f[k].l, dcol);
      dcol = dot_over_quads<R>(cur.f[k].p, s);
      X[pick(rr == 0 && j < L, j * NV + qq, sink)] = (double)dcol;
      if constexpr (NV > 4) {
        const R nsd = cur.sp * (sb - cur.sl * dpp_mov<DPP_ROR1>(sd));
        sd = l == k ? nsd : sd;
      }
    }
    if constexpr (NV > 4)
      if (rowl && j0 + l < L) X[(size_t)(j0 + l) * NV + 4] = (double)sd;
  };
  {
    static_assert(LN_RING == 4, "the ring below has four buffers");
    Fwd A, B, Cb, D;
    load_fwd(0, A);
    load_fwd(1, B);
    load_fwd(2, Cb);
    for (int c = 0; c < nch; c += LN_RING) {   // past the chains' ends every lane is masked: no traffic
      load_fwd(c + 3, D);
      fwd_chunk(c, A);
      load_fwd(c + 4, A);
      fwd_chunk(c + 1, B);
      load_fwd(c + 5, B);
      fwd_chunk(c + 2, Cb);
      load_fwd(c + 6, Cb);
      fwd_chunk(c + 3, D);
    }
  }

  // ---- backward: x_j = d_j - C_j x_{j+1}, from the top of the chain
  struct Bwd {
    R fc[LN_CH];               // C_j entries of the chunk's rows (transposed on even steps)
    R sc;                      // SA c of row jt - l
    double u[NV];              // state of chunk m - 1's row for lane l (dq_out)
    double za[ACC ? NV : 1];   // its z (accumulate)
    int cell;                  // its cell (-1: none)
    int ghost;                 // whether that cell is a ghost (dq_out)
  };
  auto load_bwd = [&](int m, Bwd& b) {
    const int jt = L - 1 - m * LN_CH, jl = jt - l, jp = jl + LN_CH;
#pragma unroll
    for (int k = 0; k < LN_CH; ++k) {
      b.fc[k] = Fc[(size_t)(s0 + max(jt - k, 0)) * 16 + e];
    }
    b.sc = R(0);
    if constexpr (NV > 4) b.sc = Fs[(size_t)(s0 + max(jl, 0)) * 3 + 2];
    const bool ok = rowl && m > 0 && jp >= 0 && jp < L;
    const int cv = cells[max(min(jp, last), 0)];
    b.cell = pick(ok, cv, -1);
    const int cl0 = pick(ok, cv, 0);
    b.ghost = 1;
    for (int v = 0; v < NV; ++v) b.u[v] = 0.0;
    if (dq_out) {   // uniform
      for (int v = 0; v < NV; ++v) b.u[v] = Uc[base + (size_t)v * nc + cl0];
      b.ghost = P.ghost[cl0];
    }
    if constexpr (ACC)
      for (int v = 0; v < NV; ++v) b.za[v] = z[base + (size_t)v * nc + cl0];
  };
  R xin = R(0), sx = R(0);   // x_{j+1}: row form before even steps, column form before odd ones
  // row L - 1 - mm * 8 - l of chunk mm is solved (flow part in LDS, SA part
  // in sx): z, the tangent primitives (to dq_out and, for the ghosts, to LDS)
  auto finish = [&](const Bwd& b, int mm) {
    if (b.cell < 0) return;
    const int jp = L - 1 - mm * LN_CH - l;
    double x[NV];
    for (int v = 0; v < 4; ++v) x[v] = X[(size_t)jp * NV + v];
    if constexpr (NV > 4) x[4] = (double)sx;
    for (int v = 0; v < NV; ++v) {
      const size_t o = base + (size_t)v * nc + b.cell;
      if constexpr (ACC) z[o] = b.za[v] + x[v];
      else z[o] = x[v];
    }
    if (dq_out) {
      double d[NV];
      tangent_primitive<NV>(b.u, x, gamma, d);
      if (!b.ghost)
        for (int v = 0; v < NV; ++v) dq_out[base + (size_t)v * nc + b.cell] = d[v];
      for (int v = 0; v < NV; ++v) X[(size_t)jp * NV + v] = d[v];
    }
  };
  auto bwd_chunk = [&](int m, const Bwd& bc) {
    finish(bc, m - 1);
    const int jt = L - 1 - m * LN_CH, jl = jt - l;
    R sdv = R(0);
    if constexpr (NV > 4) {
      if (rowl && jl >= 0) sdv = (R)X[(size_t)jl * NV + 4];
      sx = dpp_bank3<DPP_ROR8>(sx);   // lane 15 <- the previous chunk's last row, lane 7
    }
    // d_j read two steps ahead (column entry on even steps, row entry on odd)
    auto ld_d = [&](int k) { return (R)X[max(jt - k, 0) * NV + (k % 2 == 0 ? qq : rr)]; };
    R dq[LN_CH];
    dq[0] = ld_d(0);
    dq[1] = ld_d(1);
#pragma unroll
    for (int k = 0; k < LN_CH; ++k) {
      const int j = jt - k;
      if (k + 2 < LN_CH) dq[k + 2] = ld_d(k + 2);
      if (k % 2 == 0) {
        xin = dq[k] - dot_over_quads<R>(bc.fc[k], xin);
        X[pick(rr == 0 && j >= 0, j * NV + qq, sink)] = (double)xin;
      } else {
        xin = dq[k] - dot_in_quad<R>(bc.fc[k], xin);
        X[pick(qq == 0 && j >= 0, j * NV + rr, sink)] = (double)xin;
      }
      if constexpr (NV > 4) {
        const R nsx = sdv - bc.sc * dpp_mov<DPP_ROR1>(sx);
        sx = l == k ? nsx : sx;
      }
    }
  };
  {
    Bwd A, B, Cb, D;
    load_bwd(0, A);
    load_bwd(1, B);
    load_bwd(2, Cb);
    for (int m = 0; m < nch; m += LN_RING) {   // below the chains' starts every lane is masked: no traffic
      load_bwd(m + 3, D);
      bwd_chunk(m, A);
      load_bwd(m + 4, A);
      bwd_chunk(m + 1, B);
      load_bwd(m + 5, B);
      bwd_chunk(m + 2, Cb);
      load_bwd(m + 6, Cb);
      bwd_chunk(m + 3, D);
    }
    finish(A, nch - 1);   // A holds chunk nch: the last chunk's rows
  }
  if (dq_out && chain >= 0) {   // the ghosts rooted on this chain: dQ_ghost = J dQ_root
    const double* Jc = gjac + (size_t)cs * P.n_final * NV * NV;
    const int d0 = P.cq_start[chain], nd = P.cq_start[chain + 1] - d0;
    for (int i = l; i < nd; i += 16) {
      const int q = P.cq_q[d0 + i];
      const double* J = Jc + (size_t)q * NV * NV;
      const double* d = X + (size_t)P.gq_row[q] * NV;
      const int dest = P.gq_dest[q];
      for (int a = 0; a < NV; ++a) {
        double acc = 0.0;
        for (int k = 0; k < NV; ++k) acc += J[a * NV + k] * d[k];
        dq_out[base + (size_t)a * nc + dest] = acc;
      }
    }
  }
}

}  // namespace jr
