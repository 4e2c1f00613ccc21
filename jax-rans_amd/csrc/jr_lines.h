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
// read transposed on the even steps counted from the top of the chain).  The
// scalar SA chain is a Thomas solve of its own: lane l holds row j0 + l of the
// current sixteen-row chunk, and the recurrence hands its value to the next
// lane with a row rotation.
//
// The factor values and right-hand sides of the next sixteen rows are loaded
// while the current sixteen run, so the sequential steps do not wait on
// memory.  The chain's rows -- right-hand side, then d, then x, then (with
// dq_out) the tangent primitives -- live in LDS, [gper][lmax][NV] doubles.
// The workgroup is one wavefront: lanes exchange LDS values in program order,
// with no barrier.
#pragma once
#include "jr_solver.h"

namespace jr {

constexpr int LN_CH = 16;   // rows per chunk = lanes per chain

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

// Dynamic LDS of one k_line_apply launch: per chain lmax rows of NV doubles,
// then lmax cell indices.
template <int NV>
__host__ __device__ inline size_t line_apply_lds(int lmax, int gper) {
  return (size_t)gper * (size_t)lmax * (NV * sizeof(double) + sizeof(int));
}

// z <- M^-1 r (ACC: z += M^-1 r) over chains order[first .. first +
// count), gper per wavefront, each at most lmax rows (a multiple of 16); the
// longest chain of each wavefront comes first (order is by length, longest
// first).  Workgroups past the chains copy the cells on no chain (the identity,
// preconditioners.py:110-115), 64 per workgroup.  With vout set, the input is
// the Arnoldi vector w / |w| (from the pass-3 partials p3), also written to
// vout.  With dq_out set, the tangent of the primitives at (U, z) is written
// on every cell, ghosts through the linearised fill (gjac).
//
// Loads run one chunk ahead: during chunk c of the forward sweep, the factor
// values and right-hand sides of chunk c + 1 and the cell indices of chunk
// c + 2; during chunk m of the backward sweep, the factor values of chunk
// m + 1 and, for its rows, the state and ghost flags (dq_out) and z (ACC,
// accumulate).  Each backward chunk finishes its own rows: z, the tangent
// primitives and dq_out, with no load of its own.
template <class R, int NV, bool ACC>
__global__ void __launch_bounds__(64) k_line_apply(Plan P, const Ctl* __restrict__ ctl, const int* __restrict__ order,
                                                   int first, int count, int lmax, int gper, const R* __restrict__ fac,
                                                   const R* __restrict__ fac_sa, const double* __restrict__ r,
                                                   double* __restrict__ z, int check_active,
                                                   double* __restrict__ vout, const double* __restrict__ p3, int nblk,
                                                   double* __restrict__ dq_out, const double* __restrict__ Uc,
                                                   double gamma, const double* __restrict__ gjac) {
  extern __shared__ double xs[];
  const int cs = blockIdx.y, lane = threadIdx.x;
  if (check_active && !ctl[cs].active) return;
  const int nc = P.nc;
  const size_t base = (size_t)cs * NV * nc;
  const double den = apply_den(p3 + (size_t)cs * nblk, nblk, vout);   // every thread: a barrier inside
  const int waves = (count + gper - 1) / gper;
  if ((int)blockIdx.x >= waves) {   // cells on no chain: the identity
    const int k = ((int)blockIdx.x - waves) * 64 + lane;
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
  const int g = lane >> 4, l = lane & 15, rr = (lane >> 2) & 3, qq = lane & 3;
  const int idx = (int)blockIdx.x * gper + g;
  const int chain = (g < gper && idx < count) ? order[first + idx] : -1;
  const int s0 = chain >= 0 ? P.l_start[chain] : 0;
  const int L = chain >= 0 ? P.l_start[chain + 1] - s0 : 0;
  const int nch = (uni(L) + LN_CH - 1) / LN_CH;   // lane 0's chain is the wavefront's longest
  // this chain's rows and their cells; a lane with no chain reads (never
  // writes) the last group's
  const int gs = g < gper ? g : gper - 1;
  double* X = xs + (size_t)gs * lmax * NV;
  int* cells = reinterpret_cast<int*>(xs + (size_t)gper * lmax * NV) + (size_t)gs * lmax;
  const R* F = fac + (size_t)cs * P.n_slots * 48;
  const R* Fc = F + (size_t)P.n_slots * 32;
  const R* Fs = fac_sa;
  if constexpr (NV > 4) Fs = fac_sa + (size_t)cs * P.n_slots * 3;

  // ---- forward: d_j = P^-1_j (b_j - L_j d_{j-1})
  struct Fwd {
    R fl[LN_CH], fp[LN_CH];   // L_j (r, q) and P^-1_j (q, r) of the chunk's rows
    R sl, sp;                 // SA l, 1/p of row j0 + l
    double x[NV];             // right-hand side of row j0 + l
    int cell;                 // its cell (-1: past the chain)
  };
  // chunk c's loads; `cell` holds the cell of row c * 16 + l on entry and
  // that of row (c + 1) * 16 + l on return (its load left in flight)
  auto load_fwd = [&](int c, Fwd& f, int& cell) {
    const int j0 = c * LN_CH;
    f.cell = cell;
    for (int v = 0; v < NV; ++v) f.x[v] = 0.0;
    if (cell >= 0)
      for (int v = 0; v < NV; ++v) f.x[v] = r[base + (size_t)v * nc + cell];
#pragma unroll
    for (int k = 0; k < LN_CH; ++k) {
      f.fl[k] = R(0);
      f.fp[k] = R(0);
      if (j0 + k < L) {
        const R* o = F + (size_t)(s0 + j0 + k) * 32;
        f.fl[k] = o[rr * 4 + qq];
        f.fp[k] = o[16 + qq * 4 + rr];
      }
    }
    f.sl = R(0);
    f.sp = R(0);
    if constexpr (NV > 4)
      if (j0 + l < L) {
        f.sl = Fs[(size_t)(s0 + j0 + l) * 3];
        f.sp = Fs[(size_t)(s0 + j0 + l) * 3 + 1];
      }
    const int i = j0 + LN_CH + l;
    cell = -1;
    if (i < L) cell = P.l_cell[s0 + i];
  };

  // ---- backward: x_j = d_j - C_j x_{j+1}, from the top of the chain.
  // Bwd(m) holds chunk m's factor values and what finishing chunk m - 1's
  // rows needs (done at the start of chunk m, so its stores overlap the steps)
  struct Bwd {
    R fc[LN_CH];      // C_j: (q, r) on even steps, (r, q) on odd
    R sc;             // SA c of row jt - l
    double u[NV];     // state of chunk m - 1's row for lane l (dq_out)
    double za[ACC ? NV : 1];   // its z (accumulate)
    int cell;         // its cell (-1: none)
    int ghost;        // whether that cell is a ghost (dq_out)
  };
  auto load_bwd = [&](int m, Bwd& b) {
    const int jt = L - 1 - m * LN_CH, jl = jt - l, jp = jl + LN_CH;
#pragma unroll
    for (int k = 0; k < LN_CH; ++k) {
      b.fc[k] = R(0);
      if (jt - k >= 0) b.fc[k] = Fc[(size_t)(s0 + jt - k) * 16 + (k % 2 == 0 ? qq * 4 + rr : rr * 4 + qq)];
    }
    b.sc = R(0);
    if constexpr (NV > 4)
      if (jl >= 0) b.sc = Fs[(size_t)(s0 + jl) * 3 + 2];
    b.cell = (m > 0 && jp >= 0 && jp < L) ? cells[jp] : -1;
    b.ghost = 1;
    for (int v = 0; v < NV; ++v) b.u[v] = 0.0;
    if (b.cell >= 0 && dq_out) {
      for (int v = 0; v < NV; ++v) b.u[v] = Uc[base + (size_t)v * nc + b.cell];
      b.ghost = P.ghost[b.cell];
    }
    if constexpr (ACC)
      for (int v = 0; v < NV; ++v) b.za[v] = b.cell >= 0 ? z[base + (size_t)v * nc + b.cell] : 0.0;
  };
  // row L - 1 - mm * 16 - l of chunk mm is solved (flow part in LDS, SA part
  // sxl): z, the tangent primitives (to dq_out and, for the ghosts, to LDS)
  auto finish = [&](const Bwd& b, int mm, R sxl) {
    if (b.cell < 0) return;
    const int jp = L - 1 - mm * LN_CH - l;
    double x[NV];
    for (int v = 0; v < 4; ++v) x[v] = X[(size_t)jp * NV + v];
    if constexpr (NV > 4) x[4] = (double)sxl;
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

  R dcol = R(0), sd = R(0);
  Fwd cur, nxt;
  Bwd bc, bn;
  {
    int cell = l < L ? P.l_cell[s0 + l] : -1;
    if (nch > 0) load_fwd(0, nxt, cell);
    for (int c = 0; c < nch; ++c) {
      cur = nxt;
      const int j0 = c * LN_CH;
      R sb = R(0);
      if (cur.cell >= 0) {   // the chunk's right-hand sides to LDS (normalised: also to vout)
        cells[j0 + l] = cur.cell;
        for (int v = 0; v < NV; ++v) {
          const double x = vout ? cur.x[v] / den : cur.x[v];
          if (vout) vout[base + (size_t)v * nc + cur.cell] = x;
          X[(size_t)(j0 + l) * NV + v] = x;
          if (v == 4) sb = (R)x;
        }
      }
      load_fwd(c + 1, nxt, cell);   // past the chains' ends: every lane masked, no traffic
      // right-hand sides read two steps ahead of their use, ahead of the
      // step's own LDS store, so their latency overlaps the recurrence
      R bq[LN_CH];
      bq[0] = (R)X[(size_t)j0 * NV + rr];
      bq[1] = (R)X[(size_t)(j0 + 1) * NV + rr];
#pragma unroll
      for (int k = 0; k < LN_CH; ++k) {
        const int j = j0 + k;
        if (k + 2 < LN_CH) bq[k + 2] = (R)X[(size_t)(j + 2) * NV + rr];
        const R s = bq[k] - dot_in_quad<R>(cur.fl[k], dcol);
        dcol = dot_over_quads<R>(cur.fp[k], s);
        if (rr == 0 && j < L) X[(size_t)j * NV + qq] = (double)dcol;
        if constexpr (NV > 4) {
          const R nsd = cur.sp * (sb - cur.sl * dpp_mov<DPP_ROR1>(sd));
          sd = l == k ? nsd : sd;
        }
      }
      if constexpr (NV > 4)
        if (j0 + l < L) X[(size_t)(j0 + l) * NV + 4] = (double)sd;
    }
  }

  R xin = R(0), sx = R(0);   // x_{j+1}: row form before even steps, column form before odd ones
  if (nch > 0) load_bwd(0, bn);
  for (int m = 0; m < nch; ++m) {
    bc = bn;
    load_bwd(m + 1, bn);   // below the chains' starts: every lane masked, no traffic
    finish(bc, m - 1, sx);
    const int jt = L - 1 - m * LN_CH, jl = jt - l;
    R sdv = R(0);
    if constexpr (NV > 4)
      if (jl >= 0) sdv = (R)X[(size_t)jl * NV + 4];
    // d_j read two steps ahead (column entry on even steps, row entry on odd)
    auto ld_d = [&](int k) { return (R)X[(size_t)(jt - k >= 0 ? jt - k : 0) * NV + (k % 2 == 0 ? qq : rr)]; };
    R dq[LN_CH];
    dq[0] = ld_d(0);
    dq[1] = ld_d(1);
#pragma unroll
    for (int k = 0; k < LN_CH; ++k) {
      const int j = jt - k;
      if (k + 2 < LN_CH) dq[k + 2] = ld_d(k + 2);
      if (k % 2 == 0) {
        xin = dq[k] - dot_over_quads<R>(bc.fc[k], xin);
        if (rr == 0 && j >= 0) X[(size_t)j * NV + qq] = (double)xin;
      } else {
        xin = dq[k] - dot_in_quad<R>(bc.fc[k], xin);
        if (qq == 0 && j >= 0) X[(size_t)j * NV + rr] = (double)xin;
      }
      if constexpr (NV > 4) {
        const R nsx = sdv - bc.sc * dpp_mov<DPP_ROR1>(sx);
        sx = l == k ? nsx : sx;
      }
    }
  }
  if (nch > 0) finish(bn, nch - 1, sx);
  if (dq_out && chain >= 0) {   // the ghosts rooted on this chain: dQ_ghost = J dQ_root
    const double* Jc = gjac + (size_t)cs * P.n_final * NV * NV;
    const int d0 = P.cq_start[chain], nd = P.cq_start[chain + 1] - d0;
    for (int i = l; i < nd; i += LN_CH) {
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
