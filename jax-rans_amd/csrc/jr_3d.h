// The 3D path: residual and tangent, wave speeds, line preconditioner
// (block Thomas on 5x5 flow blocks + the scalar SA chain), relaxation,
// accept test and forces on a 3D canvas (src/jaxrans/mesh/atlas3d.py layout:
// flat cell (i NJ + j) NK + k, faces of family a on the cell grid with one
// more along a, concatenated i | j | k).  Same algorithms as the 2D kernels,
// restated for three face families and three velocity components.
#pragma once
#include "jr_krylov.h"
#include "jr_physics3d.h"

namespace jr {

constexpr int MAX_OPS = 4;     // boundary conditions composed by one ghost cell (corners: 3)

struct Plan3 {
  int NI, NJ, NK, nc;
  int nf[3], nft;                  // faces per family, total
  const double* vol;               // [nc]
  const double* S[3];              // family a: [3][nf[a]] area vectors
  const double* cen;               // [3][(NI+2)(NJ+2)(NK+2)] centres, edge-padded by one
  const double* dist;              // [nc]
  const uint8_t* interior;         // [nc]
  const uint8_t* ghost;            // [nc] written by the fill
  const int* nbr;                  // [nc] face-neighbour map
  int n_ghost;                     // composed fill: ghost g = ops of its root's primitives
  const int* g_dest;
  const int* g_root;
  const int* g_nops;
  const int* g_ops;                // [n_ghost][MAX_OPS] task ids
  const int* t_kind;               // [n_tasks]
  const double* t_norm;            // [3][n_tasks]
  int n_tasks;
  int n_lines, n_slots, max_len;
  const int* l_start;              // [n_lines + 1]
  const int* l_cell;               // [n_slots]
  const int* l_up;                 // concatenated face index of the link to the next slot
  const int8_t* l_ups;
  const int* l_dn;
  const int8_t* l_dns;
  int n_free;
  const int* free_cell;
  int n_walls;
  const int* w_cell;
  const int* w_fam;
  const int* w_face;
  const double* w_into;
  const double* w_mid;             // [3][n_walls]
  const double* dg;                // [10][nft] each face's diamond: d2 x d3, d3 x d1, d1 x d2, det (viscous)
  // The freestream's tangent along a flow parameter, dfree[case][8] =
  // (dq_inf[0..5], dmu_inf, unused); read only by the parameter-tangent kernels.
  const double* dfree;
};

__device__ __forceinline__ void face_vec3(const Plan3& P, int f, double* s) {
  int a = 0;
  if (f >= P.nf[0]) { f -= P.nf[0]; a = 1; if (f >= P.nf[1]) { f -= P.nf[1]; a = 2; } }
  const double* S = P.S[a];
  s[0] = S[f]; s[1] = S[(size_t)P.nf[a] + f]; s[2] = S[(size_t)2 * P.nf[a] + f];
}

// --- residual ----------------------------------------------------------------

// Primitives with the ghost fill in one launch (residual.py:312-322): threads
// [0, nc) convert the cells no stage writes; thread nc + g evaluates ghost g
// from its root cell through its boundary conditions in order.
// PT (with T = Dual): the freestream carries its tangent along a flow
// parameter (P.dfree), so the tangent is that of the fill along (dU, dq_inf).
template <class T, int NV, bool PT = false>
__global__ void __launch_bounds__(BLOCK) k_fill3(Plan3 P, double gamma, const Ctl* __restrict__ ctl,
                                                 const double* __restrict__ U, const double* __restrict__ dU,
                                                 double* __restrict__ Q, double* __restrict__ dQ, int check_active) {
  int t = blockIdx.x * BLOCK + threadIdx.x;
  const int cs = blockIdx.y;
  if (skip_case(ctl, cs, check_active)) return;
  const int nc = P.nc;
  const size_t base = (size_t)cs * NV * nc;
  typename Free<PT>::type qfree[6];
  for (int k = 0; k < 6; ++k) qfree[k] = free_val<PT>(ctl[cs].q_inf[k], P.dfree, cs, k);
  int src, dst, g = -1;
  if (t < nc) {
    if (P.ghost[t]) return;
    src = dst = t;
  } else {
    g = t - nc;
    if (g >= P.n_ghost) return;
    src = P.g_root[g];
    dst = P.g_dest[g];
  }
  T u[NV], q[NV];
  for (int k = 0; k < NV; ++k) u[k] = Io<T>::ld(U, dU, base + (size_t)k * nc + src);
  primitive3<T, NV>(u, gamma, q);
  if (g >= 0) {
    const int nops = P.g_nops[g];
    for (int o = 0; o < nops; ++o) {
      const int task = P.g_ops[(size_t)g * MAX_OPS + o];
      const double n[3] = {P.t_norm[task], P.t_norm[(size_t)P.n_tasks + task], P.t_norm[(size_t)2 * P.n_tasks + task]};
      T gq[NV];
      bc_state3<T, NV>(P.t_kind[task], q, n, qfree, gamma, gq);
      for (int k = 0; k < NV; ++k) q[k] = gq[k];
    }
  }
  for (int k = 0; k < NV; ++k) Io<T>::st(Q, dQ, base + (size_t)k * nc + dst, q[k]);
}

template <class T, int NV>
__device__ __forceinline__ void load_q3(const double* Q, const double* dQ, size_t base, int nc, int c, T* q) {
  for (int k = 0; k < NV; ++k) q[k] = Io<T>::ld(Q, dQ, base + (size_t)k * nc + c);
}

// Padded-centre index of canvas cell (i, j, k).
__device__ __forceinline__ size_t pc3(const Plan3& P, int i, int j, int k) {
  return ((size_t)(i + 1) * (P.NJ + 2) + (j + 1)) * (P.NK + 2) + (k + 1);
}

// Mean of the 8 values round a node, halving along i, then j, then k
// (src/jaxrans/numerics/gradients.py:31-40).
template <class T>
__device__ __forceinline__ T node_mean(const T* v) {   // v[(di*2 + dj)*2 + dk]
  T a00 = 0.5 * (v[0] + v[4]), a01 = 0.5 * (v[1] + v[5]), a10 = 0.5 * (v[2] + v[6]), a11 = 0.5 * (v[3] + v[7]);
  T b0 = 0.5 * (a00 + a10), b1 = 0.5 * (a01 + a11);
  return 0.5 * (b0 + b1);
}

// Face fluxes of the three families, one thread per face (concatenated
// numbering).  Faces not adjacent to an interior cell are zero.
// phi = (u, v, w, T[, nu_hat]) at the nodes of the diamond gradients: node
// (ni, nj, nk) is the mean of canvas cells (ni-1..ni, nj-1..nj, nk-1..nk),
// halving along i, then j, then k (gradients.py:31-40).  One thread per node
// of the node grid (NI+1)(NJ+1)(NK+1); nodes on the canvas edge are never
// read and are skipped.  Computing the nodes once per evaluation (instead of
// 4 x 8 cell loads in every face thread) keeps the face kernel in registers.
template <class T, int NV>
__global__ void __launch_bounds__(BLOCK) k_nodes3(Plan3 P, double gamma, const Ctl* __restrict__ ctl,
                                                  const double* __restrict__ Q, const double* __restrict__ dQ,
                                                  double* __restrict__ Nd, double* __restrict__ dNd,
                                                  int check_active) {
  constexpr int NPHI = NV > 5 ? 5 : 4;
  const int n = blockIdx.x * BLOCK + threadIdx.x;
  const int cs = blockIdx.y;
  if (skip_case(ctl, cs, check_active)) return;
  const int NI = P.NI, NJ = P.NJ, NK = P.NK, nc = P.nc;
  const int nn = (NI + 1) * (NJ + 1) * (NK + 1);
  if (n >= nn) return;
  const int ni = n / ((NJ + 1) * (NK + 1)), rem = n - ni * (NJ + 1) * (NK + 1);
  const int nj = rem / (NK + 1), nk = rem - nj * (NK + 1);
  if (ni < 1 || nj < 1 || nk < 1 || ni > NI - 1 || nj > NJ - 1 || nk > NK - 1) return;
  const size_t base = (size_t)cs * NV * nc;
  auto phi_at = [&](int i, int j, int k, T* out) {
    T q[NV];
    load_q3<T, NV>(Q, dQ, base, nc, (i * NJ + j) * NK + k, q);
    out[0] = q[1]; out[1] = q[2]; out[2] = q[3]; out[3] = temperature3(q, gamma);
    if constexpr (NPHI > 4) out[4] = q[5];
  };
  T b[2][NPHI];
  for (int dk2 = 0; dk2 < 2; ++dk2) {
    T a0[NPHI];
    for (int dj2 = 0; dj2 < 2; ++dj2) {
      T p0[NPHI], p1[NPHI];
      phi_at(ni - 1, nj - 1 + dj2, nk - 1 + dk2, p0);
      phi_at(ni, nj - 1 + dj2, nk - 1 + dk2, p1);
      for (int k = 0; k < NPHI; ++k) {
        const T pair = 0.5 * (p0[k] + p1[k]);
        if (dj2 == 0) a0[k] = pair;
        else b[dk2][k] = 0.5 * (a0[k] + pair);
      }
    }
  }
  const size_t nb = (size_t)cs * NPHI * nn + n;
  for (int k = 0; k < NPHI; ++k) Io<T>::st(Nd, dNd, nb + (size_t)k * nn, 0.5 * (b[0][k] + b[1][k]));
}

// The diamond of one face of family a (gradients.py:59-95): the cell line
// d1 from the left to the right cell centre and the node diagonals d2 =
// node(1,1) - node(0,0), d3 = node(0,1) - node(1,0) over the face's running
// axes, node (ni, nj, nk) the mean of canvas centres (ni-1..ni, nj-1..nj,
// nk-1..nk); out = d2 x d3, d3 x d1, d1 x d2 and det = d1 . (d2 x d3).
template <int a>
__device__ __forceinline__ void diamond3(const Plan3& P, int fi, int fj, int fk, int cL, double* out) {
  const int NJ = P.NJ, NK = P.NK, NI = P.NI;
  const int iL = cL / (NJ * NK), jL = (cL / NK) % NJ, kL = cL % NK;
  double d1[3], d2[3], d3[3];
  const size_t cstride = (size_t)(NI + 2) * (NJ + 2) * (NK + 2);
  for (int c = 0; c < 3; ++c)
    d1[c] = P.cen[c * cstride + pc3(P, fi, fj, fk)] - P.cen[c * cstride + pc3(P, iL, jL, kL)];
  const int r1 = a == 0 ? 1 : 0, r2 = a == 2 ? 1 : 2;
  auto node_at = [&](int s1, int s2, int* nd) {
    nd[0] = fi; nd[1] = fj; nd[2] = fk;
    nd[r1] += s1;
    nd[r2] += s2;
  };
  auto node_centre = [&](const int* nd, int c) -> double {
    double v[8];
    for (int di = 0; di < 2; ++di)
      for (int dj2 = 0; dj2 < 2; ++dj2)
        for (int dk2 = 0; dk2 < 2; ++dk2)
          v[(di * 2 + dj2) * 2 + dk2] = P.cen[c * cstride + pc3(P, nd[0] - 1 + di, nd[1] - 1 + dj2, nd[2] - 1 + dk2)];
    return node_mean<double>(v);
  };
  int n00[3], n11[3], n10[3], n01[3];
  node_at(0, 0, n00); node_at(1, 1, n11); node_at(1, 0, n10); node_at(0, 1, n01);
  for (int c = 0; c < 3; ++c) {
    d2[c] = node_centre(n11, c) - node_centre(n00, c);
    d3[c] = node_centre(n01, c) - node_centre(n10, c);
  }
  out[0] = d2[1] * d3[2] - d2[2] * d3[1];
  out[1] = d2[2] * d3[0] - d2[0] * d3[2];
  out[2] = d2[0] * d3[1] - d2[1] * d3[0];
  out[3] = d3[1] * d1[2] - d3[2] * d1[1];
  out[4] = d3[2] * d1[0] - d3[0] * d1[2];
  out[5] = d3[0] * d1[1] - d3[1] * d1[0];
  out[6] = d1[1] * d2[2] - d1[2] * d2[1];
  out[7] = d1[2] * d2[0] - d1[0] * d2[2];
  out[8] = d1[0] * d2[1] - d1[1] * d2[0];
  out[9] = (d1[0] * out[0] + d1[1] * out[1]) + d1[2] * out[2];
}

template <int a>
__device__ __forceinline__ void diamond3_face(const Plan3& P, double* __restrict__ dg, int f, int ff) {
  const int NI = P.NI, NJ = P.NJ, NK = P.NK;
  const int dj = NJ + (a == 1), dk = NK + (a == 2);
  const int fi = ff / (dj * dk), rem = ff - fi * dj * dk, fj = rem / dk, fk = rem - fj * dk;
  const int idx[3] = {fi, fj, fk};
  const int ext[3] = {NI, NJ, NK};
  const int stride[3] = {NJ * NK, NK, 1};
  double out[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (idx[a] >= 1 && idx[a] <= ext[a] - 1) {
    const int cR = (fi * NJ + fj) * NK + fk, cL = cR - stride[a];
    if (P.interior[cL] || P.interior[cR]) diamond3<a>(P, fi, fj, fk, cL, out);
  }
  for (int k = 0; k < 10; ++k) dg[(size_t)k * P.nft + f] = out[k];
}

// The diamonds of every face, once per context (the geometry is fixed).
__global__ void __launch_bounds__(BLOCK) k_diamond3(Plan3 P, double* __restrict__ dg) {
  const int f = blockIdx.x * BLOCK + threadIdx.x;
  if (f >= P.nft) return;
  if (f < P.nf[0]) diamond3_face<0>(P, dg, f, f);
  else if (f < P.nf[0] + P.nf[1]) diamond3_face<1>(P, dg, f, f - P.nf[0]);
  else diamond3_face<2>(P, dg, f, f - P.nf[0] - P.nf[1]);
}

// The family a is a template parameter so that every index below is known
// at compile time (a runtime family index sent the index arrays to scratch).
template <class T, int NV, bool VISC, int a, bool PT = false>
__device__ __forceinline__ void face3(const Plan3& P, const Physics& ph, int first_order, const Ctl* __restrict__ ctl,
                                      const double* __restrict__ Q, const double* __restrict__ dQ,
                                      const double* __restrict__ Nd, const double* __restrict__ dNd,
                                      double* __restrict__ F, int f, int ff, int cs) {
  const int NI = P.NI, NJ = P.NJ, NK = P.NK, nc = P.nc;
  const int dj = NJ + (a == 1), dk = NK + (a == 2);
  const int fi = ff / (dj * dk), rem = ff - fi * dj * dk, fj = rem / dk, fk = rem - fj * dk;
  const int idx[3] = {fi, fj, fk};
  const int ext[3] = {NI, NJ, NK};
  const int stride[3] = {NJ * NK, NK, 1};
  const size_t fbase = (size_t)cs * NV * P.nft + f;
  const bool inside = idx[a] >= 1 && idx[a] <= ext[a] - 1;
  int cR = 0, cL = 0;
  if (inside) {
    cR = (fi * NJ + fj) * NK + fk;
    cL = cR - stride[a];
  }
  if (!inside || !(P.interior[cL] || P.interior[cR])) {
    for (int k = 0; k < NV; ++k) F[fbase + (size_t)k * P.nft] = 0.0;
    return;
  }
  const double s[3] = {P.S[a][ff], P.S[a][(size_t)P.nf[a] + ff], P.S[a][(size_t)2 * P.nf[a] + ff]};
  const size_t base = (size_t)cs * NV * nc;
  const double gamma = ph.gamma;
  const int dL = stride[a];
  T qm[NV], qp[NV];
  load_q3<T, NV>(Q, dQ, base, nc, cL, qm);
  load_q3<T, NV>(Q, dQ, base, nc, cR, qp);
  T qL[5], qR[5];
  if (first_order) {
    for (int k = 0; k < 5; ++k) { qL[k] = qm[k]; qR[k] = qp[k]; }
  } else {
    T qmm[NV], qpp[NV];
    load_q3<T, NV>(Q, dQ, base, nc, cL - dL, qmm);
    load_q3<T, NV>(Q, dQ, base, nc, cR + dL, qpp);
    if (ph.kappa != 0.0) {
      double lo = 0.25 * (1.0 - ph.kappa), hi = 0.25 * (1.0 + ph.kappa);
      for (int k = 0; k < 5; ++k) {
        qL[k] = qm[k] + lo * (qm[k] - qmm[k]) + hi * (qp[k] - qm[k]);
        qR[k] = qp[k] - lo * (qpp[k] - qp[k]) - hi * (qp[k] - qm[k]);
      }
    } else {
      for (int k = 0; k < 5; ++k) {
        qL[k] = qm[k] + 0.5 * limit<T>(ph.limiter, qm[k] - qmm[k], qp[k] - qm[k]);
        qR[k] = qp[k] - 0.5 * limit<T>(ph.limiter, qp[k] - qm[k], qpp[k] - qp[k]);
      }
    }
  }
  T Fl[NV];
  if (ph.flux == FLUX_ROE) roe3<T>(qL, qR, s, gamma, Fl);
  else hllc3<T>(qL, qR, s, gamma, Fl);
  if constexpr (NV > 5) Fl[5] = lift<T>(0.0);

  if (VISC) {
    const auto mu_inf = free_val<PT>(ctl[cs].mu_inf, P.dfree, cs, 6);
    // the diamond's geometry (state independent, k_diamond3): the cross
    // products of the cell line d1 and the two node diagonals d2, d3
    double c23[3], c31[3], c12[3];
    for (int b = 0; b < 3; ++b) {
      c23[b] = P.dg[(size_t)b * P.nft + f];
      c31[b] = P.dg[(size_t)(3 + b) * P.nft + f];
      c12[b] = P.dg[(size_t)(6 + b) * P.nft + f];
    }
    const double det = P.dg[(size_t)9 * P.nft + f];
    // the face's running axes in ascending order; shifts (s1, s2) along them
    const int r1 = a == 0 ? 1 : 0, r2 = a == 2 ? 1 : 2;
    auto node_at = [&](int s1, int s2, int* nd) {
      nd[0] = fi; nd[1] = fj; nd[2] = fk;
      nd[r1] += s1;
      nd[r2] += s2;
    };
    int n00[3], n11[3], n10[3], n01[3];
    node_at(0, 0, n00); node_at(1, 1, n11); node_at(1, 0, n10); node_at(0, 1, n01);
    // phi = (u, v, w, T[, nu_hat]); node values streamed in the halving
    // order of node_mean (pairs along i, then j, then k) so that no more than
    // two cells and two partial sums are live at once
    constexpr int NPHI = NV > 5 ? 5 : 4;
    auto phi_of = [&](const T* q, T* out) {
      out[0] = q[1]; out[1] = q[2]; out[2] = q[3]; out[3] = temperature3(q, gamma);
      if constexpr (NPHI > 4) out[4] = q[5];
    };
    const int nn = (NI + 1) * (NJ + 1) * (NK + 1);
    const size_t nbase = (size_t)cs * NPHI * nn;
    auto node_phi = [&](const int* nd, T* out) {
      const size_t o = nbase + ((size_t)nd[0] * (NJ + 1) + nd[1]) * (NK + 1) + nd[2];
      for (int k = 0; k < NPHI; ++k) out[k] = Io<T>::ld(Nd, dNd, o + (size_t)k * nn);
    };
    T dp1[NPHI], dp2[NPHI], dp3[NPHI];
    {
      T pl[NPHI], pr[NPHI];
      phi_of(qm, pl);
      phi_of(qp, pr);
      for (int k = 0; k < NPHI; ++k) dp1[k] = pr[k] - pl[k];
    }
    {
      T pa[NPHI], pb[NPHI];
      node_phi(n00, pa);
      node_phi(n11, pb);
      for (int k = 0; k < NPHI; ++k) dp2[k] = pb[k] - pa[k];
      node_phi(n10, pa);
      node_phi(n01, pb);
      for (int k = 0; k < NPHI; ++k) dp3[k] = pb[k] - pa[k];
    }
    T gr[3][NPHI];                      // gr[b][k] = d phi_k / d x_b
    for (int k = 0; k < NPHI; ++k)
      for (int bb = 0; bb < 3; ++bb) gr[bb][k] = (dp1[k] * c23[bb] + dp2[k] * c31[bb] + dp3[k] * c12[bb]) / det;
    T gv[3][3], gT[3];
    for (int c = 0; c < 3; ++c) {
      for (int b = 0; b < 3; ++b) gv[c][b] = gr[b][c];
      gT[c] = gr[c][3];
    }
    T vel[3] = {0.5 * (qm[1] + qp[1]), 0.5 * (qm[2] + qp[2]), 0.5 * (qm[3] + qp[3])};
    T Tf = 0.5 * (temperature3(qm, gamma) + temperature3(qp, gamma));
    T mu = viscosity<T>(Tf, mu_inf, ph.sutherland);
    T kc = mu / (PR * (gamma - 1.0));
    T extra[3][3];
    bool has_extra = false;      // two inlined calls below keep extra in registers
    if constexpr (NV > 5) {
      T face[6];
      for (int k = 0; k < 6; ++k) face[k] = 0.5 * (qm[k] + qp[k]);
      T muf = viscosity<T>(temperature3(face, gamma), mu_inf, ph.sutherland);
      T mut = eddy_viscosity<T>(face[0], face[5] * mu_inf, muf / face[0]);
      mu = mu + mut;
      kc = kc + mut / (PR_T * (gamma - 1.0));
      if (ph.c_cr1 != 0.0) { qcr_stress3<T>(gv, mut, ph.c_cr1, extra); has_extra = true; }
      T up = val(Fl[0]) > 0.0 ? qm[5] : qp[5];
      T coef = diffusion_coefficient<T>(face[0], face[5] * mu_inf, muf);
      T gn[3] = {gr[0][4], gr[1][4], gr[2][4]};
      Fl[5] = Fl[0] * up - coef * dot3(gn, s);
    }
    T fv[5];
    if (has_extra) viscous_flux3<T>(vel, gv, gT, s, mu, kc, extra, fv);
    else viscous_flux3<T>(vel, gv, gT, s, mu, kc, nullptr, fv);
    for (int k = 0; k < 5; ++k) Fl[k] = Fl[k] - fv[k];
  }
  for (int k = 0; k < NV; ++k) {
    if (sizeof(T) == sizeof(double)) F[fbase + (size_t)k * P.nft] = val(Fl[k]);
    else F[fbase + (size_t)k * P.nft] = tan_of(Fl[k]);
  }
}

// PT (with T = Dual): mu_inf carries its tangent along a flow parameter.
template <class T, int NV, bool VISC, bool PT = false>
__global__ void __launch_bounds__(BLOCK, 2) k_faces3(Plan3 P, Physics ph, int first_order, const Ctl* __restrict__ ctl,
                                                    const double* __restrict__ Q, const double* __restrict__ dQ,
                                                    const double* __restrict__ Nd, const double* __restrict__ dNd,
                                                    double* __restrict__ F, int check_active) {
  const int f = blockIdx.x * BLOCK + threadIdx.x;
  const int cs = blockIdx.y;
  if (skip_case(ctl, cs, check_active)) return;
  if (f >= P.nft) return;
  if (f < P.nf[0]) face3<T, NV, VISC, 0, PT>(P, ph, first_order, ctl, Q, dQ, Nd, dNd, F, f, f, cs);
  else if (f < P.nf[0] + P.nf[1])
    face3<T, NV, VISC, 1, PT>(P, ph, first_order, ctl, Q, dQ, Nd, dNd, F, f, f - P.nf[0], cs);
  else face3<T, NV, VISC, 2, PT>(P, ph, first_order, ctl, Q, dQ, Nd, dNd, F, f, f - P.nf[0] - P.nf[1], cs);
}

// Per cell: R = div F (- V * SA source) at interior cells, U - U0 elsewhere;
// tangent mode: out = vdt dU + dR.
// Partials of the SA source at the state, once per residual (the tangent
// of every Krylov step then needs no dual-number source): src depends on
// the cell's density, pressure and nu_hat, on |curl v| and on grad nu_hat
// (turbulence.py:140-151).  Per cell, SALIN values: d src / d(rho, p,
// nu_hat); d src / d curl_i (through |curl v|, zero where curl v = 0, as
// the dual-number square root's guard); d src / d (grad nu_hat)_b.
constexpr int SALIN = 9;

__global__ void __launch_bounds__(BLOCK) k_sa_lin3(Plan3 P, Physics ph, const Ctl* __restrict__ ctl,
                                                   const double* __restrict__ Q, double* __restrict__ salin,
                                                   int check_active) {
  const int c = blockIdx.x * BLOCK + threadIdx.x;
  const int cs = blockIdx.y;
  if (c >= P.nc || skip_case(ctl, cs, check_active)) return;
  const int NJ = P.NJ, NK = P.NK, nc = P.nc;
  double* L = salin + (size_t)cs * SALIN * nc + c;
  if (!P.interior[c]) {
    for (int k = 0; k < SALIN; ++k) L[(size_t)k * nc] = 0.0;
    return;
  }
  const size_t base = (size_t)cs * 6 * nc;
  const int i = c / (NJ * NK), j = (c / NK) % NJ, k = c % NK;
  const int lo0 = (i * NJ + j) * NK + k, hi0 = lo0 + NJ * NK;
  const int lo1 = P.nf[0] + (i * (NJ + 1) + j) * NK + k, hi1 = lo1 + NK;
  const int lo2 = P.nf[0] + P.nf[1] + (i * NJ + j) * (NK + 1) + k, hi2 = lo2 + 1;
  const double gamma = ph.gamma, nu_inf = ctl[cs].mu_inf;
  const int nb[6] = {c - NJ * NK, c + NJ * NK, c - NK, c + NK, c - 1, c + 1};
  const int fl[6] = {lo0, hi0, lo1, hi1, lo2, hi2};
  double qc[6];
  load_q3<double, 6>(Q, nullptr, base, nc, c, qc);
  double fa[6][4];
  for (int e = 0; e < 6; ++e) {
    double qn[6];
    load_q3<double, 6>(Q, nullptr, base, nc, nb[e], qn);
    const bool low = (e & 1) == 0;
    for (int m = 0; m < 3; ++m) fa[e][m] = low ? 0.5 * (qn[1 + m] + qc[1 + m]) : 0.5 * (qc[1 + m] + qn[1 + m]);
    fa[e][3] = low ? 0.5 * (qn[5] + qc[5]) : 0.5 * (qc[5] + qn[5]);
  }
  double sv[6][3];
  for (int e = 0; e < 6; ++e) face_vec3(P, fl[e], sv[e]);
  const double V = P.vol[c];
  auto gg = [&](int m, int b) -> double {
    return (((fa[1][m] * sv[1][b] - fa[0][m] * sv[0][b]) + (fa[3][m] * sv[3][b] - fa[2][m] * sv[2][b])) +
            (fa[5][m] * sv[5][b] - fa[4][m] * sv[4][b])) / V;
  };
  double g[3][3];
  for (int m = 0; m < 3; ++m)
    for (int b = 0; b < 3; ++b) g[b][m] = gg(m, b);
  const double curl[3] = {g[1][2] - g[2][1], g[2][0] - g[0][2], g[0][1] - g[1][0]};
  const double sq = dot3(curl, curl);
  const double omega = sq > 0.0 ? sqrt(sq) : 0.0;
  const double gn[3] = {gg(3, 0), gg(3, 1), gg(3, 2)};
  const double gn2 = dot3(gn, gn);
  // seeds: density, pressure, nu_hat, |curl v|
  double part[4];
  for (int sd = 0; sd < 4; ++sd) {
    Dual q[6];
    for (int v = 0; v < 6; ++v) q[v] = mkd(qc[v], 0.0);
    if (sd == 0) q[0].d = 1.0;
    if (sd == 1) q[4].d = 1.0;
    if (sd == 2) q[5].d = 1.0;
    const Dual om = mkd(omega, sd == 3 ? 1.0 : 0.0);
    const Dual rho = q[0];
    const Dual mu = viscosity<Dual>(temperature3(q, gamma), nu_inf, ph.sutherland);
    Dual prod, destr;
    sa_source<Dual>(q[5] * nu_inf, mu / rho, om, P.dist[c], prod, destr);
    const Dual src = rho * ((prod - destr) / nu_inf + CB2 / SIGMA * nu_inf * gn2);
    part[sd] = src.d;
  }
  L[0] = part[0];
  L[(size_t)nc] = part[1];
  L[(size_t)2 * nc] = part[2];
  const double w = sq > 0.0 ? part[3] / omega : 0.0;   // d|curl| / d curl_i = curl_i / |curl|
  for (int a = 0; a < 3; ++a) L[(size_t)(3 + a) * nc] = w * curl[a];
  const double rg = qc[0] * (CB2 / SIGMA * nu_inf);
  for (int b = 0; b < 3; ++b) L[(size_t)(6 + b) * nc] = rg * (2.0 * gn[b]);
}

// PT (with T = Dual): the freestream and mu_inf carry their tangent along a
// flow parameter -- U0 moves with it, and the SA source is differentiated in
// full (its partials at the state hold dU only).
template <class T, int NV, bool PT = false>
__global__ void __launch_bounds__(BLOCK, 2) k_cells3(Plan3 P, Physics ph, const Ctl* __restrict__ ctl,
                                                    const double* __restrict__ U, const double* __restrict__ dU,
                                                    const double* __restrict__ Q, const double* __restrict__ dQ,
                                                    const double* __restrict__ F, const double* __restrict__ vdt,
                                                    const double* __restrict__ salin, double* __restrict__ out,
                                                    int check_active) {
  const int c = blockIdx.x * BLOCK + threadIdx.x;
  const int cs = blockIdx.y;
  if (skip_case(ctl, cs, check_active)) return;
  if (c >= P.nc) return;
  const int NJ = P.NJ, NK = P.NK, nc = P.nc;
  const size_t base = (size_t)cs * NV * nc;
  constexpr bool tangent = sizeof(T) != sizeof(double);
  const double vd = vdt ? vdt[(size_t)cs * nc + c] : 0.0;
  if (!P.interior[c]) {
    if (!tangent) {
      double u0[6];
      conservative3<double, NV>(ctl[cs].q_inf, ph.gamma, u0);
      for (int k = 0; k < NV; ++k) out[base + (size_t)k * nc + c] = U[base + (size_t)k * nc + c] - u0[k];
    } else if constexpr (PT) {
      Dual qf[6], u0[6];
      for (int k = 0; k < 6; ++k) qf[k] = free_val<true>(ctl[cs].q_inf[k], P.dfree, cs, k);
      conservative3<Dual, NV>(qf, ph.gamma, u0);
      for (int k = 0; k < NV; ++k) out[base + (size_t)k * nc + c] = dU[base + (size_t)k * nc + c] - u0[k].d;
    } else {
      for (int k = 0; k < NV; ++k) {
        double d = dU[base + (size_t)k * nc + c];
        out[base + (size_t)k * nc + c] = vdt ? vd * d + d : d;
      }
    }
    return;
  }
  const int i = c / (NJ * NK), j = (c / NK) % NJ, k = c % NK;
  // the cell's low and high face of each family (concatenated numbering)
  const int lo0 = (i * NJ + j) * NK + k, hi0 = lo0 + NJ * NK;
  const int lo1 = P.nf[0] + (i * (NJ + 1) + j) * NK + k, hi1 = lo1 + NK;
  const int lo2 = P.nf[0] + P.nf[1] + (i * NJ + j) * (NK + 1) + k, hi2 = lo2 + 1;
  const size_t fb = (size_t)cs * NV * P.nft;
  double r[NV];
  for (int v = 0; v < NV; ++v) {
    const double* Fv = F + fb + (size_t)v * P.nft;
    r[v] = ((Fv[hi0] - Fv[lo0]) + (Fv[hi1] - Fv[lo1])) + (Fv[hi2] - Fv[lo2]);
  }
  if constexpr (NV > 5) {
   if constexpr (tangent && !PT) {
    // the source's tangent from its partials at the state (k_sa_lin3): the
    // face-average gradients are linear in dq, the rest is a dot product
    const int nb[6] = {c - NJ * NK, c + NJ * NK, c - NK, c + NK, c - 1, c + 1};
    const int fl[6] = {lo0, hi0, lo1, hi1, lo2, hi2};
    const double* dq = dQ + base;
    double dqc[6];
    for (int v = 0; v < 6; ++v) dqc[v] = dq[(size_t)v * nc + c];
    double fa[6][4];
    for (int e = 0; e < 6; ++e) {
      const bool low = (e & 1) == 0;
      for (int m = 0; m < 3; ++m) {
        const double qn = dq[(size_t)(1 + m) * nc + nb[e]];
        fa[e][m] = low ? 0.5 * (qn + dqc[1 + m]) : 0.5 * (dqc[1 + m] + qn);
      }
      const double qn = dq[(size_t)5 * nc + nb[e]];
      fa[e][3] = low ? 0.5 * (qn + dqc[5]) : 0.5 * (dqc[5] + qn);
    }
    double sv[6][3];
    for (int e = 0; e < 6; ++e) face_vec3(P, fl[e], sv[e]);
    const double V = P.vol[c];
    auto gg = [&](int m, int b) -> double {
      return (((fa[1][m] * sv[1][b] - fa[0][m] * sv[0][b]) + (fa[3][m] * sv[3][b] - fa[2][m] * sv[2][b])) +
              (fa[5][m] * sv[5][b] - fa[4][m] * sv[4][b])) / V;
    };
    double g[3][3];
    for (int m = 0; m < 3; ++m)
      for (int b = 0; b < 3; ++b) g[b][m] = gg(m, b);
    const double dcurl[3] = {g[1][2] - g[2][1], g[2][0] - g[0][2], g[0][1] - g[1][0]};
    const double* L = salin + (size_t)cs * SALIN * nc + c;
    double ds = (L[0] * dqc[0] + L[(size_t)nc] * dqc[4]) + L[(size_t)2 * nc] * dqc[5];
    for (int i = 0; i < 3; ++i) ds += L[(size_t)(3 + i) * nc] * dcurl[i];
    for (int b = 0; b < 3; ++b) ds += L[(size_t)(6 + b) * nc] * gg(3, b);
    r[5] = r[5] - ds * V;
   } else {
    // SA source (turbulence.py:140-151): vorticity and grad nu_hat by
    // Green-Gauss over the face averages with the six neighbours
    const double gamma = ph.gamma;
    const auto nu_inf = free_val<PT>(ctl[cs].mu_inf, P.dfree, cs, 6);
    const int nb[6] = {c - NJ * NK, c + NJ * NK, c - NK, c + NK, c - 1, c + 1};
    const int fl[6] = {lo0, hi0, lo1, hi1, lo2, hi2};
    T qc[NV];
    load_q3<T, NV>(Q, dQ, base, nc, c, qc);
    T fa[6][4];                       // face averages of (u, v, w, nu_hat)
    for (int e = 0; e < 6; ++e) {
      T qn[NV];
      load_q3<T, NV>(Q, dQ, base, nc, nb[e], qn);
      const bool low = (e & 1) == 0;
      for (int m = 0; m < 3; ++m) fa[e][m] = low ? 0.5 * (qn[1 + m] + qc[1 + m]) : 0.5 * (qc[1 + m] + qn[1 + m]);
      fa[e][3] = low ? 0.5 * (qn[5] + qc[5]) : 0.5 * (qc[5] + qn[5]);
    }
    double sv[6][3];
    for (int e = 0; e < 6; ++e) face_vec3(P, fl[e], sv[e]);
    const double V = P.vol[c];
    auto gg = [&](int m, int b) -> T {
      return (((fa[1][m] * sv[1][b] - fa[0][m] * sv[0][b]) + (fa[3][m] * sv[3][b] - fa[2][m] * sv[2][b])) +
              (fa[5][m] * sv[5][b] - fa[4][m] * sv[4][b])) / V;
    };
    T g[3][3];                        // g[b][m] = d v_m / d x_b
    for (int m = 0; m < 3; ++m)
      for (int b = 0; b < 3; ++b) g[b][m] = gg(m, b);
    T omega = vorticity3<T>(g);
    T gn[3] = {gg(3, 0), gg(3, 1), gg(3, 2)};
    T rho = qc[0];
    T mu = viscosity<T>(temperature3(qc, gamma), nu_inf, ph.sutherland);
    T prod, destr;
    sa_source<T>(qc[5] * nu_inf, mu / rho, omega, P.dist[c], prod, destr);
    T src = rho * ((prod - destr) / nu_inf + CB2 / SIGMA * nu_inf * dot3(gn, gn));
    T svv = src * V;
    r[5] = r[5] - (tangent ? tan_of(svv) : val(svv));
   }
  }
  for (int v = 0; v < NV; ++v) {
    if (tangent && vdt) r[v] = vd * dU[base + (size_t)v * nc + c] + r[v];
    out[base + (size_t)v * nc + c] = r[v];
  }
}

// --- wave speeds ------------------------------------------------------------

template <int NV>
__device__ __forceinline__ void prim_at3(const double* U, size_t base, int nc, int c, double gamma, double* q) {
  double u[NV];
  for (int k = 0; k < NV; ++k) u[k] = U[base + (size_t)k * nc + c];
  primitive3<double, NV>(u, gamma, q);
}

// Face lambda (residual.py:169-192) from neighbour-filled face averages; SA:
// the face mass flux and diffusion of the scalar chains.
template <int NV>
__global__ void __launch_bounds__(BLOCK) k_face_speeds3(Plan3 P, Physics ph, const Ctl* __restrict__ ctl,
                                                        const double* __restrict__ U, double* __restrict__ lam,
                                                        double* __restrict__ sam, double* __restrict__ sad) {
  const int f = blockIdx.x * BLOCK + threadIdx.x;
  const int cs = blockIdx.y;
  if (f >= P.nft || ctl[cs].parked) return;   // a parked serving slot: nothing to do this step
  const int NI = P.NI, NJ = P.NJ, NK = P.NK, nc = P.nc;
  int a = 0, ff = f;
  if (ff >= P.nf[0]) { ff -= P.nf[0]; a = 1; if (ff >= P.nf[1]) { ff -= P.nf[1]; a = 2; } }
  const int dj = NJ + (a == 1), dk = NK + (a == 2);
  int idx[3];
  idx[0] = ff / (dj * dk);
  const int rem = ff - idx[0] * dj * dk;
  idx[1] = rem / dk;
  idx[2] = rem - idx[1] * dk;
  const int ext[3] = {NI, NJ, NK};
  int lo[3] = {idx[0], idx[1], idx[2]}, hi[3] = {idx[0], idx[1], idx[2]};
  lo[a] = idx[a] > 0 ? idx[a] - 1 : 0;                  // edge padding at the canvas ends
  hi[a] = idx[a] < ext[a] ? idx[a] : ext[a] - 1;
  const int c1 = P.nbr[(lo[0] * NJ + lo[1]) * NK + lo[2]];
  const int c2 = P.nbr[(hi[0] * NJ + hi[1]) * NK + hi[2]];
  double s[3];
  face_vec3(P, f, s);
  const size_t base = (size_t)cs * NV * nc;
  const double gamma = ph.gamma;
  double q1[NV], q2[NV], qf[NV];
  prim_at3<NV>(U, base, nc, c1, gamma, q1);
  prim_at3<NV>(U, base, nc, c2, gamma, q2);
  for (int k = 0; k < NV; ++k) qf[k] = 0.5 * (q1[k] + q2[k]);
  const double ss = dot3(s, s);
  double l = fabs(dot3(qf + 1, s)) + sqrt(gamma * qf[4] / qf[0]) * sqrt(ss);
  const size_t fo = (size_t)cs * P.nft + f;
  if (ph.viscous) {
    const double mu_inf = ctl[cs].mu_inf;
    const double vol = 0.5 * (P.vol[c1] + P.vol[c2]);
    double mu = viscosity<double>(gamma * qf[4] / qf[0], mu_inf, ph.sutherland);
    const double mul = mu;
    if constexpr (NV > 5) mu = mu + eddy_viscosity<double>(qf[0], qf[5] * mu_inf, mu / qf[0]);
    l = l + 2.0 * gamma / PR * mu * ss / (qf[0] * vol);
    if constexpr (NV > 5) {
      const double kap = diffusion_coefficient<double>(qf[0], qf[5] * mu_inf, mul);
      sam[fo] = qf[0] * dot3(qf + 1, s);
      sad[fo] = kap * ss / vol;
    }
  }
  lam[fo] = l;
}

// Per cell: sum of face wave speeds, V/dt, and the SA chain diagonal
// (preconditioners.py:259-275).
template <int NV>
__global__ void __launch_bounds__(BLOCK) k_cell_speeds3(Plan3 P, Physics ph, const Ctl* __restrict__ ctl,
                                                        const double* __restrict__ U, const double* __restrict__ lam,
                                                        const double* __restrict__ sam, const double* __restrict__ sad,
                                                        double* __restrict__ lsum, double* __restrict__ vdt,
                                                        double* __restrict__ sa_diag) {
  const int c = blockIdx.x * BLOCK + threadIdx.x;
  const int cs = blockIdx.y;
  if (c >= P.nc || ctl[cs].parked) return;   // a parked serving slot: nothing to do this step
  const int NJ = P.NJ, NK = P.NK, nc = P.nc;
  const int i = c / (NJ * NK), j = (c / NK) % NJ, k = c % NK;
  const int lo0 = (i * NJ + j) * NK + k, hi0 = lo0 + NJ * NK;
  const int lo1 = P.nf[0] + (i * (NJ + 1) + j) * NK + k, hi1 = lo1 + NK;
  const int lo2 = P.nf[0] + P.nf[1] + (i * NJ + j) * (NK + 1) + k, hi2 = lo2 + 1;
  const double* L = lam + (size_t)cs * P.nft;
  const double s = ((((L[lo0] + L[hi0]) + L[lo1]) + L[hi1]) + L[lo2]) + L[hi2];
  const double cfl = ctl[cs].cfl;
  lsum[(size_t)cs * nc + c] = s;
  vdt[(size_t)cs * nc + c] = s / cfl;
  if constexpr (NV > 5) {
    const double* m = sam + (size_t)cs * P.nft;
    const double* d = sad + (size_t)cs * P.nft;
    const double gamma = ph.gamma, mu_inf = ctl[cs].mu_inf;
    double diag = (fmax(m[hi0], 0.0) + d[hi0]) + (fmax(-m[lo0], 0.0) + d[lo0]);
    diag = diag + (fmax(m[hi1], 0.0) + d[hi1]) + (fmax(-m[lo1], 0.0) + d[lo1]);
    diag = diag + (fmax(m[hi2], 0.0) + d[hi2]) + (fmax(-m[lo2], 0.0) + d[lo2]);
    const size_t base = (size_t)cs * NV * nc;
    double q[NV];
    prim_at3<NV>(U, base, nc, c, gamma, q);
    diag = diag / q[0];
    double g[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    if (P.interior[c]) {
      const int nb[6] = {c - NJ * NK, c + NJ * NK, c - NK, c + NK, c - 1, c + 1};
      const int fl[6] = {lo0, hi0, lo1, hi1, lo2, hi2};
      double fa[6][3], sv[6][3];
      for (int e = 0; e < 6; ++e) {
        double qn[NV];
        prim_at3<NV>(U, base, nc, P.nbr[nb[e]], gamma, qn);
        const bool low = (e & 1) == 0;
        for (int mm = 0; mm < 3; ++mm) fa[e][mm] = low ? 0.5 * (qn[1 + mm] + q[1 + mm]) : 0.5 * (q[1 + mm] + qn[1 + mm]);
        face_vec3(P, fl[e], sv[e]);
      }
      const double V = P.vol[c];
      for (int mm = 0; mm < 3; ++mm)
        for (int b = 0; b < 3; ++b)
          g[b][mm] = (((fa[1][mm] * sv[1][b] - fa[0][mm] * sv[0][b]) + (fa[3][mm] * sv[3][b] - fa[2][mm] * sv[2][b])) +
                      (fa[5][mm] * sv[5][b] - fa[4][mm] * sv[4][b])) / V;
    }
    const double omega = vorticity3<double>(g);
    const double nu = viscosity<double>(gamma * q[4] / q[0], mu_inf, ph.sutherland) / q[0];
    const double dw = P.dist[c], nh = q[5];
    double destr = 0.0;
    if (nh > 0.0) {
      const double st = s_tilde<double>(omega, nh * mu_inf, nu, dw);
      const double fw_ = f_w<double>(r_of<double>(nh * mu_inf, st, dw));
      destr = 2.0 * CW1 * fw_ * mu_inf * nh / (dw * dw);
    }
    sa_diag[(size_t)cs * nc + c] = diag + destr * P.vol[c] + s / cfl;
  }
}

// --- explicit multistage Runge-Kutta (solvers/explicit.py:48-71) ------------
// Per canvas cell dt / V as k_local_dt with three face families: the faces'
// spectral radii |v . S| + a |S| at the cell's own state (lower then upper
// face of each family), plus, when viscous, the diffusive limit
// 4 (gamma / Pr) (mu / rho) (S_i^2 + S_j^2 + S_k^2) / V, S_a the mean area
// of the two faces normal to a (numerics/residual.py:142-166).  Gap and
// ghost rows get a zero factor (their R = U - U0 stays at its start).
template <int NV, bool VISC>
__global__ void __launch_bounds__(BLOCK) k_local_dt3(Plan3 P, Physics ph, const Ctl* __restrict__ ctl,
                                                     const double* __restrict__ U, double* __restrict__ fac) {
  const int c = blockIdx.x * BLOCK + threadIdx.x;
  const int cs = blockIdx.y;
  if (c >= P.nc) return;
  const int NJ = P.NJ, NK = P.NK;
  const int i = c / (NJ * NK), j = (c / NK) % NJ, k = c % NK;
  const int lo0 = (i * NJ + j) * NK + k, hi0 = lo0 + NJ * NK;
  const int lo1 = P.nf[0] + (i * (NJ + 1) + j) * NK + k, hi1 = lo1 + NK;
  const int lo2 = P.nf[0] + P.nf[1] + (i * NJ + j) * (NK + 1) + k, hi2 = lo2 + 1;
  const int fl[6] = {lo0, hi0, lo1, hi1, lo2, hi2};
  const double gamma = ph.gamma;
  double q[NV];
  prim_at3<NV>(U, (size_t)cs * NV * P.nc, P.nc, c, gamma, q);
  const double a = sqrt(gamma * q[4] / q[0]);
  double total = 0.0, mean[3];
  for (int fam = 0; fam < 3; ++fam) {
    double sl[3], sh[3];
    face_vec3(P, fl[2 * fam], sl);
    face_vec3(P, fl[2 * fam + 1], sh);
    const double ml = sqrt(dot3(sl, sl)), mh = sqrt(dot3(sh, sh));
    const double ll = fabs(dot3(q + 1, sl)) + a * ml, lh = fabs(dot3(q + 1, sh)) + a * mh;
    total = fam == 0 ? ll + lh : (total + ll) + lh;
    mean[fam] = 0.5 * (ml + mh);
  }
  const double V = P.vol[c];
  if constexpr (VISC) {
    const double squares = (mean[0] * mean[0] + mean[1] * mean[1]) + mean[2] * mean[2];
    const double mu_inf = ctl[cs].mu_inf;
    double mu = viscosity<double>(gamma * q[4] / q[0], mu_inf, ph.sutherland);
    if constexpr (NV > 5) mu = mu + eddy_viscosity<double>(q[0], q[5] * mu_inf, mu / q[0]);
    const double factor = gamma / PR * mu / q[0];
    total = total + 4.0 * factor * squares / V;
  }
  const double dt = ctl[cs].cfl * V / fmax(total, 1e-300);
  fac[(size_t)cs * P.nc + c] = P.interior[c] ? dt / V : 0.0;
}

// --- line preconditioner: 5x5 block Thomas, one thread per chain -----------

// d(F.S)/dU of the analytic flux by forward differentiation (euler.py:55-71).
__device__ __forceinline__ void flux_jacobian3(const double* q, const double* s, double gamma, double A[5][5]) {
  double U[5];
  conservative3<double, 5>(q, gamma, U);
  for (int k = 0; k < 5; ++k) {
    Dual Ud[5], qd[5], f[5];
    for (int c = 0; c < 5; ++c) Ud[c] = mkd(U[c], c == k ? 1.0 : 0.0);
    primitive3<Dual, 5>(Ud, gamma, qd);
    normal_flux3<Dual>(qd, s, gamma, f);
    for (int r = 0; r < 5; ++r) A[r][k] = f[r].d;
  }
}

// blk3[slot] = [d, Lo(25), Up(25), sa_lo, sa_d, sa_up]
constexpr int BLKW3 = 54;

// Stored in the factor's precision R (FP32 in the mixed path).
template <class R, int NV>
__global__ void __launch_bounds__(BLOCK) k_blocks3(Plan3 P, Physics ph, const Ctl* __restrict__ ctl,
                                                   const double* __restrict__ U, const double* __restrict__ lam,
                                                   const double* __restrict__ lsum, const double* __restrict__ sam,
                                                   const double* __restrict__ sad, const double* __restrict__ sa_diag,
                                                   R* __restrict__ blk) {
  // thread t: step k = t / n_lines of chain ch = t mod n_lines, so that the
  // block components are written (and read by k_thomas3) chain-minor
  const int t = blockIdx.x * BLOCK + threadIdx.x;
  const int cs = blockIdx.y;
  if (ctl[cs].parked) return;   // a parked serving slot: nothing to do this step
  const int nl = P.n_lines;
  if (t >= P.max_len * nl) return;
  const int k = t / nl, ch = t - k * nl;
  const int s0 = P.l_start[ch];
  if (k >= P.l_start[ch + 1] - s0) return;
  const int s = s0 + k;
  const int nc = P.nc;
  const size_t base = (size_t)cs * NV * nc;
  const double gamma = ph.gamma, cfl = ctl[cs].cfl;
  const double* L = lam + (size_t)cs * P.nft;
  const int cell = P.l_cell[s];
  // component e of this step at o[e * cstride]
  const size_t cstride = (size_t)P.max_len * nl;
  R* o = blk + (size_t)cs * BLKW3 * cstride + (size_t)k * nl + ch;
  double blk_lo[25], blk_up[25];
  o[0] = (R)(lsum[(size_t)cs * nc + cell] * (1.0 / cfl + 0.5));
  auto link = [&](int nb_cell, int fce, double sg, double* dst) {
    double sv[3];
    face_vec3(P, fce, sv);
    for (int c = 0; c < 3; ++c) sv[c] *= sg;
    double q[5];
    prim_at3<5>(U, base, nc, nb_cell, gamma, q);
    double A[5][5];
    flux_jacobian3(q, sv, gamma, A);
    for (int r = 0; r < 5; ++r)
      for (int c = 0; c < 5; ++c) dst[r * 5 + c] = 0.5 * A[r][c] - (r == c ? 0.5 * L[fce] : 0.0);
  };
  const bool has_lo = s > 0 && P.l_dns[s - 1] != 0;
  const bool has_up = P.l_ups[s] != 0;
  if (has_lo) link(P.l_cell[s - 1], P.l_dn[s - 1], (double)P.l_dns[s - 1], blk_lo);
  else for (int e = 0; e < 25; ++e) blk_lo[e] = 0.0;
  if (has_up) link(P.l_cell[s + 1], P.l_up[s], (double)P.l_ups[s], blk_up);
  else for (int e = 0; e < 25; ++e) blk_up[e] = 0.0;
  for (int e = 0; e < 25; ++e) {
    o[(1 + e) * cstride] = (R)blk_lo[e];
    o[(26 + e) * cstride] = (R)blk_up[e];
  }
  if constexpr (NV > 5) {
    const double* mj = sam + (size_t)cs * P.nft;
    const double* dj = sad + (size_t)cs * P.nft;
    double lo = 0.0, up = 0.0;
    if (has_lo) {
      const int fce = P.l_dn[s - 1];
      const double sg = P.l_dns[s - 1];
      lo = -(fmax(-sg * mj[fce], 0.0) + dj[fce]) / U[base + P.l_cell[s - 1]];
    }
    if (has_up) {
      const int fce = P.l_up[s];
      const double sg = P.l_ups[s];
      up = -(fmax(-sg * mj[fce], 0.0) + dj[fce]) / U[base + P.l_cell[s + 1]];
    }
    o[51 * cstride] = (R)lo;
    o[52 * cstride] = (R)sa_diag[(size_t)cs * nc + cell];
    o[53 * cstride] = (R)up;
  }
}

// Unpivoted Gauss-Jordan inverse of a 5x5 block (linalg.py:49-77).
template <class R>
__device__ __forceinline__ void invert5(R A[5][5], R X[5][5]) {
  R W[5][10];
  for (int r = 0; r < 5; ++r)
    for (int c = 0; c < 5; ++c) { W[r][c] = A[r][c]; W[r][5 + c] = r == c ? R(1) : R(0); }
  for (int k = 0; k < 5; ++k) {
    const R piv = W[k][k];
    for (int c = 0; c < 10; ++c) W[k][c] = W[k][c] / piv;
    for (int r = 0; r < 5; ++r) {
      if (r == k) continue;
      const R f = W[r][k];
      for (int c = 0; c < 10; ++c) W[r][c] = W[r][c] - f * W[k][c];
    }
  }
  for (int r = 0; r < 5; ++r)
    for (int c = 0; c < 5; ++c) X[r][c] = W[r][5 + c];
}

// Factor layout (per case): value v of chain slot position k of chain ch at
// [(k * 75 + v) * n_lines + ch] (L 0..24, P^-1 25..49, C 50..74), so the
// threads of a wave (consecutive chains) read consecutive addresses;
// SA: [(k * 3 + v) * n_lines + ch] (l, 1/p, c).
// The 3D factor in tiles of 64 chains, each chain's 75 values of a step
// packed W to a 16-byte vector (W = 2 doubles or 4 floats, padded to 76
// values): value v of step k of lane l in its tile at
// [(k * (76 / W) + v / W) * 64 + l] * W + v mod W.  A wave then reads a
// step with 16-byte loads, 1 KB per instruction, so that the 63 loads a
// wait can name keep twice the bytes in flight.
constexpr int FAC3_VALUES = 76;

template <class R> struct Vec16;
template <> struct Vec16<double> {
  using T = double2;
  static constexpr int W = 2;
  static __device__ __forceinline__ void get(const T& t, double* o) { o[0] = t.x; o[1] = t.y; }
};
template <> struct Vec16<float> {
  using T = float4;
  static constexpr int W = 4;
  static __device__ __forceinline__ void get(const T& t, float* o) { o[0] = t.x; o[1] = t.y; o[2] = t.z; o[3] = t.w; }
};

// values [v0, v0 + NVAL) of one step into out (v0 a multiple of W is not
// required: the covering vectors are loaded and the range picked out)
template <class R, int V0, int NVAL>
__device__ __forceinline__ void fac3_values(const R* __restrict__ step_base, R* out) {
  constexpr int W = Vec16<R>::W;
  constexpr int Q0 = V0 / W, Q1 = (V0 + NVAL - 1) / W;
  R buf[(Q1 - Q0 + 1) * W];
#pragma unroll
  for (int q = Q0; q <= Q1; ++q) {
    const typename Vec16<R>::T t = *reinterpret_cast<const typename Vec16<R>::T*>(step_base + (size_t)q * 64 * W);
    Vec16<R>::get(t, buf + (q - Q0) * W);
  }
#pragma unroll
  for (int e = 0; e < NVAL; ++e) out[e] = buf[V0 - Q0 * W + e];
}

template <class R, int NV>
__global__ void __launch_bounds__(64) k_thomas3(Plan3 P, const Ctl* __restrict__ ctl, const R* __restrict__ blk,
                                                R* __restrict__ fac, R* __restrict__ fac_sa) {
  const int ch = blockIdx.x * 64 + threadIdx.x, cs = blockIdx.y;
  if (ch >= P.n_lines || ctl[cs].parked) return;   // a parked serving slot: nothing to do this step
  const int s0 = P.l_start[ch], n = P.l_start[ch + 1] - s0, nl = P.n_lines;
  const size_t cstride = (size_t)P.max_len * nl;
  const R* B = blk + (size_t)cs * BLKW3 * cstride + ch;   // component e of step k at B[e cstride + k nl]
  // factor tiles of 64 chains, values packed W to a vector (FAC3_VALUES);
  // SA [case][chain / 64][step][3][64]
  constexpr int W = Vec16<R>::W;
  const size_t ngrp = ((size_t)nl + 63) / 64;
  R* F = fac + (((size_t)cs * ngrp + blockIdx.x) * P.max_len * FAC3_VALUES * 64) + (size_t)threadIdx.x * W;
  R* Fs = fac_sa + ((size_t)cs * ngrp + blockIdx.x) * P.max_len * 3 * 64 + threadIdx.x;
  auto slot = [&](int v) { return (size_t)(v / W) * 64 * W + v % W; };
  R Cp[5][5];
  for (int r = 0; r < 5; ++r)
    for (int q = 0; q < 5; ++q) Cp[r][q] = R(0);
  R cps = R(0);
  for (int k = 0; k < n; ++k) {
    // the step's 54 block values, all loaded before the first factor store
    const R* b0 = B + (size_t)k * nl;
    R bk[BLKW3];
#pragma unroll
    for (int e = 0; e < BLKW3; ++e) bk[e] = b0[(size_t)e * cstride];
    auto b = [&](int e) { return bk[e]; };
    const R d = (R)b(0);
    R Pm[5][5], Pi[5][5];
    for (int r = 0; r < 5; ++r)
      for (int q = 0; q < 5; ++q) {
        R acc = R(0);
        for (int t = 0; t < 5; ++t) acc += (R)b(1 + r * 5 + t) * Cp[t][q];
        Pm[r][q] = (r == q ? d : R(0)) - acc;
      }
    invert5<R>(Pm, Pi);
    R* o = F + (size_t)k * FAC3_VALUES * 64;
    for (int r = 0; r < 5; ++r)
      for (int q = 0; q < 5; ++q) {
        R acc = R(0);
        for (int t = 0; t < 5; ++t) acc += Pi[r][t] * (R)b(26 + t * 5 + q);
        o[slot(r * 5 + q)] = (R)b(1 + r * 5 + q);
        o[slot(25 + r * 5 + q)] = Pi[r][q];
        o[slot(50 + r * 5 + q)] = acc;
        Cp[r][q] = acc;
      }
    if constexpr (NV > 5) {
      const R lo = (R)b(51), ds = (R)b(52), up = (R)b(53);
      const R pinv = R(1) / (ds - lo * cps);
      const R cn = pinv * up;
      R* os = Fs + (size_t)k * 3 * 64;
      os[0] = lo;
      os[64] = pinv;
      os[2 * 64] = cn;
      cps = cn;
    }
  }
}

// z = M^-1 r (z += M^-1 r when accumulating), one thread per chain
// (linalg.py:103-132): the forward sweep writes d to tmp, the backward sweep
// reads it back and writes x.  With vout the input is the Arnoldi vector
// w/|w|, also written to vout.  Workgroups past the chains copy the cells on
// no chain (identity there).
//
// A chain is a dependent sequence, so each step's time is its load latency
// unless the loads run ahead.  Two register stages alternate: while one
// step computes from its stage, the loads of the next step (its factor
// blocks, the right-hand side of its cell, the SA scalars) are in flight
// into the other, and the cell index one step further is loaded too.  A
// scheduling barrier after each stage's loads keeps the compiler from
// moving the previous stage's arithmetic (and its waits) above them.
template <class R, int NV>
struct FwdStage {
  R L[25], Pm[25];
  double rhs[NV];
  R ls, ps;
  int cell;
};

template <class R, int NV>
struct BwdStage {
  R C[25];
  double d[NV], z[NV];
  R cs;
  int cell;
};

template <class R, int NV>
__device__ __forceinline__ void fwd_load(FwdStage<R, NV>& s, const R* __restrict__ o, const R* __restrict__ os,
                                         const double* __restrict__ rb, int cell, int vs, int nc) {
  R lp[50];
  fac3_values<R, 0, 50>(o, lp);
#pragma unroll
  for (int e = 0; e < 25; ++e) {
    s.L[e] = lp[e];
    s.Pm[e] = lp[25 + e];
  }
#pragma unroll
  for (int v = 0; v < NV; ++v) s.rhs[v] = rb[(size_t)v * nc + cell];
  if constexpr (NV > 5) {
    s.ls = os[0];
    s.ps = os[vs];
  }
  s.cell = cell;
}

// t = r_k - L_k d_{k-1}, d_k = P_k^-1 t (and the SA scalar), d_k to D.
template <class R, int NV>
__device__ __forceinline__ void fwd_step(FwdStage<R, NV>& s, R (&dp)[5], R& dps, double* __restrict__ D,
                                         double* __restrict__ vout, size_t base, double den, int nc) {
  if (vout) {
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const double x = s.rhs[v] / den;
      vout[base + (size_t)v * nc + s.cell] = x;
      s.rhs[v] = x;
    }
  }
  R t[5];
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    R acc = (R)s.rhs[i];
#pragma unroll
    for (int q = 0; q < 5; ++q) acc -= s.L[i * 5 + q] * dp[q];
    t[i] = acc;
  }
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    R acc = R(0);
#pragma unroll
    for (int q = 0; q < 5; ++q) acc += s.Pm[i * 5 + q] * t[q];
    dp[i] = acc;
  }
#pragma unroll
  for (int i = 0; i < 5; ++i) D[(size_t)i * nc + s.cell] = (double)dp[i];
  if constexpr (NV > 5) {
    dps = s.ps * ((R)s.rhs[5] - s.ls * dps);
    D[(size_t)5 * nc + s.cell] = (double)dps;
  }
}

template <bool ACC, class R, int NV>
__device__ __forceinline__ void bwd_load(BwdStage<R, NV>& s, const R* __restrict__ o, const R* __restrict__ os,
                                         const double* __restrict__ D, const double* __restrict__ zb, int cell,
                                         int vs, int nc) {
  fac3_values<R, 50, 25>(o, s.C);
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    s.d[v] = D[(size_t)v * nc + cell];
    s.z[v] = ACC ? zb[(size_t)v * nc + cell] : 0.0;
  }
  if constexpr (NV > 5) s.cs = os[2 * vs];
  s.cell = cell;
}

// x_k = d_k - C_k x_{k+1}, written (or added) to z.
template <bool ACC, class R, int NV>
__device__ __forceinline__ void bwd_step(const BwdStage<R, NV>& s, R (&xn)[5], R& xns, double* __restrict__ zb,
                                         int nc) {
  R x[5];
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    R acc = (R)s.d[i];
#pragma unroll
    for (int q = 0; q < 5; ++q) acc -= s.C[i * 5 + q] * xn[q];
    x[i] = acc;
  }
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    xn[i] = x[i];
    zb[(size_t)i * nc + s.cell] = ACC ? s.z[i] + (double)x[i] : (double)x[i];
  }
  if constexpr (NV > 5) {
    xns = (R)s.d[5] - s.cs * xns;
    zb[(size_t)5 * nc + s.cell] = ACC ? s.z[5] + (double)xns : (double)xns;
  }
}

template <class R, int NV, bool ACC>
__global__ void __launch_bounds__(64) k_apply3(Plan3 P, const Ctl* __restrict__ ctl, const R* __restrict__ fac,
                                               const R* __restrict__ fac_sa, const double* __restrict__ r,
                                               double* __restrict__ z, double* __restrict__ tmp, int check_active,
                                               double* __restrict__ vout, const double* __restrict__ p3, int nblk) {
  const int cs = blockIdx.y;
  if (skip_case(ctl, cs, check_active)) return;
  const int nc = P.nc, nl = P.n_lines;
  const size_t base = (size_t)cs * NV * nc;
  const double den = apply_den(p3 + (size_t)cs * nblk, nblk, vout);
  const int chain_groups = (nl + 63) / 64;
  if ((int)blockIdx.x >= chain_groups) {
    const int e = ((int)blockIdx.x - chain_groups) * 64 + threadIdx.x;
    if (e >= P.n_free) return;
    const int cell = P.free_cell[e];
    for (int v = 0; v < NV; ++v) {
      const size_t o = base + (size_t)v * nc + cell;
      const double x = apply_input(r, vout, o, den);
      z[o] = ACC ? z[o] + x : x;
    }
    return;
  }
  const int ch = blockIdx.x * 64 + threadIdx.x;
  if (ch >= nl) return;
  const int s0 = P.l_start[ch], n = P.l_start[ch + 1] - s0;
  if (n <= 0) return;
  const int* __restrict__ lc = P.l_cell + s0;
  // factor tiles of 64 chains (k_thomas3): one step of a wave's 64 chains is
  // 75 x 64 consecutive values, read at value stride 64
  const size_t ngrp = ((size_t)nl + 63) / 64;
  const R* F = fac + ((size_t)cs * ngrp + (ch >> 6)) * P.max_len * FAC3_VALUES * 64 + (size_t)(ch & 63) * Vec16<R>::W;
  const R* Fs = fac_sa + ((size_t)cs * ngrp + (ch >> 6)) * P.max_len * 3 * 64 + (ch & 63);
  const double* rb = r + base;
  double* D = tmp + base;
  double* zb = z + base;
  const size_t step = (size_t)FAC3_VALUES * 64, sstep = (size_t)3 * 64;

  R dp[5] = {R(0), R(0), R(0), R(0), R(0)};
  R dps = R(0);
  // Loads are unconditional (indices clamped to the chain) so that every
  // stage is loaded on every path and the next step's loads stay in flight
  // across the current step's arithmetic.
  {
    FwdStage<R, NV> A, B;
    fwd_load(A, F, Fs, rb, lc[0], 64, nc);
    int nxt = lc[min(1, n - 1)];
    int k = 0;
    for (; k + 1 < n; k += 2) {
      const int nxt2 = lc[min(k + 2, n - 1)];
      fwd_load(B, F + (size_t)(k + 1) * step, Fs + (size_t)(k + 1) * sstep, rb, nxt, 64, nc);
      __builtin_amdgcn_sched_barrier(0);
      fwd_step(A, dp, dps, D, vout, base, den, nc);
      const int k2 = min(k + 2, n - 1);
      nxt = lc[min(k + 3, n - 1)];
      fwd_load(A, F + (size_t)k2 * step, Fs + (size_t)k2 * sstep, rb, nxt2, 64, nc);
      __builtin_amdgcn_sched_barrier(0);
      fwd_step(B, dp, dps, D, vout, base, den, nc);
    }
    if (k < n) fwd_step(A, dp, dps, D, vout, base, den, nc);
  }
  R xn[5] = {R(0), R(0), R(0), R(0), R(0)};
  R xns = R(0);
  {
    BwdStage<R, NV> A, B;
    const R* C0 = F;   // the C values (50..74) are picked from the step's vectors
    bwd_load<ACC>(A, C0 + (size_t)(n - 1) * step, Fs + (size_t)(n - 1) * sstep, D, zb, lc[n - 1], 64, nc);
    int nxt = lc[max(n - 2, 0)];
    int k = n - 1;
    for (; k > 0; k -= 2) {
      const int nxt2 = lc[max(k - 2, 0)];
      bwd_load<ACC>(B, C0 + (size_t)(k - 1) * step, Fs + (size_t)(k - 1) * sstep, D, zb, nxt, 64, nc);
      __builtin_amdgcn_sched_barrier(0);
      bwd_step<ACC>(A, xn, xns, zb, nc);
      const int k2 = max(k - 2, 0);
      nxt = lc[max(k - 3, 0)];
      bwd_load<ACC>(A, C0 + (size_t)k2 * step, Fs + (size_t)k2 * sstep, D, zb, nxt2, 64, nc);
      __builtin_amdgcn_sched_barrier(0);
      bwd_step<ACC>(B, xn, xns, zb, nc);
    }
    if (k == 0) bwd_step<ACC>(A, xn, xns, zb, nc);
  }
}

// The same solve with three register stages, so two steps' loads are in
// flight while one computes: a chain's wave is latency-bound (one wave per
// 64 chains, about one wave per CU on the wing), and the extra stage doubles
// the bytes it keeps in flight.  The chain's cell indices are staged in LDS
// first (64 x max_len ints), so that no index load sits behind a stage's
// loads in the in-order memory counter.  For chains of at most
// APPLY3_LDS_LEN cells.
constexpr int APPLY3_LDS_LEN = 192;

template <class R, int NV, bool ACC>
__global__ void __launch_bounds__(64) k_apply3_s3(Plan3 P, const Ctl* __restrict__ ctl, const R* __restrict__ fac,
                                                  const R* __restrict__ fac_sa, const double* __restrict__ r,
                                                  double* __restrict__ z, double* __restrict__ tmp, int check_active,
                                                  double* __restrict__ vout, const double* __restrict__ p3,
                                                  int nblk) {
  extern __shared__ int cl[];          // cl[k * 64 + lane]: cell of step k of this lane's chain
  const int cs = blockIdx.y;
  if (skip_case(ctl, cs, check_active)) return;
  const int nc = P.nc, nl = P.n_lines;
  const size_t base = (size_t)cs * NV * nc;
  const double den = apply_den(p3 + (size_t)cs * nblk, nblk, vout);
  const int chain_groups = (nl + 63) / 64;
  if ((int)blockIdx.x >= chain_groups) {
    const int e = ((int)blockIdx.x - chain_groups) * 64 + threadIdx.x;
    if (e >= P.n_free) return;
    const int cell = P.free_cell[e];
    for (int v = 0; v < NV; ++v) {
      const size_t o = base + (size_t)v * nc + cell;
      const double x = apply_input(r, vout, o, den);
      z[o] = ACC ? z[o] + x : x;
    }
    return;
  }
  const int lane = threadIdx.x;
  const int ch = blockIdx.x * 64 + lane;
  if (ch >= nl) return;
  const int s0 = P.l_start[ch], n = P.l_start[ch + 1] - s0;
  if (n <= 0) return;
  for (int k = 0; k < n; ++k) cl[k * 64 + lane] = P.l_cell[s0 + k];
  auto cell_at = [&](int k) { return cl[k * 64 + lane]; };
  // factor tiles of 64 chains (k_thomas3): one step of a wave's 64 chains is
  // 75 x 64 consecutive values, read at value stride 64
  const size_t ngrp = ((size_t)nl + 63) / 64;
  const R* F = fac + ((size_t)cs * ngrp + (ch >> 6)) * P.max_len * FAC3_VALUES * 64 + (size_t)(ch & 63) * Vec16<R>::W;
  const R* Fs = fac_sa + ((size_t)cs * ngrp + (ch >> 6)) * P.max_len * 3 * 64 + (ch & 63);
  const double* rb = r + base;
  double* D = tmp + base;
  double* zb = z + base;
  const size_t step = (size_t)FAC3_VALUES * 64, sstep = (size_t)3 * 64;

  R dp[5] = {R(0), R(0), R(0), R(0), R(0)};
  R dps = R(0);
  {
    FwdStage<R, NV> A, B, C;
    auto load = [&](FwdStage<R, NV>& S, int k) {
      fwd_load(S, F + (size_t)k * step, Fs + (size_t)k * sstep, rb, cell_at(k), 64, nc);
    };
    load(A, 0);
    load(B, min(1, n - 1));
    int k = 0;
    for (; k + 2 < n; k += 3) {
      load(C, k + 2);
      __builtin_amdgcn_sched_barrier(0);
      fwd_step(A, dp, dps, D, vout, base, den, nc);
      load(A, min(k + 3, n - 1));
      __builtin_amdgcn_sched_barrier(0);
      fwd_step(B, dp, dps, D, vout, base, den, nc);
      load(B, min(k + 4, n - 1));
      __builtin_amdgcn_sched_barrier(0);
      fwd_step(C, dp, dps, D, vout, base, den, nc);
    }
    if (k < n) fwd_step(A, dp, dps, D, vout, base, den, nc);
    if (k + 1 < n) fwd_step(B, dp, dps, D, vout, base, den, nc);
  }
  R xn[5] = {R(0), R(0), R(0), R(0), R(0)};
  R xns = R(0);
  {
    BwdStage<R, NV> A, B, C;
    const R* C0 = F;   // the C values (50..74) are picked from the step's vectors
    auto load = [&](BwdStage<R, NV>& S, int k) {
      bwd_load<ACC>(S, C0 + (size_t)k * step, Fs + (size_t)k * sstep, D, zb, cell_at(k), 64, nc);
    };
    load(A, n - 1);
    load(B, max(n - 2, 0));
    int k = n - 1;
    for (; k - 2 >= 0; k -= 3) {
      load(C, k - 2);
      __builtin_amdgcn_sched_barrier(0);
      bwd_step<ACC>(A, xn, xns, zb, nc);
      load(A, max(k - 3, 0));
      __builtin_amdgcn_sched_barrier(0);
      bwd_step<ACC>(B, xn, xns, zb, nc);
      load(B, max(k - 4, 0));
      __builtin_amdgcn_sched_barrier(0);
      bwd_step<ACC>(C, xn, xns, zb, nc);
    }
    if (k >= 0) bwd_step<ACC>(A, xn, xns, zb, nc);
    if (k - 1 >= 0) bwd_step<ACC>(B, xn, xns, zb, nc);
  }
}

// --- Newton update: relaxation, accept test ----------------------------------

template <int NV>
__global__ void __launch_bounds__(BLOCK) k_relax3(Plan3 P, double gamma, Ctl* __restrict__ ctl,
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
    primitive3<double, NV>(u, gamma, q);
    double de = d[4];
    for (int k = 0; k < 3; ++k) de = de - q[1 + k] * d[1 + k];
    const double dp = (gamma - 1.0) * (de + 0.5 * dot3(q + 1, q + 1) * d[0]);
    double w = fmax(fabs(d[0]) / q[0], fabs(dp) / q[4]);
    if constexpr (NV > 5) {
      const double nh = u[5] / q[0];
      w = fmax(w, fabs((d[5] - nh * d[0]) / q[0]) / fmax(fabs(nh), 1.0));
    }
    worst = fmax(worst, w);
  }
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

template <int NV>
__global__ void __launch_bounds__(BLOCK) k_check3(Plan3 P, double gamma, Ctl* __restrict__ ctl,
                                                  const double* __restrict__ U, const double* __restrict__ X,
                                                  const double* __restrict__ Act, const double* __restrict__ R,
                                                  Red red, Ramp ramp) {
  __shared__ double sh[2][4];
  const int cs = blockIdx.y;
  if (ctl[cs].parked) return;   // a parked serving slot: nothing to do this step
  const int nc = P.nc;
  const size_t base = (size_t)cs * NV * nc;
  const double omega = ramp.relaxed ? ctl[cs].omega : 1.0;
  double v[2] = {0.0, 0.0};
  for (int c = blockIdx.x * BLOCK + threadIdx.x; c < nc; c += gridDim.x * BLOCK) {
    double cand[NV];
    bool finite = true;
    for (int k = 0; k < NV; ++k) {
      const size_t o = base + (size_t)k * nc + c;
      const double dx = omega * X[o];
      finite = finite && isfinite(dx);
      cand[k] = U[o] + dx;
      const double lr = omega * Act[o] + R[o];
      v[1] += lr * lr;
    }
    double q[NV];
    primitive3<double, NV>(cand, gamma, q);
    const bool phys = q[0] > 1e-10 && q[4] > 1e-10 && isfinite(q[0]) && isfinite(q[4]);
    if (!(phys && finite)) v[0] += 1.0;
  }
  __shared__ double tot[2];
  if (finish_sum<2>(v, red.partials + (size_t)cs * red.stride, red.counters + cs * 4, sh, tot)) {
    if (threadIdx.x == 0) {
      Ctl& c = ctl[cs];
      const double lin_red = sqrt(tot[1]) / fmax(c.bnorm, 1e-300);
      const bool accept = tot[0] == 0.0 && lin_red < 1.0;
      const double reduction = c.res_prev / fmax(c.norm, 1e-300);
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

// --- forces (forces.py:48-161, 3D) ------------------------------------------

// Laminar viscous force of wall face `fce` (concatenated numbering) from the
// ghost-filled primitives, the 3D diamond as in k_faces3.
// T = Dual: the shear's tangent, from dQ (the fill's tangent) and a mu_inf
// that may carry a parameter tangent (TM = Dual).
template <int NV, class T = double, class TM = double>
__device__ void wall_shear3(const Plan3& P, const Physics& ph, const double* Q, const double* dQ, size_t base,
                            TM mu_inf, int fce, T* out) {
  const int NI = P.NI, NJ = P.NJ, NK = P.NK, nc = P.nc;
  const double gamma = ph.gamma;
  int a = 0, ff = fce;
  if (ff >= P.nf[0]) { ff -= P.nf[0]; a = 1; if (ff >= P.nf[1]) { ff -= P.nf[1]; a = 2; } }
  const int dj = NJ + (a == 1), dk = NK + (a == 2);
  const int fi = ff / (dj * dk), rem = ff - fi * dj * dk, fj = rem / dk, fk = rem - fj * dk;
  const int stride[3] = {NJ * NK, NK, 1};
  const int cR = (fi * NJ + fj) * NK + fk, cL = cR - stride[a];
  const int iL = cL / (NJ * NK), jL = (cL / NK) % NJ, kL = cL % NK;
  double s[3];
  face_vec3(P, fce, s);
  const size_t cstride = (size_t)(NI + 2) * (NJ + 2) * (NK + 2);
  double d1[3], d2[3], d3[3];
  for (int c = 0; c < 3; ++c) d1[c] = P.cen[c * cstride + pc3(P, fi, fj, fk)] - P.cen[c * cstride + pc3(P, iL, jL, kL)];
  const int r1 = a == 0 ? 1 : 0, r2 = a == 2 ? 1 : 2;
  int nds[4][3];
  const int sh[4][2] = {{0, 0}, {1, 1}, {1, 0}, {0, 1}};
  for (int e = 0; e < 4; ++e) {
    nds[e][0] = fi; nds[e][1] = fj; nds[e][2] = fk;
    nds[e][r1] += sh[e][0];
    nds[e][r2] += sh[e][1];
  }
  auto node_centre = [&](const int* nd, int c) -> double {
    double v[8];
    for (int di = 0; di < 2; ++di)
      for (int dj2 = 0; dj2 < 2; ++dj2)
        for (int dk2 = 0; dk2 < 2; ++dk2)
          v[(di * 2 + dj2) * 2 + dk2] = P.cen[c * cstride + pc3(P, nd[0] - 1 + di, nd[1] - 1 + dj2, nd[2] - 1 + dk2)];
    return node_mean<double>(v);
  };
  for (int c = 0; c < 3; ++c) {
    d2[c] = node_centre(nds[1], c) - node_centre(nds[0], c);
    d3[c] = node_centre(nds[3], c) - node_centre(nds[2], c);
  }
  double c23[3] = {d2[1] * d3[2] - d2[2] * d3[1], d2[2] * d3[0] - d2[0] * d3[2], d2[0] * d3[1] - d2[1] * d3[0]};
  double c31[3] = {d3[1] * d1[2] - d3[2] * d1[1], d3[2] * d1[0] - d3[0] * d1[2], d3[0] * d1[1] - d3[1] * d1[0]};
  double c12[3] = {d1[1] * d2[2] - d1[2] * d2[1], d1[2] * d2[0] - d1[0] * d2[2], d1[0] * d2[1] - d1[1] * d2[0]};
  const double det = (d1[0] * c23[0] + d1[1] * c23[1]) + d1[2] * c23[2];
  auto phi_at = [&](int cell, T* o) {
    T q[5];
    for (int k = 0; k < 5; ++k) q[k] = Io<T>::ld(Q, dQ, base + (size_t)k * nc + cell);
    o[0] = q[1]; o[1] = q[2]; o[2] = q[3]; o[3] = gamma * q[4] / q[0];
  };
  T pn[4][4];
  for (int e = 0; e < 4; ++e) {
    T v[8][4];
    for (int di = 0; di < 2; ++di)
      for (int dj2 = 0; dj2 < 2; ++dj2)
        for (int dk2 = 0; dk2 < 2; ++dk2)
          phi_at(((nds[e][0] - 1 + di) * NJ + nds[e][1] - 1 + dj2) * NK + nds[e][2] - 1 + dk2, v[(di * 2 + dj2) * 2 + dk2]);
    for (int k = 0; k < 4; ++k) {
      T w[8];
      for (int m = 0; m < 8; ++m) w[m] = v[m][k];
      pn[e][k] = node_mean<T>(w);
    }
  }
  T pl[4], pr[4];
  phi_at(cL, pl);
  phi_at(cR, pr);
  T gr[3][4];
  for (int k = 0; k < 4; ++k) {
    const T dp1 = pr[k] - pl[k], dp2 = pn[1][k] - pn[0][k], dp3 = pn[3][k] - pn[2][k];
    for (int b = 0; b < 3; ++b) gr[b][k] = (dp1 * c23[b] + dp2 * c31[b] + dp3 * c12[b]) / det;
  }
  T gv[3][3], gT[3];
  for (int c = 0; c < 3; ++c) {
    for (int b = 0; b < 3; ++b) gv[c][b] = gr[b][c];
    gT[c] = gr[c][3];
  }
  T vel[3] = {0.5 * (pl[0] + pr[0]), 0.5 * (pl[1] + pr[1]), 0.5 * (pl[2] + pr[2])};
  const T Tf = 0.5 * (pl[3] + pr[3]);
  const T mu = viscosity<T>(Tf, mu_inf, ph.sutherland);
  const T kc = mu / (PR * (gamma - 1.0));
  T f[5];
  viscous_flux3<T>(vel, gv, gT, s, mu, kc, nullptr, f);
  out[0] = f[1]; out[1] = f[2]; out[2] = f[3];
}

// out[case][8] = CL, CD, CM, CD friction, Fx, Fy, Fz, CY: lift normal to the
// freestream in the x-z plane, per q_inf * area; moment about y through
// (cx, cz), pressure and friction, per q_inf * area * length.
template <int NV>
__global__ void __launch_bounds__(BLOCK) k_forces3(Plan3 P, Physics ph, const Ctl* __restrict__ ctl,
                                                   const double* __restrict__ Q, double area, double length,
                                                   double cx, double cz, double* __restrict__ out,
                                                   double* __restrict__ loads) {
  __shared__ double sh[6][4];
  const int cs = blockIdx.x;
  const size_t base = (size_t)cs * NV * P.nc;
  const double gamma = ph.gamma;
  const double* qi = ctl[cs].q_inf;
  const double speed = sqrt(dot3(qi + 1, qi + 1));
  const double dr[3] = {qi[1] / speed, qi[2] / speed, qi[3] / speed};
  const double li[3] = {-dr[2], 0.0, dr[0]};
  const double q_area = 0.5 * qi[0] * speed * speed * area;
  double v[6] = {0, 0, 0, 0, 0, 0};   // lift, drag, moment, shear x, y, z
  double fs[6] = {0, 0, 0, 0, 0, 0};  // force x, y, z
  for (int w = threadIdx.x; w < P.n_walls; w += BLOCK) {
    const int cell = P.w_cell[w], fam = P.w_fam[w];
    const int fce = (fam > 0 ? P.nf[0] : 0) + (fam > 1 ? P.nf[1] : 0) + P.w_face[w];
    const double into = P.w_into[w];
    double s[3];
    face_vec3(P, fce, s);
    for (int c = 0; c < 3; ++c) s[c] *= into;
    const double p = Q[base + (size_t)4 * P.nc + cell] - 1.0 / gamma;
    double pf[3] = {-p * s[0], -p * s[1], -p * s[2]};
    double t[3] = {0.0, 0.0, 0.0};
    if (ph.viscous) {
      wall_shear3<NV>(P, ph, Q, nullptr, base, ctl[cs].mu_inf, fce, t);
      for (int c = 0; c < 3; ++c) t[c] *= into;
    }
    if (loads) {   // per wall face: gauge pressure, viscous force on the body (x, y, z)
      double* lw = loads + (size_t)cs * 4 * P.n_walls + w;
      lw[0] = p;
      for (int c = 0; c < 3; ++c) lw[(size_t)(1 + c) * P.n_walls] = t[c];
    }
    double ld[3] = {pf[0] + t[0], pf[1] + t[1], pf[2] + t[2]};
    const double rx = P.w_mid[w] - cx, rz = P.w_mid[(size_t)2 * P.n_walls + w] - cz;
    const double moment = rz * ld[0] - rx * ld[2];
    v[0] += dot3(li, ld) / q_area;
    v[1] += dot3(dr, ld) / q_area;
    v[2] += (moment / length) / q_area;
    v[3] += t[0]; v[4] += t[1]; v[5] += t[2];
    fs[0] += ld[0]; fs[1] += ld[1]; fs[2] += ld[2];
  }
  block_sum<6>(v, sh);
  block_sum<6>(fs, sh);
  if (threadIdx.x == 0) {
    double* o = out + (size_t)cs * 8;
    o[0] = v[0];
    o[1] = v[1];
    o[2] = v[2];
    o[3] = dot3(dr, v + 3) / q_area;
    o[4] = fs[0];
    o[5] = fs[1];
    o[6] = fs[2];
    o[7] = fs[1] / q_area;
  }
}

// The tangent of k_forces3's coefficients along (dU, dq_inf, dmu_inf): the
// same integration on dual numbers over the parameter-tangent fill's dQ and
// the freestream's tangent P.dfree (the output JVP of the sensitivity,
// src/jaxrans/solvers/sensitivity.py:63).  out[case][8] = d(CL, CD, CM,
// CD friction, Fx, Fy, Fz, CY).
// dloads (may be null): the tangent of jr_wall_loads' per-face values,
// [case][4][n_walls] = d(gauge pressure, viscous force x, y, z).
template <int NV>
__global__ void __launch_bounds__(BLOCK) k_forces3_tangent(Plan3 P, Physics ph, const Ctl* __restrict__ ctl,
                                                           const double* __restrict__ Q, const double* __restrict__ dQ,
                                                           double area, double length, double cx, double cz,
                                                           double* __restrict__ out, double* __restrict__ dloads) {
  __shared__ double sh[18][4];
  const int cs = blockIdx.x;
  const size_t base = (size_t)cs * NV * P.nc;
  const double gamma = ph.gamma;
  Dual qi[5];
  for (int k = 0; k < 5; ++k) qi[k] = free_val<true>(ctl[cs].q_inf[k], P.dfree, cs, k);
  const Dual mu_inf = free_val<true>(ctl[cs].mu_inf, P.dfree, cs, 6);
  const Dual speed = jsqrt(dot3(qi + 1, qi + 1));
  const Dual dr[3] = {qi[1] / speed, qi[2] / speed, qi[3] / speed};
  const Dual li[3] = {-dr[2], mkd(0.0, 0.0), dr[0]};
  const Dual q_area = 0.5 * qi[0] * speed * speed * area;
  Dual acc[9];   // lift, drag, moment, shear x, y, z, force x, y, z
  for (int k = 0; k < 9; ++k) acc[k] = mkd(0.0, 0.0);
  for (int w = threadIdx.x; w < P.n_walls; w += BLOCK) {
    const int cell = P.w_cell[w], fam = P.w_fam[w];
    const int fce = (fam > 0 ? P.nf[0] : 0) + (fam > 1 ? P.nf[1] : 0) + P.w_face[w];
    const double into = P.w_into[w];
    double s[3];
    face_vec3(P, fce, s);
    for (int c = 0; c < 3; ++c) s[c] *= into;
    const Dual p = Io<Dual>::ld(Q, dQ, base + (size_t)4 * P.nc + cell) - 1.0 / gamma;
    const Dual pf[3] = {-p * s[0], -p * s[1], -p * s[2]};
    Dual t[3] = {mkd(0.0, 0.0), mkd(0.0, 0.0), mkd(0.0, 0.0)};
    if (ph.viscous) {
      wall_shear3<NV, Dual, Dual>(P, ph, Q, dQ, base, mu_inf, fce, t);
      for (int c = 0; c < 3; ++c) t[c] = t[c] * into;
    }
    if (dloads) {
      double* lw = dloads + (size_t)cs * 4 * P.n_walls + w;
      lw[0] = p.d;
      for (int c = 0; c < 3; ++c) lw[(size_t)(1 + c) * P.n_walls] = t[c].d;
    }
    const Dual ld[3] = {pf[0] + t[0], pf[1] + t[1], pf[2] + t[2]};
    const double rx = P.w_mid[w] - cx, rz = P.w_mid[(size_t)2 * P.n_walls + w] - cz;
    const Dual moment = rz * ld[0] - rx * ld[2];
    acc[0] = acc[0] + dot3(li, ld) / q_area;
    acc[1] = acc[1] + dot3(dr, ld) / q_area;
    acc[2] = acc[2] + (moment / length) / q_area;
    for (int c = 0; c < 3; ++c) {
      acc[3 + c] = acc[3 + c] + t[c];
      acc[6 + c] = acc[6 + c] + ld[c];
    }
  }
  double v[18];
  for (int k = 0; k < 9; ++k) { v[k] = acc[k].v; v[9 + k] = acc[k].d; }
  block_sum<18>(v, sh);
  if (threadIdx.x == 0) {
    Dual sv[3];
    for (int c = 0; c < 3; ++c) sv[c] = mkd(v[3 + c], v[12 + c]);
    const Dual friction = dot3(dr, sv) / q_area;
    const Dual cy = mkd(v[7], v[16]) / q_area;
    double* o = out + (size_t)cs * 8;
    o[0] = v[9];
    o[1] = v[10];
    o[2] = v[11];
    o[3] = friction.d;
    o[4] = v[15];
    o[5] = v[16];
    o[6] = v[17];
    o[7] = cy.d;
  }
}

}  // namespace jr
