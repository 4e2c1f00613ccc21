// Device-side plan and per-case control structures shared by the kernels and
// the C-ABI layer.  Plain structs of device pointers: a plan is uploaded once
// per grid, a case block per solve.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace jr {

constexpr int BLOCK = 256;     // threads per workgroup of the streaming kernels
constexpr int MMAX = 64;       // largest Krylov subspace supported

// The canvas and its static plans (one grid, shared by every case of a context).
struct Plan {
  int NI, NJ, nc;                  // canvas shape, cells
  int nfi, nfj;                    // i-faces (NI+1)*NJ, j-faces NI*(NJ+1)
  const double* area;              // [nc]
  const double* fi;                // [2][nfi] i-face area vectors
  const double* fj;                // [2][nfj] j-face area vectors
  const double* cen;               // [2][(NI+2)*(NJ+2)] centres, edge-padded by one
  const double* dist;              // [nc] wall distance
  const double* dg;                // [4][nfi + nfj] viscous diamond of each face: d1x, d1y, d2x, d2y
  const uint8_t* interior;         // [nc]
  const int* nbr;                  // [nc] face-neighbour map
  int n_task[2];                   // ghost-fill stages
  const int* t_dest[2];
  const int* t_src[2];
  const int* t_kind[2];
  const double* t_norm[2];         // [2][n]
  const uint8_t* ghost;            // [nc] 1: written by a ghost-fill stage
  const uint8_t* t_skip0;          // [n_task[0]] 1: the stage-2 fill overwrites this stage-1 destination
  const int* t_prev1;              // [n_task[1]] stage-1 task whose destination is this task's source, or -1
  // Linearised ghost fill of the tangent: the final fill tasks (stage-1 tasks
  // that stage 2 keeps, and every stage-2 task), each a function of one
  // non-ghost root cell; per cell, the final tasks rooted there.
  int n_final;                     // 0: not available (the tangent runs the fill kernel)
  const int* gq_task;              // [n_final] stage-1 t, or n_task[0] + stage-2 t
  const int* gq_dest;              // [n_final]
  const int* gq_root;              // [n_final]
  const int* gq_row;               // [n_final] row of the root in its chain, -1 if on no chain
  const int* cq_start;             // [n_lines + 1] CSR over chains: final tasks rooted on the chain
  const int* cq_q;                 // [n_final]
  int max_dep;                     // most final tasks rooted on one chain
  const int* gdep_start;           // [nc + 1] CSR over root cells on no chain
  const int* gdep_q;               // [n_final] final task indices, grouped by root
  int n_lines, n_slots, max_len;   // line chains
  const int* l_start;              // [n_lines+1]
  const int* l_cell;               // [n_slots]
  const int* l_up;                 // j-face of the link to the next slot
  const int8_t* l_ups;
  const int* l_dn;                 // j-face the next slot couples across
  const int8_t* l_dns;
  int n_free;                      // canvas cells on no chain
  const int* free_cell;
  int n_walls;
  const int* w_cell;
  const int* w_fam;
  const int* w_face;
  const double* w_into;
  const double* w_mid;             // [2][n_walls]
  // Geometry per case: case cs reads area + cs * gs_area, and so on (all 0
  // when the cases share the canvas geometry).  Per-case geometry lets one
  // batch hold different grids of one topology (the same index plans).
  size_t gs_area, gs_fi, gs_fj, gs_cen, gs_dist, gs_tn0, gs_tn1, gs_wmid, gs_dg;
  // Point-vortex far field: far-field tasks take the vortex of each case's
  // current circulation circ[case][2] (value, tangent), centred at (vx, vy),
  // at their face midpoints t_mid[stage] ([2][n]).
  int vortex;
  double vx, vy;
  const double* t_mid[2];
  const double* circ;
  // The freestream's tangent along a flow parameter, dfree[case][8] =
  // (dq_inf[0..5], dmu_inf, unused); read only by the parameter-tangent kernels.
  const double* dfree;
};

// The geometry of case cs (identity when the cases share it).
__device__ __forceinline__ void bind_case(Plan& P, int cs) {
  P.area += cs * P.gs_area;
  P.fi += cs * P.gs_fi;
  P.fj += cs * P.gs_fj;
  P.cen += cs * P.gs_cen;
  P.dist += cs * P.gs_dist;
  P.dg += cs * P.gs_dg;
  P.t_norm[0] += cs * P.gs_tn0;
  P.t_norm[1] += cs * P.gs_tn1;
  P.w_mid += cs * P.gs_wmid;
}

struct Physics {
  double gamma;
  double kappa;        // MUSCL kappa (0: limited)
  int flux;            // Flux
  int limiter;         // Limiter
  int nv;              // 4 Euler/laminar, 5 SA
  int viscous;
  int sa;
  double sutherland;   // <= 0: constant viscosity
  double c_cr1;        // QCR2000 coefficient (0: off)
};

// Per-case solver control, resident on the device.
struct Ctl {
  double q_inf[6];     // rho, u, v, p, nu_hat ratio (3D: rho, u, v, w, p, nu_hat ratio)
  double mu_inf;
  double cfl;
  double res_prev;     // |R| of the last accepted step (inf at start)
  double norm;         // |R(U)| reported for this step (density row, volume weighted)
  double bnorm;        // ||R||_2 over the whole canvas vector
  double rnorm;        // ||b - A x|| of the current GMRES iterate
  double omega;        // relaxation factor
  double lin_red;
  double beta;         // ||r|| at the start of the current GMRES cycle
  int cycles;          // GMRES cycles done this step
  int active;          // GMRES cycle running
  int accept;          // last step accepted
  int steps;           // Newton steps taken
  int rejected;
  int cycles_total;    // GMRES cycles run since the context was made (measurement)
  int parked;          // an empty serving slot: every pass of a Newton step skips it (skip_case, parked checks)
};

// Whether a kernel skips case cs, by its check_active argument: 0 none, 1 a
// case whose GMRES cycle is not running, 2 a parked serving slot (the passes
// of a Newton step, which leave a parked slot's state and control as they
// are).
__device__ __forceinline__ bool skip_case(const Ctl* ctl, int cs, int check) {
  return check == 1 ? !ctl[cs].active : check == 2 && ctl[cs].parked;
}

}  // namespace jr
