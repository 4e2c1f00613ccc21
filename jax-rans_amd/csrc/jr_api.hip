// C-ABI of the Newton-Krylov hot path: context set-up, plan upload, and the
// launch sequences of the residual, its tangent, the line preconditioner and
// whole Newton steps.  Declarations and the reference interfaces each entry
// point replaces: include/jrhip.h.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/jrhip.h"
#include "jr_krylov.h"
#include "jr_cr.h"
#include "jr_lines.h"
#include "jr_3d.h"

using namespace jr;

static thread_local std::string g_error;

static int fail(const std::string& msg) {
  g_error = msg;
  return -1;
}

#define HIPCHK(x)                                                                                  \
  do {                                                                                             \
    hipError_t e_ = (x);                                                                           \
    if (e_ != hipSuccess) return fail(std::string(#x) + ": " + hipGetErrorString(e_));             \
  } while (0)

struct jr_ctx {
  Plan P{};
  int dim = 2;       // 3: the 3D path (P3; P holds nc, interior and volumes for the shared vector kernels)
  Plan3 P3{};
  double* F3 = nullptr;     // 3D face fluxes [case][var][faces of the three families]
  double* dsw = nullptr;    // 3D line apply: forward-sweep workspace, one field
  double* salin = nullptr;  // SA: partials of the source at the state (SALIN2 / SALIN fields)
  double *Nd = nullptr, *dNd = nullptr;   // 3D viscous: node values of (u, v, w, T[, nu_hat]) and their tangents
  Physics ph{};
  int nv = 4, n_cases = 1;
  int m = 10, restarts = 2, exact = 0, mixed = 0, thomas = 0, cr_threads = 64;
  double ktol = 1e-2, max_rel = -1.0;
  Ramp ramp{};
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  size_t N = 0;      // entries of one case's field (nv * nc)
  int nblk = 1;      // workgroups per case of the vector kernels
  std::vector<void*> allocs;
  double *U = nullptr, *R = nullptr, *Q = nullptr, *dQ = nullptr, *Fi = nullptr, *Fj = nullptr;
  double *X = nullptr, *Act = nullptr, *Rv = nullptr, *Z = nullptr, *W = nullptr, *T = nullptr, *V = nullptr;
  double *tmpA = nullptr, *tmpB = nullptr;
  double *lam = nullptr, *sam = nullptr, *sad = nullptr, *lsum = nullptr, *vdt = nullptr, *sa_diag = nullptr;
  void *fac = nullptr, *fac_sa = nullptr, *ws = nullptr, *ws_sa = nullptr;
  double* blk = nullptr;
  Ctl* ctl = nullptr;
  double *hv = nullptr, *h2v = nullptr, *G = nullptr, *H = nullptr, *y = nullptr, *hy = nullptr, *hist = nullptr, *coef = nullptr;
  int hist_cap = 0;
  int prof_rep = 0;  // launch counter of jr_profile
  // A block of Newton steps replayed from a HIP graph (captured on the first
  // block of a given length; the history buffer is part of the capture).
  bool use_graph = false;
  // Asynchronous blocks (serving): no host synchronisation inside the block,
  // the residual history lands in pinned host memory behind ev_block.
  bool async_mode = false;
  int pending_steps = 0;
  bool pending_forces = false;
  double* hist_pinned = nullptr;   // [pinned_cap][n_cases] history, then [n_cases][8] coefficients
  int pinned_cap = 0;
  hipEvent_t ev_block = nullptr;
  size_t bytes = 0;         // device memory this context holds
  int graph_steps = 0;
  hipGraphExec_t graph = nullptr;
  Red red{};
  Cg cg{};           // CGS2 partial sums
  double* gjac = nullptr;   // [n_cases][n_final][nv][nv] linearised ghost fill
  bool dq_fused = false;    // the line apply writes the tangent input (fill linearised)
  // Event probe: when armed, launches of the selected kernels are bracketed
  // by event pairs from a preallocated pool (measurement only).
  unsigned probe_mask = 0;
  std::vector<hipEvent_t> probe_ev;
  std::vector<int> probe_id;      // kernel id of each recorded pair
  int probe_used = 0;
  bool has_q = false;
  // Linear solve with a given right-hand side (jr_linear_solve): while set,
  // a Newton step solves its system for b = solve_b and stops before the update.
  const double* solve_b = nullptr;
  double* solve_buf = nullptr;
  bool per_case_geo = false;   // geometry arrays per case (jr_set_geometry)
  uint64_t topo_hash = 0;      // hash of the index plans: grids of one topology share them
  int device = 0;              // the HIP device current at creation (its stream's device)
  double* loads = nullptr;     // jr_wall_loads: per wall face [case][1 + dim][n_walls]
  double* circ = nullptr;      // point-vortex far field: [case][2] circulation and its tangent
  double* dfree = nullptr;     // parameter tangents: [case][8] the freestream's tangent (Plan::dfree)
  double* zero_field = nullptr;   // parameter tangents at dU = 0: a zero state tangent
  double* dloads = nullptr;    // jr_wall_loads_tangent: [case][4][n_walls]
  bool tile_tangent = false;   // the Krylov tangent by LDS tiles (k_tangent_tile) where it applies
  // 2D block Thomas (k_thomas, k_line_apply): the chains by length, longest
  // first, and their length classes
  const int* lorder = nullptr;
  LineLaunch lines{};
};

// FNV-1a over the index plans of a grid (stages, chains, walls, neighbours,
// interior): two grids with the same hash share every index plan, so one
// batch can hold both with per-case geometry.
static uint64_t topology_hash(const jr_grid_desc* g) {
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](const void* p, size_t bytes) {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < bytes; ++i) h = (h ^ b[i]) * 1099511628211ull;
  };
  const size_t nc = (size_t)g->NI * g->NJ;
  mix(&g->NI, sizeof(g->NI));
  mix(&g->NJ, sizeof(g->NJ));
  mix(g->interior, nc);
  mix(g->neighbours, nc * sizeof(int64_t));
  for (int s = 0; s < 2; ++s) {
    mix(&g->n_task[s], sizeof(int32_t));
    mix(g->task_dest[s], g->n_task[s] * sizeof(int64_t));
    mix(g->task_src[s], g->n_task[s] * sizeof(int64_t));
    mix(g->task_kind[s], g->n_task[s] * sizeof(int32_t));
  }
  mix(&g->n_lines, sizeof(int32_t));
  mix(&g->n_slots, sizeof(int32_t));
  mix(g->line_start, ((size_t)g->n_lines + 1) * sizeof(int64_t));
  for (const int64_t* a : {g->line_cells, g->line_up_face, g->line_up_sign, g->line_down_face, g->line_down_sign})
    mix(a, (size_t)g->n_slots * sizeof(int64_t));
  mix(&g->n_walls, sizeof(int32_t));
  for (const int64_t* a : {g->wall_cell, g->wall_family, g->wall_face}) mix(a, (size_t)g->n_walls * sizeof(int64_t));
  mix(g->wall_into, (size_t)g->n_walls * sizeof(double));
  return h;
}

template <class T>
static int dalloc(jr_ctx* c, T** p, size_t count) {
  void* q = nullptr;
  if (count == 0) count = 1;
  HIPCHK(hipMalloc(&q, count * sizeof(T)));
  HIPCHK(hipMemset(q, 0, count * sizeof(T)));
  c->allocs.push_back(q);
  c->bytes += count * sizeof(T);
  *p = static_cast<T*>(q);
  return 0;
}

template <class D, class S>
static int upload(jr_ctx* c, const D** dst, const S* src, size_t count) {
  if (!src) return fail("a required plan array is null");
  std::vector<D> h(count);
  for (size_t i = 0; i < count; ++i) h[i] = static_cast<D>(src[i]);
  D* d = nullptr;
  if (dalloc(c, &d, count)) return -1;
  HIPCHK(hipMemcpy(d, h.data(), count * sizeof(D), hipMemcpyHostToDevice));
  *dst = d;
  return 0;
}

#define TRY(x)          \
  do {                  \
    if ((x) != 0) {     \
      return -1;        \
    }                   \
  } while (0)

static dim3 grid1(size_t n, int per, int cases) { return dim3((unsigned)((n + per - 1) / per), (unsigned)cases); }

// Kernel ids of the event probe (jr_probe_arm).
enum ProbeId { PR_APPLY = 0, PR_FACES = 1, PR_FILL = 2, PR_CELLS = 3, PR_DOTS = 4, PR_UPDATE_DOTS = 5,
               PR_UPDATE_NORM = 6, PR_FACTOR = 7, PR_N = 8 };

// Event pair for the next launch when kernel `id` is probed and the pool has
// room (-1 otherwise).  The probed launch goes through hipExtLaunchKernelGGL,
// whose events are stamped by the dispatch itself: kernel start to kernel end.
static int probe_slot(jr_ctx* c, int id) {
  if (!(c->probe_mask & (1u << id)) || 2 * (c->probe_used + 1) > (int)c->probe_ev.size()) return -1;
  const int k = c->probe_used++;
  c->probe_id[k] = id;
  return k;
}

#define JR_LAUNCH(ID, KERNEL, GRID, BLOCKDIM, SHMEM, ...)                                                     \
  do {                                                                                                       \
    const int pk_ = probe_slot(c, ID);                                                                       \
    if (pk_ < 0)                                                                                             \
      hipLaunchKernelGGL(KERNEL, GRID, BLOCKDIM, SHMEM, c->stream, __VA_ARGS__);                             \
    else                                                                                                     \
      hipExtLaunchKernelGGL(KERNEL, GRID, BLOCKDIM, SHMEM, c->stream, c->probe_ev[2 * pk_],                  \
                            c->probe_ev[2 * pk_ + 1], 0, __VA_ARGS__);                                       \
  } while (0)

// --- residual and tangent launch sequences -----------------------------------

// check: which cases the kernels skip (skip_case; 2 in a Newton step: the
// parked slots).
template <int NV, bool VISC>
static void launch_residual(jr_ctx* c, const double* U, double* out, int first_order, int check = 0) {
  const Plan& P = c->P;
  if (P.vortex)   // the far-field reference moves with the current lift
    hipLaunchKernelGGL((k_circulation<double, NV>), dim3(c->n_cases), dim3(BLOCK), 0, c->stream, P, c->ph.gamma,
                       c->ctl, U, nullptr, c->circ, check);
  hipLaunchKernelGGL((k_fill<double, NV>), grid1((size_t)P.nc + P.n_task[0] + P.n_task[1], BLOCK, c->n_cases),
                     dim3(BLOCK), 0, c->stream, P, c->ph.gamma, c->ctl, U, nullptr, c->Q, nullptr, check);
  if (VISC)
    hipLaunchKernelGGL((k_nodes<double, NV>), grid1((size_t)(P.NI + 1) * (P.NJ + 1), BLOCK, c->n_cases), dim3(BLOCK),
                       0, c->stream, P, c->ph.gamma, c->ctl, c->Q, nullptr, c->Nd, nullptr, check);
  hipLaunchKernelGGL((k_faces<double, NV, VISC>), grid1(P.nfi + P.nfj, BLOCK, c->n_cases), dim3(BLOCK), 0,
                     c->stream, P, c->ph, first_order, c->ctl, c->Q, nullptr, c->Nd, nullptr, c->Fi, c->Fj, check);
  hipLaunchKernelGGL((k_cells<double, NV>), grid1(P.nc, BLOCK, c->n_cases), dim3(BLOCK), 0, c->stream, P, c->ph,
                     c->ctl, U, nullptr, c->Q, nullptr, c->Fi, c->Fj, nullptr, nullptr, out, check);
  if (NV > 4)   // the SA source's partials at this state, for the tangents that follow
    hipLaunchKernelGGL(k_sa_lin, grid1(P.nc, BLOCK, c->n_cases), dim3(BLOCK), 0, c->stream, P, c->ph, c->ctl, c->Q,
                       c->salin, check);
}

// Requires c->Q (primitives with ghosts) for the same U.
// dq_ready: the preceding line apply already wrote dQ on every cell (ghosts
// through the linearised fill), so the fill kernel is not needed.
template <int NV, bool VISC>
static void launch_tangent(jr_ctx* c, const double* U, const double* dU, double* out, int first_order,
                           const double* vdt, int check_active, bool dq_ready = false) {
  const Plan& P = c->P;
  if (P.vortex && !dq_ready)   // the circulation's tangent for the far-field ghosts
    hipLaunchKernelGGL((k_circulation<Dual, NV>), dim3(c->n_cases), dim3(BLOCK), 0, c->stream, P, c->ph.gamma,
                       (const Ctl*)c->ctl, U, dU, c->circ, check_active);
  if (!dq_ready)
    JR_LAUNCH(PR_FILL, (k_fill<Dual, NV>), grid1((size_t)P.nc + P.n_task[0] + P.n_task[1], BLOCK, c->n_cases),
              dim3(BLOCK), 0, P, c->ph.gamma, c->ctl, U, dU, (double*)nullptr, c->dQ, check_active);
  if (dq_ready && first_order && c->tile_tangent) {
    // nodes, faces and cells in one pass over LDS tiles
    const unsigned tiles = (unsigned)(((P.NI + TT_I - 1) / TT_I) * ((P.NJ + TT_J - 1) / TT_J));
    JR_LAUNCH(PR_FACES, (k_tangent_tile<NV, VISC>), dim3(tiles, c->n_cases), dim3(TT_THREADS), 0, P, c->ph,
              (const Ctl*)c->ctl, dU, (const double*)c->Q, (const double*)c->dQ, vdt, (const double*)c->salin, out,
              check_active);
    return;
  }
  if (VISC)
    hipLaunchKernelGGL((k_nodes<Dual, NV>), grid1((size_t)(P.NI + 1) * (P.NJ + 1), BLOCK, c->n_cases), dim3(BLOCK), 0,
                       c->stream, P, c->ph.gamma, (const Ctl*)c->ctl, (const double*)c->Q, (const double*)c->dQ,
                       (double*)nullptr, c->dNd, check_active);
  JR_LAUNCH(PR_FACES, (k_faces<Dual, NV, VISC>), grid1(P.nfi + P.nfj, BLOCK, c->n_cases), dim3(BLOCK), 0, P, c->ph,
            first_order, (const Ctl*)c->ctl, (const double*)c->Q, (const double*)c->dQ, (const double*)c->Nd,
            (const double*)c->dNd, c->Fi, c->Fj, check_active);
  JR_LAUNCH(PR_CELLS, (k_cells<Dual, NV>), grid1(P.nc, BLOCK, c->n_cases), dim3(BLOCK), 0, P, c->ph,
            (const Ctl*)c->ctl, U, dU, (const double*)c->Q, (const double*)c->dQ, (const double*)c->Fi,
            (const double*)c->Fj, vdt, (const double*)c->salin, out, check_active);
}

template <int NV>
static void launch_speeds(jr_ctx* c, const double* U) {
  const Plan& P = c->P;
  hipLaunchKernelGGL((k_face_speeds<NV>), grid1(P.nfi + P.nfj, BLOCK, c->n_cases), dim3(BLOCK), 0, c->stream, P, c->ph,
                     c->ctl, U, c->lam, c->sam, c->sad);
  hipLaunchKernelGGL((k_cell_speeds<NV>), grid1(P.nc, BLOCK, c->n_cases), dim3(BLOCK), 0, c->stream, P, c->ph, c->ctl,
                     U, c->lam, c->sam, c->sad, c->lsum, c->vdt, c->sa_diag);
}

template <class R, int NV>
static void launch_factor(jr_ctx* c, const double* U) {
  const Plan& P = c->P;
  // the chain blocks in the factor's precision (the buffer holds FP64 blocks, so FP32 ones fit)
  R* blk = reinterpret_cast<R*>(c->blk);
  hipLaunchKernelGGL((k_blocks<R, NV>), grid1(P.n_slots, BLOCK, c->n_cases), dim3(BLOCK), 0, c->stream, P, c->ph,
                     c->ctl, U, c->lam, c->lsum, c->sam, c->sad, c->sa_diag, blk);
  if (c->thomas)
    JR_LAUNCH(PR_FACTOR, (k_thomas<R, NV>), dim3((P.n_lines + 63) / 64, c->n_cases), dim3(64), 0, P,
              (const Ctl*)c->ctl, (const int*)c->lorder, (const R*)blk, (R*)c->fac, (R*)c->fac_sa);
  else
    JR_LAUNCH(PR_FACTOR, (k_cr_factor<R, NV>), dim3(P.n_lines, c->n_cases), dim3(c->cr_threads), 0, P,
              (const Ctl*)c->ctl, (const R*)blk, (R*)c->fac, (R*)c->fac_sa, (R*)c->ws, (R*)c->ws_sa);
}

template <class R, int NV>
static void launch_apply(jr_ctx* c, const double* r, double* z, int accumulate, int check_active,
                         double* vout = nullptr, double* dq_out = nullptr) {
  const Plan& P = c->P;
  if (c->thomas) {
    const unsigned groups = (unsigned)(c->lines.wave0[c->lines.n] + (P.n_free + 63) / 64);
    if (groups > 0) {
      auto kern = accumulate ? k_line_apply<R, NV, true> : k_line_apply<R, NV, false>;
      JR_LAUNCH(PR_APPLY, kern, dim3(groups, c->n_cases), dim3(64), line_apply_lds<NV>(c->lines.rows), P,
                (const Ctl*)c->ctl, (const int*)c->lorder, c->lines, (const R*)c->fac, (const R*)c->fac_sa, r, z,
                check_active, vout, (const double*)c->cg.p3, c->cg.nblk, dq_out, (const double*)c->U, c->ph.gamma,
                (const double*)c->gjac);
    }
  } else {
    const int T = c->cr_threads;
    const unsigned groups = (unsigned)(P.n_lines + (P.n_free + T - 1) / T);
    const size_t lds_cr = cr_apply_lds<R, NV>(P.max_len, P.max_dep, dq_out != nullptr);
    JR_LAUNCH(PR_APPLY, (k_cr_apply<R, NV>), dim3(groups, c->n_cases), dim3(T), lds_cr, P, (const Ctl*)c->ctl,
              (const R*)c->fac, (const R*)c->fac_sa, r, z, accumulate, check_active, vout, (const double*)c->cg.p3,
              c->cg.nblk, dq_out, (const double*)c->U, c->ph.gamma, (const double*)c->gjac);
  }
}

static int get_ctl(jr_ctx* c, std::vector<Ctl>& h);

// Whether any case's GMRES cycle is active (reads the control blocks: a
// stream synchronisation).  On a read error it answers true.
static bool any_active(jr_ctx* c) {
  std::vector<Ctl> h;
  if (get_ctl(c, h) != 0) return true;
  for (const Ctl& x : h)
    if (x.active) return true;
  return false;
}

// --- CGS2 of Arnoldi column j, instantiated for j / KC + 1 basis chunks --------

template <int NV, int NCH, int PW>
static void cgs2_passes_pw(jr_ctx* c, int j, int passes) {
  const Plan& P = c->P;
  const dim3 vg(c->nblk, c->n_cases);
  const size_t N = c->N, vstride = (size_t)c->n_cases * N;
  if (passes & 1)
    JR_LAUNCH(PR_DOTS, (k_dots<NV, NCH, PW>), vg, dim3(BLOCK), 0, P.nc, (const Ctl*)c->ctl, (const double*)c->V,
              vstride, N, (const double*)c->W, j, c->cg, (const double*)c->hv, (const double*)c->h2v, c->H, c->G,
              c->m);
  if (passes & 2)
    JR_LAUNCH(PR_UPDATE_DOTS, (k_update_dots<NV, NCH, PW>), vg, dim3(BLOCK), 0, P.nc, (const Ctl*)c->ctl,
              (const double*)c->V, vstride, N, c->W, j, c->hv, c->cg);
  if (passes & 4)
    JR_LAUNCH(PR_UPDATE_NORM, (k_update_norm<NV, NCH, PW>), vg, dim3(BLOCK), 0, P.nc, (const Ctl*)c->ctl,
              (const double*)c->V, vstride, N, c->W, j, c->h2v, c->cg);
}

// Two entries per thread and loop step (16-byte loads) when the field length
// is even (every case offset then keeps the alignment) and j < 16: for the
// later columns the doubled batch of basis values leaves one wave per SIMD
// (256 VGPRs), and one entry per step at two or three waves per SIMD is
// faster (the 32-case polar's CGS2 passes: 26.7 -> 25.3 ms per Newton step).
template <int NV, int NCH>
static void cgs2_passes(jr_ctx* c, int j, int passes) {
  if constexpr (NCH <= 2) {
    if (c->N % 2 == 0) {
      cgs2_passes_pw<NV, NCH, 2>(c, j, passes);
      return;
    }
  }
  cgs2_passes_pw<NV, NCH, 1>(c, j, passes);
}

template <int NV>
static void cgs2_column(jr_ctx* c, int j, int passes = 7) {
  switch (j / KC) {
    case 0: cgs2_passes<NV, 1>(c, j, passes); break;
    case 1: cgs2_passes<NV, 2>(c, j, passes); break;
    case 2: cgs2_passes<NV, 3>(c, j, passes); break;
    case 3: cgs2_passes<NV, 4>(c, j, passes); break;
    case 4: cgs2_passes<NV, 5>(c, j, passes); break;
    case 5: cgs2_passes<NV, 6>(c, j, passes); break;
    case 6: cgs2_passes<NV, 7>(c, j, passes); break;
    default: cgs2_passes<NV, 8>(c, j, passes); break;
  }
}
static_assert(MMAX <= 8 * KC, "cgs2_column covers j < 64");

// --- one Newton step ----------------------------------------------------------
// Krylov basis layout: V[k][case][N], so V_k of every case is one contiguous
// field batch (case stride N) that the apply and tangent kernels take as is.

template <int NV, bool VISC, int MM, class R>
static void gmres_and_update(jr_ctx* c) {
  const Plan& P = c->P;
  const dim3 vg(c->nblk, c->n_cases);
  const size_t N = c->N, vstride = (size_t)c->n_cases * N;
  const int fo = c->exact ? 0 : 1;
  for (int cyc = 0; cyc < c->restarts; ++cyc) {
    hipLaunchKernelGGL((k_cycle_begin<NV>), vg, dim3(BLOCK), 0, c->stream, P.nc, c->ctl, c->Rv, c->V, N, c->restarts,
                       c->ktol, c->G);
    // A later cycle runs only if some case has not met its tolerance
    // (krylov.py:85-87); when none has, its ~5m launches would all exit at
    // once, so the host reads the flags and skips them (not under capture).
    if (cyc > 0 && !c->use_graph && !c->async_mode && !any_active(c)) break;
    for (int j = 0; j < c->m; ++j) {
      // z = M^-1 V_j; for j > 0 the apply also forms V_j = w / |w| from the
      // previous column's w
      // the cyclic-reduction apply also writes the tangent's dQ (ghost cells
      // through the linearised fill), so the tangent needs no fill launch
      double* dq = c->dq_fused ? c->dQ : nullptr;
      if (j == 0) launch_apply<R, NV>(c, c->V, c->Z, 0, 1, nullptr, dq);
      else launch_apply<R, NV>(c, c->W, c->Z, 0, 1, c->V + (size_t)j * vstride, dq);
      launch_tangent<NV, VISC>(c, c->U, c->Z, c->W, fo, c->vdt, 1, dq != nullptr);
      cgs2_column<NV>(c, j);
    }
    hipLaunchKernelGGL((k_normalize<NV>), vg, dim3(BLOCK), 0, c->stream, P.nc, c->ctl, c->W,
                       c->V + (size_t)c->m * vstride, N, c->cg);
    hipLaunchKernelGGL(k_lstsq, dim3(c->n_cases), dim3(64), 0, c->stream, c->n_cases, c->ctl, c->H, c->G, c->m, c->y,
                       c->hy, c->cg, c->hv, c->h2v);
    hipLaunchKernelGGL((k_combine<NV, MM>), vg, dim3(BLOCK), 0, c->stream, P.nc, c->ctl, c->V, vstride, N, c->y, c->hy,
                       c->m, c->T, c->Act, c->Rv, c->R, c->red);
    launch_apply<R, NV>(c, c->T, c->X, 1, 1);
  }
  if (c->solve_b) return;   // a linear solve: X is the answer, U stays
  const dim3 cg(c->nblk, c->n_cases);
  if (c->max_rel > 0.0)
    hipLaunchKernelGGL((k_relax<NV>), cg, dim3(BLOCK), 0, c->stream, P, c->ph.gamma, c->ctl, c->U, c->X, c->red,
                       c->max_rel);
  Ramp ramp = c->ramp;
  ramp.relaxed = c->max_rel > 0.0;
  hipLaunchKernelGGL((k_check<NV>), cg, dim3(BLOCK), 0, c->stream, P, c->ph.gamma, c->ctl, c->U, c->X, c->Act, c->R,
                     c->red, ramp);
  hipLaunchKernelGGL((k_update<NV>), vg, dim3(BLOCK), 0, c->stream, P.nc, c->ctl, c->U, c->X, ramp.relaxed);
}

template <int NV, bool VISC, int MM, class R>
static void newton_step(jr_ctx* c, double* hist_row) {
  const Plan& P = c->P;
  const dim3 vg(c->nblk, c->n_cases);
  launch_residual<NV, VISC>(c, c->U, c->R, 0, 2);
  if (c->dq_fused)
    hipLaunchKernelGGL((k_ghost_jac<NV>), grid1(P.n_final, BLOCK, c->n_cases), dim3(BLOCK), 0, c->stream, P,
                       c->ph.gamma, c->ctl, c->Q, c->gjac);
  if (c->solve_b)   // the GMRES start r = -R is then b
    hipLaunchKernelGGL(k_negate, grid1((size_t)c->n_cases * c->N, BLOCK, 1), dim3(BLOCK), 0, c->stream,
                       (size_t)c->n_cases * c->N, c->solve_b, c->R);
  hipLaunchKernelGGL((k_norms<NV>), vg, dim3(BLOCK), 0, c->stream, P, c->ctl, c->R, c->X, c->Act, c->Rv, c->red,
                     hist_row, c->n_cases);
  launch_speeds<NV>(c, c->U);
  launch_factor<R, NV>(c, c->U);
  gmres_and_update<NV, VISC, MM, R>(c);
}

// One explicit multistage step (solvers/explicit.py:48-71): dt / V frozen at
// U^n, the first stage's residual reported, U^(k) = U^n - alpha_k dt/V
// R(U^(k-1)); the intermediate stages live in W, the last writes U.
template <int NV, bool VISC>
static void rk_step(jr_ctx* c, double* hist_row, const double* alphas, int n) {
  const Plan& P = c->P;
  const dim3 vg(c->nblk, c->n_cases);
  const dim3 eg((unsigned)(((size_t)NV * P.nc + BLOCK - 1) / BLOCK), (unsigned)c->n_cases);
  hipLaunchKernelGGL((k_local_dt<NV, VISC>), grid1(P.nc, BLOCK, c->n_cases), dim3(BLOCK), 0, c->stream, P, c->ph,
                     c->ctl, c->U, c->vdt);
  launch_residual<NV, VISC>(c, c->U, c->R, 0);
  hipLaunchKernelGGL((k_norms<NV>), vg, dim3(BLOCK), 0, c->stream, P, c->ctl, c->R, (double*)nullptr, (double*)nullptr,
                     (double*)nullptr, c->red, hist_row, c->n_cases);   // norms only: no GMRES start
  for (int k = 0; k < n; ++k) {
    if (k > 0) launch_residual<NV, VISC>(c, c->W, c->R, 0);
    double* out = k + 1 < n ? c->W : c->U;
    hipLaunchKernelGGL((k_rk_stage<NV>), eg, dim3(BLOCK), 0, c->stream, P.nc, (const double*)c->U,
                       (const double*)c->vdt, (const double*)c->R, alphas[k], out);
  }
}


// --- the 3D path ---------------------------------------------------------------

template <int NV, bool VISC>
static void launch_residual3(jr_ctx* c, const double* U, double* out, int first_order, int check = 0) {
  const Plan3& P = c->P3;
  hipLaunchKernelGGL((k_fill3<double, NV>), grid1((size_t)P.nc + P.n_ghost, BLOCK, c->n_cases), dim3(BLOCK), 0,
                     c->stream, P, c->ph.gamma, c->ctl, U, nullptr, c->Q, nullptr, check);
  const size_t nn = (size_t)(P.NI + 1) * (P.NJ + 1) * (P.NK + 1);
  if (VISC)
    hipLaunchKernelGGL((k_nodes3<double, NV>), grid1(nn, BLOCK, c->n_cases), dim3(BLOCK), 0, c->stream, P,
                       c->ph.gamma, c->ctl, c->Q, nullptr, c->Nd, nullptr, check);
  hipLaunchKernelGGL((k_faces3<double, NV, VISC>), grid1(P.nft, BLOCK, c->n_cases), dim3(BLOCK), 0, c->stream, P,
                     c->ph, first_order, c->ctl, c->Q, nullptr, c->Nd, nullptr, c->F3, check);
  hipLaunchKernelGGL((k_cells3<double, NV>), grid1(P.nc, BLOCK, c->n_cases), dim3(BLOCK), 0, c->stream, P, c->ph,
                     c->ctl, U, nullptr, c->Q, nullptr, c->F3, nullptr, nullptr, out, check);
  if (NV > 5)   // the SA source's partials at this state, for the tangents that follow
    hipLaunchKernelGGL(k_sa_lin3, grid1(P.nc, BLOCK, c->n_cases), dim3(BLOCK), 0, c->stream, P, c->ph, c->ctl, c->Q,
                       c->salin, check);
}

template <int NV, bool VISC>
static void launch_tangent3(jr_ctx* c, const double* U, const double* dU, double* out, int first_order,
                            const double* vdt, int check_active) {
  const Plan3& P = c->P3;
  JR_LAUNCH(PR_FILL, (k_fill3<Dual, NV>), grid1((size_t)P.nc + P.n_ghost, BLOCK, c->n_cases), dim3(BLOCK), 0, P,
            c->ph.gamma, (const Ctl*)c->ctl, U, dU, (double*)nullptr, c->dQ, check_active);
  const size_t nn = (size_t)(P.NI + 1) * (P.NJ + 1) * (P.NK + 1);
  if (VISC)
    hipLaunchKernelGGL((k_nodes3<Dual, NV>), grid1(nn, BLOCK, c->n_cases), dim3(BLOCK), 0, c->stream, P, c->ph.gamma,
                       (const Ctl*)c->ctl, (const double*)c->Q, (const double*)c->dQ, (double*)nullptr, c->dNd,
                       check_active);
  JR_LAUNCH(PR_FACES, (k_faces3<Dual, NV, VISC>), grid1(P.nft, BLOCK, c->n_cases), dim3(BLOCK), 0, P, c->ph,
            first_order, (const Ctl*)c->ctl, (const double*)c->Q, (const double*)c->dQ, (const double*)c->Nd,
            (const double*)c->dNd, c->F3, check_active);
  JR_LAUNCH(PR_CELLS, (k_cells3<Dual, NV>), grid1(P.nc, BLOCK, c->n_cases), dim3(BLOCK), 0, P, c->ph,
            (const Ctl*)c->ctl, U, dU, (const double*)c->Q, (const double*)c->dQ, (const double*)c->F3, vdt,
            (const double*)c->salin, out, check_active);
}

template <int NV>
static void launch_speeds3(jr_ctx* c, const double* U) {
  const Plan3& P = c->P3;
  hipLaunchKernelGGL((k_face_speeds3<NV>), grid1(P.nft, BLOCK, c->n_cases), dim3(BLOCK), 0, c->stream, P, c->ph,
                     c->ctl, U, c->lam, c->sam, c->sad);
  hipLaunchKernelGGL((k_cell_speeds3<NV>), grid1(P.nc, BLOCK, c->n_cases), dim3(BLOCK), 0, c->stream, P, c->ph,
                     c->ctl, U, c->lam, c->sam, c->sad, c->lsum, c->vdt, c->sa_diag);
}

template <class R, int NV>
static void launch_factor3(jr_ctx* c, const double* U) {
  const Plan3& P = c->P3;
  R* blk = reinterpret_cast<R*>(c->blk);   // blocks in the factor's precision
  hipLaunchKernelGGL((k_blocks3<R, NV>), grid1((size_t)P.max_len * P.n_lines, BLOCK, c->n_cases), dim3(BLOCK), 0,
                     c->stream, P, c->ph, c->ctl, U, c->lam, c->lsum, c->sam, c->sad, c->sa_diag, blk);
  JR_LAUNCH(PR_FACTOR, (k_thomas3<R, NV>), dim3((P.n_lines + 63) / 64, c->n_cases), dim3(64), 0, P,
            (const Ctl*)c->ctl, (const R*)blk, (R*)c->fac, (R*)c->fac_sa);
}

template <class R, int NV>
static void launch_apply3(jr_ctx* c, const double* r, double* z, int accumulate, int check_active,
                          double* vout = nullptr) {
  const Plan3& P = c->P3;
  const unsigned groups = (unsigned)((P.n_lines + 63) / 64 + (P.n_free + 63) / 64);
  if (P.max_len <= APPLY3_LDS_LEN) {   // three stages, the chains' cells in LDS
    const size_t lds = (size_t)64 * std::max(1, P.max_len) * sizeof(int);
    if (accumulate)
      JR_LAUNCH(PR_APPLY, (k_apply3_s3<R, NV, true>), dim3(groups, c->n_cases), dim3(64), lds, P,
                (const Ctl*)c->ctl, (const R*)c->fac, (const R*)c->fac_sa, r, z, c->dsw, check_active, vout,
                (const double*)c->cg.p3, c->cg.nblk);
    else
      JR_LAUNCH(PR_APPLY, (k_apply3_s3<R, NV, false>), dim3(groups, c->n_cases), dim3(64), lds, P,
                (const Ctl*)c->ctl, (const R*)c->fac, (const R*)c->fac_sa, r, z, c->dsw, check_active, vout,
                (const double*)c->cg.p3, c->cg.nblk);
    return;
  }
  if (accumulate)
    JR_LAUNCH(PR_APPLY, (k_apply3<R, NV, true>), dim3(groups, c->n_cases), dim3(64), 0, P, (const Ctl*)c->ctl,
              (const R*)c->fac, (const R*)c->fac_sa, r, z, c->dsw, check_active, vout, (const double*)c->cg.p3,
              c->cg.nblk);
  else
    JR_LAUNCH(PR_APPLY, (k_apply3<R, NV, false>), dim3(groups, c->n_cases), dim3(64), 0, P, (const Ctl*)c->ctl,
              (const R*)c->fac, (const R*)c->fac_sa, r, z, c->dsw, check_active, vout, (const double*)c->cg.p3,
              c->cg.nblk);
}

template <int NV, bool VISC, int MM, class R>
static void gmres_and_update3(jr_ctx* c) {
  const Plan& P = c->P;
  const dim3 vg(c->nblk, c->n_cases);
  const size_t N = c->N, vstride = (size_t)c->n_cases * N;
  const int fo = c->exact ? 0 : 1;
  for (int cyc = 0; cyc < c->restarts; ++cyc) {
    hipLaunchKernelGGL((k_cycle_begin<NV>), vg, dim3(BLOCK), 0, c->stream, P.nc, c->ctl, c->Rv, c->V, N, c->restarts,
                       c->ktol, c->G);
    if (cyc > 0 && !c->use_graph && !c->async_mode && !any_active(c)) break;
    for (int j = 0; j < c->m; ++j) {
      if (j == 0) launch_apply3<R, NV>(c, c->V, c->Z, 0, 1);
      else launch_apply3<R, NV>(c, c->W, c->Z, 0, 1, c->V + (size_t)j * vstride);
      launch_tangent3<NV, VISC>(c, c->U, c->Z, c->W, fo, c->vdt, 1);
      cgs2_column<NV>(c, j);
    }
    hipLaunchKernelGGL((k_normalize<NV>), vg, dim3(BLOCK), 0, c->stream, P.nc, c->ctl, c->W,
                       c->V + (size_t)c->m * vstride, N, c->cg);
    hipLaunchKernelGGL(k_lstsq, dim3(c->n_cases), dim3(64), 0, c->stream, c->n_cases, c->ctl, c->H, c->G, c->m, c->y,
                       c->hy, c->cg, c->hv, c->h2v);
    hipLaunchKernelGGL((k_combine<NV, MM>), vg, dim3(BLOCK), 0, c->stream, P.nc, c->ctl, c->V, vstride, N, c->y, c->hy,
                       c->m, c->T, c->Act, c->Rv, c->R, c->red);
    launch_apply3<R, NV>(c, c->T, c->X, 1, 1);
  }
  if (c->solve_b) return;   // a linear solve: X is the answer, U stays
  const dim3 cg(c->nblk, c->n_cases);
  if (c->max_rel > 0.0)
    hipLaunchKernelGGL((k_relax3<NV>), cg, dim3(BLOCK), 0, c->stream, c->P3, c->ph.gamma, c->ctl, c->U, c->X, c->red,
                       c->max_rel);
  Ramp ramp = c->ramp;
  ramp.relaxed = c->max_rel > 0.0;
  hipLaunchKernelGGL((k_check3<NV>), cg, dim3(BLOCK), 0, c->stream, c->P3, c->ph.gamma, c->ctl, c->U, c->X, c->Act,
                     c->R, c->red, ramp);
  hipLaunchKernelGGL((k_update<NV>), vg, dim3(BLOCK), 0, c->stream, P.nc, c->ctl, c->U, c->X, ramp.relaxed);
}

template <int NV, bool VISC, int MM, class R>
static void newton_step3(jr_ctx* c, double* hist_row) {
  const Plan& P = c->P;
  const dim3 vg(c->nblk, c->n_cases);
  launch_residual3<NV, VISC>(c, c->U, c->R, 0, 2);
  if (c->solve_b)   // the GMRES start r = -R is then b
    hipLaunchKernelGGL(k_negate, grid1((size_t)c->n_cases * c->N, BLOCK, 1), dim3(BLOCK), 0, c->stream,
                       (size_t)c->n_cases * c->N, c->solve_b, c->R);
  hipLaunchKernelGGL((k_norms<NV>), vg, dim3(BLOCK), 0, c->stream, P, c->ctl, c->R, c->X, c->Act, c->Rv, c->red,
                     hist_row, c->n_cases);
  launch_speeds3<NV>(c, c->U);
  launch_factor3<R, NV>(c, c->U);
  gmres_and_update3<NV, VISC, MM, R>(c);
}

template <int NV, bool VISC>
static void step3_m(jr_ctx* c, double* hist_row) {
  if (c->m <= 16) {
    if (c->mixed) newton_step3<NV, VISC, 16, float>(c, hist_row); else newton_step3<NV, VISC, 16, double>(c, hist_row);
  } else if (c->m <= 32) {
    if (c->mixed) newton_step3<NV, VISC, 32, float>(c, hist_row); else newton_step3<NV, VISC, 32, double>(c, hist_row);
  } else {
    if (c->mixed) newton_step3<NV, VISC, 64, float>(c, hist_row); else newton_step3<NV, VISC, 64, double>(c, hist_row);
  }
}

static void step3_dispatch(jr_ctx* c, double* hist_row) {
  if (c->nv == 6) step3_m<6, true>(c, hist_row);
  else if (c->ph.viscous) step3_m<5, true>(c, hist_row);
  else step3_m<5, false>(c, hist_row);
}

// The same step on a 3D canvas.
template <int NV, bool VISC>
static void rk_step3(jr_ctx* c, double* hist_row, const double* alphas, int n) {
  const Plan& P = c->P;
  const dim3 vg(c->nblk, c->n_cases);
  const dim3 eg((unsigned)(((size_t)NV * P.nc + BLOCK - 1) / BLOCK), (unsigned)c->n_cases);
  hipLaunchKernelGGL((k_local_dt3<NV, VISC>), grid1(P.nc, BLOCK, c->n_cases), dim3(BLOCK), 0, c->stream, c->P3,
                     c->ph, c->ctl, c->U, c->vdt);
  launch_residual3<NV, VISC>(c, c->U, c->R, 0);
  hipLaunchKernelGGL((k_norms<NV>), vg, dim3(BLOCK), 0, c->stream, P, c->ctl, c->R, (double*)nullptr, (double*)nullptr,
                     (double*)nullptr, c->red, hist_row, c->n_cases);   // norms only: no GMRES start
  for (int k = 0; k < n; ++k) {
    if (k > 0) launch_residual3<NV, VISC>(c, c->W, c->R, 0);
    double* out = k + 1 < n ? c->W : c->U;
    hipLaunchKernelGGL((k_rk_stage<NV>), eg, dim3(BLOCK), 0, c->stream, P.nc, (const double*)c->U,
                       (const double*)c->vdt, (const double*)c->R, alphas[k], out);
  }
}

static void rk_dispatch(jr_ctx* c, double* hist_row, const double* alphas, int n) {
  if (c->dim == 3) {
    if (c->nv == 6) rk_step3<6, true>(c, hist_row, alphas, n);
    else if (c->ph.viscous) rk_step3<5, true>(c, hist_row, alphas, n);
    else rk_step3<5, false>(c, hist_row, alphas, n);
    return;
  }
  if (c->nv == 5) rk_step<5, true>(c, hist_row, alphas, n);
  else if (c->ph.viscous) rk_step<4, true>(c, hist_row, alphas, n);
  else rk_step<4, false>(c, hist_row, alphas, n);
}

// --- dispatch over the compile-time variants ------------------------------------

template <int NV, bool VISC, int MM>
static void step_dispatch_r(jr_ctx* c, double* hist_row) {
  if (c->mixed) newton_step<NV, VISC, MM, float>(c, hist_row);
  else newton_step<NV, VISC, MM, double>(c, hist_row);
}

template <int NV, bool VISC>
static void step_dispatch_m(jr_ctx* c, double* hist_row) {
  if (c->m <= 16) step_dispatch_r<NV, VISC, 16>(c, hist_row);
  else if (c->m <= 32) step_dispatch_r<NV, VISC, 32>(c, hist_row);
  else step_dispatch_r<NV, VISC, 64>(c, hist_row);
}

static void step_dispatch(jr_ctx* c, double* hist_row) {
  if (c->dim == 3) return step3_dispatch(c, hist_row);
  if (c->nv == 5) step_dispatch_m<5, true>(c, hist_row);
  else if (c->ph.viscous) step_dispatch_m<4, true>(c, hist_row);
  else step_dispatch_m<4, false>(c, hist_row);
}

static void residual_dispatch(jr_ctx* c, const double* U, double* out, int fo) {
  if (c->dim == 3) {
    if (c->nv == 6) launch_residual3<6, true>(c, U, out, fo);
    else if (c->ph.viscous) launch_residual3<5, true>(c, U, out, fo);
    else launch_residual3<5, false>(c, U, out, fo);
    return;
  }
  if (c->nv == 5) launch_residual<5, true>(c, U, out, fo);
  else if (c->ph.viscous) launch_residual<4, true>(c, U, out, fo);
  else launch_residual<4, false>(c, U, out, fo);
}

static void tangent_dispatch(jr_ctx* c, const double* U, const double* dU, double* out, int fo, const double* vdt) {
  if (c->dim == 3) {
    if (c->nv == 6) launch_tangent3<6, true>(c, U, dU, out, fo, vdt, 0);
    else if (c->ph.viscous) launch_tangent3<5, true>(c, U, dU, out, fo, vdt, 0);
    else launch_tangent3<5, false>(c, U, dU, out, fo, vdt, 0);
    return;
  }
  if (c->nv == 5) launch_tangent<5, true>(c, U, dU, out, fo, vdt, 0);
  else if (c->ph.viscous) launch_tangent<4, true>(c, U, dU, out, fo, vdt, 0);
  else launch_tangent<4, false>(c, U, dU, out, fo, vdt, 0);
}

static void speeds_dispatch(jr_ctx* c, const double* U) {
  if (c->dim == 3) {
    if (c->nv == 6) launch_speeds3<6>(c, U); else launch_speeds3<5>(c, U);
    return;
  }
  if (c->nv == 5) launch_speeds<5>(c, U);
  else launch_speeds<4>(c, U);
}

static void factor_dispatch(jr_ctx* c, const double* U) {
  if (c->dim == 3) {
    if (c->nv == 6) { if (c->mixed) launch_factor3<float, 6>(c, U); else launch_factor3<double, 6>(c, U); }
    else { if (c->mixed) launch_factor3<float, 5>(c, U); else launch_factor3<double, 5>(c, U); }
    return;
  }
  if (c->nv == 5) { if (c->mixed) launch_factor<float, 5>(c, U); else launch_factor<double, 5>(c, U); }
  else { if (c->mixed) launch_factor<float, 4>(c, U); else launch_factor<double, 4>(c, U); }
}

static void apply_dispatch(jr_ctx* c, const double* r, double* z) {
  if (c->dim == 3) {
    if (c->nv == 6) { if (c->mixed) launch_apply3<float, 6>(c, r, z, 0, 0); else launch_apply3<double, 6>(c, r, z, 0, 0); }
    else { if (c->mixed) launch_apply3<float, 5>(c, r, z, 0, 0); else launch_apply3<double, 5>(c, r, z, 0, 0); }
    return;
  }
  if (c->nv == 5) { if (c->mixed) launch_apply<float, 5>(c, r, z, 0, 0); else launch_apply<double, 5>(c, r, z, 0, 0); }
  else { if (c->mixed) launch_apply<float, 4>(c, r, z, 0, 0); else launch_apply<double, 4>(c, r, z, 0, 0); }
}

static int forces3_dispatch(jr_ctx* c, double area, double length, double cx, double cz,
                            double* loads = nullptr) {
  const Plan3& P = c->P3;
  if (c->nv == 6) {
    launch_residual3<6, true>(c, c->U, c->tmpA, 0);
    hipLaunchKernelGGL((k_forces3<6>), dim3(c->n_cases), dim3(BLOCK), 0, c->stream, P, c->ph, c->ctl, c->Q, area,
                       length, cx, cz, c->coef, loads);
  } else {
    if (c->ph.viscous) launch_residual3<5, true>(c, c->U, c->tmpA, 0);
    else launch_residual3<5, false>(c, c->U, c->tmpA, 0);
    hipLaunchKernelGGL((k_forces3<5>), dim3(c->n_cases), dim3(BLOCK), 0, c->stream, P, c->ph, c->ctl, c->Q, area,
                       length, cx, cz, c->coef, loads);
  }
  return 0;
}

static int forces_dispatch(jr_ctx* c, double length, double cx, double cy, double* loads = nullptr) {
  const Plan& P = c->P;
  if (c->nv == 5) {
    launch_residual<5, true>(c, c->U, c->tmpA, 0);
    hipLaunchKernelGGL((k_forces<5>), dim3(c->n_cases), dim3(BLOCK), 0, c->stream, P, c->ph, c->ctl, c->Q, length, cx,
                       cy, c->coef, loads);
  } else {
    if (c->ph.viscous) launch_residual<4, true>(c, c->U, c->tmpA, 0);
    else launch_residual<4, false>(c, c->U, c->tmpA, 0);
    hipLaunchKernelGGL((k_forces<4>), dim3(c->n_cases), dim3(BLOCK), 0, c->stream, P, c->ph, c->ctl, c->Q, length, cx,
                       cy, c->coef, loads);
  }
  return 0;
}

// --- host helpers -----------------------------------------------------------------

static int check_launch() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(std::string("kernel launch: ") + hipGetErrorString(e));
  return 0;
}

static size_t field_bytes(const jr_ctx* c) { return (size_t)c->n_cases * c->N * sizeof(double); }

// --- parameter tangents (sensitivity) -----------------------------------------

// The exact residual's tangent along (dU, dq_inf, dmu_inf) at U: the
// parameter-tangent kernels (PT) over the primal primitives, node values and
// circulation of a residual evaluation at U.
template <int NV, bool VISC>
static void launch_param_tangent(jr_ctx* c, const double* U, const double* dU, double* out) {
  const Plan& P = c->P;
  launch_residual<NV, VISC>(c, U, c->R, 0);
  if (P.vortex)
    hipLaunchKernelGGL((k_circulation<Dual, NV, true>), dim3(c->n_cases), dim3(BLOCK), 0, c->stream, P, c->ph.gamma,
                       (const Ctl*)c->ctl, U, dU, c->circ, 0);
  hipLaunchKernelGGL((k_fill<Dual, NV, true>), grid1((size_t)P.nc + P.n_task[0] + P.n_task[1], BLOCK, c->n_cases),
                     dim3(BLOCK), 0, c->stream, P, c->ph.gamma, (const Ctl*)c->ctl, U, dU, (double*)nullptr, c->dQ, 0);
  if (VISC)
    hipLaunchKernelGGL((k_nodes<Dual, NV>), grid1((size_t)(P.NI + 1) * (P.NJ + 1), BLOCK, c->n_cases), dim3(BLOCK), 0,
                       c->stream, P, c->ph.gamma, (const Ctl*)c->ctl, (const double*)c->Q, (const double*)c->dQ,
                       (double*)nullptr, c->dNd, 0);
  hipLaunchKernelGGL((k_faces<Dual, NV, VISC, true>), grid1(P.nfi + P.nfj, BLOCK, c->n_cases), dim3(BLOCK), 0,
                     c->stream, P, c->ph, 0, (const Ctl*)c->ctl, (const double*)c->Q, (const double*)c->dQ,
                     (const double*)c->Nd, (const double*)c->dNd, c->Fi, c->Fj, 0);
  hipLaunchKernelGGL((k_cells<Dual, NV, true>), grid1(P.nc, BLOCK, c->n_cases), dim3(BLOCK), 0, c->stream, P, c->ph,
                     (const Ctl*)c->ctl, U, dU, (const double*)c->Q, (const double*)c->dQ, (const double*)c->Fi,
                     (const double*)c->Fj, (const double*)nullptr, (const double*)c->salin, out, 0);
}

// The force coefficients' tangent along (dU, dq_inf, dmu_inf) at U.
template <int NV, bool VISC>
static void launch_forces_tangent(jr_ctx* c, const double* U, const double* dU, double length, double cx, double cy,
                                  double* out) {
  const Plan& P = c->P;
  launch_residual<NV, VISC>(c, U, c->R, 0);
  if (P.vortex)
    hipLaunchKernelGGL((k_circulation<Dual, NV, true>), dim3(c->n_cases), dim3(BLOCK), 0, c->stream, P, c->ph.gamma,
                       (const Ctl*)c->ctl, U, dU, c->circ, 0);
  hipLaunchKernelGGL((k_fill<Dual, NV, true>), grid1((size_t)P.nc + P.n_task[0] + P.n_task[1], BLOCK, c->n_cases),
                     dim3(BLOCK), 0, c->stream, P, c->ph.gamma, (const Ctl*)c->ctl, U, dU, (double*)nullptr, c->dQ, 0);
  hipLaunchKernelGGL((k_forces_tangent<NV>), dim3(c->n_cases), dim3(BLOCK), 0, c->stream, P, c->ph,
                     (const Ctl*)c->ctl, (const double*)c->Q, (const double*)c->dQ, length, cx, cy, out);
}

template <int NV, bool VISC>
static void launch_param_tangent3(jr_ctx* c, const double* U, const double* dU, double* out) {
  const Plan3& P = c->P3;
  launch_residual3<NV, VISC>(c, U, c->R, 0);
  hipLaunchKernelGGL((k_fill3<Dual, NV, true>), grid1((size_t)P.nc + P.n_ghost, BLOCK, c->n_cases), dim3(BLOCK), 0,
                     c->stream, P, c->ph.gamma, (const Ctl*)c->ctl, U, dU, (double*)nullptr, c->dQ, 0);
  const size_t nn = (size_t)(P.NI + 1) * (P.NJ + 1) * (P.NK + 1);
  if (VISC)
    hipLaunchKernelGGL((k_nodes3<Dual, NV>), grid1(nn, BLOCK, c->n_cases), dim3(BLOCK), 0, c->stream, P, c->ph.gamma,
                       (const Ctl*)c->ctl, (const double*)c->Q, (const double*)c->dQ, (double*)nullptr, c->dNd, 0);
  hipLaunchKernelGGL((k_faces3<Dual, NV, VISC, true>), grid1(P.nft, BLOCK, c->n_cases), dim3(BLOCK), 0, c->stream, P,
                     c->ph, 0, (const Ctl*)c->ctl, (const double*)c->Q, (const double*)c->dQ, (const double*)c->Nd,
                     (const double*)c->dNd, c->F3, 0);
  hipLaunchKernelGGL((k_cells3<Dual, NV, true>), grid1(P.nc, BLOCK, c->n_cases), dim3(BLOCK), 0, c->stream, P, c->ph,
                     (const Ctl*)c->ctl, U, dU, (const double*)c->Q, (const double*)c->dQ, (const double*)c->F3,
                     (const double*)nullptr, (const double*)c->salin, out, 0);
}

template <int NV, bool VISC>
static void launch_forces3_tangent(jr_ctx* c, const double* U, const double* dU, double area, double length,
                                   double cx, double cz, double* out, double* dloads = nullptr) {
  const Plan3& P = c->P3;
  launch_residual3<NV, VISC>(c, U, c->R, 0);
  hipLaunchKernelGGL((k_fill3<Dual, NV, true>), grid1((size_t)P.nc + P.n_ghost, BLOCK, c->n_cases), dim3(BLOCK), 0,
                     c->stream, P, c->ph.gamma, (const Ctl*)c->ctl, U, dU, (double*)nullptr, c->dQ, 0);
  hipLaunchKernelGGL((k_forces3_tangent<NV>), dim3(c->n_cases), dim3(BLOCK), 0, c->stream, P, c->ph,
                     (const Ctl*)c->ctl, (const double*)c->Q, (const double*)c->dQ, area, length, cx, cz, out, dloads);
}

extern "C" {

const char* jr_last_error(void) { return g_error.c_str(); }

int jr_device_count(int* n) {
  HIPCHK(hipGetDeviceCount(n));
  return 0;
}

int jr_device_alloc(size_t bytes, void** out) {
  HIPCHK(hipMalloc(out, bytes ? bytes : 1));
  return 0;
}

int jr_device_free(void* p) {
  HIPCHK(hipFree(p));
  return 0;
}

int jr_copy_to_device(void* dst, const void* src, size_t bytes) {
  HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return 0;
}

int jr_copy_to_host(void* dst, const void* src, size_t bytes) {
  HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return 0;
}

int jr_ctx_create(const jr_grid_desc* g, const jr_numerics* num, int n_cases, jr_ctx** out) {
  if (!g || !num || !out) return fail("jr_ctx_create: null argument");
  if (n_cases < 1) return fail("jr_ctx_create: n_cases must be >= 1");
  if (g->NI < 5 || g->NJ < 5) return fail("jr_ctx_create: canvas too small");
  if (num->krylov_subspace < 1 || num->krylov_subspace > MMAX)
    return fail("jr_ctx_create: krylov_subspace must lie in [1, " + std::to_string(MMAX) + "]");
  if (num->krylov_restarts < 1) return fail("jr_ctx_create: krylov_restarts must be >= 1");
  if (num->flux < 0 || num->flux > 1) return fail("jr_ctx_create: unknown flux");
  if (num->limiter < 0 || num->limiter > 3) return fail("jr_ctx_create: unknown limiter");
  if (num->sa && !num->viscous) return fail("jr_ctx_create: SA needs a viscous flow");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail("jr_ctx_create: no HIP device");
  jr_ctx* c = new jr_ctx();
  (void)hipGetDevice(&c->device);
  auto bail = [&](int) {
    jr_ctx_destroy(c);
    return -1;
  };
  Plan& P = c->P;
  P.NI = g->NI;
  P.NJ = g->NJ;
  P.nc = g->NI * g->NJ;
  P.nfi = (g->NI + 1) * g->NJ;
  P.nfj = g->NI * (g->NJ + 1);
  const size_t nc = P.nc;
  c->topo_hash = topology_hash(g);
  if (upload(c, &P.area, g->area, nc) || upload(c, &P.fi, g->face_i, 2 * (size_t)P.nfi) ||
      upload(c, &P.fj, g->face_j, 2 * (size_t)P.nfj) ||
      upload(c, &P.cen, g->centres, 2 * (size_t)(g->NI + 2) * (g->NJ + 2)) ||
      upload(c, &P.dist, g->distance, nc) || upload(c, &P.interior, g->interior, nc) ||
      upload(c, &P.nbr, g->neighbours, nc))
    return bail(0);
  // validate indices on the host before any kernel sees them
  for (size_t i = 0; i < nc; ++i)
    if (g->neighbours[i] < 0 || g->neighbours[i] >= (int64_t)nc) {
      fail("jr_ctx_create: neighbour index out of range");
      return bail(0);
    }
  for (int s = 0; s < 2; ++s) {
    P.n_task[s] = g->n_task[s];
    for (int t = 0; t < g->n_task[s]; ++t) {
      if (g->task_dest[s][t] < 0 || g->task_dest[s][t] >= (int64_t)nc || g->task_src[s][t] < 0 ||
          g->task_src[s][t] >= (int64_t)nc || g->task_kind[s][t] < 0 || g->task_kind[s][t] > 6) {
        fail("jr_ctx_create: ghost-fill task out of range");
        return bail(0);
      }
    }
    if (upload(c, &P.t_dest[s], g->task_dest[s], g->n_task[s]) ||
        upload(c, &P.t_src[s], g->task_src[s], g->n_task[s]) ||
        upload(c, &P.t_kind[s], g->task_kind[s], g->n_task[s]) ||
        upload(c, &P.t_norm[s], g->task_normal[s], 2 * (size_t)g->n_task[s]))
      return bail(0);
  }
  if (num->far_field != 0 && num->far_field != 1) {
    fail("jr_ctx_create: unknown far-field model");
    return bail(0);
  }
  P.vortex = num->far_field == 1;
  if (P.vortex) {
    if (!g->task_mid[0] || !g->task_mid[1]) {
      fail("jr_ctx_create: the point-vortex far field needs the task midpoints");
      return bail(0);
    }
    P.vx = g->vortex_x;
    P.vy = g->vortex_y;
    for (int s = 0; s < 2; ++s)
      if (upload(c, &P.t_mid[s], g->task_mid[s], 2 * (size_t)g->n_task[s])) return bail(0);
    double* circ = nullptr;
    if (dalloc(c, &circ, (size_t)n_cases * 2)) return bail(0);
    c->circ = circ;
    P.circ = circ;
  }
  std::vector<int> fq_task, fq_dest, fq_root;   // final fill tasks (below)
  bool roots_ok = true;
  {
    // Maps of the one-launch fill: which cells a stage writes, which stage-1
    // destinations stage 2 overwrites, and the stage-1 task behind a stage-2
    // source.  A stage writing one cell twice would be a race: refused.
    std::vector<uint8_t> ghost(nc, 0), skip0((size_t)std::max(1, g->n_task[0]), 0);
    std::vector<int> dest0(nc, -1), prev1((size_t)std::max(1, g->n_task[1]), -1);
    std::vector<uint8_t> seen1(nc, 0);
    for (int t = 0; t < g->n_task[0]; ++t) {
      const int64_t d = g->task_dest[0][t];
      if (dest0[d] >= 0) {
        fail("jr_ctx_create: a ghost-fill stage writes one cell twice");
        return bail(0);
      }
      dest0[d] = t;
      ghost[d] = 1;
    }
    for (int t = 0; t < g->n_task[1]; ++t) {
      const int64_t d = g->task_dest[1][t];
      if (seen1[d]) {
        fail("jr_ctx_create: a ghost-fill stage writes one cell twice");
        return bail(0);
      }
      seen1[d] = 1;
      ghost[d] = 1;
      if (dest0[d] >= 0) skip0[dest0[d]] = 1;
      prev1[t] = dest0[g->task_src[1][t]];
    }
    if (upload(c, &P.ghost, ghost.data(), nc) || upload(c, &P.t_skip0, skip0.data(), skip0.size()) ||
        upload(c, &P.t_prev1, prev1.data(), prev1.size()))
      return bail(0);
    // Final fill tasks (the last writer of each ghost cell) and their root
    // cells: a stage-1 task's source, a stage-2 task's source, or -- when that
    // is a stage-1 ghost -- the source of that stage-1 task.  Every root must
    // be a non-ghost cell for the linearised fill; otherwise it stays off.
    for (int t = 0; t < g->n_task[0]; ++t) {
      if (skip0[t]) continue;
      fq_task.push_back(t);
      fq_dest.push_back((int)g->task_dest[0][t]);
      fq_root.push_back((int)g->task_src[0][t]);
    }
    for (int t = 0; t < g->n_task[1]; ++t) {
      fq_task.push_back(g->n_task[0] + t);
      fq_dest.push_back((int)g->task_dest[1][t]);
      fq_root.push_back(prev1[t] >= 0 ? (int)g->task_src[0][prev1[t]] : (int)g->task_src[1][t]);
    }
    for (int r : fq_root)
      if (ghost[r]) roots_ok = false;
  }
  P.n_lines = g->n_lines;
  P.n_slots = g->n_slots;
  int max_len = 0;
  std::vector<uint8_t> on_chain(nc, 0);
  for (int l = 0; l < g->n_lines; ++l) {
    int64_t a = g->line_start[l], b = g->line_start[l + 1];
    if (a < 0 || b < a || b > g->n_slots) {
      fail("jr_ctx_create: line starts out of range");
      return bail(0);
    }
    max_len = std::max<int>(max_len, (int)(b - a));
    for (int64_t s = a; s < b; ++s) {
      int64_t cell = g->line_cells[s];
      if (cell < 0 || cell >= (int64_t)nc || on_chain[cell]) {
        fail("jr_ctx_create: chain cell out of range or repeated");
        return bail(0);
      }
      on_chain[cell] = 1;
      bool up = g->line_up_sign[s] != 0, dn = g->line_down_sign[s] != 0;
      if (up != dn || (up && (s + 1 >= b || g->line_up_face[s] < 0 || g->line_up_face[s] >= P.nfj ||
                              g->line_down_face[s] < 0 || g->line_down_face[s] >= P.nfj))) {
        fail("jr_ctx_create: chain link out of range");
        return bail(0);
      }
    }
    if (b > a && g->line_up_sign[b - 1] != 0) {
      fail("jr_ctx_create: the last slot of a chain must have no link");
      return bail(0);
    }
  }
  P.max_len = max_len;
  {
    // Linearised-fill dependents: a final task rooted on a chain goes to that
    // chain's list (with the root's row in the chain: the apply's workgroup
    // has the root's tangent); one rooted on no chain to its root cell's list.
    const int nf = (int)fq_task.size();
    std::vector<int> chain_of(nc, -1), row_of(nc, -1);
    for (int l = 0; l < g->n_lines; ++l)
      for (int64_t sl = g->line_start[l]; sl < g->line_start[l + 1]; ++sl) {
        chain_of[g->line_cells[sl]] = l;
        row_of[g->line_cells[sl]] = (int)(sl - g->line_start[l]);
      }
    std::vector<int> cq_start(g->n_lines + 1, 0), cq_q(std::max(1, nf)), gq_row(std::max(1, nf), -1);
    std::vector<int> dep_start(nc + 1, 0), dep_q(std::max(1, nf));
    for (int q = 0; q < nf; ++q) {
      const int r = fq_root[q];
      if (chain_of[r] >= 0) {
        cq_start[chain_of[r] + 1] += 1;
        gq_row[q] = row_of[r];
      } else {
        dep_start[r + 1] += 1;
      }
    }
    int max_dep = 0;
    for (int l = 0; l < g->n_lines; ++l) {
      max_dep = std::max(max_dep, cq_start[l + 1]);
      cq_start[l + 1] += cq_start[l];
    }
    for (size_t i = 0; i < nc; ++i) dep_start[i + 1] += dep_start[i];
    {
      std::vector<int> fc(cq_start.begin(), cq_start.end() - 1), fd(dep_start.begin(), dep_start.end() - 1);
      for (int q = 0; q < nf; ++q) {
        const int r = fq_root[q];
        if (chain_of[r] >= 0) cq_q[fc[chain_of[r]]++] = q;
        else dep_q[fd[r]++] = q;
      }
    }
    P.n_final = roots_ok ? nf : 0;
    P.max_dep = max_dep;
    if (nf == 0) {
      fq_task.push_back(0);
      fq_dest.push_back(0);
      fq_root.push_back(0);
    }
    if (upload(c, &P.gq_task, fq_task.data(), fq_task.size()) ||
        upload(c, &P.gq_dest, fq_dest.data(), fq_dest.size()) ||
        upload(c, &P.gq_root, fq_root.data(), fq_root.size()) ||
        upload(c, &P.gq_row, gq_row.data(), gq_row.size()) ||
        upload(c, &P.cq_start, cq_start.data(), cq_start.size()) ||
        upload(c, &P.cq_q, cq_q.data(), cq_q.size()) ||
        upload(c, &P.gdep_start, dep_start.data(), dep_start.size()) ||
        upload(c, &P.gdep_q, dep_q.data(), dep_q.size()))
      return bail(0);
  }
  // the longest chain k_line_apply holds in 64 KiB of LDS, one per wavefront
  // rows of one k_line_apply wavefront: whole multiples of 16 within 64 KiB
  const int line_rows_cap = (int)((64 * 1024 - line_apply_lds<5>(0)) / line_rows_lds<5>(16)) * 16;
  if ((max_len + 15) / 16 * 16 > line_rows_cap) {
    fail("jr_ctx_create: a line chain longer than " + std::to_string(line_rows_cap) + " cells is not supported");
    return bail(0);
  }
  std::vector<int> freec;
  for (size_t i = 0; i < nc; ++i)
    if (!on_chain[i]) freec.push_back((int)i);
  P.n_free = (int)freec.size();
  if (upload(c, &P.l_start, g->line_start, (size_t)g->n_lines + 1) ||
      upload(c, &P.l_cell, g->line_cells, g->n_slots) || upload(c, &P.l_up, g->line_up_face, g->n_slots) ||
      upload(c, &P.l_ups, g->line_up_sign, g->n_slots) || upload(c, &P.l_dn, g->line_down_face, g->n_slots) ||
      upload(c, &P.l_dns, g->line_down_sign, g->n_slots) || upload(c, &P.free_cell, freec.data(), freec.size()))
    return bail(0);
  P.n_walls = g->n_walls;
  for (int w = 0; w < g->n_walls; ++w) {
    int64_t fam = g->wall_family[w], f = g->wall_face[w], cell = g->wall_cell[w];
    if ((fam != 0 && fam != 1) || f < 0 || f >= (fam == 0 ? P.nfi : P.nfj) || cell < 0 || cell >= (int64_t)nc) {
      fail("jr_ctx_create: wall face out of range");
      return bail(0);
    }
  }
  if (upload(c, &P.w_cell, g->wall_cell, g->n_walls) || upload(c, &P.w_fam, g->wall_family, g->n_walls) ||
      upload(c, &P.w_face, g->wall_face, g->n_walls) || upload(c, &P.w_into, g->wall_into, g->n_walls) ||
      upload(c, &P.w_mid, g->wall_mid, 2 * (size_t)g->n_walls))
    return bail(0);

  {
    // the viscous faces' diamond geometry, once per grid
    double* dg = nullptr;
    if (dalloc(c, &dg, num->viscous ? 4 * ((size_t)P.nfi + P.nfj) : 1)) return bail(0);
    P.dg = dg;
    if (num->viscous) {
      hipLaunchKernelGGL(k_diamond, dim3((unsigned)((P.nfi + P.nfj + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, 0, P, 0,
                         dg);
      if (hipDeviceSynchronize() != hipSuccess) {
        fail("jr_ctx_create: the diamond geometry kernel failed");
        return bail(0);
      }
    }
  }
  Physics& ph = c->ph;
  ph.gamma = num->gamma;
  ph.kappa = num->kappa;
  ph.flux = num->flux;
  ph.limiter = num->limiter;
  ph.nv = num->sa ? 5 : 4;
  ph.viscous = num->viscous;
  ph.sa = num->sa;
  ph.sutherland = num->sutherland;
  ph.c_cr1 = num->c_cr1;
  c->nv = ph.nv;
  c->n_cases = n_cases;
  c->m = num->krylov_subspace;
  c->restarts = num->krylov_restarts;
  c->exact = num->jacobian_exact;
  c->mixed = num->mixed;
  // cyclic reduction takes chains of up to CR_MAX_LEN cells (one thread per
  // row); longer chains are solved by block Thomas, the same exact solve
  c->thomas = num->line_solver == 1 || P.max_len > CR_MAX_LEN;
  c->cr_threads = std::max(64, (P.max_len + 63) / 64 * 64);
  if (c->thomas) {
    // length classes: lengths in (lmax / 2, lmax], lmax a multiple of 16; one
    // wavefront's LDS holds `rows` rows, four chains of the most populous
    // class or one longest chain, whichever is more, within 64 KiB; each class
    // takes rows / lmax chains per wavefront, at most four
    std::vector<int> ord(g->n_lines);
    for (int l = 0; l < g->n_lines; ++l) ord[l] = l;
    auto len = [&](int l) { return (int)(g->line_start[l + 1] - g->line_start[l]); };
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return len(a) > len(b); });
    LineLaunch& LL = c->lines;
    LL.n = 0;
    for (int i = 0; i < g->n_lines;) {
      const int lmax = std::max(LN_CH, (len(ord[i]) + LN_CH - 1) / LN_CH * LN_CH);
      int j = i;
      while (j < g->n_lines && (lmax <= LN_CH || 2 * len(ord[j]) > lmax)) ++j;
      if (LL.n == LN_MAX_CLASSES) {
        fail("jr_ctx_create: too many chain length classes");
        return bail(0);
      }
      LL.first[LL.n] = i;
      LL.count[LL.n] = j - i;
      LL.lmax[LL.n] = lmax;
      ++LL.n;
      i = j;
    }
    const int cap = line_rows_cap;
    int most = 0;
    for (int k = 1; k < LL.n; ++k)
      if (LL.count[k] > LL.count[most]) most = k;
    LL.rows = LL.n ? std::min(cap, std::max(LL.lmax[0], 4 * LL.lmax[most])) : LN_CH;
    LL.wave0[0] = 0;
    for (int k = 0; k < LL.n; ++k) {
      LL.gper[k] = std::max(1, std::min(4, LL.rows / LL.lmax[k]));
      LL.wave0[k + 1] = LL.wave0[k] + (LL.count[k] + LL.gper[k] - 1) / LL.gper[k];
    }
    if (ord.empty()) ord.push_back(0);
    if (upload(c, &c->lorder, ord.data(), ord.size())) return bail(0);
  }
  {
    const size_t need = c->nv == 5 ? cr_apply_lds<double, 5>(P.max_len, P.max_dep, true)
                                   : cr_apply_lds<double, 4>(P.max_len, P.max_dep, true);
    // the point-vortex ghosts also move with the circulation, a sum over all
    // walls that no chain's apply can form: the tangent then runs the fill
    c->dq_fused = P.n_final > 0 && !P.vortex && (c->thomas || need <= 60 * 1024);
  }
  c->ktol = num->krylov_tol;
  c->max_rel = num->max_relative_update;
  c->ramp.cfl_growth = num->cfl_growth;
  c->ramp.cfl_cutback = num->cfl_cutback;
  c->ramp.cfl_min = num->cfl_min;
  c->ramp.cfl_max = num->cfl_max;
  c->ramp.krylov_success = num->krylov_success;
  c->N = (size_t)c->nv * nc;
  c->nblk = (int)std::min<size_t>(CG_BLK, std::max<size_t>(1, (c->N + BLOCK * 2 - 1) / (BLOCK * 2)));

  const size_t F = (size_t)n_cases * c->N;
  const size_t nf = (size_t)P.nfi + P.nfj;
  if (dalloc(c, &c->U, F) || dalloc(c, &c->R, F) || dalloc(c, &c->Q, F) || dalloc(c, &c->dQ, F) ||
      dalloc(c, &c->Fi, (size_t)n_cases * c->nv * P.nfi) || dalloc(c, &c->Fj, (size_t)n_cases * c->nv * P.nfj) ||
      dalloc(c, &c->Nd, (size_t)n_cases * 4 * (P.NI + 1) * (P.NJ + 1)) ||
      dalloc(c, &c->salin, (size_t)n_cases * (num->sa ? SALIN2 : 1) * nc) ||
      dalloc(c, &c->dNd, (size_t)n_cases * 4 * (P.NI + 1) * (P.NJ + 1)) ||
      dalloc(c, &c->X, F) || dalloc(c, &c->Act, F) || dalloc(c, &c->Rv, F) || dalloc(c, &c->Z, F) ||
      dalloc(c, &c->W, F) || dalloc(c, &c->T, F) || dalloc(c, &c->V, F * (c->m + 1)) || dalloc(c, &c->tmpA, F) ||
      dalloc(c, &c->tmpB, F) || dalloc(c, &c->lam, (size_t)n_cases * nf) || dalloc(c, &c->sam, (size_t)n_cases * nf) ||
      dalloc(c, &c->sad, (size_t)n_cases * nf) || dalloc(c, &c->lsum, (size_t)n_cases * nc) ||
      dalloc(c, &c->vdt, (size_t)n_cases * nc) || dalloc(c, &c->sa_diag, (size_t)n_cases * nc))
    return bail(0);
  const size_t rsz = c->mixed ? sizeof(float) : sizeof(double);
  {
    void* p;
    // per chain slot: factor CR_FAC (+CR_FAC_SA) values and the cyclic-reduction
    // workspace 96 (+6 SA); block Thomas keeps 48 (+3) and needs no workspace
    const size_t slots = (size_t)n_cases * P.n_slots;
    const bool cr = !c->thomas;
    if (dalloc(c, (char**)&p, slots * (cr ? CR_FAC : 48) * rsz)) return bail(0);
    c->fac = p;
    if (dalloc(c, (char**)&p, slots * (cr ? CR_FAC_SA : 3) * rsz)) return bail(0);
    c->fac_sa = p;
    if (dalloc(c, (char**)&p, (cr ? slots * 96 : 1) * rsz)) return bail(0);
    c->ws = p;
    if (dalloc(c, (char**)&p, (cr ? slots * 6 : 1) * rsz)) return bail(0);
    c->ws_sa = p;
    if (dalloc(c, &c->blk, (size_t)n_cases * P.n_slots * BLKW)) return bail(0);
  }
  if (dalloc(c, &c->ctl, (size_t)n_cases) || dalloc(c, &c->hv, (size_t)n_cases * (MMAX + 1)) ||
      dalloc(c, &c->h2v, (size_t)n_cases * (MMAX + 1)) || dalloc(c, &c->H, (size_t)n_cases * (MMAX + 1) * MMAX) ||
      dalloc(c, &c->G, (size_t)n_cases * GIV) || dalloc(c, &c->y, (size_t)n_cases * (MMAX + 1)) ||
      dalloc(c, &c->hy, (size_t)n_cases * (MMAX + 1)) || dalloc(c, &c->coef, (size_t)n_cases * 8))
    return bail(0);
  if (dalloc(c, &c->gjac, (size_t)n_cases * std::max(1, P.n_final) * c->nv * c->nv)) return bail(0);
  c->cg.nblk = c->nblk;
  if (dalloc(c, &c->cg.p1, (size_t)n_cases * MMAX * c->nblk) ||
      dalloc(c, &c->cg.p2, (size_t)n_cases * MMAX * c->nblk) || dalloc(c, &c->cg.p3, (size_t)n_cases * c->nblk))
    return bail(0);
  c->red.stride = 1024 * (MMAX + 2);
  if (dalloc(c, &c->red.partials, (size_t)n_cases * c->red.stride) ||
      dalloc(c, &c->red.counters, (size_t)n_cases * 4))
    return bail(0);
  // Every device buffer that a launch takes exists before the first launch.
  for (const void* p : {(const void*)c->U, (const void*)c->R, (const void*)c->Q, (const void*)c->dQ,
                        (const void*)c->Fi, (const void*)c->Fj, (const void*)c->X, (const void*)c->Act,
                        (const void*)c->Rv, (const void*)c->Z, (const void*)c->W, (const void*)c->T,
                        (const void*)c->V, (const void*)c->tmpA, (const void*)c->tmpB, (const void*)c->lam,
                        (const void*)c->sam, (const void*)c->sad, (const void*)c->lsum, (const void*)c->vdt,
                        (const void*)c->sa_diag, (const void*)c->ctl, (const void*)c->hv, (const void*)c->h2v,
                        (const void*)c->G, (const void*)c->H, (const void*)c->y, (const void*)c->hy,
                        (const void*)c->coef, (const void*)c->red.partials, (const void*)c->red.counters,
                        (const void*)c->cg.p1, (const void*)c->cg.p2, (const void*)c->cg.p3,
                        (const void*)c->gjac})
    if (!p) {
      fail("jr_ctx_create: a device buffer was not allocated");
      return bail(0);
    }
  if (!c->fac || !c->fac_sa || !c->ws || !c->ws_sa || !c->blk) {
    fail("jr_ctx_create: a line-preconditioner buffer was not allocated");
    return bail(0);
  }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    fail("jr_ctx_create: stream");
    return bail(0);
  }
  if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    fail("jr_ctx_create: events");
    return bail(0);
  }
  // Control blocks: freestream unset, CFL 1.
  std::vector<Ctl> h(n_cases);
  for (auto& x : h) {
    std::memset(&x, 0, sizeof(Ctl));
    x.q_inf[0] = 1.0;
    x.q_inf[3] = 1.0 / ph.gamma;
    x.cfl = 1.0;
    x.res_prev = INFINITY;
  }
  if (hipMemcpy(c->ctl, h.data(), sizeof(Ctl) * n_cases, hipMemcpyHostToDevice) != hipSuccess) {
    fail("jr_ctx_create: control upload");
    return bail(0);
  }
  // every zero fill and upload (null stream) is complete before the first
  // launch on the context's non-blocking stream
  if (hipDeviceSynchronize() != hipSuccess) {
    fail("jr_ctx_create: synchronize");
    return bail(0);
  }
  *out = c;
  return 0;
}

int jr_ctx_create3(const jr_grid_desc3* g, const jr_numerics* num, int n_cases, jr_ctx** out) {
  if (!g || !num || !out) return fail("jr_ctx_create3: null argument");
  if (n_cases < 1) return fail("jr_ctx_create3: n_cases must be >= 1");
  if (g->NI < 5 || g->NJ < 5 || g->NK < 5) return fail("jr_ctx_create3: canvas too small");
  if ((int64_t)g->NI * g->NJ * g->NK > (int64_t)1 << 30) return fail("jr_ctx_create3: canvas too large");
  if (num->krylov_subspace < 1 || num->krylov_subspace > MMAX)
    return fail("jr_ctx_create3: krylov_subspace must lie in [1, " + std::to_string(MMAX) + "]");
  if (num->krylov_restarts < 1) return fail("jr_ctx_create3: krylov_restarts must be >= 1");
  if (num->flux < 0 || num->flux > 1) return fail("jr_ctx_create3: unknown flux");
  if (num->limiter < 0 || num->limiter > 3) return fail("jr_ctx_create3: unknown limiter");
  if (num->sa && !num->viscous) return fail("jr_ctx_create3: SA needs a viscous flow");
  if (num->far_field != 0) return fail("jr_ctx_create3: the point-vortex far field is defined in 2D only");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail("jr_ctx_create3: no HIP device");
  jr_ctx* c = new jr_ctx();
  c->dim = 3;
  (void)hipGetDevice(&c->device);
  auto bail = [&](int) {
    jr_ctx_destroy(c);
    return -1;
  };
  Plan3& P = c->P3;
  P.NI = g->NI; P.NJ = g->NJ; P.NK = g->NK;
  const size_t nc = (size_t)g->NI * g->NJ * g->NK;
  P.nc = (int)nc;
  P.nf[0] = (g->NI + 1) * g->NJ * g->NK;
  P.nf[1] = g->NI * (g->NJ + 1) * g->NK;
  P.nf[2] = g->NI * g->NJ * (g->NK + 1);
  P.nft = P.nf[0] + P.nf[1] + P.nf[2];
  const size_t npc = (size_t)(g->NI + 2) * (g->NJ + 2) * (g->NK + 2);
  if (upload(c, &P.vol, g->volume, nc) || upload(c, &P.S[0], g->face[0], 3 * (size_t)P.nf[0]) ||
      upload(c, &P.S[1], g->face[1], 3 * (size_t)P.nf[1]) || upload(c, &P.S[2], g->face[2], 3 * (size_t)P.nf[2]) ||
      upload(c, &P.cen, g->centres, 3 * npc) || upload(c, &P.dist, g->distance, nc) ||
      upload(c, &P.interior, g->interior, nc) || upload(c, &P.nbr, g->neighbours, nc))
    return bail(0);
  for (size_t i = 0; i < nc; ++i)
    if (g->neighbours[i] < 0 || g->neighbours[i] >= (int64_t)nc) {
      fail("jr_ctx_create3: neighbour index out of range");
      return bail(0);
    }
  // The composed fill: every ghost written once, from an interior root,
  // through valid tasks.
  std::vector<uint8_t> ghost(nc, 0);
  std::vector<int> ops((size_t)std::max(1, g->n_ghost) * MAX_OPS, -1);
  for (int t = 0; t < g->n_tasks; ++t)
    if (g->task_kind[t] < 0 || g->task_kind[t] > 6) {
      fail("jr_ctx_create3: unknown ghost-fill task kind");
      return bail(0);
    }
  for (int q = 0; q < g->n_ghost; ++q) {
    const int64_t d = g->ghost_dest[q], r = g->ghost_root[q], n = g->ghost_nops[q];
    if (d < 0 || d >= (int64_t)nc || r < 0 || r >= (int64_t)nc || !g->interior[r] || g->interior[d] || ghost[d] ||
        n < 0 || n > MAX_OPS) {
      fail("jr_ctx_create3: composed ghost fill out of range");
      return bail(0);
    }
    ghost[d] = 1;
    for (int o = 0; o < n; ++o) {
      const int64_t t = g->ghost_ops[(size_t)q * MAX_OPS + o];
      if (t < 0 || t >= g->n_tasks) {
        fail("jr_ctx_create3: ghost-fill task out of range");
        return bail(0);
      }
      ops[(size_t)q * MAX_OPS + o] = (int)t;
    }
  }
  P.n_ghost = g->n_ghost;
  P.n_tasks = g->n_tasks;
  if (upload(c, &P.ghost, ghost.data(), nc) || upload(c, &P.g_dest, g->ghost_dest, std::max(1, g->n_ghost)) ||
      upload(c, &P.g_root, g->ghost_root, std::max(1, g->n_ghost)) ||
      upload(c, &P.g_nops, g->ghost_nops, std::max(1, g->n_ghost)) || upload(c, &P.g_ops, ops.data(), ops.size()) ||
      upload(c, &P.t_kind, g->task_kind, std::max(1, g->n_tasks)) ||
      upload(c, &P.t_norm, g->task_normal, 3 * (size_t)std::max(1, g->n_tasks)))
    return bail(0);
  // Chains.
  P.n_lines = g->n_lines;
  P.n_slots = g->n_slots;
  int max_len = 0;
  std::vector<uint8_t> on_chain(nc, 0);
  for (int l = 0; l < g->n_lines; ++l) {
    const int64_t a = g->line_start[l], b = g->line_start[l + 1];
    if (a < 0 || b < a || b > g->n_slots) {
      fail("jr_ctx_create3: line starts out of range");
      return bail(0);
    }
    max_len = std::max<int>(max_len, (int)(b - a));
    for (int64_t sl = a; sl < b; ++sl) {
      const int64_t cell = g->line_cells[sl];
      if (cell < 0 || cell >= (int64_t)nc || on_chain[cell]) {
        fail("jr_ctx_create3: chain cell out of range or repeated");
        return bail(0);
      }
      on_chain[cell] = 1;
      const bool up = g->line_up_sign[sl] != 0, dn = g->line_down_sign[sl] != 0;
      if (up != dn || (up && (sl + 1 >= b || g->line_up_face[sl] < 0 || g->line_up_face[sl] >= P.nft ||
                              g->line_down_face[sl] < 0 || g->line_down_face[sl] >= P.nft))) {
        fail("jr_ctx_create3: chain link out of range");
        return bail(0);
      }
    }
    if (b > a && g->line_up_sign[b - 1] != 0) {
      fail("jr_ctx_create3: the last slot of a chain must have no link");
      return bail(0);
    }
  }
  P.max_len = max_len;
  std::vector<int> freec;
  for (size_t i = 0; i < nc; ++i)
    if (!on_chain[i]) freec.push_back((int)i);
  P.n_free = (int)freec.size();
  if (upload(c, &P.l_start, g->line_start, (size_t)g->n_lines + 1) ||
      upload(c, &P.l_cell, g->line_cells, std::max(1, g->n_slots)) ||
      upload(c, &P.l_up, g->line_up_face, std::max(1, g->n_slots)) ||
      upload(c, &P.l_ups, g->line_up_sign, std::max(1, g->n_slots)) ||
      upload(c, &P.l_dn, g->line_down_face, std::max(1, g->n_slots)) ||
      upload(c, &P.l_dns, g->line_down_sign, std::max(1, g->n_slots)) ||
      upload(c, &P.free_cell, freec.data(), std::max<size_t>(1, freec.size())))
    return bail(0);
  P.n_walls = g->n_walls;
  for (int w = 0; w < g->n_walls; ++w) {
    const int64_t fam = g->wall_family[w], f = g->wall_face[w], cell = g->wall_cell[w];
    if (fam < 0 || fam > 2 || f < 0 || f >= P.nf[fam] || cell < 0 || cell >= (int64_t)nc || !g->interior[cell]) {
      fail("jr_ctx_create3: wall face out of range");
      return bail(0);
    }
  }
  if (upload(c, &P.w_cell, g->wall_cell, std::max(1, g->n_walls)) ||
      upload(c, &P.w_fam, g->wall_family, std::max(1, g->n_walls)) ||
      upload(c, &P.w_face, g->wall_face, std::max(1, g->n_walls)) ||
      upload(c, &P.w_into, g->wall_into, std::max(1, g->n_walls)) ||
      upload(c, &P.w_mid, g->wall_mid, 3 * (size_t)std::max(1, g->n_walls)))
    return bail(0);
  // The shared vector kernels read the cell count, the interior mask and the
  // cell volumes from the 2D plan.
  c->P.nc = P.nc;
  c->P.interior = P.interior;
  c->P.area = P.vol;

  Physics& ph = c->ph;
  ph.gamma = num->gamma;
  ph.kappa = num->kappa;
  ph.flux = num->flux;
  ph.limiter = num->limiter;
  ph.nv = num->sa ? 6 : 5;
  ph.viscous = num->viscous;
  ph.sa = num->sa;
  ph.sutherland = num->sutherland;
  ph.c_cr1 = num->c_cr1;
  c->nv = ph.nv;
  c->n_cases = n_cases;
  c->m = num->krylov_subspace;
  c->restarts = num->krylov_restarts;
  c->exact = num->jacobian_exact;
  c->mixed = num->mixed;
  c->thomas = 1;
  c->dq_fused = false;
  c->ktol = num->krylov_tol;
  c->max_rel = num->max_relative_update;
  c->ramp.cfl_growth = num->cfl_growth;
  c->ramp.cfl_cutback = num->cfl_cutback;
  c->ramp.cfl_min = num->cfl_min;
  c->ramp.cfl_max = num->cfl_max;
  c->ramp.krylov_success = num->krylov_success;
  c->N = (size_t)c->nv * nc;
  c->nblk = (int)std::min<size_t>(CG_BLK, std::max<size_t>(1, (c->N + BLOCK * 2 - 1) / (BLOCK * 2)));
  const size_t F = (size_t)n_cases * c->N;
  const size_t nf = (size_t)P.nft;
  if (dalloc(c, &c->U, F) || dalloc(c, &c->R, F) || dalloc(c, &c->Q, F) || dalloc(c, &c->dQ, F) ||
      dalloc(c, &c->F3, (size_t)n_cases * c->nv * nf) || dalloc(c, &c->X, F) || dalloc(c, &c->Act, F) ||
      dalloc(c, &c->Rv, F) || dalloc(c, &c->Z, F) || dalloc(c, &c->W, F) || dalloc(c, &c->T, F) ||
      dalloc(c, &c->V, F * (c->m + 1)) || dalloc(c, &c->tmpA, F) || dalloc(c, &c->tmpB, F) ||
      dalloc(c, &c->dsw, F) || dalloc(c, &c->salin, (size_t)n_cases * (num->sa ? SALIN : 1) * nc) ||
      dalloc(c, &c->Nd, (size_t)n_cases * 5 * (g->NI + 1) * (g->NJ + 1) * (g->NK + 1)) ||
      dalloc(c, &c->dNd, (size_t)n_cases * 5 * (g->NI + 1) * (g->NJ + 1) * (g->NK + 1)) ||
      dalloc(c, &c->lam, (size_t)n_cases * nf) || dalloc(c, &c->sam, (size_t)n_cases * nf) ||
      dalloc(c, &c->sad, (size_t)n_cases * nf) || dalloc(c, &c->lsum, (size_t)n_cases * nc) ||
      dalloc(c, &c->vdt, (size_t)n_cases * nc) || dalloc(c, &c->sa_diag, (size_t)n_cases * nc))
    return bail(0);
  c->Fi = c->Fj = c->F3;
  {
    const size_t rsz = c->mixed ? sizeof(float) : sizeof(double);
    // factor tiles of 64 chains (k_thomas3, k_apply3): chains rounded up to whole tiles
    const size_t rows = (size_t)n_cases * std::max(1, P.max_len) * (((size_t)std::max(1, P.n_lines) + 63) / 64 * 64);
    void* p;
    if (dalloc(c, (char**)&p, rows * FAC3_VALUES * rsz)) return bail(0);
    c->fac = p;
    if (dalloc(c, (char**)&p, rows * 3 * rsz)) return bail(0);
    c->fac_sa = p;
    c->ws = c->ws_sa = p;
    // chain blocks component-major, chain-minor: [case][BLKW3][max_len][n_lines]
    if (dalloc(c, &c->blk, (size_t)n_cases * std::max<size_t>(1, (size_t)P.max_len * P.n_lines) * BLKW3))
      return bail(0);
  }
  if (dalloc(c, &c->ctl, (size_t)n_cases) || dalloc(c, &c->hv, (size_t)n_cases * (MMAX + 1)) ||
      dalloc(c, &c->h2v, (size_t)n_cases * (MMAX + 1)) || dalloc(c, &c->H, (size_t)n_cases * (MMAX + 1) * MMAX) ||
      dalloc(c, &c->G, (size_t)n_cases * GIV) || dalloc(c, &c->y, (size_t)n_cases * (MMAX + 1)) ||
      dalloc(c, &c->hy, (size_t)n_cases * (MMAX + 1)) || dalloc(c, &c->coef, (size_t)n_cases * 8) ||
      dalloc(c, &c->gjac, 1))
    return bail(0);
  c->cg.nblk = c->nblk;
  if (dalloc(c, &c->cg.p1, (size_t)n_cases * MMAX * c->nblk) ||
      dalloc(c, &c->cg.p2, (size_t)n_cases * MMAX * c->nblk) || dalloc(c, &c->cg.p3, (size_t)n_cases * c->nblk))
    return bail(0);
  c->red.stride = 1024 * (MMAX + 2);
  if (dalloc(c, &c->red.partials, (size_t)n_cases * c->red.stride) ||
      dalloc(c, &c->red.counters, (size_t)n_cases * 4))
    return bail(0);
  for (const void* p : {(const void*)c->U, (const void*)c->R, (const void*)c->Q, (const void*)c->dQ,
                        (const void*)c->F3, (const void*)c->X, (const void*)c->Act, (const void*)c->Rv,
                        (const void*)c->Z, (const void*)c->W, (const void*)c->T, (const void*)c->V,
                        (const void*)c->tmpA, (const void*)c->tmpB, (const void*)c->dsw, (const void*)c->Nd,
                        (const void*)c->dNd, (const void*)c->lam,
                        (const void*)c->sam, (const void*)c->sad, (const void*)c->lsum, (const void*)c->vdt,
                        (const void*)c->sa_diag, (const void*)c->ctl, (const void*)c->hv, (const void*)c->h2v,
                        (const void*)c->G, (const void*)c->H, (const void*)c->y, (const void*)c->hy,
                        (const void*)c->coef, (const void*)c->red.partials, (const void*)c->red.counters,
                        (const void*)c->cg.p1, (const void*)c->cg.p2, (const void*)c->cg.p3, (const void*)c->fac, (const void*)c->fac_sa,
                        (const void*)c->blk})
    if (!p) {
      fail("jr_ctx_create3: a device buffer was not allocated");
      return bail(0);
    }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    fail("jr_ctx_create3: stream");
    return bail(0);
  }
  if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    fail("jr_ctx_create3: events");
    return bail(0);
  }
  if (num->viscous) {
    // the diamonds of the viscous face gradients depend on the grid only
    double* dg = nullptr;
    if (dalloc(c, &dg, (size_t)10 * P.nft)) return bail(0);
    P.dg = dg;
    // the zero fills of dalloc run on the null stream, which the context's
    // non-blocking stream does not wait for
    if (hipDeviceSynchronize() != hipSuccess) {
      fail("jr_ctx_create3: synchronize");
      return bail(0);
    }
    hipLaunchKernelGGL(k_diamond3, dim3((P.nft + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, c->stream, P, dg);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess) {
      fail("jr_ctx_create3: face diamonds");
      return bail(0);
    }
  }
  std::vector<Ctl> h(n_cases);
  for (auto& x : h) {
    std::memset(&x, 0, sizeof(Ctl));
    x.q_inf[0] = 1.0;
    x.q_inf[4] = 1.0 / ph.gamma;
    x.cfl = 1.0;
    x.res_prev = INFINITY;
  }
  if (hipMemcpy(c->ctl, h.data(), sizeof(Ctl) * n_cases, hipMemcpyHostToDevice) != hipSuccess) {
    fail("jr_ctx_create3: control upload");
    return bail(0);
  }
  // every zero fill and upload (null stream) is complete before the first
  // launch on the context's non-blocking stream
  if (hipDeviceSynchronize() != hipSuccess) {
    fail("jr_ctx_create: synchronize");
    return bail(0);
  }
  *out = c;
  return 0;
}

int jr_ctx_destroy(jr_ctx* c) {
  if (!c) return 0;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (void* p : c->allocs) (void)hipFree(p);
  if (c->hist) (void)hipFree(c->hist);
  if (c->graph) (void)hipGraphExecDestroy(c->graph);
  for (hipEvent_t e : c->probe_ev) (void)hipEventDestroy(e);
  if (c->hist_pinned) (void)hipHostFree(c->hist_pinned);
  if (c->ev_block) (void)hipEventDestroy(c->ev_block);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

static void drop_graph(jr_ctx* c) {
  if (c->graph) (void)hipGraphExecDestroy(c->graph);
  c->graph = nullptr;
  c->graph_steps = 0;
}

static int get_ctl(jr_ctx* c, std::vector<Ctl>& h) {
  h.resize(c->n_cases);
  HIPCHK(hipMemcpyAsync(h.data(), c->ctl, sizeof(Ctl) * c->n_cases, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

static int put_ctl(jr_ctx* c, const std::vector<Ctl>& h) {
  HIPCHK(hipMemcpyAsync(c->ctl, h.data(), sizeof(Ctl) * c->n_cases, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int jr_set_freestream(jr_ctx* c, const double* q_inf, const double* mu_inf) {
  if (!c || !q_inf) return fail("jr_set_freestream: null argument");
  if (c->ph.viscous && !mu_inf) return fail("jr_set_freestream: a viscous context needs mu_inf");
  std::vector<Ctl> h;
  TRY(get_ctl(c, h));
  const int nq = c->dim == 3 ? 6 : 5, ip = c->dim == 3 ? 4 : 3;
  for (int k = 0; k < c->n_cases; ++k) {
    for (int v = 0; v < nq; ++v) h[k].q_inf[v] = q_inf[k * nq + v];
    h[k].mu_inf = mu_inf ? mu_inf[k] : 0.0;
    if (!(h[k].q_inf[0] > 0.0) || !(h[k].q_inf[ip] > 0.0)) return fail("jr_set_freestream: non-physical freestream");
  }
  c->has_q = true;
  return put_ctl(c, h);
}

int jr_set_state(jr_ctx* c, const double* U) {
  if (!c || !U) return fail("jr_set_state: null argument");
  HIPCHK(hipMemcpyAsync(c->U, U, field_bytes(c), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int jr_set_state_case(jr_ctx* c, int k, const double* U) {
  if (!c || !U) return fail("jr_set_state_case: null argument");
  if (k < 0 || k >= c->n_cases) return fail("jr_set_state_case: case out of range");
  HIPCHK(hipMemcpyAsync(c->U + (size_t)k * c->N, U, c->N * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

// Geometry of case k from a grid of the context's topology.  The first call
// gives every case its own copy of the geometry arrays (filled with the
// context's grid), so the other cases keep theirs.
int jr_set_geometry(jr_ctx* c, int k, const jr_grid_desc* g) {
  if (!c || !g) return fail("jr_set_geometry: null argument");
  if (c->dim != 2) return fail("jr_set_geometry: 2D contexts only");
  if (k < 0 || k >= c->n_cases) return fail("jr_set_geometry: case out of range");
  if (g->NI != c->P.NI || g->NJ != c->P.NJ || topology_hash(g) != c->topo_hash)
    return fail("jr_set_geometry: the grid's index plans differ from the context's (another topology)");
  if (c->pending_steps) return fail("jr_set_geometry: an asynchronous block is pending");
  if (c->P.vortex) return fail("jr_set_geometry: per-case geometry with the point-vortex far field is not built");
  Plan& P = c->P;
  const size_t nc = P.nc;
  const size_t n_area = nc, n_fi = 2 * (size_t)P.nfi, n_fj = 2 * (size_t)P.nfj,
               n_cen = 2 * (size_t)(P.NI + 2) * (P.NJ + 2), n_tn0 = 2 * (size_t)std::max(1, P.n_task[0]),
               n_tn1 = 2 * (size_t)std::max(1, P.n_task[1]), n_wm = 2 * (size_t)std::max(1, P.n_walls);
  HIPCHK(hipStreamSynchronize(c->stream));
  if (!c->per_case_geo) {
    // one copy per case, each starting as the shared geometry
    auto widen = [&](const double** field, size_t count, size_t* stride) -> int {
      double* d = nullptr;
      if (dalloc(c, &d, count * c->n_cases)) return -1;
      for (int q = 0; q < c->n_cases; ++q)
        HIPCHK(hipMemcpy(d + q * count, *field, count * sizeof(double), hipMemcpyDeviceToDevice));
      *field = d;
      *stride = count;
      return 0;
    };
    if (widen(&P.area, n_area, &P.gs_area) || widen(&P.fi, n_fi, &P.gs_fi) || widen(&P.fj, n_fj, &P.gs_fj) ||
        widen(&P.cen, n_cen, &P.gs_cen) || widen(&P.dist, n_area, &P.gs_dist) ||
        (c->ph.viscous && widen(&P.dg, 4 * ((size_t)P.nfi + P.nfj), &P.gs_dg)) ||
        widen(&P.t_norm[0], n_tn0, &P.gs_tn0) || widen(&P.t_norm[1], n_tn1, &P.gs_tn1) ||
        widen(&P.w_mid, n_wm, &P.gs_wmid))
      return -1;
    c->per_case_geo = true;
    drop_graph(c);
  }
  std::vector<double> dist(nc);
  for (size_t i = 0; i < nc; ++i) dist[i] = g->distance[i];
  auto put = [&](const double* field, size_t stride, const double* src, size_t count) -> int {
    if (!src) return fail("jr_set_geometry: a geometry array is null");
    HIPCHK(hipMemcpy((double*)field + (size_t)k * stride, src, count * sizeof(double), hipMemcpyHostToDevice));
    return 0;
  };
  TRY(put(P.area, P.gs_area, g->area, n_area));
  TRY(put(P.fi, P.gs_fi, g->face_i, n_fi));
  TRY(put(P.fj, P.gs_fj, g->face_j, n_fj));
  TRY(put(P.cen, P.gs_cen, g->centres, n_cen));
  TRY(put(P.dist, P.gs_dist, dist.data(), n_area));
  if (P.n_task[0]) TRY(put(P.t_norm[0], P.gs_tn0, g->task_normal[0], 2 * (size_t)P.n_task[0]));
  if (P.n_task[1]) TRY(put(P.t_norm[1], P.gs_tn1, g->task_normal[1], 2 * (size_t)P.n_task[1]));
  if (P.n_walls) TRY(put(P.w_mid, P.gs_wmid, g->wall_mid, 2 * (size_t)P.n_walls));
  if (c->ph.viscous)   // this case's diamonds from its centres
    hipLaunchKernelGGL(k_diamond, dim3((unsigned)((P.nfi + P.nfj + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, 0, P, k,
                       (double*)P.dg + (size_t)k * P.gs_dg);
  // the fills and copies above ran on the null stream; the context's stream
  // does not wait for it
  HIPCHK(hipDeviceSynchronize());
  return 0;
}

int jr_get_state(jr_ctx* c, double* U) {
  if (!c || !U) return fail("jr_get_state: null argument");
  HIPCHK(hipMemcpyAsync(U, c->U, field_bytes(c), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int jr_residual_dev(jr_ctx* c, const double* U, double* R, int first_order) {
  if (!c || !U || !R) return fail("jr_residual_dev: null argument");
  if (!c->has_q) return fail("jr_residual_dev: set the freestream first");
  residual_dispatch(c, U, R, first_order);
  return check_launch();
}

int jr_residual(jr_ctx* c, const double* U, double* R, int first_order) {
  if (!c || !U || !R) return fail("jr_residual: null argument");
  HIPCHK(hipMemcpyAsync(c->tmpA, U, field_bytes(c), hipMemcpyHostToDevice, c->stream));
  TRY(jr_residual_dev(c, c->tmpA, c->tmpB, first_order));
  HIPCHK(hipMemcpyAsync(R, c->tmpB, field_bytes(c), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int jr_tangent_dev(jr_ctx* c, const double* U, const double* dU, double* out, int first_order, int with_vdt) {
  if (!c || !U || !dU || !out) return fail("jr_tangent_dev: null argument");
  if (!c->has_q) return fail("jr_tangent_dev: set the freestream first");
  // primal primitives with ghosts at U, then the tangent
  residual_dispatch(c, U, c->R, first_order);
  if (with_vdt) speeds_dispatch(c, U);
  tangent_dispatch(c, U, dU, out, first_order, with_vdt ? c->vdt : nullptr);
  return check_launch();
}

int jr_tangent(jr_ctx* c, const double* U, const double* dU, double* out, int first_order, int with_vdt) {
  if (!c || !U || !dU || !out) return fail("jr_tangent: null argument");
  HIPCHK(hipMemcpyAsync(c->tmpA, U, field_bytes(c), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->Z, dU, field_bytes(c), hipMemcpyHostToDevice, c->stream));
  TRY(jr_tangent_dev(c, c->tmpA, c->Z, c->tmpB, first_order, with_vdt));
  HIPCHK(hipMemcpyAsync(out, c->tmpB, field_bytes(c), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int jr_wave_speeds(jr_ctx* c, const double* U, double* lam, double* lsum) {
  if (!c || !U) return fail("jr_wave_speeds: null argument");
  if (!c->has_q) return fail("jr_wave_speeds: set the freestream first");
  HIPCHK(hipMemcpyAsync(c->tmpA, U, field_bytes(c), hipMemcpyHostToDevice, c->stream));
  speeds_dispatch(c, c->tmpA);
  TRY(check_launch());
  const size_t nf = c->dim == 3 ? (size_t)c->P3.nft : (size_t)c->P.nfi + c->P.nfj;
  if (lam) HIPCHK(hipMemcpyAsync(lam, c->lam, (size_t)c->n_cases * nf * 8, hipMemcpyDeviceToHost, c->stream));
  if (lsum) HIPCHK(hipMemcpyAsync(lsum, c->lsum, (size_t)c->n_cases * c->P.nc * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int jr_precond_factor(jr_ctx* c, const double* U) {
  if (!c || !U) return fail("jr_precond_factor: null argument");
  if (!c->has_q) return fail("jr_precond_factor: set the freestream first");
  HIPCHK(hipMemcpyAsync(c->tmpA, U, field_bytes(c), hipMemcpyHostToDevice, c->stream));
  speeds_dispatch(c, c->tmpA);
  factor_dispatch(c, c->tmpA);
  TRY(check_launch());
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int jr_precond_apply(jr_ctx* c, const double* r, double* z) {
  if (!c || !r || !z) return fail("jr_precond_apply: null argument");
  HIPCHK(hipMemcpyAsync(c->tmpA, r, field_bytes(c), hipMemcpyHostToDevice, c->stream));
  apply_dispatch(c, c->tmpA, c->tmpB);
  TRY(check_launch());
  HIPCHK(hipMemcpyAsync(z, c->tmpB, field_bytes(c), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int jr_newton_init(jr_ctx* c, const double* cfl0) {
  if (!c || !cfl0) return fail("jr_newton_init: null argument");
  std::vector<Ctl> h;
  TRY(get_ctl(c, h));
  for (int k = 0; k < c->n_cases; ++k) {
    if (!(cfl0[k] > 0.0)) return fail("jr_newton_init: cfl must be positive");
    h[k].cfl = cfl0[k];
    h[k].res_prev = INFINITY;
    h[k].steps = 0;
    h[k].rejected = 0;
    h[k].accept = 0;
    h[k].omega = 1.0;
  }
  return put_ctl(c, h);
}

int jr_newton_init_case(jr_ctx* c, int k, double cfl0) {
  if (!c) return fail("jr_newton_init_case: null context");
  if (k < 0 || k >= c->n_cases) return fail("jr_newton_init_case: case out of range");
  if (!(cfl0 > 0.0)) return fail("jr_newton_init_case: cfl must be positive");
  std::vector<Ctl> h;
  TRY(get_ctl(c, h));
  h[k].cfl = cfl0;
  h[k].res_prev = INFINITY;
  h[k].steps = 0;
  h[k].rejected = 0;
  h[k].accept = 0;
  h[k].omega = 1.0;
  return put_ctl(c, h);
}

int jr_park_case(jr_ctx* c, int k, int parked) {
  if (!c) return fail("jr_park_case: null context");
  if (k < 0 || k >= c->n_cases) return fail("jr_park_case: case out of range");
  if (c->pending_steps) return fail("jr_park_case: an asynchronous block is pending: wait for it first");
  std::vector<Ctl> h;
  TRY(get_ctl(c, h));
  h[k].parked = parked != 0;
  return put_ctl(c, h);
}

int jr_set_cfl(jr_ctx* c, const double* cfl) {
  if (!c || !cfl) return fail("jr_set_cfl: null argument");
  std::vector<Ctl> h;
  TRY(get_ctl(c, h));
  for (int k = 0; k < c->n_cases; ++k) h[k].cfl = cfl[k];
  return put_ctl(c, h);
}

int jr_set_tile_tangent(jr_ctx* c, int enable) {
  if (!c) return fail("jr_set_tile_tangent: null context");
  if (c->pending_steps) return fail("jr_set_tile_tangent: an asynchronous block is pending: wait for it first");
  HIPCHK(hipStreamSynchronize(c->stream));
  drop_graph(c);   // captured launches hold the old kernel path
  c->tile_tangent = enable != 0;
  return 0;
}

int jr_set_graphs(jr_ctx* c, int enable) {
  if (!c) return fail("jr_set_graphs: null context");
  HIPCHK(hipStreamSynchronize(c->stream));
  drop_graph(c);
  c->use_graph = enable != 0;
  return 0;
}

int jr_probe_arm(jr_ctx* c, unsigned mask, int max_launches) {
  if (!c) return fail("jr_probe_arm: null context");
  if (max_launches < 0 || max_launches > (1 << 20)) return fail("jr_probe_arm: max_launches out of range");
  if (mask && c->use_graph) return fail("jr_probe_arm: the probe needs kernel-by-kernel launches (graphs off)");
  HIPCHK(hipStreamSynchronize(c->stream));
  while ((int)c->probe_ev.size() < 2 * max_launches) {
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    c->probe_ev.push_back(e);
  }
  c->probe_id.assign(c->probe_ev.size() / 2, -1);
  c->probe_mask = mask & ((1u << PR_N) - 1);
  c->probe_used = 0;
  return 0;
}

int jr_probe_read(jr_ctx* c, double* ms_total, int* launches) {
  if (!c || !ms_total || !launches) return fail("jr_probe_read: null argument");
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int k = 0; k < PR_N; ++k) {
    ms_total[k] = 0.0;
    launches[k] = 0;
  }
  for (int k = 0; k < c->probe_used; ++k) {
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, c->probe_ev[2 * k], c->probe_ev[2 * k + 1]));
    ms_total[c->probe_id[k]] += ms;
    launches[c->probe_id[k]] += 1;
  }
  c->probe_mask = 0;
  c->probe_used = 0;
  return 0;
}

int jr_newton_block_async(jr_ctx* c, int n_steps, const double* force_ref) {
  if (!c) return fail("jr_newton_block_async: null context");
  if (n_steps < 1) return fail("jr_newton_block_async: n_steps must be >= 1");
  if (!c->has_q) return fail("jr_newton_block_async: set the freestream first");
  if (c->pending_steps) return fail("jr_newton_block_async: a block is already pending on this context");
  if (c->use_graph) return fail("jr_newton_block_async: graphs are for jr_newton_block");
  if (force_ref && c->dim == 3) return fail("jr_newton_block_async: forces of a 3D context come from jr_forces3");
  if (n_steps > c->hist_cap) {
    HIPCHK(hipStreamSynchronize(c->stream));
    drop_graph(c);
    if (c->hist) (void)hipFree(c->hist);
    c->hist = nullptr;
    HIPCHK(hipMalloc(&c->hist, (size_t)n_steps * c->n_cases * sizeof(double)));
    c->hist_cap = n_steps;
  }
  if (n_steps > c->pinned_cap) {
    if (c->hist_pinned) (void)hipHostFree(c->hist_pinned);
    c->hist_pinned = nullptr;
    HIPCHK(hipHostMalloc((void**)&c->hist_pinned, (size_t)(n_steps + 8) * c->n_cases * sizeof(double)));
    c->pinned_cap = n_steps;
  }
  if (!c->ev_block) HIPCHK(hipEventCreateWithFlags(&c->ev_block, hipEventDisableTiming));
  c->async_mode = true;
  for (int s = 0; s < n_steps; ++s) step_dispatch(c, c->hist + (size_t)s * c->n_cases);
  c->async_mode = false;
  TRY(check_launch());
  HIPCHK(hipMemcpyAsync(c->hist_pinned, c->hist, (size_t)n_steps * c->n_cases * sizeof(double),
                        hipMemcpyDeviceToHost, c->stream));
  if (force_ref) {
    forces_dispatch(c, force_ref[0], force_ref[1], force_ref[2]);
    TRY(check_launch());
    HIPCHK(hipMemcpyAsync(c->hist_pinned + (size_t)c->pinned_cap * c->n_cases, c->coef,
                          (size_t)c->n_cases * 8 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(hipEventRecord(c->ev_block, c->stream));
  c->pending_steps = n_steps;
  c->pending_forces = force_ref != nullptr;
  return 0;
}

int jr_newton_block_poll(jr_ctx* c) {
  if (!c) return fail("jr_newton_block_poll: null context");
  if (!c->pending_steps) return 1;
  const hipError_t e = hipEventQuery(c->ev_block);
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) return 0;
  return fail(std::string("jr_newton_block_poll: ") + hipGetErrorString(e));
}

int jr_newton_block_wait(jr_ctx* c, double* hist_host, double* forces_host) {
  if (!c) return fail("jr_newton_block_wait: null context");
  if (!c->pending_steps) return fail("jr_newton_block_wait: no block pending");
  if (forces_host && !c->pending_forces) return fail("jr_newton_block_wait: the block was started without forces");
  HIPCHK(hipEventSynchronize(c->ev_block));
  if (hist_host) std::memcpy(hist_host, c->hist_pinned, (size_t)c->pending_steps * c->n_cases * sizeof(double));
  if (forces_host)
    std::memcpy(forces_host, c->hist_pinned + (size_t)c->pinned_cap * c->n_cases, (size_t)c->n_cases * 8 * sizeof(double));
  c->pending_steps = 0;
  c->pending_forces = false;
  return 0;
}

int jr_ctx_bytes(jr_ctx* c, size_t* bytes) {
  if (!c || !bytes) return fail("jr_ctx_bytes: null argument");
  *bytes = c->bytes + (size_t)c->hist_cap * c->n_cases * sizeof(double);
  return 0;
}

int jr_mem_info(size_t* free_bytes, size_t* total_bytes) {
  if (!free_bytes || !total_bytes) return fail("jr_mem_info: null argument");
  HIPCHK(hipMemGetInfo(free_bytes, total_bytes));
  return 0;
}

int jr_newton_block(jr_ctx* c, int n_steps, double* hist_host) {
  if (!c) return fail("jr_newton_block: null context");
  if (c->pending_steps) return fail("jr_newton_block: an asynchronous block is pending: wait for it first");
  if (n_steps < 1) return fail("jr_newton_block: n_steps must be >= 1");
  if (!c->has_q) return fail("jr_newton_block: set the freestream first");
  if (n_steps > c->hist_cap) {
    HIPCHK(hipStreamSynchronize(c->stream));
    drop_graph(c);   // the captured launches hold the old history pointer
    if (c->hist) (void)hipFree(c->hist);
    c->hist = nullptr;
    HIPCHK(hipMalloc(&c->hist, (size_t)n_steps * c->n_cases * sizeof(double)));
    c->hist_cap = n_steps;
  }
  if (c->use_graph) {
    if (!c->graph || c->graph_steps != n_steps) {
      drop_graph(c);
      hipGraph_t g = nullptr;
      HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      for (int s = 0; s < n_steps; ++s) step_dispatch(c, c->hist + (size_t)s * c->n_cases);
      const hipError_t launched = hipGetLastError();
      HIPCHK(hipStreamEndCapture(c->stream, &g));
      if (launched != hipSuccess) {
        (void)hipGraphDestroy(g);
        return fail(std::string("jr_newton_block: launch during capture: ") + hipGetErrorString(launched));
      }
      const hipError_t made = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (made != hipSuccess) {
        c->graph = nullptr;
        return fail(std::string("jr_newton_block: graph instantiation: ") + hipGetErrorString(made));
      }
      c->graph_steps = n_steps;
    }
    HIPCHK(hipGraphLaunch(c->graph, c->stream));
  } else {
    for (int s = 0; s < n_steps; ++s) {
      step_dispatch(c, c->hist + (size_t)s * c->n_cases);
      TRY(check_launch());
    }
  }
  if (hist_host) {
    HIPCHK(hipMemcpyAsync(hist_host, c->hist, (size_t)n_steps * c->n_cases * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

// --- A block of Newton steps that allocates nothing --------------------------
// The residual history of jr_newton_block_noalloc is staged in this array
// instead of the context's lazily grown history buffer.  Every device the
// library runs on holds its own copy (512 KiB) until the process ends; host
// code reaches it only through the address hipGetSymbolAddress returns for
// that device, never through the array's name.
static __device__ __align__(256) double scratch[JR_NOALLOC_HISTORY_MAX];

static std::mutex g_scratch_mu;
static double* g_scratch_ptr[JR_NOALLOC_MAX_DEVICES] = {};

// The device address of `scratch` on `device`, looked up once per device with
// that device current (the caller's current device is restored).
static int scratch_on(int device, double** out) {
  if (device < 0 || device >= JR_NOALLOC_MAX_DEVICES)
    return fail("jr_newton_block_noalloc: device " + std::to_string(device) + " out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int current = 0;
    HIPCHK(hipGetDevice(&current));
    if (current != device) HIPCHK(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (current != device) HIPCHK(hipSetDevice(current));
    if (e != hipSuccess || !p)
      return fail(std::string("jr_newton_block_noalloc: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[device];
  return 0;
}

__global__ void __launch_bounds__(BLOCK) k_fill_zero(size_t n, double* __restrict__ p) {
  const size_t e = (size_t)blockIdx.x * BLOCK + threadIdx.x;
  if (e < n) p[e] = 0.0;
}

int jr_newton_block_noalloc(jr_ctx* c, int n_steps, double* hist_host) {
  if (!c) return fail("jr_newton_block_noalloc: null context");
  if (c->pending_steps) return fail("jr_newton_block_noalloc: an asynchronous block is pending: wait for it first");
  if (n_steps < 1) return fail("jr_newton_block_noalloc: n_steps must be >= 1");
  if (!c->has_q) return fail("jr_newton_block_noalloc: set the freestream first");
  const size_t rows = (size_t)n_steps * c->n_cases;
  if (rows > (size_t)JR_NOALLOC_HISTORY_MAX)
    return fail("jr_newton_block_noalloc: n_steps x n_cases = " + std::to_string(rows) + " exceeds the " +
                std::to_string(JR_NOALLOC_HISTORY_MAX) + " history entries of the static staging array");
  double* hist = nullptr;
  TRY(scratch_on(c->device, &hist));
  // The staging array keeps the previous call's history: clear the rows this
  // block writes (a fill kernel, not a memset, so a capture could hold it).
  hipLaunchKernelGGL(k_fill_zero, grid1(rows, BLOCK, 1), dim3(BLOCK), 0, c->stream, rows, hist);
  TRY(check_launch());
  // Every GMRES cycle is launched (a cycle no case needs exits on the device,
  // as on the asynchronous path): no host read of the control blocks, and the
  // same results as jr_newton_block's kernel-by-kernel launches.
  c->async_mode = true;
  for (int s = 0; s < n_steps; ++s) {
    step_dispatch(c, hist + (size_t)s * c->n_cases);
    if (check_launch() != 0) {
      c->async_mode = false;
      return -1;
    }
  }
  c->async_mode = false;
  if (hist_host)
    HIPCHK(hipMemcpyAsync(hist_host, hist, rows * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int jr_get_scalars(jr_ctx* c, double* out) {
  if (!c || !out) return fail("jr_get_scalars: null argument");
  std::vector<Ctl> h;
  TRY(get_ctl(c, h));
  for (int k = 0; k < c->n_cases; ++k) {
    double* o = out + k * 9;
    o[0] = h[k].cfl;
    o[1] = h[k].res_prev;
    o[2] = h[k].norm;
    o[3] = h[k].lin_red;
    o[4] = h[k].omega;
    o[5] = h[k].accept;
    o[6] = h[k].steps;
    o[7] = h[k].rejected;
    o[8] = h[k].cycles_total;
  }
  return 0;
}

// One solve of the Newton step's linear system at the context's state and
// CFL, for the right-hand side b: GMRES (the context's subspace, restarts and
// tolerance, line-preconditioned) on V/dt + dR~/dU, started from zero.  The
// state is left unchanged.  The step's residual, wave speeds and line factor
// are recomputed at the state first, as a Newton step does.
int jr_linear_solve(jr_ctx* c, const double* b, double* x) {
  if (!c || !b || !x) return fail("jr_linear_solve: null argument");
  if (c->pending_steps) return fail("jr_linear_solve: an asynchronous block is pending: wait for it first");
  if (!c->has_q) return fail("jr_linear_solve: set the freestream first");
  if (!c->solve_buf) {
    TRY(dalloc(c, &c->solve_buf, (size_t)c->n_cases * c->N));
    HIPCHK(hipDeviceSynchronize());   // dalloc clears on the null stream
  }
  HIPCHK(hipMemcpyAsync(c->solve_buf, b, field_bytes(c), hipMemcpyHostToDevice, c->stream));
  c->solve_b = c->solve_buf;
  step_dispatch(c, nullptr);
  c->solve_b = nullptr;
  TRY(check_launch());
  HIPCHK(hipMemcpyAsync(x, c->X, field_bytes(c), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

// Stage U, dU (null: zero) and the freestream's tangent for a parameter
// tangent: U in tmpA, dU in Z (or the zero field), dq_inf / dmu_inf into
// dfree.  Returns the device dU.
static int stage_param_tangent(jr_ctx* c, const char* who, const double* U, const double* dU, const double* dq_inf,
                               const double* dmu_inf, const double** dU_dev) {
  if (!c || !U || !dq_inf) return fail(std::string(who) + ": null argument");
  if (c->pending_steps) return fail(std::string(who) + ": an asynchronous block is pending: wait for it first");
  if (!c->has_q) return fail(std::string(who) + ": set the freestream first");
  if (!c->dfree) {
    TRY(dalloc(c, &c->dfree, (size_t)c->n_cases * 8));
    HIPCHK(hipDeviceSynchronize());   // dalloc clears on the null stream
    c->P.dfree = c->dfree;
    c->P3.dfree = c->dfree;
  }
  const int nq = c->dim == 3 ? 6 : 5;   // the layout of jr_set_freestream's q_inf
  std::vector<double> h((size_t)c->n_cases * 8, 0.0);
  for (int k = 0; k < c->n_cases; ++k) {
    for (int v = 0; v < nq; ++v) h[(size_t)k * 8 + v] = dq_inf[k * nq + v];
    h[(size_t)k * 8 + 6] = dmu_inf ? dmu_inf[k] : 0.0;
    for (int v = 0; v < 7; ++v)
      if (!std::isfinite(h[(size_t)k * 8 + v])) return fail(std::string(who) + ": non-finite freestream tangent");
  }
  HIPCHK(hipMemcpyAsync(c->dfree, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->tmpA, U, field_bytes(c), hipMemcpyHostToDevice, c->stream));
  if (dU) {
    HIPCHK(hipMemcpyAsync(c->Z, dU, field_bytes(c), hipMemcpyHostToDevice, c->stream));
    *dU_dev = c->Z;
  } else {
    if (!c->zero_field) {
      TRY(dalloc(c, &c->zero_field, (size_t)c->n_cases * c->N));
      HIPCHK(hipDeviceSynchronize());
    }
    *dU_dev = c->zero_field;
  }
  return 0;
}

int jr_tangent_freestream(jr_ctx* c, const double* U, const double* dU, const double* dq_inf, const double* dmu_inf,
                          double* out) {
  if (!out) return fail("jr_tangent_freestream: null argument");
  const double* dUd = nullptr;
  TRY(stage_param_tangent(c, "jr_tangent_freestream", U, dU, dq_inf, dmu_inf, &dUd));
  if (c->dim == 3) {
    if (c->nv == 6) launch_param_tangent3<6, true>(c, c->tmpA, dUd, c->tmpB);
    else if (c->ph.viscous) launch_param_tangent3<5, true>(c, c->tmpA, dUd, c->tmpB);
    else launch_param_tangent3<5, false>(c, c->tmpA, dUd, c->tmpB);
  } else if (c->nv == 5) {
    launch_param_tangent<5, true>(c, c->tmpA, dUd, c->tmpB);
  } else if (c->ph.viscous) {
    launch_param_tangent<4, true>(c, c->tmpA, dUd, c->tmpB);
  } else {
    launch_param_tangent<4, false>(c, c->tmpA, dUd, c->tmpB);
  }
  TRY(check_launch());
  HIPCHK(hipMemcpyAsync(out, c->tmpB, field_bytes(c), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int jr_forces_tangent(jr_ctx* c, const double* U, const double* dU, const double* dq_inf, const double* dmu_inf,
                      double length, double cx, double cy, double* out) {
  if (!out) return fail("jr_forces_tangent: null argument");
  if (!(length > 0.0)) return fail("jr_forces_tangent: length must be positive");
  if (c && c->dim == 3) return fail("jr_forces_tangent: a 3D context needs jr_forces3_tangent (a reference area)");
  const double* dUd = nullptr;
  TRY(stage_param_tangent(c, "jr_forces_tangent", U, dU, dq_inf, dmu_inf, &dUd));
  if (c->nv == 5) launch_forces_tangent<5, true>(c, c->tmpA, dUd, length, cx, cy, c->coef);
  else if (c->ph.viscous) launch_forces_tangent<4, true>(c, c->tmpA, dUd, length, cx, cy, c->coef);
  else launch_forces_tangent<4, false>(c, c->tmpA, dUd, length, cx, cy, c->coef);
  TRY(check_launch());
  HIPCHK(hipMemcpyAsync(out, c->coef, (size_t)c->n_cases * 8 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int jr_forces3_tangent(jr_ctx* c, const double* U, const double* dU, const double* dq_inf, const double* dmu_inf,
                       double area, double length, double cx, double cz, double* out) {
  if (!out) return fail("jr_forces3_tangent: null argument");
  if (c && c->dim != 3) return fail("jr_forces3_tangent: not a 3D context");
  if (!(area > 0.0) || !(length > 0.0)) return fail("jr_forces3_tangent: area and length must be positive");
  const double* dUd = nullptr;
  TRY(stage_param_tangent(c, "jr_forces3_tangent", U, dU, dq_inf, dmu_inf, &dUd));
  if (c->nv == 6) launch_forces3_tangent<6, true>(c, c->tmpA, dUd, area, length, cx, cz, c->coef);
  else if (c->ph.viscous) launch_forces3_tangent<5, true>(c, c->tmpA, dUd, area, length, cx, cz, c->coef);
  else launch_forces3_tangent<5, false>(c, c->tmpA, dUd, area, length, cx, cz, c->coef);
  TRY(check_launch());
  HIPCHK(hipMemcpyAsync(out, c->coef, (size_t)c->n_cases * 8 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int jr_wall_loads_tangent(jr_ctx* c, const double* U, const double* dU, const double* dq_inf, const double* dmu_inf,
                          double* out) {
  if (!out) return fail("jr_wall_loads_tangent: null argument");
  if (c && c->dim != 3) return fail("jr_wall_loads_tangent: 3D contexts only");
  const double* dUd = nullptr;
  TRY(stage_param_tangent(c, "jr_wall_loads_tangent", U, dU, dq_inf, dmu_inf, &dUd));
  const int nw = c->P3.n_walls;
  const size_t n = (size_t)c->n_cases * 4 * std::max(1, nw);
  if (!c->dloads) {
    TRY(dalloc(c, &c->dloads, n));
    HIPCHK(hipDeviceSynchronize());   // dalloc clears on the null stream
  }
  if (c->nv == 6) launch_forces3_tangent<6, true>(c, c->tmpA, dUd, 1.0, 1.0, 0.0, 0.0, c->coef, c->dloads);
  else if (c->ph.viscous) launch_forces3_tangent<5, true>(c, c->tmpA, dUd, 1.0, 1.0, 0.0, 0.0, c->coef, c->dloads);
  else launch_forces3_tangent<5, false>(c, c->tmpA, dUd, 1.0, 1.0, 0.0, 0.0, c->coef, c->dloads);
  TRY(check_launch());
  HIPCHK(hipMemcpyAsync(out, c->dloads, (size_t)c->n_cases * 4 * nw * sizeof(double), hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

// n_steps explicit multistage steps on the device state (2D contexts):
// alphas[n_alphas] the stage coefficients, the CFL each case's control
// value (jr_newton_init / jr_set_cfl); hist_host[n_steps][n_cases] receives
// |R(U)| before each step.
int jr_rk_block(jr_ctx* c, int n_steps, const double* alphas, int n_alphas, double* hist_host) {
  if (!c || !alphas) return fail("jr_rk_block: null argument");
  if (c->pending_steps) return fail("jr_rk_block: an asynchronous block is pending: wait for it first");
  if (n_steps < 1) return fail("jr_rk_block: n_steps must be >= 1");
  if (n_alphas < 1 || n_alphas > 16) return fail("jr_rk_block: between 1 and 16 stage coefficients");
  if (!c->has_q) return fail("jr_rk_block: set the freestream first");
  if (n_steps > c->hist_cap) {
    HIPCHK(hipStreamSynchronize(c->stream));
    drop_graph(c);   // the captured launches hold the old history pointer
    if (c->hist) (void)hipFree(c->hist);
    c->hist = nullptr;
    HIPCHK(hipMalloc(&c->hist, (size_t)n_steps * c->n_cases * sizeof(double)));
    c->hist_cap = n_steps;
  }
  for (int s = 0; s < n_steps; ++s) {
    rk_dispatch(c, c->hist + (size_t)s * c->n_cases, alphas, n_alphas);
    TRY(check_launch());
  }
  if (hist_host)
    HIPCHK(hipMemcpyAsync(hist_host, c->hist, (size_t)n_steps * c->n_cases * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int jr_forces(jr_ctx* c, double length, double cx, double cy, double* out) {
  if (!c || !out) return fail("jr_forces: null argument");
  if (!c->has_q) return fail("jr_forces: set the freestream first");
  if (c->dim == 3) return fail("jr_forces: a 3D context needs jr_forces3 (a reference area)");
  forces_dispatch(c, length, cx, cy);
  TRY(check_launch());
  HIPCHK(hipMemcpyAsync(out, c->coef, (size_t)c->n_cases * 8 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int jr_wall_loads(jr_ctx* c, double* out) {
  if (!c || !out) return fail("jr_wall_loads: null argument");
  if (!c->has_q) return fail("jr_wall_loads: set the freestream first");
  const int nw = c->dim == 3 ? c->P3.n_walls : c->P.n_walls;
  const size_t n = (size_t)c->n_cases * (c->dim + 1) * std::max(1, nw);
  if (!c->loads) {
    TRY(dalloc(c, &c->loads, n));
    HIPCHK(hipDeviceSynchronize());   // dalloc clears on the null stream
  }
  if (c->dim == 3) forces3_dispatch(c, 1.0, 1.0, 0.0, 0.0, c->loads);
  else forces_dispatch(c, 1.0, 0.0, 0.0, c->loads);
  TRY(check_launch());
  HIPCHK(hipMemcpyAsync(out, c->loads, (size_t)c->n_cases * (c->dim + 1) * nw * sizeof(double),
                        hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int jr_forces3(jr_ctx* c, double area, double length, double cx, double cz, double* out) {
  if (!c || !out) return fail("jr_forces3: null argument");
  if (!c->has_q) return fail("jr_forces3: set the freestream first");
  if (c->dim != 3) return fail("jr_forces3: not a 3D context");
  if (!(area > 0.0) || !(length > 0.0)) return fail("jr_forces3: area and length must be positive");
  forces3_dispatch(c, area, length, cx, cz);
  TRY(check_launch());
  HIPCHK(hipMemcpyAsync(out, c->coef, (size_t)c->n_cases * 8 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

}  // extern "C"

template <int NV, bool VISC, int MM, class R>
static void profile_launch(jr_ctx* c, int which) {
  const Plan& P = c->P;
  switch (which) {
    case 0: launch_residual<NV, VISC>(c, c->U, c->R, 0); break;
    case 1: launch_tangent<NV, VISC>(c, c->U, c->Z, c->W, c->exact ? 0 : 1, c->vdt, 0); break;
    case 2: launch_apply<R, NV>(c, c->Rv, c->Z, 0, 0); break;
    case 3:
      hipLaunchKernelGGL((k_faces<double, NV, VISC>), grid1(P.nfi + P.nfj, BLOCK, c->n_cases), dim3(BLOCK), 0,
                         c->stream, P, c->ph, 0, c->ctl, c->Q, nullptr, c->Nd, nullptr, c->Fi, c->Fj, 0);
      break;
    case 4:
      hipLaunchKernelGGL((k_faces<Dual, NV, VISC>), grid1(P.nfi + P.nfj, BLOCK, c->n_cases), dim3(BLOCK), 0,
                         c->stream, P, c->ph, c->exact ? 0 : 1, c->ctl, c->Q, c->dQ, c->Nd, c->dNd, c->Fi, c->Fj, 0);
      break;
    case 5: {
      const int j = c->prof_rep++ % c->m;   // cycles through the Arnoldi columns
      cgs2_column<NV>(c, j, 2);
      break;
    }
    case 6: launch_factor<R, NV>(c, c->U); break;
    case 7:
      hipLaunchKernelGGL((k_cells<Dual, NV>), grid1(P.nc, BLOCK, c->n_cases), dim3(BLOCK), 0, c->stream, P, c->ph,
                         c->ctl, c->U, c->Z, c->Q, c->dQ, c->Fi, c->Fj, c->vdt, c->salin, c->W, 0);
      break;
    default: break;
  }
}

template <int NV, bool VISC, int MM>
static void profile_r(jr_ctx* c, int which) {
  if (c->mixed) profile_launch<NV, VISC, MM, float>(c, which);
  else profile_launch<NV, VISC, MM, double>(c, which);
}
template <int NV, bool VISC>
static void profile_m(jr_ctx* c, int which) {
  if (c->m <= 16) profile_r<NV, VISC, 16>(c, which);
  else if (c->m <= 32) profile_r<NV, VISC, 32>(c, which);
  else profile_r<NV, VISC, 64>(c, which);
}
static void profile_dispatch(jr_ctx* c, int which) {
  if (c->nv == 5) profile_m<5, true>(c, which);
  else if (c->ph.viscous) profile_m<4, true>(c, which);
  else profile_m<4, false>(c, which);
}

extern "C" int jr_profile(jr_ctx* c, int which, int reps, float* ms) {
  if (!c || !ms) return fail("jr_profile: null argument");
  if (which < 0 || which > 7 || reps < 1) return fail("jr_profile: bad kernel id or reps");
  if (c->dim == 3) return fail("jr_profile: the per-kernel profile covers the 2D path; use jr_probe_arm");
  if (!c->has_q) return fail("jr_profile: set the freestream first");
  // inputs: the residual at the device state (fills Q), wave speeds, factor,
  // a Krylov-like vector in Z and Rv, V filled with it
  residual_dispatch(c, c->U, c->R, 0);
  speeds_dispatch(c, c->U);
  factor_dispatch(c, c->U);
  HIPCHK(hipMemcpyAsync(c->Z, c->R, field_bytes(c), hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->Rv, c->R, field_bytes(c), hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->W, c->R, field_bytes(c), hipMemcpyDeviceToDevice, c->stream));
  for (int k = 0; k <= c->m; ++k)
    HIPCHK(hipMemcpyAsync(c->V + (size_t)k * c->n_cases * c->N, c->R, field_bytes(c), hipMemcpyDeviceToDevice,
                          c->stream));
  tangent_dispatch(c, c->U, c->Z, c->W, c->exact ? 0 : 1, c->vdt);   // fills dQ for id 4, 7
  {
    std::vector<Ctl> h;
    TRY(get_ctl(c, h));
    for (auto& x : h) x.active = 1;
    TRY(put_ctl(c, h));
  }
  profile_dispatch(c, which);   // warm
  c->prof_rep = 0;
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  for (int r = 0; r < reps; ++r) profile_dispatch(c, which);
  HIPCHK(hipEventRecord(c->ev1, c->stream));
  HIPCHK(hipEventSynchronize(c->ev1));
  TRY(check_launch());
  float t = 0.0f;
  HIPCHK(hipEventElapsedTime(&t, c->ev0, c->ev1));
  *ms = t / reps;
  return 0;
}

extern "C" {

int jr_timer_start(jr_ctx* c) {
  if (!c) return fail("jr_timer_start: null context");
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  return 0;
}

int jr_timer_stop(jr_ctx* c, float* ms) {
  if (!c || !ms) return fail("jr_timer_stop: null argument");
  HIPCHK(hipEventRecord(c->ev1, c->stream));
  HIPCHK(hipEventSynchronize(c->ev1));
  HIPCHK(hipEventElapsedTime(ms, c->ev0, c->ev1));
  return 0;
}

int jr_synchronize(jr_ctx* c) {
  if (!c) return fail("jr_synchronize: null context");
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

}  // extern "C"
