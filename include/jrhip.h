/* jrhip -- C ABI of the MI355X-native Newton-Krylov hot path.
 *
 * One context holds one canvas (a grid's plans, uploaded once) and
 * n_cases cases on it (a single run, or a batched polar).  The library owns
 * every device buffer; the state U lives on the device between calls.
 * Return codes: 0 OK, < 0 error (text in jr_last_error()).  A non-physical
 * or non-finite state is not an error: it is a rejected step, visible in the
 * residual history and the per-case scalars.
 *
 * Layout of every canvas field: (case, variable, i, j), j fastest, FP64 --
 * the reference's (nvar, NI, NJ) canvas layout with a case axis in front.
 *
 * Each entry point names the reference interface it replaces (paths
 * relative to the reference repository).  Threading: one context is bound to
 * one HIP stream and must not be used from two host threads at once.
 */
#ifndef JRHIP_H
#define JRHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct jr_ctx jr_ctx;

/* The atlas plans (src/jaxrans/mesh/atlas.py:97-228), host arrays. */
typedef struct {
  int32_t NI, NJ;
  const double* area;          /* [NI*NJ] */
  const double* face_i;        /* [2][(NI+1)*NJ] */
  const double* face_j;        /* [2][NI*(NJ+1)] */
  const double* centres;       /* [2][(NI+2)*(NJ+2)] one-ghost edge-padded cell centres */
  const double* distance;      /* [NI*NJ] wall distance (SA) */
  const uint8_t* interior;     /* [NI*NJ] */
  const int64_t* neighbours;   /* [NI*NJ] */
  int32_t n_task[2];           /* ghost-fill stages */
  const int64_t* task_dest[2];
  const int64_t* task_src[2];
  const int32_t* task_kind[2]; /* 0 copy, 1 slip, 2 no-slip, 3 symmetry, 4 far field, 5 inflow, 6 outflow */
  const double* task_normal[2];/* [2][n] */
  int32_t n_lines, n_slots;    /* line chains */
  const int64_t* line_start;   /* [n_lines+1] */
  const int64_t* line_cells;   /* [n_slots] */
  const int64_t* line_up_face;
  const int64_t* line_up_sign;
  const int64_t* line_down_face;
  const int64_t* line_down_sign;
  int32_t n_walls;
  const int64_t* wall_cell;
  const int64_t* wall_family;
  const int64_t* wall_face;
  const double* wall_into;
  const double* wall_mid;      /* [2][n_walls] */
  /* Point-vortex far field only (jr_numerics.far_field = 1; NULL otherwise):
   * face midpoints of the ghost-fill tasks, [2][n_task[s]] per stage, and the
   * vortex centre, the quarter chord of the longest body
   * (Atlas.quarter_chord, src/jaxrans/mesh/atlas.py:492-507). */
  const double* task_mid[2];
  double vortex_x, vortex_y;
} jr_grid_desc;

/* The 3D atlas (src/jaxrans/mesh/atlas3d.py:119-287), host arrays.  Canvas
 * cell (i, j, k) at flat index (i NJ + j) NK + k; the faces of family a lie
 * on the cell grid with one more along a, flat in the same order; chain
 * links and the composed ghost fill index the families concatenated
 * (i | j | k).  The fill is given composed: ghost g takes the primitives of
 * interior cell ghost_root[g] through boundary tasks ghost_ops[g][0 ..
 * ghost_nops[g]) in order (the reference's staged fill, evaluated per ghost). */
typedef struct {
  int32_t NI, NJ, NK;
  const double* volume;        /* [NI*NJ*NK] */
  const double* face[3];       /* family a: [3][nf_a] area vectors */
  const double* centres;       /* [3][(NI+2)*(NJ+2)*(NK+2)] one-ghost edge-padded cell centres */
  const double* distance;      /* [NI*NJ*NK] wall distance */
  const uint8_t* interior;     /* [NI*NJ*NK] */
  const int64_t* neighbours;   /* [NI*NJ*NK] */
  int32_t n_ghost;
  const int64_t* ghost_dest;
  const int64_t* ghost_root;
  const int64_t* ghost_nops;
  const int64_t* ghost_ops;    /* [n_ghost][4] task ids, -1 unused */
  int32_t n_tasks;
  const int32_t* task_kind;    /* as jr_grid_desc */
  const double* task_normal;   /* [3][n_tasks] outward unit normals */
  int32_t n_lines, n_slots;
  const int64_t* line_start;   /* [n_lines+1] */
  const int64_t* line_cells;   /* [n_slots] */
  const int64_t* line_up_face; /* concatenated face index */
  const int64_t* line_up_sign;
  const int64_t* line_down_face;
  const int64_t* line_down_sign;
  int32_t n_walls;
  const int64_t* wall_cell;
  const int64_t* wall_family;  /* normal axis 0, 1, 2 */
  const int64_t* wall_face;    /* flat index in that family */
  const double* wall_into;
  const double* wall_mid;      /* [3][n_walls] */
} jr_grid_desc3;

/* Discretization and solver settings (src/jaxrans/core/config.py:112-212). */
typedef struct {
  double gamma;
  double kappa;                /* MUSCL kappa, 0: limited */
  int32_t flux;                /* 0 hllc, 1 roe */
  int32_t limiter;             /* 0 minmod, 1 none, 2 van_albada, 3 van_albada_smooth */
  int32_t viscous;
  int32_t sa;                  /* Spalart-Allmaras row */
  double sutherland;           /* S/T_inf, <= 0: constant viscosity */
  double c_cr1;                /* QCR2000, 0: off */
  int32_t jacobian_exact;      /* 0: first-order Jacobian (defect correction) */
  int32_t mixed;               /* 1: FP32 line-preconditioner factor and solve */
  int32_t line_solver;         /* 0: block cyclic reduction (linalg.py:159-200), 1: block Thomas */
  int32_t krylov_subspace;
  int32_t krylov_restarts;
  double krylov_tol;
  double krylov_success;
  double cfl_growth, cfl_cutback, cfl_min, cfl_max;
  double max_relative_update;  /* <= 0: none */
  int32_t far_field;           /* FAR_FIELDS (src/jaxrans/numerics/boundary.py:243-246): 0 characteristic,
                                  1 characteristic_point_vortex (2D: the far-field reference is the
                                  vortex of the current wall-pressure lift) */
} jr_numerics;

/* Context (replaces make_residual + make_newton_step set-up,
 * src/jaxrans/numerics/residual.py:273-334, solvers/implicit.py:84-98). */
int jr_ctx_create(const jr_grid_desc* grid, const jr_numerics* num, int n_cases, jr_ctx** out);

/* A context on a 3D canvas (the reference builds the same operators from a 3D
 * atlas, src/jaxrans/numerics/residual.py:273-334 with mesh/atlas3d.py).
 * Every entry point below works on it, fields [n_cases][nvar][NI][NJ][NK]
 * with nvar 5 (rho, rho u, rho v, rho w, E) or 6 (+ rho nu_hat); the line
 * preconditioner solves its chains by block Thomas (5x5 flow blocks). */
int jr_ctx_create3(const jr_grid_desc3* grid, const jr_numerics* num, int n_cases, jr_ctx** out);
int jr_ctx_destroy(jr_ctx* ctx);

/* Per-case freestream primitives q_inf[n_cases][5] (rho, u, v, p, nu_hat ratio;
 * 3D: q_inf[n_cases][6], rho, u, v, w, p, nu_hat ratio)
 * and mu_inf[n_cases] (NULL: inviscid) (src/jaxrans/core/state.py:176-192). */
int jr_set_freestream(jr_ctx* ctx, const double* q_inf, const double* mu_inf);

/* State on the device, host copies in/out: [n_cases][nvar][NI][NJ];
 * jr_set_state_case writes one case's [nvar][NI][NJ]. */
int jr_set_state(jr_ctx* ctx, const double* U_host);
int jr_set_state_case(jr_ctx* ctx, int case_index, const double* U_host);
int jr_get_state(jr_ctx* ctx, double* U_host);

/* A different grid of the same topology (identical index plans: stages,
 * chains, walls, neighbours; only coordinates differ) for one case of a 2D
 * context: its areas, face vectors, centres, wall distance, boundary normals
 * and wall midpoints.  The first call gives every case its own geometry.
 * One batch then steps different geometries in shared launches, as the
 * reference's slot scheduler steps different freestreams on one grid
 * (src/jaxrans/sweep.py:172-327).  Refuses a grid of another topology. */
int jr_set_geometry(jr_ctx* ctx, int case_index, const jr_grid_desc* grid);

/* R(U) (ResidualOperator.residual, src/jaxrans/numerics/residual.py:324-327);
 * first_order selects first_order_faces (reconstruction.py:71-80). */
int jr_residual(jr_ctx* ctx, const double* U_host, double* R_host, int first_order);
int jr_residual_dev(jr_ctx* ctx, const double* U_dev, double* R_dev, int first_order);

/* J dU (+ V/dt dU if with_vdt, at the context's current CFL): the Krylov
 * operator of src/jaxrans/solvers/implicit.py:105-116. */
int jr_tangent(jr_ctx* ctx, const double* U_host, const double* dU_host, double* out_host, int first_order,
               int with_vdt);
int jr_tangent_dev(jr_ctx* ctx, const double* U_dev, const double* dU_dev, double* out_dev, int first_order,
                   int with_vdt);

/* Face wave speeds lam[n_cases][(NI+1)*NJ + NI*(NJ+1)] and their cell sums
 * lsum[n_cases][NI*NJ] (face_wave_speeds + cell_wave_speed_sum,
 * src/jaxrans/numerics/residual.py:169-203). */
int jr_wave_speeds(jr_ctx* ctx, const double* U_host, double* lam_host, double* lsum_host);

/* Line preconditioner: factor at U and the current CFL, then z = M^-1 r
 * (line_implicit_j, src/jaxrans/solvers/preconditioners.py:290-306). */
int jr_precond_factor(jr_ctx* ctx, const double* U_host);
int jr_precond_apply(jr_ctx* ctx, const double* r_host, double* z_host);

/* Newton control: start every case at cfl0[n_cases] with |R|_prev = inf
 * (ImplicitState, src/jaxrans/solvers/implicit.py:35-51). */
int jr_newton_init(jr_ctx* ctx, const double* cfl0);
/* The same for one case (a slot refilled between blocks, sweep.py:229-264). */
int jr_newton_init_case(jr_ctx* ctx, int case_index, double cfl0);

/* Park (parked = 1) or release (0) case k of a batched context: every pass
 * of a Newton step (residual, wave speeds, chain blocks and factor, GMRES,
 * relaxation, acceptance test, update) skips a parked case, whose state and
 * control stay as they are and whose history entries repeat its last norm;
 * jr_wave_speeds and jr_precond_factor skip it too.  An empty serving slot
 * then costs only its workgroups' exit.  A released case continues from its
 * state; refill it first with jr_set_geometry, jr_set_state_case and
 * jr_newton_init_case.  The slot mask of the reference's slot scheduler
 * (masked, src/jaxrans/sweep.py:203-207), which keeps an idle slot's state
 * and whose batch shrink (:313-320) stops paying for idle slots. */
int jr_park_case(jr_ctx* ctx, int case_index, int parked);

/* n_steps pseudo-transient Newton-Krylov steps on the device state
 * (one block of run_to_steady, src/jaxrans/solvers/driver.py:87-104);
 * hist_host[n_steps][n_cases] receives |R(U)| before each update. */
int jr_newton_block(jr_ctx* ctx, int n_steps, double* hist_host);

/* jr_newton_block that allocates nothing: no hipMalloc, hipHostMalloc or
 * hipFree, no graph, no host read inside the block.  No counterpart in the
 * reference; it is for callers that cannot allocate device memory once their
 * contexts exist (a serving loop's steady state, code under HIP graph
 * capture).  The residual history is staged in a static device array of
 * JR_NOALLOC_HISTORY_MAX doubles (512 KiB, one per device, held until the
 * process ends), so n_steps * n_cases must not exceed JR_NOALLOC_HISTORY_MAX
 * (larger blocks return an error).  Every GMRES cycle is launched, cycles no
 * case needs exit on the device: results equal jr_newton_block's.  All calls
 * on one device share the staging array: calls that overlap on one device
 * (two contexts on two host threads) race and must be serialised by the
 * caller. */
#define JR_NOALLOC_HISTORY_MAX 65536
#define JR_NOALLOC_MAX_DEVICES 64
int jr_newton_block_noalloc(jr_ctx* ctx, int n_steps, double* hist_host);

/* One solve of the Newton step's linear system, (V/dt + dR~/dU) x = b, at
 * the device state and the current CFL (set it with jr_set_cfl): GMRES from
 * zero with the context's subspace, restarts and tolerance, right-
 * preconditioned by the line solve; R~ is the first-order residual unless the
 * context takes the exact Jacobian.  b_host and x_host are fields
 * [n_cases][nv][cells]; the state is not changed.  The approximate inverse
 * of the sensitivity's defect correction: newton_inverse,
 * src/jaxrans/derivatives.py:94-120. */
int jr_linear_solve(jr_ctx* ctx, const double* b_host, double* x_host);

/* Parameter tangents of the sensitivity solve.  The flow parameters reach
 * the residual and the force coefficients only through the freestream
 * primitives and mu_inf (flow_inputs, src/jaxrans/derivatives.py:65-70), so
 * a parameter direction is the freestream's tangent: dq_inf_host[n_cases][5]
 * (3D: [6]; the layout of jr_set_freestream's q_inf) and
 * dmu_inf_host[n_cases] (null: zero).  dU_host may be null (zero).
 *
 * jr_tangent_freestream: d/de R(U + e dU; q_inf + e dq_inf, mu_inf + e dmu_inf)
 * at e = 0, the exact (second-order) residual -- jax.jvp of the residual
 * over the parameters (src/jaxrans/solvers/sensitivity.py:60) and, with dU,
 * over the state as well.  Ghost states, the U - U0 rows off the interior,
 * the viscous and SA terms and the point-vortex reference all carry the
 * freestream's tangent.  out_host: a field [n_cases][nv][cells].
 *
 * jr_forces_tangent: the tangent of jr_forces' coefficients at U along the
 * same direction, the output JVP of src/jaxrans/solvers/sensitivity.py:63
 * (force_coefficients over the state and the freestream,
 * src/jaxrans/derivatives.py:82-84).  out_host[n_cases][8] =
 * d(cl, cd, cm, cd_friction, fx, fy), 0, 0. */
int jr_tangent_freestream(jr_ctx* ctx, const double* U_host, const double* dU_host, const double* dq_inf_host,
                          const double* dmu_inf_host, double* out_host);
int jr_forces_tangent(jr_ctx* ctx, const double* U_host, const double* dU_host, const double* dq_inf_host,
                      const double* dmu_inf_host, double length, double centre_x, double centre_y, double* out_host);
/* ... and of jr_forces3's coefficients (3D contexts): out_host[n_cases][8] =
 * d(cl, cd, cm, cd_friction, fx, fy, fz, cy). */
int jr_forces3_tangent(jr_ctx* ctx, const double* U_host, const double* dU_host, const double* dq_inf_host,
                       const double* dmu_inf_host, double area, double length, double centre_x, double centre_z,
                       double* out_host);
/* ... and of jr_wall_loads' per-face values (3D contexts), out_host
 * [n_cases][4][n_walls] = d(gauge pressure, viscous force x, y, z): the
 * swept (2.5D) section forms its coefficients from them on the host
 * (force_coefficients of a swept flow, src/jaxrans/post/forces.py:136-160). */
int jr_wall_loads_tangent(jr_ctx* ctx, const double* U_host, const double* dU_host, const double* dq_inf_host,
                          const double* dmu_inf_host, double* out_host);

/* n_steps explicit multistage Runge-Kutta steps on the device state, 2D
 * contexts: U^(k) = U^n - alphas[k] dt/V R(U^(k-1)) with the local time step
 * dt frozen at U^n (CFL: each case's control value, set by jr_newton_init or
 * jr_set_cfl); hist_host[n_steps][n_cases] receives |R(U^n)| of each step.
 * make_rk_step, src/jaxrans/solvers/explicit.py:48-71, with local_time_step,
 * src/jaxrans/numerics/residual.py:142-166. */
int jr_rk_block(jr_ctx* ctx, int n_steps, const double* alphas, int n_alphas, double* hist_host);

/* The Krylov operator's first-order tangent by LDS tiles (k_tangent_tile:
 * node values, face fluxes and cell divergence in one pass, enable = 1) or
 * by the separate node, face and cell kernels (enable = 0, the default: the
 * faster on MI355X), for 2D contexts whose line apply writes the tangent
 * primitives (block Thomas, or cyclic reduction within its LDS budget).  The
 * two agree to round-off.  A switch for measurement and tests; no reference
 * counterpart. */
int jr_set_tile_tangent(jr_ctx* ctx, int enable);

/* Execution option: replay a block of Newton steps from a HIP graph captured
 * on the first block of that length (1), or launch every kernel on the
 * context stream (0, the default: faster on MI355X with ROCm 7).  Results are
 * identical.  No reference counterpart: execution. */
int jr_set_graphs(jr_ctx* ctx, int enable);

/* Event probe of the hot kernels during real Newton steps: arm with a mask
 * of kernel ids (bit 0 line apply, 1 tangent face kernel, 2 tangent fill,
 * 3 tangent cell kernel, 4 CGS2 pass 1, 5 pass 2, 6 pass 3, 7 line factor)
 * and room for max_launches bracketed launches; read returns per id the
 * summed device milliseconds and the launch count, and disarms.  Needs graphs
 * off.  No reference counterpart: measurement. */
int jr_probe_arm(jr_ctx* ctx, unsigned mask, int max_launches);
int jr_probe_read(jr_ctx* ctx, double* ms_total /*[8]*/, int* launches /*[8]*/);

/* Asynchronous block of n_steps Newton steps (serving many contexts from one
 * host thread): enqueued on the context stream with no host synchronisation
 * (every GMRES cycle is launched; cycles no case needs exit on the device,
 * so results equal jr_newton_block's); poll returns 1 when the block is done
 * (0 not yet, < 0 error); wait blocks until then and copies the residual
 * history [n_steps][n_cases].  With force_ref = {length, x_ref, y_ref} the
 * block ends with jr_forces' coefficients [n_cases][8], which wait copies to
 * forces_host (NULL to skip; force_ref NULL and forces_host non-NULL is an
 * error).  One pending block per context.  The blocks of
 * driver.py:87-104, split so the host can keep several solves in flight.
 * No reference counterpart for the split itself: execution. */
int jr_newton_block_async(jr_ctx* ctx, int n_steps, const double* force_ref);
int jr_newton_block_poll(jr_ctx* ctx);
int jr_newton_block_wait(jr_ctx* ctx, double* hist_host, double* forces_host);

/* Device memory held by a context, and the device's free / total memory
 * (admission of served solves by free HBM).  No reference counterpart. */
int jr_ctx_bytes(jr_ctx* ctx, size_t* bytes);
int jr_mem_info(size_t* free_bytes, size_t* total_bytes);

/* Per-case scalars [n_cases][9]: cfl, |R|_prev, |R|, linear reduction,
 * omega, accepted(last step), steps, rejected, GMRES cycles run so far. */
int jr_get_scalars(jr_ctx* ctx, double* out_host);
int jr_set_cfl(jr_ctx* ctx, const double* cfl);

/* Forces of the device state (force_coefficients, src/jaxrans/post/forces.py:108-161):
 * out[n_cases][8] = cl, cd, cm, cd_friction, fx, fy, 0, 0. */
int jr_forces(jr_ctx* ctx, double length, double centre_x, double centre_y, double* out_host);

/* Forces of a 3D context (force_coefficients, src/jaxrans/post/forces.py:108-161):
 * per q_inf * area; lift normal to the freestream in the x-z plane; pitching
 * moment about the y axis through (centre_x, centre_z), pressure and
 * friction, per q_inf * area * length.  out[n_cases][8] = cl, cd, cm,
 * cd_friction, fx, fy, fz, cy. */
int jr_forces3(jr_ctx* ctx, double area, double length, double centre_x, double centre_z, double* out_host);

/* Per wall face of the device state, in the atlas's wall order (wall_loads,
 * src/jaxrans/post/forces.py:48-68): out[n_cases][1 + dim][n_walls] = the
 * gauge pressure p - 1/gamma of the wall cell, then the viscous force on the
 * body (into-fluid sign times the viscous flux of the velocity rows at the
 * face; zero when inviscid), per component.  The wall distributions of
 * Result.surface / _surface.dat are formed from these on the host. */
int jr_wall_loads(jr_ctx* ctx, double* out_host);

/* Average device time of one launch of a hot kernel on the device state,
 * timed with HIP events on the context stream: 0 residual (all launches),
 * 1 tangent (all launches), 2 line-preconditioner apply, 3 residual face
 * kernel, 4 tangent face kernel, 5 CGS2 update+dots (launch r at Arnoldi
 * column j = r mod m: with reps a multiple of m, the average over a cycle),
 * 6 line factor, 7 tangent cell kernel.  No reference counterpart: measurement. */
int jr_profile(jr_ctx* ctx, int kernel, int reps, float* ms_per_launch);

/* Device timing of the context's stream (HIP events): start, then stop
 * returns the elapsed milliseconds.  No reference counterpart: measurement. */
int jr_timer_start(jr_ctx* ctx);
int jr_timer_stop(jr_ctx* ctx, float* ms);
int jr_synchronize(jr_ctx* ctx);

/* Device buffers for the *_dev entry points. */
int jr_device_alloc(size_t bytes, void** out);
int jr_device_free(void* p);
int jr_copy_to_device(void* dst, const void* src, size_t bytes);
int jr_copy_to_host(void* dst, const void* src, size_t bytes);
int jr_device_count(int* n);

const char* jr_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
