/* swm_mi355x.h -- C ABI of the MI355X (gfx950) dynamical core of the CSU
 * geodesic-grid shallow-water model.
 *
 * The library replaces the reference's hot path, the time step
 *   step::step_dt, step::wrap_prognostics   (swm/source/step.F:114-212, :438-493)
 * and the operator / halo / solver modules it calls, behind entry points that
 * a Fortran ISO_C_BINDING shim binds one for one.  Each declaration names the
 * reference interface it replaces (file:line, paths relative to the
 * reference's doe_ca/ directory).
 *
 * Conventions
 *  - Host arrays use the reference's Fortran layout: column-major, first index
 *    fastest, the 10 rhombic panels ("subdomains") last, a one-cell halo on
 *    every panel: cell fields (iim,jjm,10), corner scalars (2,iim,jjm,10),
 *    corner vectors (2,2,iim,jjm,10) = (space, corner, i, j, panel), edge
 *    fields (3,iim,jjm,10), with iim = jjm = 2 + 2^g.
 *  - Device data lives in the context; it is laid out as one plane per
 *    component (structure of arrays) and never crosses PCIe inside a step.
 *  - Every function returns swm_status; SWM_OK is 0.  The reference has no
 *    return codes and stops the program on fatal errors (utilities/layer_1/
 *    abort.F:26-49); a Fortran shim maps a non-zero status to that behaviour.
 *  - One host thread per context; calls on one context are serialised on its
 *    stream.  Contexts are independent.
 */
#ifndef SWM_MI355X_H
#define SWM_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int swm_status;
enum {
  SWM_OK = 0,
  SWM_ERR_ARG = 1,            /* bad argument or unknown name */
  SWM_ERR_DIMS = 2,           /* array extents do not match the grid */
  SWM_ERR_HIP = 3,            /* HIP runtime error (message in swm_last_error) */
  SWM_ERR_NOT_CONVERGED = 4,  /* elliptic solve failed: elliptic_solver.F:176-178 */
  SWM_ERR_NAN = 5,            /* non-finite value detected in the state */
  SWM_ERR_NO_DEVICE = 6,      /* no usable gfx950 device */
  SWM_ERR_STATE = 7           /* call out of order (e.g. step before init) */
};

/* Text of the last error on this thread (never NULL). */
const char* swm_last_error(void);

/* ------------------------------------------------------------------------
 * Host-side grid set-up (needs no device)
 * ------------------------------------------------------------------------ */
typedef struct swm_grid swm_grid;

/* Cell centres of the recursively bisected icosahedron, level g (1..10),
 * written as (3, 10*4^g+2) in the reference's global order (NP, SP, rows).
 * Replaces utilities/grid_generation/generate_grid.F:69-130,327-463. */
swm_status swm_grid_points(int g, double* p_list);

/* Builds the grid: panel index maps, halo tables, masks, metrics, Coriolis
 * parameter and the multigrid hierarchy.  p_list (3, max_ig) may be NULL to
 * use swm_grid_points.  radius and omega are the sphere's radius (m) and
 * rotation rate (1/s).
 * Replaces initialize_wrap (utilities/layer_2/wrap_data.F:34-157),
 * set_grid_center_mask / set_grid_corner_mask (utilities/layer_2/grid_mask.F:
 * 60-333), generate_grid_metrics (utilities/layer_3/grid_metrics.F:55-219),
 * initialize_multigrid (utilities/layer_3/multigrid_solver.F:61-474) and
 * set_coriolis_parameter (swm/source/initialize.F:286-321). */
swm_status swm_grid_create(int g, const double* p_list, double radius, double omega,
                           swm_grid** out);
void swm_grid_destroy(swm_grid* grid);

/* Grid extents: level, panel extents, number of cells. */
swm_status swm_grid_dims(const swm_grid* grid, int* g, int* iim, int* jjm, int* max_ig);

/* Named set-up arrays, copied out in the reference layout.  Names (element
 * type): "p","c","area","area_inv","corner_area","corner_area_inv",
 * "sub_area","dc","normal","tangent","v_weights","s_weights","l_weights","f",
 * "grid_center_mask","grid_corner_mask","polygon_type","tot_area" (double);
 * "grid_center_index","grid_corner_index" (int32);
 * halo tables "face_dst","face_src","vrtx_dst","vrtx_src","edge_dst",
 * "edge_src" (int32, flat (component, slot) indices, source -1 = zero);
 * multigrid weights "lap_w<l>" for level l = 1..g (double, (7, 2+2^l, 2+2^l, 10)).
 * swm_grid_array_len returns the element count, or -1 for an unknown name. */
int64_t swm_grid_array_len(const swm_grid* grid, const char* name);
swm_status swm_grid_array(const swm_grid* grid, const char* name, void* out);

/* ------------------------------------------------------------------------
 * Device model (one context = one model state on one GPU)
 * ------------------------------------------------------------------------ */
typedef struct swm_ctx swm_ctx;

/* SWM_FP64: the reference's arithmetic, operation for operation.
 * SWM_MIXED: FP32 state, metrics and stencils (thickness stored less its
 * initial area-weighted mean, vorticity less the Coriolis parameter; the
 * zero-sum gradient and Laplacian stencils formed on differences), FP32
 * multigrid cycles inside an FP64 defect-correction loop on FP64
 * streamfunction / velocity potential, FP64 global sums.  No counterpart in
 * the reference. */
enum { SWM_FP64 = 0, SWM_MIXED = 1 };

/* Advection of mass and tracers in step_dt (step.F:131-135, namelist flags
 * l_fct / l_centered / l_upwind / l_remap): flux-corrected transport
 * (mass_fct.F, tracer_fct.F), centred / upstream fluxes with ABM3 stepping
 * of the thickness and the mass-weighted tracers (step.F:221-327, :495-602),
 * or incremental remapping (get_remap_flux, step.F:687-771, over
 * utilities/layer_2/remap.F; FP64 single-member contexts only). */
enum { SWM_ADV_FCT = 0, SWM_ADV_CENTERED = 1, SWM_ADV_UPWIND = 2, SWM_ADV_REMAP = 3 };

typedef struct swm_config {
  int ntracer;       /* passive tracers (the reference compiles 5: swm_parameters.F:66) */
  double dt;         /* time step (s) */
  double grav;       /* gravity (m/s^2) */
  int precision;     /* SWM_FP64 or SWM_MIXED */
  int device;        /* HIP device ordinal */
  int advection;     /* SWM_ADV_FCT (0, the default of a zeroed struct), SWM_ADV_CENTERED, SWM_ADV_UPWIND,
                        SWM_ADV_REMAP */
} swm_config;

/* Checks, once per device, that the kernels read their workgroup size from
 * the implicit kernel arguments where the code-object v5 layout puts it
 * (the build pins version 5); swm_create runs it.  SWM_ERR_HIP if not.  No
 * counterpart in the reference. */
swm_status swm_self_test(int device);

/* Allocates the device state for a grid (the grid may be destroyed after). */
swm_status swm_create(const swm_grid* grid, const swm_config* cfg, swm_ctx** out);
void swm_destroy(swm_ctx* ctx);

/* ------------------------------------------------------------------------
 * Ensembles (no counterpart in the reference, which runs one model per
 * process).  A context of nmem members holds nmem independent model states
 * on the same grid with the same settings; swm_step advances every member in
 * the same launches (one grid row of workgroups per member; the metrics are
 * read once for all).  Host-side calls (initialisation, fields, budgets,
 * restart files, the single operators) address the member chosen by
 * swm_select_member (member 0 by default).  Every member must be initialised
 * before the first step; initialising a member restarts the step counter of
 * the context (the next step is a forward Euler step for every member).  In
 * the mixed path the members share the thickness offset of the member
 * initialised first, and in both paths the surface height.  The linearised
 * model is for single-member contexts.  When a member is initialised again
 * after steps, the others continue from their latest state: their next step
 * is a forward Euler step from it, as after swm_restart_clock.
 * ------------------------------------------------------------------------ */
swm_status swm_create_members(const swm_grid* grid, const swm_config* cfg, int nmem, swm_ctx** out);
int swm_members(const swm_ctx* ctx);
swm_status swm_select_member(swm_ctx* ctx, int member);

/* Fixed schedule of the mixed solve: ncycles > 0 correction cycles per solve
 * with no host-side convergence test, so that steps can be captured in a HIP
 * graph; 0 restores the test.  The solves after an initialisation whose
 * warm start has not yet its full order (from zero, then the linear
 * extrapolation while only two solutions are held) take max(ncycles,
 * nstart).  The largest mean squared residuals of the two
 * fields and the number of solves that ended above the (absolute) criterion
 * are recorded per member on the device: swm_solver_track gives them for the
 * selected member (out[0..2]), counted from its last initialisation.
 * FP64 contexts: every two-field solve of the step runs exactly ncycles
 * V-cycles from the warm start (nstart unused), which makes the model a
 * smooth function of its initial state (no cycle count that switches with
 * the residual), e.g. for finite-difference checks; the linear solves of
 * the tangent-linear and adjoint model keep their own criterion. */
swm_status swm_set_fixed_cycles(swm_ctx* ctx, int ncycles, int nstart);
swm_status swm_solver_track(swm_ctx* ctx, double* out);

/* An ensemble: groups (contexts, each of one or more members, e.g. at
 * different grid levels) advanced together.  One ensemble step advances
 * group k by substeps[k] of its own time steps, so a level g+1 group with
 * half the time step takes 2.  swm_ensemble_step captures the launches of
 * steps_per_graph ensemble steps of every group into a HIP graph once per
 * phase of the groups (ABM3 start and time levels, solver buffers, launch
 * choice) and replays it with one graph launch after that; every group must
 * run the mixed step with a fixed schedule (swm_set_fixed_cycles).  The
 * groups' streams are synchronised before and the ensemble's stream after
 * the steps, so the contexts' own calls see the results.
 * swm_ensemble_tune times candidate launch choices (one stream, or one per
 * group forked inside the graph; 1, 3 or 6 ensemble steps per graph;
 * workgroups of 64, 128 or 256 for the stencil and for the multigrid
 * kernels), each over `rounds` x 6 ensemble steps, keeps the fastest, and
 * restores the groups' state and clocks; trials[0..max_trials) receive the
 * candidates with their time per ensemble step. */
typedef struct swm_ensemble swm_ensemble;
typedef struct swm_ensemble_choice {
  int concurrent;        /* 0: groups one after another on one stream; 1: a stream per group */
  int steps_per_graph;   /* ensemble steps per captured graph */
  int blk_step;          /* workgroup size of the step's stencil kernels and halo fills */
  int blk_mg;            /* workgroup size of the multigrid smoother and transfers */
  double ms_per_step;    /* measured ms per ensemble step (output) */
} swm_ensemble_choice;
swm_status swm_ensemble_create(swm_ctx* const* groups, const int* substeps, int ngroups, swm_ensemble** out);
void swm_ensemble_destroy(swm_ensemble* ens);
swm_status swm_ensemble_set(swm_ensemble* ens, const swm_ensemble_choice* choice);
swm_status swm_ensemble_get(const swm_ensemble* ens, swm_ensemble_choice* choice);
swm_status swm_ensemble_step(swm_ensemble* ens, int nsteps);
swm_status swm_ensemble_tune(swm_ensemble* ens, int rounds, swm_ensemble_choice* trials, int max_trials,
                             int* ntrials);

/* Initial state from thickness h and surface height hs (iim,jjm,10), corner
 * velocity u (2,2,iim,jjm,10) and tracers (iim,jjm,ntracer,10); vorticity,
 * divergence and kinetic energy are derived from u on the device as the
 * reference does after its initial conditions (swm/source/initialize.F:
 * 418-468, utilities/layer_4/operators_hl.F:153-266), then the halos are
 * filled as by the first wrap_prognostics (swm/source/driver.F:78). */
swm_status swm_init_state(swm_ctx* ctx, const double* h, const double* hs, const double* u,
                          const double* tracer);

/* The reference's built-in initial state, shallow-water test case 5 (zonal
 * flow over a mountain), initial_conditions_swtc5 (swm/source/initialize.F:
 * 330-513, default build): h, hs (iim,jjm,10), u (2,2,iim,jjm,10), tracer
 * (iim,jjm,ntracer,10) on the host, halos filled, ready for swm_init_state.
 * Tracer 1 is hs, tracer 2 a plateau of 1000 over the mountain, the rest the
 * Coriolis parameter; ntracer >= 2 as the reference requires.  Needs no
 * device. */
swm_status swm_swtc5_state(const swm_grid* grid, int ntracer, double grav, double* h, double* hs, double* u,
                           double* tracer);

/* The reference driver's run settings, data/namelist as model_control.F:
 * 98-139 (input) reads it, line by line: tau_end_day, tau_end_hour, dt
 * (F7.2), l_restart, l_upwind, l_centered, l_fct, l_remap (L1), n_restart,
 * n_output, n_diagnostics (list-directed).  Logicals are 0 / 1.  Needs no
 * device. */
typedef struct swm_namelist {
  double tau_end_day, tau_end_hour, dt;
  int l_restart, l_upwind, l_centered, l_fct, l_remap;
  int n_restart, n_output, n_diagnostics;
} swm_namelist;
swm_status swm_read_namelist(const char* path, swm_namelist* out);

/* Advances nsteps time steps: for each, the time-level pointers and ABM3
 * weights (swm/source/model_control.F:28-89), step_dt and wrap_prognostics
 * (swm/source/step.F:114-212, :438-493; loop of swm/source/driver.F:90-130).
 * Returns SWM_ERR_NOT_CONVERGED if an elliptic solve fails, SWM_ERR_NAN if
 * its residual norms (read back by every solve of the convergence test) are
 * not finite: the state holds NaN or Inf. */
swm_status swm_step(swm_ctx* ctx, int nsteps);

/* Halo shadow check (debugging): while enabled, every halo fill of the step
 * (wrap_data.F:1790-2607 on the device) is checked against the same fill
 * done on the host from the host copy of the tables: the fields of every
 * member are copied to the host before and after the device fill and
 * compared bit for bit.  A step with a mismatch returns SWM_ERR_STATE;
 * swm_halo_shadow_stats gives the fields checked and the values that
 * differed since the context was created.  Synchronises at every fill (slow);
 * refused inside graph launches.  No counterpart in the reference. */
swm_status swm_set_halo_shadow(swm_ctx* ctx, int enable);
swm_status swm_halo_shadow_stats(const swm_ctx* ctx, long long* checks, long long* mismatches);

/* Restarts the step counter of every member without changing its state:
 * the latest time level is copied into both levels and the next step is a
 * forward Euler step, as after a restart read (swm/source/swm_restart.F:
 * 76-87 restarts the ABM3 from Euler).  The model time goes on.  No
 * counterpart as a call in the reference. */
swm_status swm_restart_clock(swm_ctx* ctx);

/* Incremental remapping (SWM_ADV_REMAP): the tracer limiter's switches
 * l_limit_tracers_min and l_limit_tracers_max of limit_gradient (remap.F:
 * 2152-2397), both 1 by default (the two-sided van Leer limit that
 * l_limit_tracers = .true. asks for).  remap_mass_and_tracers (remap.F:
 * 1072-1604) gives the two switches no value when the caller does not pass
 * them, and get_remap_flux does not: each build of the reference uses
 * whatever its stack holds (observed: a level-4 build on in every step, a
 * level-3 build on in the first step only).  swm_step returns
 * SWM_ERR_STATE when a thickness below -1e-10 reaches the remapping
 * (remap_prep stops the reference there). */
swm_status swm_set_remap_limiter(swm_ctx* ctx, int limit_min, int limit_max);

/* Copies a field of the current time level to the host (reference layout):
 * "h","eta","div","relative","ke","psi","chi" and the surface height "hs"
 * (iim,jjm,10); "u"
 * (2,2,iim,jjm,10); "tracer" (iim,jjm,ntracer,10). */
swm_status swm_get_field(swm_ctx* ctx, const char* name, double* out);

/* Waits for all work queued on the context's stream. */
swm_status swm_synchronize(swm_ctx* ctx);

/* Steps taken and V-cycles run by the last elliptic solve. */
int swm_step_count(const swm_ctx* ctx);
int swm_last_vcycles(const swm_ctx* ctx);

/* Convergence settings of the elliptic solve.  The reference's are the
 * defaults: mean(r^2) over real cells <= 1e-20 (relative = 0, tol = 1e-20)
 * within 12 V-cycles, else SWM_ERR_NOT_CONVERGED where the reference stops
 * (elliptic_solver.F:120-180).  relative = 1 tests sqrt(mean(r^2) /
 * mean(rhs^2)) <= tol instead; relative = 1 with tol = 0 solves to the
 * FP64 floor: until a cycle no longer halves the residual.  max_cycles
 * counts all cycles (>= 2).  No counterpart in the reference, where these
 * are compile-time constants. */
swm_status swm_set_solver(swm_ctx* ctx, double tol, int relative, int max_cycles);

/* Mixed path only: the start of each elliptic solve, the quadratic
 * extrapolation 3 x_n - 3 x_(n-1) + x_(n-2) of the last three solutions
 * (extrapolate = 2, the default; the linear one while fewer are held), the
 * linear 2 x_n - x_(n-1) (1), or the last solution (0, the reference's warm
 * start), and the number of correction cycles before the criterion is
 * first tested (min_cycles, default 1; the reference runs 2 from its warm
 * start).  No counterpart in the reference. */
swm_status swm_set_mixed_start(swm_ctx* ctx, int extrapolate, int min_cycles);

/* Mixed path: the shape of the FP32 V-cycles, which only precondition the
 * FP64 defect correction (the solution meets the same criterion with any
 * shape; a cheaper cycle reduces the residual less).  fine_sweeps: sweeps
 * before and after the coarse correction on the finest level (1 to 4; 2 is
 * the reference's count).  half_weights (default on): the finest level's
 * smoother and restriction read its Laplacian weights in half precision,
 * scaled, in difference form.  No counterpart in the reference. */
swm_status swm_set_mixed_smoother(swm_ctx* ctx, int fine_sweeps, int half_weights);

/* Stream overlap in the FCT step (default 3): the tracer transport, which
 * reads only the new mass and the mass flux, runs on a second stream of the
 * context beside the vorticity, divergence and elliptic solve, and the step
 * waits for it before the final halo fill.  mode 0: off; 1: started as soon
 * as the new mass is formed; n >= 2: started by the step's first V-cycle
 * when its down leg reaches level g - n + 1, where the cycle's launches
 * become latency-bound and leave bandwidth free (below level 4 it starts as
 * in mode 1).  The arithmetic is the same in every mode; at level 9 with one
 * tracer mode 3 saves about 2 % of either step, mode 1 under 1 %. */
swm_status swm_set_overlap(swm_ctx* ctx, int mode);

/* Mixed path: the coarse correction as one dense FP32 solve, alpha = M beta
 * with M the inverse of the level's Laplacian bordered by ones, formed once
 * per context in FP64, instead of the reference's V-cycle
 * (multigrid_solver.F:476-591):
 *   direct = 1 (default): of level 3 (642 cells, 1.6 MB; formed on the
 *              host);
 *   direct = 2: of level 4 (2562 cells, 26 MB; formed on the device at
 *              set-up), where level 5 exists, else as direct = 1; at level
 *              9 it takes the same cycles and time as direct = 1;
 *   direct = 0: the reference's V-cycle below level 4.
 * It only preconditions the FP64 defect correction, whose criterion is
 * unchanged.  Refused for FP64 contexts, whose multigrid cycle is the
 * reference's. */
swm_status swm_set_mixed_coarse(swm_ctx* ctx, int direct);

/* Kernel fusions of the multigrid cycle and the step, both paths (default
 * SWM_FUSE_ALL).
 * SWM_FUSE_INJECT_SWEEP: the restriction of the residual to a coarser level
 * that runs its own sweeps also performs that level's first smoothing sweep,
 * which starts from zero and so needs the injected right-hand side at its
 * own slot only; one launch per level and cycle fewer.  The results are the
 * same bit for bit with or without; the switch is there to show it.  No
 * counterpart in the reference (multigrid_solver.F:476-591 clears alpha,
 * sweeps, and injects, level by level).
 * SWM_FUSE_HALO_PUSH (FCT step): the kernels that form a field which the
 * step then fills the halo of also store its halo copies, from an inverted
 * fill table (the slot whose value a halo slot holds writes it there), so
 * that the step's halo fill launches are left out; the tracers' halo fill
 * runs at the end of their transport.  The same values land in the same
 * slots.  No counterpart in the reference (wrap_data.F:1790-2607 fills each
 * halo after the routine that formed the field).
 * SWM_FUSE_MASS_ANTI (FCT step): the final mass kernel corrects the
 * antidiffusive fluxes it reads (set_corrected_mss_anti, mass_fct.F:410-470)
 * instead of a pass of its own between the limiter and the final mass,
 * where the set-up finds the fill tables copy each halo edge from the
 * slot whose face value the halo holds (it does on the icosahedral grids).
 * The same values land in the same slots. */
#define SWM_FUSE_INJECT_SWEEP 1
#define SWM_FUSE_HALO_PUSH 2
#define SWM_FUSE_MASS_ANTI 4
#define SWM_FUSE_ALL 7
swm_status swm_set_fusion(swm_ctx* ctx, int flags);

/* Mean squared residuals at the end of the last elliptic solve, and mean
 * squared right-hand sides (relative criterion only), per field (psi, chi). */
swm_status swm_last_residual(const swm_ctx* ctx, double resid_ms[2], double rhs_ms[2]);

/* Elliptic solves, V-cycles (mixed path: correction cycles) and residual
 * checks since the last initialisation. */
swm_status swm_solver_stats(const swm_ctx* ctx, long long* solves, long long* vcycles, long long* checks);

/* Device time of the last swm_step call and of its hottest kernel, measured
 * with HIP events on the context's stream (ms), and that kernel's mean
 * duration and algorithmic bytes per launch. */
swm_status swm_timing(const swm_ctx* ctx, double* step_ms, double* kernel_ms_mean,
                      double* kernel_bytes, int* kernel_launches);

/* HIP-event timing of the finest-level multigrid smoother inside swm_step,
 * read back by swm_timing: two back-to-back sweeps (a level's smoothing
 * pair, neither from zero) between one pair of events, every such pair (1),
 * every n-th (n > 1: the event records cost the step a few microseconds
 * each), or off (0); the mean is per launch. */
swm_status swm_profile(swm_ctx* ctx, int enable);

/* Microbench of single reference operators on the device-resident state of
 * an initialised FP64 context (no host copies): op 0 div_of_an_edge_flux
 * (operators.F:270-392) of the mass flux, 1 curl_of_a_vector (operators.F:
 * 47-147) of the corner velocity, 2 laplacian_of_a_scalar (operators.F:
 * 402-518) of the thickness, 3 the face halo fill (wrap_data.F:1790-2607)
 * of one cell field; 4 and 5 the remapping's transport (locate_triangles and
 * transport_integrals, remap.F:2903-3963) and field (construct_fields,
 * remap.F:1803-2397) kernels of a remapping context after a step; 6 one
 * application of the dissipation operator (after swm_init_diss_op) to the
 * velocity.  Runs reps back-to-back launches of the kernel the operator
 * entry points run, between HIP events on the context's stream; returns the
 * mean microseconds per launch and the algorithmic bytes per launch (40,
 * 144 and 72 B per real cell; 16 B per ghost slot; 1128 + 48 ntracer,
 * 288 + 32 ntracer and 368 B per real cell).  The state is not changed.  No
 * counterpart in the reference. */
swm_status swm_operator_bench(swm_ctx* ctx, int op, int reps, double* us_per_launch, double* bytes_per_launch);

/* Seeded synthetic initial state built on the device: harmonic sums of
 * degree 1..lmax (coefficients N(0,1) l^-1.5, splitmix64 streams seeded
 * 20261203 + 100*seed_index + 1..4) for vorticity, divergence (10 % of its
 * RMS) and thickness (h0 plus a perturbation of RMS h_rms), corner velocity
 * from the model's discrete gradient scaled to max |u| = umax, tracers as
 * sums of nbumps Gaussian bumps of width bump_radius (m).  Then derives the
 * remaining fields as swm_init_state does.  No counterpart in the reference
 * (it runs fixed test cases, swm/source/initialize.F:330-696). */
swm_status swm_init_synthetic(swm_ctx* ctx, int seed_index, int lmax, double umax, double h0,
                              double h_rms, int nbumps, double bump_radius);

/* swm_init_synthetic plus a member perturbation: vorticity harmonics of the
 * same degrees from stream 20261203 + 100*seed_index + 1000 + member_seed,
 * scaled to vort_rel times the RMS of the vorticity and added through the
 * streamfunction before the velocity is formed and scaled to umax. */
swm_status swm_init_synthetic_member(swm_ctx* ctx, int seed_index, int lmax, double umax, double h0,
                                     double h_rms, int nbumps, double bump_radius, int member_seed,
                                     double vort_rel);

/* Global averages of budgets.F:31-247 compute_budgets at the current time
 * level, summed over real cells in FP64 with a compensated final sum:
 * out[0] mass, [1] kinetic energy (h ke), [2] potential energy (g h^2 / 2),
 * [3] topographic term (g h hs), [4] total energy ([1] + [2] + [3]),
 * [5] potential enstrophy (eta^2 / h), [6] absolute vorticity, [7]
 * divergence, [8 + q] mass-weighted variance of tracer q (the reference
 * reports tracers 1 and 2), each per unit area.  nout >= 8 + ntracer
 * (at most 16 tracers).  The reference's derived outputs (fractional
 * energy / enstrophy change, variance ratios) are taken against the caller's
 * t = 0 values. */
swm_status swm_budgets(swm_ctx* ctx, double* out, int nout);

/* Compensated (Neumaier) FP64 global mass sum of the current thickness,
 * sum(h * area) over real cells (swm/source/budgets.F:108-109). */
swm_status swm_total_mass(swm_ctx* ctx, double* mass);

/* ------------------------------------------------------------------------
 * Tangent-linear and adjoint model of a window of FP64 steps (step.F:
 * 114-212) with any of the three advection schemes: centred, upstream, or
 * flux-corrected transport (mass_fct.F) linearised with the limiter's
 * coefficients and the upstream directions frozen at the base trajectory.
 * Tracers are carried by the *_tracer entry points below.  Hand-written TL and
 * AD kernels cover every operator of the step; halo fills transpose to
 * accumulations into their sources; the elliptic solve's transpose reuses
 * the solver through the symmetry of area times Laplacian.  The trajectory
 * is kept as snapshots every `every` steps and recomputed one segment at a
 * time.  No counterpart in the reference.
 * ------------------------------------------------------------------------ */
/* Initial state from relative vorticity zeta, divergence delta and
 * thickness h (iim,jjm,10): (psi, chi) solve L psi = zeta, L chi = delta
 * (means removed), u = velocity(psi, chi) (operators.F:1402-1514), then the
 * state as swm_init_state derives it; hs stays as set.  The solves (and
 * those of the linearised steps) go to a relative residual of 1e-13 or to
 * the FP64 floor where that is higher (a cycle no longer halves the
 * residual). */
swm_status swm_init_vorticity(swm_ctx* ctx, const double* zeta, const double* delta, const double* h);

/* Runs nsteps steps from a freshly initialised state (step count 0),
 * keeping snapshots every `every` steps: the base of swm_tl_run /
 * swm_adj_run.  FP64, single-member contexts. */
swm_status swm_adj_forward(swm_ctx* ctx, int nsteps, int every);

/* Tangent-linear model of the window.  mode 0: in0..in3 the perturbation
 * (h, eta, div (iim,jjm,10), u (2,2,iim,jjm,10)) of the state read by the
 * first step; mode 1: in0..in2 the perturbation of the control (zeta,
 * delta, h) of swm_init_vorticity, in3 unused.  Out: the perturbation of the
 * state after the last step. */
swm_status swm_tl_run(swm_ctx* ctx, int mode, const double* in0, const double* in1, const double* in2,
                      const double* in3, double* out_h, double* out_eta, double* out_div, double* out_u);

/* Adjoint of the window: lh, leta, ldiv, lu, the adjoint (gradient) of the
 * state after the last step; mode 0: out0..out3 the adjoint of the state
 * read by the first step (h, eta, div, u); mode 1: out0..out2 the gradient
 * with respect to the control (zeta, delta, h), out3 unused. */
swm_status swm_adj_run(swm_ctx* ctx, int mode, const double* lh, const double* leta, const double* ldiv,
                       const double* lu, double* out0, double* out1, double* out2, double* out3);

/* What the linearised model differentiates in the elliptic solves of the
 * window's steps.  exact = 1 (default): the forward model as it runs, each
 * solve being the base step's V-cycles (their number as the base step's
 * convergence test chose it, elliptic_solver.F:120-180) from the previous
 * step's solution; the tangent-linear runs the same cycles on the
 * perturbations and carries the warm start's perturbation from step to step,
 * the adjoint runs the transposed cycle kernels in reverse order.  The
 * gradient is then the model's own, whatever its criterion (the reference's
 * takes 2-3 cycles).  exact = 0: the solves are taken as exact inverses of
 * the Laplacian, solved to a relative residual of 1e-13 (or the FP64 floor)
 * in both sweeps; the derivative of a forward that solves to that floor.
 * Levels 1..3 always use exact = 0. */
swm_status swm_set_linearisation(swm_ctx* ctx, int exact);

/* FCT windows: how many edges had a frozen limiter coefficient below 1
 * (the mass limiter of mass_fct.F:403-532 acting, where the frozen
 * linearisation is not the derivative), summed over the base steps that the
 * last swm_tl_run / swm_adj_run recomputed, and the number of edge-steps
 * (three edges per interior slot and step).  Both 0 for other schemes. */
swm_status swm_adj_limiter_stats(swm_ctx* ctx, long long* below_one, long long* edges);

/* The tracers in the linearised model.  Tracers do not feed back on the
 * dynamics, but depend on it through the mass flux and the thickness; these
 * entry points carry them through the window: the tracer transport of each
 * scheme (step.F:611-678 with tracer_fct.F, its limiter's coefficients
 * frozen at the base trajectory as for the mass; step.F:278-327 and the
 * tracer fluxes of operators.F:1841-1991 for the centred and upstream
 * schemes).  Tracer arrays are (iim,jjm,ntracer,10); the context must hold
 * tracers.  The plain swm_tl_run / swm_adj_run leave the tracers out. */
/* swm_init_vorticity with the initial tracers (both time levels) set to
 * `tracer` (NULL: zero, as swm_init_vorticity). */
swm_status swm_init_vorticity_tracer(swm_ctx* ctx, const double* zeta, const double* delta, const double* h,
                                     const double* tracer);
/* swm_tl_run with in_tr the tracers' perturbation (mode 0: the tracers read
 * by the first step; mode 1: the initial tracers of the control) and out_tr
 * the tracers' perturbation after the last step. */
swm_status swm_tl_run_tracer(swm_ctx* ctx, int mode, const double* in0, const double* in1, const double* in2,
                             const double* in3, const double* in_tr, double* out_h, double* out_eta,
                             double* out_div, double* out_u, double* out_tr);
/* swm_adj_run with ltr the adjoint of the tracers after the last step and
 * out_tr the adjoint of the tracers read by the first step (mode 0) or the
 * gradient with respect to the initial tracers (mode 1). */
swm_status swm_adj_run_tracer(swm_ctx* ctx, int mode, const double* lh, const double* leta, const double* ldiv,
                              const double* lu, const double* ltr, double* out0, double* out1, double* out2,
                              double* out3, double* out_tr);
/* FCT tracer runs: edges with a frozen tracer limiter coefficient below 1
 * (tracer_fct.F:539-593 acting: there the frozen linearisation is not the
 * derivative) over the base steps of the last tracer run, and the number
 * of edge-steps (three per interior slot, tracer and step). */
swm_status swm_adj_tracer_limiter_stats(swm_ctx* ctx, long long* below_one, long long* edges);

/* ------------------------------------------------------------------------
 * Restart files: the direct-access format of utilities/layer_2/io_utilities.F
 * as swm/source/swm_restart.F writes and reads it (FP64 values, one record
 * per global cell holding H, TRACER(ntracer) and U(2,2), a time stamp
 * record), interchangeable with the reference's files in both directions.
 * ------------------------------------------------------------------------ */
/* Model time in whole seconds, nsecond_dyn of time_manager_module.F: zero
 * after swm_init_state, the file's stamp after swm_read_restart, and
 * advanced by dt, truncated to an integer, every step
 * (time_manager_module.F:96).  -1 for a null context. */
long long swm_time_seconds(const swm_ctx* ctx);

/* write_swm_restart (swm/source/swm_restart.F:113-175): the current time
 * level and the model time.  path NULL writes the reference's file name for
 * the model time, "restarts/swm_restart.dDDDDDD.hh:mm:ss" (the directory
 * must exist). */
swm_status swm_write_restart(swm_ctx* ctx, const char* path);

/* read_swm_restart (swm/source/swm_restart.F:25-81): replaces the state with
 * the file's and derives vorticity, divergence and kinetic energy as
 * swm_init_state does; the surface height stays as set before (the
 * reference sets it from its initial conditions, swm/source/initialize.F:
 * 222-231), the model time takes the file's stamp and the step counter
 * restarts at zero, so the next step is a forward Euler step as in the
 * reference.  path NULL reads "restarts/swm_restart". */
swm_status swm_read_restart(swm_ctx* ctx, const char* path);

/* write_swm_restart's file name for a model time (swm_restart.F:122-145),
 * NUL-terminated into buf (len >= 38). */
swm_status swm_restart_name(long long nsecond, char* buf, int len);

/* The file codec alone, on host arrays in the reference layout (h
 * (iim,jjm,10), u (2,2,iim,jjm,10), tracer (iim,jjm,ntracer,10)); needs no
 * device.  The writer takes each global cell's values from its real slot
 * (grid_center_mask 1); the reader fills every slot, halos included, from
 * its global cell's record (grid_center_index), and accepts FP32 (real_bytes
 * 4) and non-restart files (time stamp 0).  A file of another grid or
 * tracer count is refused with SWM_ERR_DIMS. */
swm_status swm_restart_file_write(const swm_grid* grid, const char* path, int ntracer, long long nsecond,
                                  const double* h, const double* u, const double* tracer);
swm_status swm_restart_file_read(const swm_grid* grid, const char* path, int ntracer, long long* nsecond,
                                 double* h, double* u, double* tracer);

/* ------------------------------------------------------------------------
 * Quick-plot output (swm/source/swm_qp_module.F over
 * utilities/layer_1/netcdf_geodesic.F and netcdf_general_tools.F): one
 * netCDF classic file per run holding, per output time, the 11 fields
 * thickness, height, relative, divergence, absolute, kinetic_energy and
 * tracer_1..tracer_5 as FP32 at the real cells (grid_center_mask > 0.5,
 * panel-major), with the cells' centre and six corner coordinates, their
 * global numbers and the reference's dimensions, variables and attributes.
 * ------------------------------------------------------------------------ */
typedef struct swm_qp swm_qp;

/* The context's panel extents (iim, jjm) and tracer count. */
swm_status swm_ctx_layout(const swm_ctx* ctx, int* iim, int* jjm, int* ntracer);

/* The reference's file name for a run starting at model time nsecond on MPI
 * task `task` (swm_qp_module.F:130-136, :176):
 * "qp_output/hswm_dDDDDDDpTTT.g2.nc", NUL-terminated into buf (len >= 32). */
swm_status swm_qp_name(long long nsecond, int task, char* buf, int len);

/* initialize_swm_qp (swm_qp_module.F:70-214) with ginit_nc
 * (netcdf_geodesic.F:58-171): creates the file with the grid's coordinates
 * and no record.  simulation_tag is the "run" attribute (character*40, blanks
 * trimmed); path NULL uses swm_qp_name(nsecond, 0), whose directory must
 * exist.  swm_qp_create derives grid_center_lat/lon and grid_corner_lat/lon
 * from the grid (grid_metrics.F:194-206); swm_qp_create_coords takes them in
 * the reference layout, (iim, jjm, 10) and (2, iim, jjm, 10) radians. */
swm_status swm_qp_create(const swm_grid* grid, const char* simulation_tag, long long nsecond, const char* path,
                         swm_qp** out);
swm_status swm_qp_create_coords(const swm_grid* grid, const char* simulation_tag, long long nsecond,
                                const char* path, const double* center_lat, const double* center_lon,
                                const double* corner_lat, const double* corner_lon, swm_qp** out);

/* pack_swm_qp (swm_qp_module.F:309-370): the context's current time level
 * into the file's fields (h, h + hs, relative vorticity, divergence, absolute
 * vorticity, kinetic energy, tracers 1..min(ntracer, 5)), each FP64 value
 * rounded to FP32.  Fields with no model counterpart stay zero. */
swm_status swm_qp_pack(swm_qp* qp, swm_ctx* ctx);

/* qpsum in replace mode (qp_utilities.F) of one field by its qp name, from a
 * host array (iim, jjm, 10). */
swm_status swm_qp_pack_field(swm_qp* qp, const char* field, const double* x);

/* write_swm_qp (swm_qp_module.F:218-300) with gwrt_nc and set_time
 * (netcdf_general_tools.F:288-352): appends one record at model time nsecond
 * (time = the middle of the period since the previous write, start_period,
 * end_period, period_length in days, char_time "mm/dd/yyyy" in the no-leap
 * calendar), then zeroes the packed fields. */
swm_status swm_qp_write(swm_qp* qp, long long nsecond);

int swm_qp_records(const swm_qp* qp);
const char* swm_qp_path(const swm_qp* qp);
void swm_qp_destroy(swm_qp* qp);

/* ------------------------------------------------------------------------
 * Reference operators (utilities/layer_2/operators.F), one entry point per
 * subroutine with the subroutine's own argument list, in its order, plus
 * nsdm: the arrays hold nsdm panels (the reference calls the subroutines
 * once per panel inside its ksdm loops, i.e. nsdm = 1; the loop may be
 * replaced by one call with nsdm = 10).  Metric arrays are arguments, as in
 * the reference; every array is (leading dims, nx, ny, nsdm) in the
 * reference layout.  Host arrays are copied in, the operator runs on the
 * device, the result is copied out.  Results are set on i = 2..nx-1,
 * j = 2..ny-1 and are zero (ghost fill for average_to_cell_corners)
 * elsewhere, as in the reference.  The fixed counts must be those of the
 * model: npad 1, ndata 3, nspace 2, ncorners 2, nedges 3, nweights 7, r3 3
 * (else SWM_ERR_DIMS).
 * ------------------------------------------------------------------------ */
/* operators.F:47-147 */
swm_status swm_op_curl_of_a_vector(swm_ctx* ctx, int npad, int nx, int ny, int nsdm, int ndata, int nspace,
                                   int ncorners, const double* area_inv, const double* v_weights,
                                   const double* xin, double* xout);
/* operators.F:155-261 */
swm_status swm_op_div_of_a_vector(swm_ctx* ctx, int npad, int nx, int ny, int nsdm, int ndata, int nspace,
                                  int ncorners, const double* area_inv, const double* v_weights,
                                  const double* xin, double* xout);
/* operators.F:270-392 */
swm_status swm_op_div_of_an_edge_flux(swm_ctx* ctx, int npad, int nx, int ny, int nsdm, int nedges,
                                      const double* area_inv, const double* xin, double* xout);
/* operators.F:402-518 */
swm_status swm_op_laplacian_of_a_scalar(swm_ctx* ctx, int npad, int nx, int ny, int nsdm, int nweights,
                                        const double* l_weights, const double* xin, double* xout);
/* operators.F:528-582 (every slot) */
swm_status swm_op_k_cross_v(swm_ctx* ctx, int nx, int ny, int nsdm, int nspace, int ncorners,
                            const double* xin, double* xout);
/* operators.F:592-682 */
swm_status swm_op_gradient(swm_ctx* ctx, int npad, int nx, int ny, int nsdm, int ndata, int nspace,
                           int ncorners, const double* corner_area_inv, const double* v_weights,
                           const double* xin, double* xout);
/* operators.F:982-1112; flag is the logical flag (non-zero = .true.);
 * min_value: -1 = absent, 1 = present and .true. (0, present and .false.,
 * leaves the reference's result undefined and is refused) */
swm_status swm_op_average_to_cell_corners(swm_ctx* ctx, int npad, int nx, int ny, int nsdm, int ndata,
                                          int ncorners, const double* s_weights, const double* xin,
                                          double* xout, int flag, int min_value);
/* operators.F:1316-1393 */
swm_status swm_op_calculate_center_kinetic_energy(swm_ctx* ctx, int npad, int nx, int ny, int nsdm, int ndata,
                                                  int nspace, int ncorners, const double* sub_area,
                                                  const double* area_inv, const double* ke_c, double* ke);
/* operators.F:1402-1514 */
swm_status swm_op_calculate_velocity(swm_ctx* ctx, int npad, int nx, int ny, int nsdm, int ndata, int nspace,
                                     int ncorners, const double* corner_area_inv, const double* v_weights,
                                     const double* psi, const double* chi, double* xout);
/* operators.F:1626-1722 */
swm_status swm_op_centered_mass_flux(swm_ctx* ctx, int npad, int nx, int ny, int nsdm, int ndata, int ncorners,
                                     int nedges, int nspace, int r3, const double* dc, const double* normal,
                                     const double* mass_c, const double* u, const double* mass, double* flux);
/* operators.F:1731-1832 */
swm_status swm_op_upstream_mass_flux(swm_ctx* ctx, int npad, int nx, int ny, int nsdm, int ndata, int ncorners,
                                     int nedges, int nspace, int r3, const double* dc, const double* normal,
                                     const double* u, const double* mass, double* flux);
/* operators.F:2000-2074 */
swm_status swm_op_centered_vorticity_flux(swm_ctx* ctx, int npad, int nx, int ny, int nsdm, int nedges,
                                          const double* mass_flux, const double* mass, const double* xin,
                                          double* xout);
/* operators.F:2083-2186 */
swm_status swm_op_centered_divergence_curl(swm_ctx* ctx, int npad, int nx, int ny, int nsdm, int ndata,
                                           int ncorners, int nedges, int nspace, int r3, const double* dc,
                                           const double* tangent, const double* mass_c, const double* eta,
                                           const double* u, const double* mass, double* flux);

/* global_reductions::global_real_sum (utilities/layer_2/global_reductions.F:
 * 77-98): sum(x * mask) over n values (mask NULL: sum(x)).  The reference
 * sums in memory order and all-reduces over ranks; here FP64 block partials
 * and a compensated fixed-order final sum give the same result on every
 * call, within about one rounding of the exact sum.  n <= 2^30.
 * swm_global_real_sum takes host arrays and returns when *out holds the sum;
 * swm_global_real_sum_device takes device arrays and writes the sum to the
 * device double *out in order on the context's stream (block partials in
 * the context's buffers). */
swm_status swm_global_real_sum(swm_ctx* ctx, const double* x, const double* mask, int64_t n, double* out);
swm_status swm_global_real_sum_device(swm_ctx* ctx, const double* x, const double* mask, int64_t n, double* out);

/* swm_global_real_sum_device without allocation and without a context: no
 * counterpart in the reference; for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops).  x, mask and
 * out are device arrays on GPU `device`; the sum is written to *out in order
 * on `stream` (a hipStream_t; NULL: the default stream), with the same
 * kernels, so the same result bit for bit.  The block partials live in a
 * file-scope device array of 1024 doubles (8 KB per GPU, held until the
 * process ends), shared by every call on that GPU: calls that overlap in
 * time on one GPU race.  n <= 2^30 (larger n is refused with SWM_ERR_ARG).
 * Allocates nothing (no hipMalloc, hipFree or allocating library call). */
swm_status swm_global_real_sum_noalloc(const double* x, const double* mask, int64_t n, double* out, int device,
                                       void* stream);

/* mass_fct::set_mass_fct (utilities/layer_2/mass_fct.F:23-539) with its
 * own argument list: the whole sphere of the context's grid (iim, jjm of the
 * grid, nsdm 10, nlm 1, ntprog 2, nttend 3, the model's fixed counts), host
 * arrays in the reference layout.  mss (iim,jjm,1,2,10): level n3 read, level
 * n4 written; msh_f (3,iim,jjm,1,3,10) the AB3 edge-flux history, slot
 * tend_index(1) written, the slots combined with tend_weights; outputs
 * mss_low (iim,jjm,1,10) and mss_flux_horz (3,iim,jjm,1,10).  The optional
 * mss_flux_forced and mask (NULL when absent; step.F passes neither) are
 * refused.  Runs the step's mass-FCT kernels and halo fills on the device. */
swm_status swm_op_set_mass_fct(swm_ctx* ctx, int iim, int jjm, int nlm, int nsdm, int ntprog, int nttend,
                               int ncorners, int ndata, int nedges, int nspace, int npad, int r3, int n3, int n4,
                               const int* tend_index, const double* tend_weights, double dt, const double* area_inv,
                               const double* area, const double* dc, const double* normal, double* mss,
                               const double* mss_c, const double* u, double* mss_low, double* mss_flux_horz,
                               double* msh_f, const double* mss_flux_forced, const double* mask);

/* tracer_fct::set_tracer_fct (utilities/layer_2/tracer_fct.F:22-670) with its
 * own argument list, on the whole sphere (iim, jjm of the grid, nsdm 10, nlm
 * 1, nedges 3): trc, trc_high, mss, mss_low, mss_horz (iim,jjm,1,10) and
 * mss_flux_horz (3,iim,jjm,1,10) in, trc_fct (iim,jjm,1,10) out; the
 * optional outputs trc_flux_horz (3,iim,jjm,1,10) and msstrc_low_order
 * (iim,jjm,1,10) are written when not NULL; the optional mask (step.F does
 * not pass it) is refused. */
swm_status swm_op_set_tracer_fct(swm_ctx* ctx, int iim, int jjm, int nlm, int nsdm, int nedges, double dt,
                                 const double* area_inv, const double* area, const double* trc, const double* trc_high,
                                 const double* mss, const double* mss_low, const double* mss_horz,
                                 const double* mss_flux_horz, double* trc_fct, double* trc_flux_horz,
                                 double* msstrc_low_order, const double* mask);

/* init_diss_op (swm/source/swm_dissipation_operator.F:50-203): the
 * operator's geometry for the context from the grid it was made with. */
swm_status swm_init_diss_op(swm_ctx* ctx, const swm_grid* grid);

/* dissipation (swm/source/swm_dissipation_operator.F:212-443): norder
 * applications of the vector Laplacian (gradient of the edge divergence plus
 * k x gradient of the edge curl) to the corner vectors xin (2,2,iim,jjm,10),
 * each after a corner halo fill, times -visc for even norder and +visc for
 * odd, into xout (zero off the interior cells).  The reference computes it
 * for output only (driver.F:108-115, viscosity builds). */
swm_status swm_op_dissipation(swm_ctx* ctx, int norder, double visc, const double* xin, double* xout);

/* remap_mass_and_tracers (utilities/layer_2/remap.F:1072-1604) under its
 * mandatory arguments, in place on mass (iim,jjm,nmass,nsdm) and trace
 * (iim,jjm,ntrace,nmass,nsdm) with corner velocities uc (2,2,iim,jjm,nsdm);
 * edgeflux (1,3,iim,jjm,nsdm, may be NULL) receives the mass transport over
 * dt.  The optional arguments take the values get_remap_flux gives them
 * (step.F:726-730): linear integrals, first-order departure points, no
 * mask, tracers independent, their gradients limited with the switches of
 * swm_set_remap_limiter.  A context made with SWM_ADV_REMAP, the whole grid
 * (10 panels) and nmass = 1 (SWM_ERR_DIMS otherwise); SWM_ERR_STATE for a
 * mass below -1e-10, where remap_prep stops the reference. */
swm_status swm_op_remap_mass_and_tracers(swm_ctx* ctx, int iim, int jjm, int nmass, int nsdm, int ntrace,
                                         double* mass, double* trace, double dt, const double* uc, double* edgeflux);

/* Halo fill in place, wrap_data::wrap (utilities/layer_2/wrap_data.F:1790-1974):
 * kind 0 = face (cell data), 1 = vrtx (corner data, nspace components first,
 * 1..4), 2 = edge (nspace 1 or 2 components first); nlayer layers of
 * (iim,jjm,10) panels each (face), or (nspace,2,iim,jjm) resp.
 * (nspace,3,iim,jjm) per layer, as the reference's arrays are laid out. */
swm_status swm_wrap(swm_ctx* ctx, int kind, int nspace, int nlayer, double* x);

/* elliptic_solver_xxx::elliptic (utilities/layer_3/elliptic_solver.F:40-195):
 * rhs and xout are (iim,jjm,nvar,10); rhs loses its area-weighted mean, xout
 * is the warm start and the result. */
swm_status swm_elliptic(swm_ctx* ctx, int nvar, double* rhs, double* xout);

/* multigrid_solver_xxx::multigrid (utilities/layer_3/multigrid_solver.F:476-591):
 * one V-cycle on (iim,jjm,nvar,10) arrays. */
swm_status swm_multigrid(swm_ctx* ctx, int nvar, const double* beta, double* alpha);

#ifdef __cplusplus
}
#endif
#endif /* SWM_MI355X_H */
