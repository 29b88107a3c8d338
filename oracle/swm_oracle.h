/* swm_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU FP64 restatement of the reference shallow-water dynamical-core step
 * (swm/source/step.F step_dt + wrap_prognostics and everything they call),
 * used as the checker of the device path and as the timed CPU baseline.
 * Arrays are in the reference's Fortran layout (column-major, 1-cell halo,
 * 10 panels); time levels are kept as separate arrays.  Nothing in the shipped
 * library links or calls this code.
 */
#ifndef SWM_ORACLE_H
#define SWM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_model orc_model;

/* Grid description handed to the oracle (arrays are copied).
 * Layouts: cell (iim,jjm,10); corner (2,iim,jjm,10); (3,2,iim,jjm,10) for
 * dc/sub_area/s_weights; (3,3,2,iim,jjm,10) normal/tangent;
 * (3,2,2,iim,jjm,10) v_weights; (7,iim,jjm,10) l_weights.
 * Halo tables: flat (component, slot) destination/source pairs, source -1
 * meaning zero; ncomp 1 (cells), 2 (corners), 3 (edges).
 * lap_w[l-1]: (7, 2+2^l, 2+2^l, 10) multigrid Laplacian weights of level l. */
/* advection schemes of step_dt (step.F:131-135): l_fct, l_centered, l_upwind,
 * l_remap */
enum { ORC_FCT = 0, ORC_CENTERED = 1, ORC_UPWIND = 2, ORC_REMAP = 3 };

typedef struct {
  int g, ntracer;
  double dt, grav;
  int scheme;   /* ORC_FCT, ORC_CENTERED or ORC_UPWIND */
  const double *area, *area_inv, *corner_area_inv, *sub_area, *dc, *normal, *tangent,
      *v_weights, *s_weights, *l_weights, *f, *hs, *center_mask;
  int nface, nvrtx, nedge;
  const int32_t *face_dst, *face_src, *vrtx_dst, *vrtx_src, *edge_dst, *edge_src;
  const double* const* lap_w;
  double tot_area;   /* total area of the real cells (grid_metrics.F) */
} orc_grid;

/* Grid quantities of incremental remapping (init_remap, remap.F:176-1063)
 * in the reference layout, slot (iim,jjm,10) last: corners c (3,2,slot),
 * edge midpoints w (3,3,2,slot) and trans_to_X1 (3,3,2,slot) of the grid
 * metrics; trans_X1_to_X4 (3,3,3,slot); trans_X4_to_X3 (2,2,4,3,slot);
 * cl/cr/tl/tr/bl/br_x4 (2,3,slot); p_x4 (2,4,3,slot); c_x3 (2,6,slot);
 * gwgt_x3 (3,7,slot); the rest (slot).  radius: the sphere's radius a. */
typedef struct {
  double radius;
  const double *c, *w, *trans_to_x1;
  const double *trans_x1_to_x4, *trans_x4_to_x3, *cl_x4, *cr_x4, *tl_x4, *tr_x4, *bl_x4, *br_x4, *p_x4;
  const double *c_x3, *area_inv_x3, *xav, *yav, *xxav, *xyav, *yyav, *gwgt_x3, *garea_inv;
} orc_remap;

orc_model* orc_create(const orc_grid* grid);

/* The remapping quantities of an ORC_REMAP model (copied), and the tracer
 * limiter's switches l_limit_tracers_min / l_limit_tracers_max, which
 * remap_mass_and_tracers (remap.F:1072-1604) leaves without a value when
 * get_remap_flux (step.F:687-771) does not pass them: 1, 1 is the
 * two-sided van Leer limit of limit_gradient (remap.F:2152-2397).
 * orc_set_remap_limiter changes the switches between steps. */
void orc_set_remap(orc_model* m, const orc_remap* r, int limit_min, int limit_max);
void orc_set_remap_limiter(orc_model* m, int limit_min, int limit_max);
/* remap_mass_and_tracers (remap.F:1072-1604) alone, with get_remap_flux's
 * arguments and the model's dt: mass (slots) and trace (iim,jjm,ntracer,10)
 * in place, edgeflux (3,slots, may be NULL) = the mass transport over dt.
 * Returns 2 for a mass below -1e-10 (nothing written). */
int orc_remap_mass_and_tracers(orc_model* m, double* mass, double* trace, const double* uc, double* edgeflux);
void orc_destroy(orc_model* m);

/* Initial state from thickness h, surface height hs is part of the grid,
 * corner velocity u (2,2,iim,jjm,10) and tracers (iim,jjm,ntracer,10); the
 * derived vorticity, divergence and kinetic energy follow the reference IC
 * flow (initialize.F:418-468) and the first wrap_prognostics. */
void orc_init(orc_model* m, const double* h, const double* u, const double* tracer);

/* Advances nsteps steps (time_manager, get_pointers, get_weights, step_dt,
 * wrap_prognostics).  Returns 0, 1 if the elliptic solve did not converge,
 * 2 if remapping met a negative thickness (remap_prep stops there,
 * remap.F:1730-1734) or an ORC_REMAP model has no remapping quantities. */
int orc_step(orc_model* m, int nsteps);

/* Current-level fields: "h","eta","div","relative","ke","psi","chi" (cell),
 * "u" (2,2,iim,jjm,10), "tracer" (iim,jjm,ntracer,10). */
int orc_get(const orc_model* m, const char* name, double* out);

/* Global averages of budgets.F:31-247 compute_budgets at the current level:
 * out[0..7] = mass, kinetic, potential, topographic, total energy, potential
 * enstrophy, absolute vorticity, divergence; out[8+q] = mass-weighted
 * variance of tracer q (the reference reports tracers 1 and 2). */
void orc_budgets(const orc_model* m, double* out);

/* Global sums of the elliptic solve (mean removal, residual norm): mode 0
 * (default) in memory order as the reference's sum(X*MASK)
 * (elliptic_solver.F:80-92, :341-358, compiled as one sequential loop);
 * mode 1 compensated (Neumaier), within about one rounding of the exact sum
 * at any size.  The order's rounding error grows with the number of cells:
 * at 2.6 million cells it moves the mean by ~1e-13 of the field's scale,
 * which the solve and the velocity gradient turn into ~1e-12 of u. */
void orc_set_sum_mode(orc_model* m, int mode);

/* Host threads of the panel-parallel loops (OpenMP; results do not depend
 * on the count). */
void orc_set_threads(int n);

/* Number of V-cycles run by the last elliptic solve (2 + extra). */
int orc_last_vcycles(const orc_model* m);

/* Single operators on all 10 panels (reference argument meaning). */
void orc_div_of_an_edge_flux(const orc_model* m, const double* xin, double* xout);
void orc_laplacian(const orc_model* m, const double* xin, double* xout);
void orc_curl_of_a_vector(const orc_model* m, const double* xin, double* xout);
void orc_div_of_a_vector(const orc_model* m, const double* xin, double* xout);
void orc_center_ke(const orc_model* m, const double* ke_c, double* ke);
void orc_average_to_corners(const orc_model* m, const double* xin, double* xout);
void orc_velocity(const orc_model* m, const double* psi, const double* chi, double* u);
void orc_centered_mass_flux(const orc_model* m, const double* mass_c, const double* u,
                            double* flux);
void orc_upstream_mass_flux(const orc_model* m, const double* u, const double* mass,
                            double* flux);
void orc_vorticity_flux(const orc_model* m, const double* mass_flux, const double* mass,
                        const double* eta, double* flux);
void orc_divergence_curl(const orc_model* m, const double* mass_c, const double* eta,
                         const double* u, const double* mass, double* flux);
/* set_mass_fct (mass_fct.F:23-539) and set_tracer_fct (tracer_fct.F:22-670)
 * under their own arguments on the whole sphere; msh_f is the edge-flux
 * history, three (3, slots) arrays one after the other, tend_index 1-based */
void orc_set_mass_fct(orc_model* m, const int* tend_index, const double* tend_weights, const double* mss_n3,
                      const double* mss_c, const double* u, double* msh_f, double* mss_low, double* mss_flux,
                      double* mss_n4);
void orc_set_tracer_fct(orc_model* m, const double* trc, const double* trc_high, const double* mss,
                        const double* mss_low, const double* mss_horz, const double* mflux, double* trc_fct);
/* init_diss_op (swm_dissipation_operator.F:50-203) from the grid's cell
 * centres p (3,slot), corners c (3,2,slot), edge midpoints w (3,3,2,slot)
 * and corner frames trans_to_X2 (3,3,2,slot), and dissipation (:212-443)
 * of corner vectors xin -> xout (2,2,iim,jjm,10); 1 without orc_set_diss. */
void orc_set_diss(orc_model* m, const double* p, const double* c, const double* w, const double* trans_to_x2,
                  double radius);
int orc_dissipation(const orc_model* m, int norder, double visc, const double* xin, double* xout);
/* One multigrid V-cycle and one full elliptic solve on (iim,jjm,nvar,10). */
void orc_multigrid(orc_model* m, int nvar, const double* beta, double* alpha);
int orc_elliptic(orc_model* m, int nvar, double* rhs, double* xout);
/* Halo fills in place: kind 0 face (cell), 1 corner (nspace components
 * interleaved first), 2 edge. */
void orc_wrap(const orc_model* m, int kind, int nspace, double* x);

#ifdef __cplusplus
}
#endif
#endif
