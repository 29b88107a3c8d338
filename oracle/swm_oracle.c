/* swm_oracle.c -- TEST INFRASTRUCTURE ONLY: the oracle.
 *
 * Plain-C FP64 restatement of the reference dynamical-core step.  Every
 * routine follows the reference routine cited in its comment loop for loop,
 * with the same floating-point operation order, so that results agree with the
 * compiled reference to round-off (bit for bit where no library call or
 * summation order differs).  Index macros are 1-based like the reference.
 */
#include "swm_oracle.h"

#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <stdlib.h>
#include <string.h>

struct orc_model {
  int g, iim, jjm, nt;
  size_t ns;             /* slots = iim*jjm*10 */
  double dt, grav;
  int scheme;
  double tot_area;
  double *area, *area_inv, *cai, *sub_area, *dc, *normal, *tangent, *vw, *sw, *lw, *f, *hs, *mask;
  int nface, nvrtx, nedge;
  int32_t *face_dst, *face_src, *vrtx_dst, *vrtx_src, *edge_dst, *edge_src;
  double** lap_w;        /* per level 1..g */
  /* prognostic state: time levels as separate arrays */
  double *h[2], *eta[2], *div[2], **tr[2];
  double *eta_f[3], *div_f[3];
  double *u[2];          /* (2,2,slots) */
  double *relative, *ke, *psi, *chi, *inv_h;
  double *xout;          /* (slots, 2): elliptic warm start */
  double *msh_f[3];      /* (3,slots) edge-flux tendency history */
  double *h_f[3];        /* thickness tendency history (centred / upwind) */
  double **tr_f[3];      /* mass-times-tracer tendency history (centred / upwind) */
  double **work_old;     /* previous-previous tracer levels */
  double *mass_flux;     /* (3,slots) */
  int step_count, last_vcycles;
  double w1, w2, w3;
  int n3, n4, n1f, n2f, n3f;
  int sum_mode;          /* elliptic sums: 0 memory order (the reference), 1 compensated */
  orc_remap rm;          /* remapping quantities (ORC_REMAP) */
  int rm_owned, rm_limit_min, rm_limit_max;
  double *ds_t, *ds_magt, *ds_magn;   /* dissipation operator (orc_set_diss) */
};

#define S(i, j, k) ((size_t)((i)-1) + (size_t)iim * ((size_t)((j)-1) + (size_t)jjm * ((k)-1)))

static double* dup(const double* x, size_t n) {
  double* y = (double*)malloc(n * sizeof(double));
  if (x) memcpy(y, x, n * sizeof(double));
  else memset(y, 0, n * sizeof(double));
  return y;
}
static int32_t* dupi(const int32_t* x, size_t n) {
  int32_t* y = (int32_t*)malloc((n ? n : 1) * sizeof(int32_t));
  if (n) memcpy(y, x, n * sizeof(int32_t));
  return y;
}
static double* zeros(size_t n) { return (double*)calloc(n, sizeof(double)); }

/* ---------------------------------------------------------------- halo --- */
/* wrap_data.F:1790-2513 as the gather it amounts to; x holds nspace values per
 * (component, slot) entry */
static void halo(const int32_t* dst, const int32_t* src, int n, double* x, int nspace) {
  for (int k = 0; k < n; ++k)
    for (int s = 0; s < nspace; ++s)
      x[(size_t)nspace * dst[k] + s] = src[k] < 0 ? 0.0 : x[(size_t)nspace * src[k] + s];
}
static void wrap_face(const orc_model* m, double* x) { halo(m->face_dst, m->face_src, m->nface, x, 1); }
static void wrap_vrtx(const orc_model* m, double* x, int nspace) { halo(m->vrtx_dst, m->vrtx_src, m->nvrtx, x, nspace); }
static void wrap_edge(const orc_model* m, double* x) { halo(m->edge_dst, m->edge_src, m->nedge, x, 1); }

void orc_wrap(const orc_model* m, int kind, int nspace, double* x) {
  if (kind == 0) wrap_face(m, x);
  else if (kind == 1) wrap_vrtx(m, x, nspace);
  else wrap_edge(m, x);
}

/* ----------------------------------------------------------- operators --- */
/* operators.F:270-392 div_of_an_edge_flux */
static void div_edge(int iim, int jjm, const double* area_inv, const double* x, double* out) {
#define X3(e, i, j, k) x[(e)-1 + 3 * S(i, j, k)]
  #pragma omp parallel for schedule(static)
  for (int k = 1; k <= 10; ++k) {
    for (int j = 1; j <= jjm; ++j)
      for (int i = 1; i <= iim; ++i) out[S(i, j, k)] = 0.0;
    for (int j = 2; j <= jjm - 1; ++j)
      for (int i = 2; i <= iim - 1; ++i)
        out[S(i, j, k)] = (X3(1, i, j, k) + X3(2, i, j, k) + X3(3, i, j, k) - X3(1, i + 1, j, k) -
                           X3(2, i + 1, j + 1, k) - X3(3, i, j + 1, k)) * area_inv[S(i, j, k)];
    int i = 1, j = jjm - 1;
    out[S(i, j, k)] = (-X3(2, i, j, k) - X3(3, i, j, k) - X3(1, i + 1, j, k) - X3(2, i + 1, j + 1, k) -
                       X3(3, i, j + 1, k)) * area_inv[S(i, j, k)];
    i = iim - 1; j = 1;
    out[S(i, j, k)] = (-X3(1, i, j, k) - X3(2, i, j, k) - X3(1, i + 1, j, k) - X3(2, i + 1, j + 1, k) -
                       X3(3, i, j + 1, k)) * area_inv[S(i, j, k)];
  }
#undef X3
}

/* operators.F:402-518 laplacian_of_a_scalar */
static void laplacian(int iim, int jjm, const double* lw, const double* x, double* out) {
#define L(n, i, j, k) lw[(n)-1 + 7 * S(i, j, k)]
#define XX(i, j, k) x[S(i, j, k)]
  #pragma omp parallel for schedule(static)
  for (int k = 1; k <= 10; ++k) {
    for (int j = 1; j <= jjm; ++j)
      for (int i = 1; i <= iim; ++i) out[S(i, j, k)] = 0.0;
    for (int j = 2; j <= jjm - 1; ++j)
      for (int i = 2; i <= iim - 1; ++i)
        out[S(i, j, k)] = +L(7, i, j, k) * XX(i, j, k) + L(1, i, j, k) * XX(i - 1, j, k) +
                          L(2, i, j, k) * XX(i - 1, j - 1, k) + L(3, i, j, k) * XX(i, j - 1, k) +
                          L(4, i, j, k) * XX(i + 1, j, k) + L(5, i, j, k) * XX(i + 1, j + 1, k) +
                          L(6, i, j, k) * XX(i, j + 1, k);
    int i = 1, j = jjm - 1;
    out[S(i, j, k)] = +L(7, i, j, k) * XX(i, j, k) + L(2, i, j, k) * XX(iim, 1, k) +
                      L(3, i, j, k) * XX(i, j - 1, k) + L(4, i, j, k) * XX(i + 1, j, k) +
                      L(5, i, j, k) * XX(i + 1, j + 1, k) + L(6, i, j, k) * XX(i, j + 1, k);
    i = iim - 1; j = 1;
    out[S(i, j, k)] = +L(7, i, j, k) * XX(i, j, k) + L(1, i, j, k) * XX(i - 1, j, k) +
                      L(2, i, j, k) * XX(1, jjm, k) + L(4, i, j, k) * XX(i + 1, j, k) +
                      L(5, i, j, k) * XX(i + 1, j + 1, k) + L(6, i, j, k) * XX(i, j + 1, k);
  }
#undef L
#undef XX
}

#define VW(kk, sp, ic, i, j, k) vw[(kk)-1 + 3 * ((sp)-1 + 2 * ((ic)-1 + 2 * S(i, j, k)))]
#define UU(sp, ic, i, j, k) xin[(sp)-1 + 2 * ((ic)-1 + 2 * S(i, j, k))]

/* operators.F:47-147 curl_of_a_vector (computed from i,j = 1 as the reference) */
static void curl_vec(int iim, int jjm, const double* area_inv, const double* vw, const double* xin,
                     double* out) {
  #pragma omp parallel for schedule(static)
  for (int k = 1; k <= 10; ++k) {
    for (int j = 1; j <= jjm; ++j)
      for (int i = 1; i <= iim; ++i) out[S(i, j, k)] = 0.0;
    for (int j = 1; j <= jjm - 1; ++j)
      for (int i = 1; i <= iim - 1; ++i)
        out[S(i, j, k)] =
            (+VW(1, 2, 1, i, j, k) * UU(1, 1, i, j, k) - VW(1, 1, 1, i, j, k) * UU(2, 1, i, j, k) +
             VW(1, 2, 2, i, j, k) * UU(1, 2, i, j, k) - VW(1, 1, 2, i, j, k) * UU(2, 2, i, j, k) +
             VW(2, 2, 1, i + 1, j, k) * UU(1, 1, i + 1, j, k) - VW(2, 1, 1, i + 1, j, k) * UU(2, 1, i + 1, j, k) +
             VW(2, 2, 2, i + 1, j + 1, k) * UU(1, 2, i + 1, j + 1, k) -
             VW(2, 1, 2, i + 1, j + 1, k) * UU(2, 2, i + 1, j + 1, k) +
             VW(3, 2, 1, i + 1, j + 1, k) * UU(1, 1, i + 1, j + 1, k) -
             VW(3, 1, 1, i + 1, j + 1, k) * UU(2, 1, i + 1, j + 1, k) +
             VW(3, 2, 2, i, j + 1, k) * UU(1, 2, i, j + 1, k) - VW(3, 1, 2, i, j + 1, k) * UU(2, 2, i, j + 1, k)) *
            area_inv[S(i, j, k)];
  }
}

/* operators.F:155-261 div_of_a_vector */
static void div_vec(int iim, int jjm, const double* area_inv, const double* vw, const double* xin,
                    double* out) {
  #pragma omp parallel for schedule(static)
  for (int k = 1; k <= 10; ++k) {
    for (int j = 1; j <= jjm; ++j)
      for (int i = 1; i <= iim; ++i) out[S(i, j, k)] = 0.0;
    for (int j = 1; j <= jjm - 1; ++j)
      for (int i = 1; i <= iim - 1; ++i)
        out[S(i, j, k)] =
            (-VW(1, 1, 1, i, j, k) * UU(1, 1, i, j, k) - VW(1, 2, 1, i, j, k) * UU(2, 1, i, j, k) -
             VW(1, 1, 2, i, j, k) * UU(1, 2, i, j, k) - VW(1, 2, 2, i, j, k) * UU(2, 2, i, j, k) -
             VW(2, 1, 1, i + 1, j, k) * UU(1, 1, i + 1, j, k) - VW(2, 2, 1, i + 1, j, k) * UU(2, 1, i + 1, j, k) -
             VW(2, 1, 2, i + 1, j + 1, k) * UU(1, 2, i + 1, j + 1, k) -
             VW(2, 2, 2, i + 1, j + 1, k) * UU(2, 2, i + 1, j + 1, k) -
             VW(3, 1, 1, i + 1, j + 1, k) * UU(1, 1, i + 1, j + 1, k) -
             VW(3, 2, 1, i + 1, j + 1, k) * UU(2, 1, i + 1, j + 1, k) -
             VW(3, 1, 2, i, j + 1, k) * UU(1, 2, i, j + 1, k) - VW(3, 2, 2, i, j + 1, k) * UU(2, 2, i, j + 1, k)) *
            area_inv[S(i, j, k)];
  }
}
#undef UU

/* operators.F:1316-1393 calculate_center_kinetic_energy */
static void center_ke(int iim, int jjm, const double* sub_area, const double* area_inv,
                      const double* kec, double* ke) {
#define SA(kk, ic, i, j, k) sub_area[(kk)-1 + 3 * ((ic)-1 + 2 * S(i, j, k))]
#define KC(ic, i, j, k) kec[(ic)-1 + 2 * S(i, j, k)]
  #pragma omp parallel for schedule(static)
  for (int k = 1; k <= 10; ++k) {
    for (int j = 1; j <= jjm; ++j)
      for (int i = 1; i <= iim; ++i) ke[S(i, j, k)] = 0.0;
    for (int j = 1; j <= jjm - 1; ++j)
      for (int i = 1; i <= iim - 1; ++i)
        ke[S(i, j, k)] = (+SA(1, 1, i, j, k) * KC(1, i, j, k) + SA(1, 2, i, j, k) * KC(2, i, j, k) +
                          SA(2, 1, i + 1, j, k) * KC(1, i + 1, j, k) +
                          SA(2, 2, i + 1, j + 1, k) * KC(2, i + 1, j + 1, k) +
                          SA(3, 1, i + 1, j + 1, k) * KC(1, i + 1, j + 1, k) +
                          SA(3, 2, i, j + 1, k) * KC(2, i, j + 1, k)) *
                         area_inv[S(i, j, k)];
  }
#undef SA
#undef KC
}

/* operators.F:982-1112 average_to_cell_corners (flag = .true.) */
static void avg_corners(int iim, int jjm, const double* sw, const double* x, double* out) {
#define SW(kk, ic, i, j, k) sw[(kk)-1 + 3 * ((ic)-1 + 2 * S(i, j, k))]
#define O2(ic, i, j, k) out[(ic)-1 + 2 * S(i, j, k)]
  #pragma omp parallel for schedule(static)
  for (int k = 1; k <= 10; ++k) {
    const double fill = x[S(2, 2, k)];
    for (int j = 1; j <= jjm; ++j)
      for (int i = 1; i <= iim; ++i)
        if (j == 1 || j >= jjm - 1 || i == 1 || i >= iim - 1) { O2(1, i, j, k) = fill; O2(2, i, j, k) = fill; }
    for (int j = 2; j <= jjm - 1; ++j)
      for (int i = 2; i <= iim - 1; ++i) {
        O2(1, i, j, k) = +SW(1, 1, i, j, k) * x[S(i, j, k)] + SW(2, 1, i, j, k) * x[S(i - 1, j, k)] +
                         SW(3, 1, i, j, k) * x[S(i - 1, j - 1, k)];
        O2(2, i, j, k) = +SW(1, 2, i, j, k) * x[S(i, j, k)] + SW(2, 2, i, j, k) * x[S(i - 1, j - 1, k)] +
                         SW(3, 2, i, j, k) * x[S(i, j - 1, k)];
      }
  }
#undef SW
#undef O2
}

/* operators.F:592-682 gradient (corner vector of a cell scalar) */
static void gradient(int iim, int jjm, const double* cai, const double* vw, const double* x, double* out) {
#define G2(sp, ic, i, j, k) out[(sp)-1 + 2 * ((ic)-1 + 2 * S(i, j, k))]
  memset(out, 0, sizeof(double) * 4 * (size_t)iim * jjm * 10);
  #pragma omp parallel for schedule(static)
  for (int k = 1; k <= 10; ++k)
    for (int j = 2; j <= jjm - 1; ++j)
      for (int i = 2; i <= iim - 1; ++i)
        for (int sp = 1; sp <= 2; ++sp) {
          G2(sp, 1, i, j, k) = (+VW(1, sp, 1, i, j, k) * x[S(i, j, k)] + VW(2, sp, 1, i, j, k) * x[S(i - 1, j, k)] +
                                VW(3, sp, 1, i, j, k) * x[S(i - 1, j - 1, k)]) *
                               cai[0 + 2 * S(i, j, k)];
          G2(sp, 2, i, j, k) = (+VW(1, sp, 2, i, j, k) * x[S(i, j, k)] + VW(2, sp, 2, i, j, k) * x[S(i - 1, j - 1, k)] +
                                VW(3, sp, 2, i, j, k) * x[S(i, j - 1, k)]) *
                               cai[1 + 2 * S(i, j, k)];
        }
#undef G2
}

/* operators.F:1402-1514 calculate_velocity: grad(chi) + k x grad(psi) */
static void velocity(int iim, int jjm, const double* cai, const double* vw, const double* psi,
                     const double* chi, double* u) {
  const size_t n = 4 * (size_t)iim * jjm * 10;
  double* w1 = zeros(n);
  gradient(iim, jjm, cai, vw, chi, u);
  gradient(iim, jjm, cai, vw, psi, w1);
  for (size_t q = 0; q < n; q += 2) {
    const double kx = -w1[q + 1], ky = +w1[q];   /* k_cross_v (operators.F:528-582) */
    u[q] = u[q] + kx;
    u[q + 1] = u[q + 1] + ky;
  }
  free(w1);
}
#undef VW

#define NRM(kk, a, ic, i, j, k) nrm[(kk)-1 + 3 * ((a)-1 + 3 * ((ic)-1 + 2 * S(i, j, k)))]
#define DC(kk, ic, i, j, k) dc[(kk)-1 + 3 * ((ic)-1 + 2 * S(i, j, k))]
#define UV(sp, ic, i, j, k) u[(sp)-1 + 2 * ((ic)-1 + 2 * S(i, j, k))]
#define MC(ic, i, j, k) mc[(ic)-1 + 2 * S(i, j, k)]
#define FL(e, i, j, k) flux[(e)-1 + 3 * S(i, j, k)]

/* the six half-edge normal (or tangential) velocity fluxes of cell (i,j) as
 * combined by centered_mass_flux (operators.F:1626-1722); mc == NULL drops the
 * corner mass factor (upstream_mass_flux, operators.F:1731-1832) */
static void half_edges(int iim, int jjm, const double* dc, const double* nrm, const double* mc,
                       const double* u, int i, int j, int k, double r[2][3]) {
#define HALF(c1, ii, jj, kkk, ic)                                                            \
  ((UV(1, ic, ii, jj, k) * NRM(kkk, 1, ic, ii, jj, k) + UV(2, ic, ii, jj, k) * NRM(kkk, 2, ic, ii, jj, k)) * \
   (mc ? MC(ic, ii, jj, k) : 1.0))
  if (mc) {
    r[0][0] = HALF(0, i, j + 1, 2, 2); r[0][0] = r[0][0] * DC(2, 2, i, j + 1, k);
    r[1][0] = HALF(0, i, j, 1, 1);     r[1][0] = r[1][0] * DC(1, 1, i, j, k);
    r[0][1] = HALF(0, i, j, 3, 1);     r[0][1] = r[0][1] * DC(3, 1, i, j, k);
    r[1][1] = HALF(0, i, j, 1, 2);     r[1][1] = r[1][1] * DC(1, 2, i, j, k);
    r[0][2] = HALF(0, i, j, 3, 2);     r[0][2] = r[0][2] * DC(3, 2, i, j, k);
    r[1][2] = HALF(0, i + 1, j, 2, 1); r[1][2] = r[1][2] * DC(2, 1, i + 1, j, k);
  } else {
#define HN(ii, jj, kkk, ic) (UV(1, ic, ii, jj, k) * NRM(kkk, 1, ic, ii, jj, k) + UV(2, ic, ii, jj, k) * NRM(kkk, 2, ic, ii, jj, k))
    r[0][0] = HN(i, j + 1, 2, 2); r[0][0] = r[0][0] * DC(2, 2, i, j + 1, k);
    r[1][0] = HN(i, j, 1, 1);     r[1][0] = r[1][0] * DC(1, 1, i, j, k);
    r[0][1] = HN(i, j, 3, 1);     r[0][1] = r[0][1] * DC(3, 1, i, j, k);
    r[1][1] = HN(i, j, 1, 2);     r[1][1] = r[1][1] * DC(1, 2, i, j, k);
    r[0][2] = HN(i, j, 3, 2);     r[0][2] = r[0][2] * DC(3, 2, i, j, k);
    r[1][2] = HN(i + 1, j, 2, 1); r[1][2] = r[1][2] * DC(2, 1, i + 1, j, k);
#undef HN
  }
#undef HALF
}

static void centered_mass_flux(int iim, int jjm, const double* dc, const double* nrm, const double* mc,
                               const double* u, double* flux) {
  memset(flux, 0, sizeof(double) * 3 * (size_t)iim * jjm * 10);
  #pragma omp parallel for schedule(static)
  for (int k = 1; k <= 10; ++k)
    for (int j = 2; j <= jjm - 1; ++j)
      for (int i = 2; i <= iim - 1; ++i) {
        double r[2][3];
        half_edges(iim, jjm, dc, nrm, mc, u, i, j, k, r);
        FL(1, i, j, k) = -r[0][0] + r[1][0];
        FL(2, i, j, k) = -r[0][1] + r[1][1];
        FL(3, i, j, k) = -r[0][2] + r[1][2];
      }
}

static void upstream_mass_flux(int iim, int jjm, const double* dc, const double* nrm, const double* u,
                               const double* mass, double* flux) {
  memset(flux, 0, sizeof(double) * 3 * (size_t)iim * jjm * 10);
  #pragma omp parallel for schedule(static)
  for (int k = 1; k <= 10; ++k)
    for (int j = 2; j <= jjm - 1; ++j)
      for (int i = 2; i <= iim - 1; ++i) {
        double r[2][3];
        half_edges(iim, jjm, dc, nrm, NULL, u, i, j, k, r);
        double v[3];
        for (int e = 0; e < 3; ++e) v[e] = r[1][e] - r[0][e];
        /* maxval/minval over (v, 0) */
        const size_t up[3] = {S(i - 1, j, k), S(i - 1, j - 1, k), S(i, j - 1, k)};
        for (int e = 0; e < 3; ++e) {
          const double mx = v[e] > 0.0 ? v[e] : 0.0, mn = v[e] < 0.0 ? v[e] : 0.0;
          FL(e + 1, i, j, k) = mx * mass[S(i, j, k)] + mn * mass[up[e]];
        }
      }
}

/* operators.F:2000-2074 centered_vorticity_flux */
static void vorticity_flux(int iim, int jjm, const double* mf, const double* mass, const double* eta,
                           double* flux) {
  const size_t ns = (size_t)iim * jjm * 10;
  double* pv = zeros(ns);
  for (size_t n = 0; n < ns; ++n) {
    const double w = mass[n] > 0.0 ? 1.0 / mass[n] : 0.0;
    pv[n] = eta[n] * w;
  }
  memset(flux, 0, sizeof(double) * 3 * ns);
  #pragma omp parallel for schedule(static)
  for (int k = 1; k <= 10; ++k)
    for (int j = 2; j <= jjm - 1; ++j)
      for (int i = 2; i <= iim - 1; ++i) {
        FL(1, i, j, k) = 0.5 * (mf[0 + 3 * S(i, j, k)]) * (pv[S(i, j, k)] + pv[S(i - 1, j, k)]);
        FL(2, i, j, k) = 0.5 * (mf[1 + 3 * S(i, j, k)]) * (pv[S(i, j, k)] + pv[S(i - 1, j - 1, k)]);
        FL(3, i, j, k) = 0.5 * (mf[2 + 3 * S(i, j, k)]) * (pv[S(i, j, k)] + pv[S(i, j - 1, k)]);
      }
  free(pv);
}

/* operators.F:2083-2186 centered_divergence_curl */
static void divergence_curl(int iim, int jjm, const double* dc, const double* tng, const double* mc,
                            const double* eta, const double* u, const double* mass, double* flux) {
  const size_t ns = (size_t)iim * jjm * 10;
  double* pv = zeros(ns);
  for (size_t n = 0; n < ns; ++n) {
    const double w = mass[n] > 0.0 ? 1.0 / mass[n] : 0.0;
    pv[n] = eta[n] * w;
  }
  memset(flux, 0, sizeof(double) * 3 * ns);
  #pragma omp parallel for schedule(static)
  for (int k = 1; k <= 10; ++k)
    for (int j = 2; j <= jjm - 1; ++j)
      for (int i = 2; i <= iim - 1; ++i) {
        double r[2][3];
        half_edges(iim, jjm, dc, tng, mc, u, i, j, k, r);
        FL(1, i, j, k) = (-r[0][0] + r[1][0]) * (pv[S(i - 1, j, k)] + pv[S(i, j, k)]) * 0.5;
        FL(2, i, j, k) = (-r[0][1] + r[1][1]) * (pv[S(i - 1, j - 1, k)] + pv[S(i, j, k)]) * 0.5;
        FL(3, i, j, k) = (-r[0][2] + r[1][2]) * (pv[S(i, j - 1, k)] + pv[S(i, j, k)]) * 0.5;
      }
  free(pv);
}
#undef NRM
#undef DC
#undef UV
#undef MC
#undef FL

/* operators.F:1841-1914 upstream_tracer_flux */
static void upstream_tracer_flux(int iim, int jjm, const double* mf, const double* x, double* out) {
  const size_t n = (size_t)iim * jjm * 10;
  for (size_t q = 0; q < 3 * n; ++q) out[q] = 0.0;
  #pragma omp parallel for schedule(static)
  for (int k = 1; k <= 10; ++k)
    for (int j = 2; j <= jjm - 1; ++j)
      for (int i = 2; i <= iim - 1; ++i) {
        const size_t s = S(i, j, k);
        const size_t up[3] = {S(i - 1, j, k), S(i - 1, j - 1, k), S(i, j - 1, k)};
        for (int e = 0; e < 3; ++e) {
          const double F = mf[e + 3 * s];
          const double mx = F > 0.0 ? F : 0.0, mn = F < 0.0 ? F : 0.0;   /* maxval / minval of (F, 0) */
          out[e + 3 * s] = mx * x[s] + mn * x[up[e]];
        }
      }
}

/* operators.F:1923-1991 centered_tracer_flux */
static void centered_tracer_flux(int iim, int jjm, const double* mf, const double* x, double* out) {
  const size_t n = (size_t)iim * jjm * 10;
  for (size_t q = 0; q < 3 * n; ++q) out[q] = 0.0;
  #pragma omp parallel for schedule(static)
  for (int k = 1; k <= 10; ++k)
    for (int j = 2; j <= jjm - 1; ++j)
      for (int i = 2; i <= iim - 1; ++i) {
        const size_t s = S(i, j, k);
        const size_t up[3] = {S(i - 1, j, k), S(i - 1, j - 1, k), S(i, j - 1, k)};
        for (int e = 0; e < 3; ++e) out[e + 3 * s] = 0.5 * (mf[e + 3 * s]) * (x[s] + x[up[e]]);
      }
}

/* ------------------------------------------------------------ mass FCT --- */
static inline double dmin(double a, double b) { return a < b ? a : b; }
static inline double dmax(double a, double b) { return a > b ? a : b; }

#define E3(a, e, i, j, k) a[(e)-1 + 3 * S(i, j, k)]

/* low-order or flux-corrected update of one cell field from edge fluxes, with
 * the pole forms (mass_fct.F:350-445 set_mass_low, :478-532
 * set_flux_corrected_mass; tracer_fct.F:251-321, :601-665) */
static void edge_update(const orc_model* m, const double* base, const double* F, double* out) {
  const int iim = m->iim, jjm = m->jjm;
  const double dt = m->dt;
  memset(out, 0, sizeof(double) * m->ns);
  #pragma omp parallel for schedule(static)
  for (int k = 1; k <= 10; ++k) {
    int i = 1, j = jjm - 1;
    out[S(i, j, k)] = base[S(i, j, k)] + dt * m->area_inv[S(i, j, k)] *
                                             (E3(F, 2, i, j, k) + E3(F, 3, i, j, k) + E3(F, 1, i + 1, j, k) +
                                              E3(F, 2, i + 1, j + 1, k) + E3(F, 3, i, j + 1, k));
    i = iim - 1; j = 1;
    out[S(i, j, k)] = base[S(i, j, k)] + dt * m->area_inv[S(i, j, k)] *
                                             (E3(F, 1, i, j, k) + E3(F, 2, i, j, k) + E3(F, 1, i + 1, j, k) +
                                              E3(F, 2, i + 1, j + 1, k) + E3(F, 3, i, j + 1, k));
    for (j = 2; j <= jjm - 1; ++j)
      for (i = 2; i <= iim - 1; ++i)
        out[S(i, j, k)] = base[S(i, j, k)] -
                          dt * m->area_inv[S(i, j, k)] *
                              ((E3(F, 1, i, j, k) + E3(F, 2, i, j, k) + E3(F, 3, i, j, k)) -
                               (E3(F, 1, i + 1, j, k) + E3(F, 2, i + 1, j + 1, k) + E3(F, 3, i, j + 1, k)));
  }
}

/* mass_fct.F:23-539 set_mass_fct; updates mss_n4, mss_low, mss_flux (=F1+A) */
static void mass_fct(orc_model* m, const double* mss_n3, const double* mss_c, const double* u,
                     double* mss_low, double* mss_flux, double* mss_n4) {
  const int iim = m->iim, jjm = m->jjm;
  const size_t ns = m->ns, ne = 3 * ns;
  const int ti[3] = {m->n3f, m->n2f, m->n1f};
  const double tw[3] = {m->w3, m->w2, m->w1};
  double *flx1 = zeros(ne), *flx2 = zeros(ne), *anti = zeros(ne), *beta = zeros(ns);
  upstream_mass_flux(iim, jjm, m->dc, m->normal, u, mss_n3, flx1);
  wrap_edge(m, flx1);
  centered_mass_flux(iim, jjm, m->dc, m->normal, mss_c, u, m->msh_f[ti[0] - 1]);
  wrap_edge(m, m->msh_f[ti[0] - 1]);
  for (int t = 0; t < 3; ++t)
    #pragma omp parallel for schedule(static)
    for (size_t q = 0; q < ne; ++q) flx2[q] = flx2[q] + tw[t] * m->msh_f[ti[t] - 1][q];
  #pragma omp parallel for schedule(static)
  for (size_t q = 0; q < ne; ++q) anti[q] = flx2[q] - flx1[q];
  edge_update(m, mss_n3, flx1, mss_low);
  /* set_beta_dn (mass_fct.F:366-402): the interior loop starts at i,j = 1
   * and so overwrites the pole values computed first */
  const double mss_min = 1.0e-6, small = 1.0e-18, dt = m->dt;
  #pragma omp parallel for schedule(static)
  for (int k = 1; k <= 10; ++k) {
    int i = 1, j = jjm - 1;
    double a_out = -(dmin(0.0, E3(anti, 2, i, j, k)) + dmin(0.0, E3(anti, 3, i, j, k)) +
                     dmin(0.0, E3(anti, 1, i + 1, j, k)) + dmin(0.0, E3(anti, 2, i + 1, j + 1, k)) +
                     dmin(0.0, E3(anti, 3, i, j + 1, k)));
    beta[S(i, j, k)] = m->area[S(i, j, k)] * (mss_low[S(i, j, k)] - mss_min) / (dt * (a_out + small));
    i = iim - 1; j = 1;
    a_out = -(dmin(0.0, E3(anti, 1, i, j, k)) + dmin(0.0, E3(anti, 2, i, j, k)) +
              dmin(0.0, E3(anti, 1, i + 1, j, k)) + dmin(0.0, E3(anti, 2, i + 1, j + 1, k)) +
              dmin(0.0, E3(anti, 3, i, j + 1, k)));
    beta[S(i, j, k)] = m->area[S(i, j, k)] * (mss_low[S(i, j, k)] - mss_min) / (dt * (a_out + small));
    for (j = 1; j <= jjm - 1; ++j)
      for (i = 1; i <= iim - 1; ++i) {
        a_out = (dmax(0.0, E3(anti, 1, i, j, k)) + dmax(0.0, E3(anti, 2, i, j, k)) +
                 dmax(0.0, E3(anti, 3, i, j, k))) -
                (dmin(0.0, E3(anti, 1, i + 1, j, k)) + dmin(0.0, E3(anti, 2, i + 1, j + 1, k)) +
                 dmin(0.0, E3(anti, 3, i, j + 1, k)));
        beta[S(i, j, k)] = m->area[S(i, j, k)] * (mss_low[S(i, j, k)] - mss_min) / (dt * (a_out + small));
      }
  }
  wrap_face(m, beta);
  /* set_corrected_mss_anti (mass_fct.F:410-470) */
  #pragma omp parallel for schedule(static)
  for (int k = 1; k <= 10; ++k)
    for (int j = 2; j <= jjm - 1; ++j)
      for (int i = 2; i <= iim - 1; ++i) {
        const double b = dmin(1.0, beta[S(i, j, k)]);
        double a = E3(anti, 1, i, j, k);
        E3(anti, 1, i, j, k) = b * dmax(0.0, a) + dmin(1.0, beta[S(i - 1, j, k)]) * dmin(0.0, a);
        a = E3(anti, 2, i, j, k);
        E3(anti, 2, i, j, k) = b * dmax(0.0, a) + dmin(1.0, beta[S(i - 1, j - 1, k)]) * dmin(0.0, a);
        a = E3(anti, 3, i, j, k);
        E3(anti, 3, i, j, k) = b * dmax(0.0, a) + dmin(1.0, beta[S(i, j - 1, k)]) * dmin(0.0, a);
      }
  wrap_edge(m, anti);
  #pragma omp parallel for schedule(static)
  for (size_t q = 0; q < ne; ++q) mss_flux[q] = flx1[q] + anti[q];
  edge_update(m, mss_low, anti, mss_n4);
  free(flx1); free(flx2); free(anti); free(beta);
}

/* ---------------------------------------------------------- tracer FCT --- */
/* tracer_fct.F:22-670 set_tracer_fct; trc and mss_horz are wrapped in place
 * as the reference does */
static void tracer_fct(orc_model* m, double* trc, const double* trc_high, const double* mss,
                       const double* mss_low, double* mss_horz, const double* mflux, double* trc_fct) {
  const int iim = m->iim, jjm = m->jjm;
  const size_t ns = m->ns, ne = 3 * ns;
  const double dt = m->dt;
  double *f1 = zeros(ne), *f2 = zeros(ne), *anti = zeros(ne);
  double *ml = zeros(ns), *mf = zeros(ns), *tmin = zeros(ns), *tmax = zeros(ns), *tlow = zeros(ns);
  double *bdn = zeros(ns), *bup = zeros(ns);
  /* set_tracer_flux_horizontal (:166-243) */
  #pragma omp parallel for schedule(static)
  for (int k = 1; k <= 10; ++k)
    for (int j = 2; j <= jjm - 1; ++j)
      for (int i = 2; i <= iim - 1; ++i) {
        const size_t up[3] = {S(i - 1, j, k), S(i - 1, j - 1, k), S(i, j - 1, k)};
        for (int e = 0; e < 3; ++e) {
          const double F = mflux[e + 3 * S(i, j, k)];
          f1[e + 3 * S(i, j, k)] = trc[S(i, j, k)] * dmax(0.0, F) + trc[up[e]] * dmin(0.0, F);
        }
      }
  wrap_edge(m, f1);
  #pragma omp parallel for schedule(static)
  for (int k = 1; k <= 10; ++k)
    for (int j = 2; j <= jjm - 1; ++j)
      for (int i = 2; i <= iim - 1; ++i) {
        const size_t up[3] = {S(i - 1, j, k), S(i - 1, j - 1, k), S(i, j - 1, k)};
        for (int e = 0; e < 3; ++e)
          f2[e + 3 * S(i, j, k)] = 0.5 * (trc_high[S(i, j, k)] + trc_high[up[e]]) * mflux[e + 3 * S(i, j, k)];
      }
  wrap_edge(m, f2);
  #pragma omp parallel for schedule(static)
  for (size_t q = 0; q < ne; ++q) anti[q] = f2[q] - f1[q];
  /* set_msstrc_low (:251-321) */
  {
    double* base = zeros(ns);
    #pragma omp parallel for schedule(static)
    for (size_t n = 0; n < ns; ++n) base[n] = mss[n] * trc[n];
    edge_update(m, base, f1, ml);
    free(base);
  }
  /* set_limiter (:329-412) */
  wrap_face(m, trc);
  wrap_face(m, mss_horz);
  wrap_face(m, ml);
  #pragma omp parallel for schedule(static)
  for (size_t n = 0; n < ns; ++n) tlow[n] = mss_horz[n] > 0.0 ? ml[n] / mss_horz[n] : 0.0;
  #pragma omp parallel for schedule(static)
  for (int k = 1; k <= 10; ++k) {
    const int p = iim - 1, q = jjm - 1;
    double v12[12];
    {
      const size_t s0 = S(1, q, k);
      const double t[12] = {trc[s0], tlow[s0], trc[S(p, 1, k)], trc[S(1, q - 1, k)], trc[S(2, q, k)],
                            trc[S(2, q + 1, k)], trc[S(1, q + 1, k)], tlow[S(p, 1, k)], tlow[S(1, q - 1, k)],
                            tlow[S(2, q, k)], tlow[S(2, q + 1, k)], tlow[S(1, q + 1, k)]};
      memcpy(v12, t, sizeof t);
      double mn = v12[0], mx = v12[0];
      for (int n = 1; n < 12; ++n) { mn = dmin(mn, v12[n]); mx = dmax(mx, v12[n]); }
      tmin[s0] = mn; tmax[s0] = mx;
    }
    {
      const size_t s0 = S(p, 1, k);
      const double t[12] = {trc[s0], tlow[s0], trc[S(p - 1, 1, k)], trc[S(1, q, k)], trc[S(p + 1, 1, k)],
                            trc[S(p + 1, 2, k)], trc[S(p, 2, k)], tlow[S(p - 1, 1, k)], tlow[S(1, q, k)],
                            tlow[S(p + 1, 1, k)], tlow[S(p + 1, 2, k)], tlow[S(p, 2, k)]};
      double mn = t[0], mx = t[0];
      for (int n = 1; n < 12; ++n) { mn = dmin(mn, t[n]); mx = dmax(mx, t[n]); }
      tmin[s0] = mn; tmax[s0] = mx;
    }
    for (int j = 2; j <= jjm - 1; ++j)
      for (int i = 2; i <= iim - 1; ++i) {
        const size_t nb[7] = {S(i, j, k), S(i - 1, j, k), S(i - 1, j - 1, k), S(i, j - 1, k),
                              S(i + 1, j, k), S(i + 1, j + 1, k), S(i, j + 1, k)};
        double mn = trc[nb[0]], mx = trc[nb[0]];
        mn = dmin(mn, tlow[nb[0]]); mx = dmax(mx, tlow[nb[0]]);
        for (int n = 1; n < 7; ++n) { mn = dmin(mn, trc[nb[n]]); mx = dmax(mx, trc[nb[n]]); }
        for (int n = 1; n < 7; ++n) { mn = dmin(mn, tlow[nb[n]]); mx = dmax(mx, tlow[nb[n]]); }
        tmin[nb[0]] = mn; tmax[nb[0]] = mx;
      }
  }
  /* set_beta_weights_trc (:420-531) */
  const double small = 1.0e-18;
  #pragma omp parallel for schedule(static)
  for (int k = 1; k <= 10; ++k) {
    int i = 1, j = jjm - 1;
    for (int pole = 0; pole < 2; ++pole) {
      if (pole == 1) { i = iim - 1; j = 1; }
      double e[5];
      if (pole == 0) {
        e[0] = E3(anti, 2, i, j, k); e[1] = E3(anti, 3, i, j, k); e[2] = E3(anti, 1, i + 1, j, k);
        e[3] = E3(anti, 2, i + 1, j + 1, k); e[4] = E3(anti, 3, i, j + 1, k);
      } else {
        e[0] = E3(anti, 1, i, j, k); e[1] = E3(anti, 2, i, j, k); e[2] = E3(anti, 1, i + 1, j, k);
        e[3] = E3(anti, 2, i + 1, j + 1, k); e[4] = E3(anti, 3, i, j + 1, k);
      }
      const double a_out = -(dmin(0.0, e[0]) + dmin(0.0, e[1]) + dmin(0.0, e[2]) + dmin(0.0, e[3]) + dmin(0.0, e[4]));
      const double a_in = (dmax(0.0, e[0]) + dmax(0.0, e[1]) + dmax(0.0, e[2]) + dmax(0.0, e[3]) + dmax(0.0, e[4]));
      const size_t s0 = S(i, j, k);
      bdn[s0] = (m->area[s0] / dt) * (ml[s0] - mss_low[s0] * tmin[s0]) / (a_out + small);
      bup[s0] = (m->area[s0] / dt) * (mss_low[s0] * tmax[s0] - ml[s0]) / (a_in + small);
    }
    for (j = 2; j <= jjm - 1; ++j)
      for (i = 2; i <= iim - 1; ++i) {
        const double o1 = E3(anti, 1, i, j, k), o2 = E3(anti, 2, i, j, k), o3 = E3(anti, 3, i, j, k);
        const double n1 = E3(anti, 1, i + 1, j, k), n2 = E3(anti, 2, i + 1, j + 1, k), n3 = E3(anti, 3, i, j + 1, k);
        const double a_out = (dmax(0.0, o1) + dmax(0.0, o2) + dmax(0.0, o3)) - (dmin(0.0, n1) + dmin(0.0, n2) + dmin(0.0, n3));
        const double a_in = -((dmin(0.0, o1) + dmin(0.0, o2) + dmin(0.0, o3)) - (dmax(0.0, n1) + dmax(0.0, n2) + dmax(0.0, n3)));
        const size_t s0 = S(i, j, k);
        bdn[s0] = (m->area[s0] / dt) * (ml[s0] - mss_low[s0] * tmin[s0]) / (a_out + small);
        bup[s0] = (m->area[s0] / dt) * (mss_low[s0] * tmax[s0] - ml[s0]) / (a_in + small);
      }
  }
  wrap_face(m, bdn);
  wrap_face(m, bup);
  /* set_corrected_trc_anti (:539-593) */
  #pragma omp parallel for schedule(static)
  for (int k = 1; k <= 10; ++k)
    for (int j = 2; j <= jjm - 1; ++j)
      for (int i = 2; i <= iim - 1; ++i) {
        const size_t s0 = S(i, j, k);
        const size_t up[3] = {S(i - 1, j, k), S(i - 1, j - 1, k), S(i, j - 1, k)};
        for (int e = 0; e < 3; ++e) {
          const double a = anti[e + 3 * s0];
          anti[e + 3 * s0] = dmin(dmin(1.0, bdn[s0]), bup[up[e]]) * dmax(0.0, a) +
                             dmin(dmin(1.0, bup[s0]), bdn[up[e]]) * dmin(0.0, a);
        }
      }
  wrap_edge(m, anti);
  edge_update(m, ml, anti, mf);
  #pragma omp parallel for schedule(static)
  for (size_t n = 0; n < ns; ++n) trc_fct[n] = mss_horz[n] != 0.0 ? mf[n] / mss_horz[n] : 0.0;
  free(f1); free(f2); free(anti); free(ml); free(mf); free(tmin); free(tmax); free(tlow); free(bdn); free(bup);
}
#undef E3

/* mass_fct.F:23-539 set_mass_fct under its own arguments: the AB3 slots and
 * weights of the edge-flux history given (tend_index, 1-based; tend_weights),
 * the history msh_f as three (3, slots) arrays one after the other (the slot
 * tend_index[0] is written), outputs the low-order mass, the total mass flux
 * and the new mass.  The model's own history and weights are left as they
 * were. */
void orc_set_mass_fct(orc_model* m, const int* tend_index, const double* tend_weights, const double* mss_n3,
                      const double* mss_c, const double* u, double* msh_f, double* mss_low, double* mss_flux,
                      double* mss_n4) {
  const size_t ne = 3 * m->ns;
  double* keep[3];
  const int f3 = m->n3f, f2 = m->n2f, f1 = m->n1f;
  const double w3 = m->w3, w2 = m->w2, w1 = m->w1;
  for (int t = 0; t < 3; ++t) {
    keep[t] = dup(m->msh_f[t], ne);
    memcpy(m->msh_f[t], msh_f + t * ne, ne * sizeof(double));
  }
  m->n3f = tend_index[0]; m->n2f = tend_index[1]; m->n1f = tend_index[2];
  m->w3 = tend_weights[0]; m->w2 = tend_weights[1]; m->w1 = tend_weights[2];
  mass_fct(m, mss_n3, mss_c, u, mss_low, mss_flux, mss_n4);
  for (int t = 0; t < 3; ++t) {
    memcpy(msh_f + t * ne, m->msh_f[t], ne * sizeof(double));
    memcpy(m->msh_f[t], keep[t], ne * sizeof(double));
    free(keep[t]);
  }
  m->n3f = f3; m->n2f = f2; m->n1f = f1;
  m->w3 = w3; m->w2 = w2; m->w1 = w1;
}

/* tracer_fct.F:22-670 set_tracer_fct under its own arguments (trc and
 * mss_horz are read as given; their halo fills act on copies) */
void orc_set_tracer_fct(orc_model* m, const double* trc, const double* trc_high, const double* mss,
                        const double* mss_low, const double* mss_horz, const double* mflux, double* trc_fct) {
  double* t = dup(trc, m->ns);
  double* h = dup(mss_horz, m->ns);
  tracer_fct(m, t, trc_high, mss, mss_low, h, mflux, trc_fct);
  free(t);
  free(h);
}

/* ------------------------------------------------------------ multigrid -- */
/* multigrid_solver.F:1105-1158 wrap_simple, closed form for 10 panels */
static void wrap_simple(int lev, double* x, int nvar) {
  const int iim = 2 + (1 << lev), jjm = iim, p = iim - 1, q = jjm - 1;
  const size_t nl = (size_t)iim * jjm * 10;
  for (int v = 0; v < nvar; ++v) {
    double* y = x + v * nl;
#define Y(i, j, k) y[S(i, j, k)]
    for (int nsd = 3; nsd <= 9; nsd += 2) Y(1, q, nsd) = Y(1, q, 1);
    for (int nsd = 2; nsd <= 8; nsd += 2) Y(p, 1, nsd) = Y(p, 1, 10);
    for (int nsd = 1; nsd <= 9; nsd += 2) {
      const int a = (nsd + 7) % 10 + 1, b = (nsd + 8) % 10 + 1, c = nsd % 10 + 1, d = (nsd + 1) % 10 + 1;
      for (int j = 1; j <= q - 1; ++j) Y(1, j, nsd) = Y(p - (j - 1), q, a);
      for (int i = 2; i <= p; ++i) Y(i, 1, nsd) = Y(i, q, b);
      for (int j = 2; j <= q; ++j) Y(p + 1, j, nsd) = Y(2, j, c);
      for (int i = 2; i <= p; ++i) Y(i, q + 1, nsd) = Y(2, q - (i - 2), d);
    }
    for (int nsd = 2; nsd <= 10; nsd += 2) {
      const int a = (nsd + 8) % 10 + 1, b = (nsd + 7) % 10 + 1, c = (nsd + 1) % 10 + 1, d = nsd % 10 + 1;
      for (int j = 2; j <= q; ++j) Y(1, j, nsd) = Y(p, j, a);
      for (int i = 1; i <= p - 1; ++i) Y(i, 1, nsd) = Y(p, q - (i - 1), b);
      for (int j = 2; j <= q; ++j) Y(p + 1, j, nsd) = Y(p - (j - 2), 2, c);
      for (int i = 2; i <= p; ++i) Y(i, q + 1, nsd) = Y(i, 2, d);
    }
    Y(1, q + 1, 1) = Y(2, q, 5);
    Y(p + 1, 1, 1) = Y(2, q, 7);
    Y(p + 1, 1, 10) = Y(p, 2, 4);
    Y(1, q + 1, 10) = Y(p, 2, 6);
#undef Y
  }
}

/* multigrid_solver.F:593-682 relax (weighted Jacobi, omega = 0.8) */
static void relax(int iters, int lev, const double* lw, const double* beta, double* alpha, int nvar) {
  const int iim = 2 + (1 << lev), jjm = iim;
  const size_t nl = (size_t)iim * jjm * 10;
  const double omega = 0.8;
  double* lwt = zeros(nl);
  double* temp = zeros(nl * nvar);
  for (size_t n = 0; n < nl; ++n) lwt[n] = lw[6 + 7 * n] != 0.0 ? 1.0 / (lw[6 + 7 * n] - 0.0) : 0.0;
#define LW(n, i, j, k) lw[(n)-1 + 7 * S(i, j, k)]
  for (int it = 0; it < iters; ++it) {
    wrap_simple(lev, alpha, nvar);
    for (int v = 0; v < nvar; ++v) {
      const double* a = alpha + v * nl;
      const double* b = beta + v * nl;
      double* t = temp + v * nl;
#define A(i, j, k) a[S(i, j, k)]
      #pragma omp parallel for schedule(static)
      for (int k = 1; k <= 10; ++k) {
        for (int j = 2; j <= jjm - 1; ++j)
          for (int i = 2; i <= iim - 1; ++i) {
            const double tt = lwt[S(i, j, k)] * (b[S(i, j, k)] - LW(1, i, j, k) * A(i - 1, j, k) -
                                                 LW(2, i, j, k) * A(i - 1, j - 1, k) - LW(3, i, j, k) * A(i, j - 1, k) -
                                                 LW(4, i, j, k) * A(i + 1, j, k) - LW(5, i, j, k) * A(i + 1, j + 1, k) -
                                                 LW(6, i, j, k) * A(i, j + 1, k));
            t[S(i, j, k)] = (1.0 - omega) * A(i, j, k) + omega * tt;
          }
        int i = 1, j = jjm - 1;
        double tt = lwt[S(i, j, k)] * (b[S(i, j, k)] - LW(2, i, j, k) * A(iim, 1, k) - LW(3, i, j, k) * A(i, j - 1, k) -
                                       LW(4, i, j, k) * A(i + 1, j, k) - LW(5, i, j, k) * A(i + 1, j + 1, k) -
                                       LW(6, i, j, k) * A(i, j + 1, k));
        t[S(i, j, k)] = (1.0 - omega) * A(i, j, k) + omega * tt;
        i = iim - 1; j = 1;
        tt = lwt[S(i, j, k)] * (b[S(i, j, k)] - LW(1, i, j, k) * A(i - 1, j, k) - LW(2, i, j, k) * A(1, jjm, k) -
                                LW(4, i, j, k) * A(i + 1, j, k) - LW(5, i, j, k) * A(i + 1, j + 1, k) -
                                LW(6, i, j, k) * A(i, j + 1, k));
        t[S(i, j, k)] = (1.0 - omega) * A(i, j, k) + omega * tt;
      }
#undef A
    }
    memcpy(alpha, temp, sizeof(double) * nl * nvar);
  }
  free(lwt);
  free(temp);
}

/* multigrid_solver.F:684-886 inject_residual: residual at the odd-odd points
 * of the fine level is the right-hand side of the next coarser level */
static void inject_residual(int lev, const double* lw, double* xa2, const double* xb2, double* x1, int nvar) {
  const int iim2 = 2 + (1 << lev), jjm2 = iim2, iim1 = 2 + (1 << (lev - 1)), jjm1 = iim1;
  const size_t nl2 = (size_t)iim2 * jjm2 * 10, nl1 = (size_t)iim1 * jjm1 * 10;
  wrap_simple(lev, xa2, nvar);
  for (int v = 0; v < nvar; ++v) {
    double* a = xa2 + v * nl2;
    const double* b = xb2 + v * nl2;
    double* c = x1 + v * nl1;
    #pragma omp parallel for schedule(static)
    for (int k = 1; k <= 10; ++k) {
      int iim = iim2, jjm = jjm2;
#define A(i, j) a[S(i, j, k)]
      for (int j2 = 3; j2 <= jjm2 - 1; j2 += 2)
        for (int i2 = 3; i2 <= iim2 - 1; i2 += 2) {
          const double r = b[S(i2, j2, k)] -
                           ((LW(7, i2, j2, k) - 0.0) * A(i2, j2) + LW(1, i2, j2, k) * A(i2 - 1, j2) +
                            LW(2, i2, j2, k) * A(i2 - 1, j2 - 1) + LW(3, i2, j2, k) * A(i2, j2 - 1) +
                            LW(4, i2, j2, k) * A(i2 + 1, j2) + LW(5, i2, j2, k) * A(i2 + 1, j2 + 1) +
                            LW(6, i2, j2, k) * A(i2, j2 + 1));
          iim = iim1; jjm = jjm1;
          c[S((i2 + 1) / 2, (j2 + 1) / 2, k)] = r;
          iim = iim2; jjm = jjm2;
        }
      int i2 = 1, j2 = jjm2 - 1;
      const double rn = b[S(i2, j2, k)] -
                        ((LW(7, i2, j2, k) - 0.0) * A(i2, j2) + LW(2, i2, j2, k) * A(iim2, 1) +
                         LW(3, i2, j2, k) * A(i2, j2 - 1) + LW(4, i2, j2, k) * A(i2 + 1, j2) +
                         LW(5, i2, j2, k) * A(i2 + 1, j2 + 1) + LW(6, i2, j2, k) * A(i2, j2 + 1));
      i2 = iim2 - 1; j2 = 1;
      const double rs = b[S(i2, j2, k)] -
                        ((LW(7, i2, j2, k) - 0.0) * A(i2, j2) + LW(1, i2, j2, k) * A(i2 - 1, j2) +
                         LW(2, i2, j2, k) * A(1, jjm2) + LW(4, i2, j2, k) * A(i2 + 1, j2) +
                         LW(5, i2, j2, k) * A(i2 + 1, j2 + 1) + LW(6, i2, j2, k) * A(i2, j2 + 1));
#undef A
      iim = iim1; jjm = jjm1;
      if (k == 1) c[S(1, jjm1 - 1, k)] = rn;
      if (k == 10) c[S(iim1 - 1, 1, k)] = rs;
    }
  }
}

/* multigrid_solver.F:888-1054 interpolate: coarse correction added to the
 * fine level at coincident points, two-point averages between them */
static void interpolate(int lev, double* x1, double* x2, int nvar) {
  const int iim1 = 2 + (1 << lev), jjm1 = iim1, iim2 = 2 + (1 << (lev + 1)), jjm2 = iim2;
  const size_t nl1 = (size_t)iim1 * jjm1 * 10, nl2 = (size_t)iim2 * jjm2 * 10;
  wrap_simple(lev, x1, nvar);
  double* temp = zeros(nl2);
  for (int v = 0; v < nvar; ++v) {
    const double* c = x1 + v * nl1;
    double* f = x2 + v * nl2;
    memset(temp, 0, sizeof(double) * nl2);
    #pragma omp parallel for schedule(static)
    for (int k = 1; k <= 10; ++k) {
      int iim = iim1, jjm = jjm1;
      for (int j1 = 1; j1 <= iim1 - 1; ++j1)
        for (int i1 = 1; i1 <= iim1 - 1; ++i1) {
          const double val = c[S(i1, j1, k)];
          iim = iim2; jjm = jjm2;
          temp[S(2 * i1 - 1, 2 * j1 - 1, k)] = val;
          iim = iim1; jjm = jjm1;
        }
    }
    int iim = iim2, jjm = jjm2;
    const int p2 = iim2 - 1, q2 = jjm2 - 1;
    #pragma omp parallel for schedule(static)
    for (int k = 1; k <= 10; ++k) {
      for (int j = 3; j <= jjm2 - 1; j += 2)
        for (int i = 3; i <= iim2 - 1; i += 2) {
          f[S(i, j, k)] = f[S(i, j, k)] + temp[S(i, j, k)];
          f[S(i - 1, j, k)] = f[S(i - 1, j, k)] + 0.5 * (temp[S(i, j, k)] + temp[S(i - 2, j, k)]);
          f[S(i - 1, j - 1, k)] = f[S(i - 1, j - 1, k)] + 0.5 * (temp[S(i, j, k)] + temp[S(i - 2, j - 2, k)]);
          f[S(i, j - 1, k)] = f[S(i, j - 1, k)] + 0.5 * (temp[S(i, j, k)] + temp[S(i, j - 2, k)]);
        }
      f[S(1, q2, k)] = f[S(1, q2, k)] + temp[S(1, q2, k)];
      f[S(p2, 1, k)] = f[S(p2, 1, k)] + temp[S(p2, 1, k)];
    }
  }
  free(temp);
}
#undef LW

/* multigrid_solver.F:476-591 one V-cycle on nvar fields of the finest level */
void orc_multigrid(orc_model* m, int nvar, const double* beta, double* alpha) {
  const int G = m->g;
  double* aw[16] = {0};
  double* bw[16] = {0};
  for (int l = 1; l <= G; ++l) {
    const size_t nl = (size_t)(2 + (1 << l)) * (2 + (1 << l)) * 10;
    aw[l] = zeros(nl * nvar);
    bw[l] = zeros(nl * nvar);
  }
  memcpy(aw[G], alpha, sizeof(double) * m->ns * nvar);
  memcpy(bw[G], beta, sizeof(double) * m->ns * nvar);
  for (int l = G; l >= 2; --l) {
    relax(2, l, m->lap_w[l - 1], bw[l], aw[l], nvar);
    inject_residual(l, m->lap_w[l - 1], aw[l], bw[l], bw[l - 1], nvar);
  }
  relax(24, 1, m->lap_w[0], bw[1], aw[1], nvar);
  for (int l = 1; l <= G - 1; ++l) {
    interpolate(l, aw[l], aw[l + 1], nvar);
    relax(2, l + 1, m->lap_w[l], bw[l + 1], aw[l + 1], nvar);
  }
  memcpy(alpha, aw[G], sizeof(double) * m->ns * nvar);
  for (int l = 1; l <= G; ++l) { free(aw[l]); free(bw[l]); }
}

/* Neumaier's compensated running sum (sum_mode 1): s + c is the sum of the
 * added terms to about one rounding, whatever their count */
static void neumaier_add(double* s, double* c, double v) {
  const double t = *s + v;
  if (fabs(*s) >= fabs(v)) *c += (*s - t) + v;
  else *c += (v - t) + *s;
  *s = t;
}

/* elliptic_solver.F:203-239 check_convergence + :341-358 l2_norm */
static double l2_resid(orc_model* m, const double* rhs, const double* x) {
  const size_t ns = m->ns;
  double* w = zeros(ns);
  laplacian(m->iim, m->jjm, m->lw, x, w);
  double num = 0.0, cnt = 0.0, comp = 0.0;
  for (size_t n = 0; n < ns; ++n) {
    const double r = (w[n] - 0.0 * x[n]) - rhs[n];
    if (m->sum_mode) neumaier_add(&num, &comp, (r * r) * m->mask[n]);
    else num = num + (r * r) * m->mask[n];
  }
  num = num + comp;
  for (size_t n = 0; n < ns; ++n) cnt = cnt + m->mask[n] * m->mask[n];
  free(w);
  const int isum = (int)cnt;
  return num / (double)isum;
}

/* elliptic_solver.F:40-195 elliptic: mean removal, two V-cycles, then up to
 * ten more for each field that has not reached mean(resid^2) <= 1e-20 */
void orc_set_sum_mode(orc_model* m, int mode) { m->sum_mode = mode ? 1 : 0; }

void orc_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n > 0 ? n : 1);
#else
  (void)n;
#endif
}

int orc_elliptic(orc_model* m, int nvar, double* rhs, double* xout) {
  const size_t ns = m->ns;
  for (int v = 0; v < nvar; ++v) { wrap_face(m, xout + v * ns); }
  for (int v = 0; v < nvar; ++v) { wrap_face(m, rhs + v * ns); }
  double total = 0.0, ct = 0.0;
  for (size_t n = 0; n < ns; ++n) {
    if (m->sum_mode) neumaier_add(&total, &ct, (m->area[n] * m->mask[n]) * 1.0);
    else total = total + (m->area[n] * m->mask[n]) * 1.0;
  }
  total = total + ct;
  for (int v = 0; v < nvar; ++v) {
    double s = 0.0, cs = 0.0;
    for (size_t n = 0; n < ns; ++n) {
      if (m->sum_mode) neumaier_add(&s, &cs, rhs[v * ns + n] * (m->area[n] * m->mask[n]));
      else s = s + rhs[v * ns + n] * (m->area[n] * m->mask[n]);
    }
    s = s + cs;
    const double xmean = s / total;
    for (size_t n = 0; n < ns; ++n) rhs[v * ns + n] = rhs[v * ns + n] - xmean;
  }
  for (int it = 0; it < 2; ++it) orc_multigrid(m, nvar, rhs, xout);
  m->last_vcycles = 2;
  for (int v = 0; v < nvar; ++v) wrap_face(m, xout + v * ns);
  /* the fields still above the criterion after the first two cycles all get
   * the further cycles together, until none is above it */
  int bad[8], nbad = 0;
  for (int v = 0; v < nvar; ++v)
    if (l2_resid(m, rhs + v * ns, xout + v * ns) > 1.0e-20) bad[nbad++] = v;
  int iter = 0, any = nbad > 0;
  while (any && iter < 10) {
    ++iter;
    for (int b = 0; b < nbad; ++b) orc_multigrid(m, 1, rhs + bad[b] * ns, xout + bad[b] * ns);
    for (int b = 0; b < nbad; ++b) wrap_face(m, xout + bad[b] * ns);
    any = 0;
    for (int b = 0; b < nbad; ++b)
      if (l2_resid(m, rhs + bad[b] * ns, xout + bad[b] * ns) > 1.0e-20) any = 1;
  }
  m->last_vcycles = 2 + iter;
  return any ? 1 : 0;
}

/* ------------------------------------------------ incremental remapping --- */
/* utilities/layer_2/remap.F:1072-1604 remap_mass_and_tracers with the
 * arguments get_remap_flux (swm/source/step.F:687-771) passes: one mass
 * field, ntracer independent tracers, linear integrals, first-order
 * departure points, tracer gradients limited, no checks. */
#define RM_EPS 1.0e-10
static const int rm_si[3][4] = {{0, 0, -1, -1}, {0, -1, -1, 0}, {0, -1, 0, 1}};   /* ishift_p */
static const int rm_sj[3][4] = {{0, 1, 0, -1}, {0, 0, -1, -1}, {0, -1, -1, 0}};   /* jshift_p */

/* operators.F:691-835 gradient_at_cell_centers (x and y components) */
static void rm_gradient(const orc_model* m, int k, const double* x, double* gx, double* gy) {
  const int iim = m->iim, jjm = m->jjm;
  const double *gw = m->rm.gwgt_x3, *ga = m->rm.garea_inv;
#define GW(c, nv, i, j) gw[(c) + 3 * ((nv) + 7 * S(i, j, k))]
  for (int j = 1; j <= jjm; ++j)
    for (int i = 1; i <= iim; ++i) { gx[S(i, j, k)] = 0.0; gy[S(i, j, k)] = 0.0; }
  for (int j = 2; j <= jjm - 1; ++j)
    for (int i = 2; i <= iim - 1; ++i)
      for (int c = 0; c < 2; ++c) {
        double* o = c == 0 ? gx : gy;
        o[S(i, j, k)] = (x[S(i, j, k)] * GW(c, 0, i, j) + x[S(i - 1, j, k)] * GW(c, 1, i, j) +
                         x[S(i - 1, j - 1, k)] * GW(c, 2, i, j) + x[S(i, j - 1, k)] * GW(c, 3, i, j) +
                         x[S(i + 1, j, k)] * GW(c, 4, i, j) + x[S(i + 1, j + 1, k)] * GW(c, 5, i, j) +
                         x[S(i, j + 1, k)] * GW(c, 6, i, j)) * ga[S(i, j, k)];
      }
  {
    const int i = 1, j = jjm - 1;   /* north pole: side 1 missing, (i-1,j-1) in (iim,1) */
    for (int c = 0; c < 2; ++c) {
      double* o = c == 0 ? gx : gy;
      o[S(i, j, k)] = (x[S(i, j, k)] * GW(c, 0, i, j) + x[S(iim, 1, k)] * GW(c, 2, i, j) +
                       x[S(i, j - 1, k)] * GW(c, 3, i, j) + x[S(i + 1, j, k)] * GW(c, 4, i, j) +
                       x[S(i + 1, j + 1, k)] * GW(c, 5, i, j) + x[S(i, j + 1, k)] * GW(c, 6, i, j)) * ga[S(i, j, k)];
    }
  }
  {
    const int i = iim - 1, j = 1;   /* south pole: side 3 missing, (i-1,j-1) in (1,jjm) */
    for (int c = 0; c < 2; ++c) {
      double* o = c == 0 ? gx : gy;
      o[S(i, j, k)] = (x[S(i, j, k)] * GW(c, 0, i, j) + x[S(i - 1, j, k)] * GW(c, 1, i, j) +
                       x[S(1, jjm, k)] * GW(c, 2, i, j) + x[S(i + 1, j, k)] * GW(c, 4, i, j) +
                       x[S(i + 1, j + 1, k)] * GW(c, 5, i, j) + x[S(i, j + 1, k)] * GW(c, 6, i, j)) * ga[S(i, j, k)];
    }
  }
#undef GW
}

static inline double rm_min(double a, double b) { return b < a ? b : a; }
static inline double rm_max(double a, double b) { return b > a ? b : a; }

/* remap.F:2152-2397 limit_gradient (l_pd false) */
static void rm_limit(const orc_model* m, int k, const double* phi, const double* phimask, const double* xs,
                     const double* ys, double* phix, double* phiy) {
  const int iim = m->iim, jjm = m->jjm;
  double* pmn = (double*)calloc((size_t)iim * jjm, sizeof(double));
  double* pmx = (double*)calloc((size_t)iim * jjm, sizeof(double));
#define L2(i, j) ((size_t)((i)-1) + (size_t)iim * ((j)-1))
#define NB(ii, jj) (phimask[S(ii, jj, k)] * phi[S(ii, jj, k)] + (1.0 - phimask[S(ii, jj, k)]) * phi[S(i, j, k)])
  for (int j = 2; j <= jjm - 1; ++j)
    for (int i = 2; i <= iim - 1; ++i) {
      const double w[7] = {NB(i - 1, j), NB(i - 1, j - 1), NB(i, j - 1), NB(i + 1, j), NB(i + 1, j + 1), NB(i, j + 1),
                           phi[S(i, j, k)]};
      double lo = w[0], hi = w[0];
      for (int q = 1; q < 7; ++q) { lo = rm_min(lo, w[q]); hi = rm_max(hi, w[q]); }
      pmn[L2(i, j)] = lo - phi[S(i, j, k)];
      pmx[L2(i, j)] = hi - phi[S(i, j, k)];
    }
  {
    const int i = 1, j = jjm - 1;
    const double w[6] = {NB(iim, 1), NB(i, j - 1), NB(i + 1, j), NB(i + 1, j + 1), NB(i, j + 1), phi[S(i, j, k)]};
    double lo = w[0], hi = w[0];
    for (int q = 1; q < 6; ++q) { lo = rm_min(lo, w[q]); hi = rm_max(hi, w[q]); }
    pmn[L2(i, j)] = lo - phi[S(i, j, k)];
    pmx[L2(i, j)] = hi - phi[S(i, j, k)];
  }
  {
    const int i = iim - 1, j = 1;
    const double w[6] = {NB(i - 1, j), NB(1, jjm), NB(i + 1, j), NB(i + 1, j + 1), NB(i, j + 1), phi[S(i, j, k)]};
    double lo = w[0], hi = w[0];
    for (int q = 1; q < 6; ++q) { lo = rm_min(lo, w[q]); hi = rm_max(hi, w[q]); }
    pmn[L2(i, j)] = lo - phi[S(i, j, k)];
    pmx[L2(i, j)] = hi - phi[S(i, j, k)];
  }
#undef NB
  for (int j = 1; j <= jjm - 1; ++j)
    for (int i = 1; i <= iim - 1; ++i) {
      const size_t n = S(i, j, k);
      if (!(phimask[n] > RM_EPS)) continue;
      const double w0 = -xs[n] * phix[n] - ys[n] * phiy[n];
      const double* cx = &m->rm.c_x3[12 * n];
      double qmn = 0.0, qmx = 0.0;
      for (int nv = 0; nv < 6; ++nv) {
        const double v = w0 + phix[n] * cx[2 * nv] + phiy[n] * cx[2 * nv + 1];
        if (nv == 0) { qmn = v; qmx = v; }
        else { qmn = rm_min(qmn, v); qmx = rm_max(qmx, v); }
      }
      const double wk1 = fabs(qmn) > 0.0 ? rm_max(0.0, pmn[L2(i, j)] / qmn) : 1.0;
      const double wk2 = fabs(qmx) > 0.0 ? rm_max(0.0, pmx[L2(i, j)] / qmx) : 1.0;
      double wk3 = rm_min(rm_min(1.0, wk1), wk2);
      if (!m->rm_limit_min) wk3 = rm_min(1.0, wk2);
      if (!m->rm_limit_max) wk3 = rm_min(1.0, wk1);
      if (!m->rm_limit_min && !m->rm_limit_max) wk3 = 1.0;
      phix[n] = wk3 * phix[n];
      phiy[n] = wk3 * phiy[n];
    }
#undef L2
  free(pmn);
  free(pmx);
}

/* remap_mass_and_tracers; mass (slots) and trace[q] (slots) are updated in
 * place, mflux (3, slots) receives the mass transport across the edges */
static int rm_remap(orc_model* m, double* mass, double** trace, const double* uc, double* mflux) {
  const int iim = m->iim, jjm = m->jjm, nt = m->nt;
  const size_t ns = m->ns;
  const double a = m->rm.radius, dt = m->dt;
  int bad = 0;
  double *mmask = zeros(ns), *mx = zeros(ns), *my = zeros(ns), *mp = zeros(ns), *xmav = zeros(ns),
         *ymav = zeros(ns), *dp = zeros(6 * ns);
  double **tx = (double**)calloc(nt ? nt : 1, sizeof(double*)), **ty = (double**)calloc(nt ? nt : 1, sizeof(double*)),
         **tp = (double**)calloc(nt ? nt : 1, sizeof(double*)), **mtflux = (double**)calloc(nt ? nt : 1, sizeof(double*));
  for (int q = 0; q < nt; ++q) { tx[q] = zeros(ns); ty[q] = zeros(ns); tp[q] = zeros(ns); mtflux[q] = zeros(3 * ns); }

  /* remap_prep (remap.F:1613-1794) */
  for (size_t n = 0; n < ns; ++n) {
    if (mass[n] < -RM_EPS) bad = 1;
    else if (mass[n] < RM_EPS) {
      mass[n] = 0.0;
      mmask[n] = 0.0;
      for (int q = 0; q < nt; ++q) trace[q][n] = 0.0;
    } else mmask[n] = 1.0;
  }
  if (bad) goto done;

  /* construct_fields (remap.F:1803-2061) */
  for (int k = 1; k <= 10; ++k) {
    rm_gradient(m, k, mass, mx, my);
    for (int q = 0; q < nt; ++q) rm_gradient(m, k, trace[q], tx[q], ty[q]);
    for (int j = 1; j <= jjm - 1; ++j)
      for (int i = 1; i <= iim - 1; ++i) {
        const size_t n = S(i, j, k);
        if (mass[n] > RM_EPS) {
          mp[n] = mass[n] - m->rm.xav[n] * mx[n] - m->rm.yav[n] * my[n];
          const double wk1 = 1.0 / mass[n];
          xmav[n] = (mx[n] * m->rm.xxav[n] + my[n] * m->rm.xyav[n] + mp[n] * m->rm.xav[n]) * wk1;
          ymav[n] = (mx[n] * m->rm.xyav[n] + my[n] * m->rm.yyav[n] + mp[n] * m->rm.yav[n]) * wk1;
        } else {
          mp[n] = 0.0; xmav[n] = 0.0; ymav[n] = 0.0;
        }
      }
    for (int q = 0; q < nt; ++q) {
      rm_limit(m, k, trace[q], mmask, xmav, ymav, tx[q], ty[q]);
      for (int j = 1; j <= jjm - 1; ++j)
        for (int i = 1; i <= iim - 1; ++i) {
          const size_t n = S(i, j, k);
          tp[q][n] = mmask[n] > RM_EPS ? trace[q][n] - tx[q][n] * xmav[n] - ty[q][n] * ymav[n] : 0.0;
        }
    }
  }
  wrap_face(m, mx); wrap_face(m, my); wrap_face(m, mp);
  for (int q = 0; q < nt; ++q) { wrap_face(m, tx[q]); wrap_face(m, ty[q]); wrap_face(m, tp[q]); }

  /* departure_points_1order (remap.F:2406-2537), after the velocity wrap */
  for (int k = 1; k <= 10; ++k)
    for (int j = 2; j <= jjm - 1; ++j)
      for (int i = 2; i <= iim - 1; ++i)
        for (int nc = 0; nc < 2; ++nc) {
          const size_t cs = 2 * S(i, j, k) + nc;
          const double u1 = uc[2 * cs], u2 = uc[2 * cs + 1];
          const double* T = &m->rm.trans_to_x1[9 * cs];
          const double* c = &m->rm.c[3 * cs];
          /* the reference's speed test uses the first component twice */
          const double wk1 = sqrt(u1 * u1 + u1 * u1);
          if (wk1 > RM_EPS) {
            double ut[3];
            for (int r = 0; r < 3; ++r) ut[r] = T[r] * u1 + T[r + 3] * u2;
            for (int r = 0; r < 3; ++r) dp[3 * cs + r] = a * c[r] - ut[r] * dt;
          } else {
            for (int r = 0; r < 3; ++r) dp[3 * cs + r] = a * c[r];
          }
        }
  wrap_vrtx(m, dp, 3);

  /* locate_triangles (remap.F:2903-3624) and transport_integrals
   * (remap.F:3633-3963), linear integrals, edge by edge */
  memset(mflux, 0, 3 * ns * sizeof(double));
  for (int k = 1; k <= 10; ++k)
    for (int j = 2; j <= jjm - 1; ++j)
      for (int i = 2; i <= iim - 1; ++i)
        for (int ne = 0; ne < 3; ++ne) {
          const size_t e = ne + 3 * S(i, j, k);
          double fluxsign[4] = {0.0, 0.0, 0.0, 0.0};
          double tp_[4][4][2];   /* [group][point 0..3][x,y] */
          int nds[4] = {0, 0, 0, 0};
          memset(tp_, 0, sizeof tp_);
          double cl[2] = {m->rm.cl_x4[2 * e], m->rm.cl_x4[2 * e + 1]}, cr[2] = {m->rm.cr_x4[2 * e], m->rm.cr_x4[2 * e + 1]};
          const double tl[2] = {m->rm.tl_x4[2 * e], m->rm.tl_x4[2 * e + 1]}, tr[2] = {m->rm.tr_x4[2 * e], m->rm.tr_x4[2 * e + 1]};
          const double bl[2] = {m->rm.bl_x4[2 * e], m->rm.bl_x4[2 * e + 1]}, br[2] = {m->rm.br_x4[2 * e], m->rm.br_x4[2 * e + 1]};
          double il[2] = {0.0, 0.0}, ir[2] = {0.0, 0.0}, ic[2] = {0.0, 0.0};
          /* departure points of the edge's ends, relative to its midpoint */
          size_t csl, csr;
          int nd0, nc0;
          if (ne == 0) { csl = 2 * S(i, j, k) + 0; csr = 2 * S(i, j + 1, k) + 1; nd0 = 0; nc0 = 0; }
          else if (ne == 1) { csl = 2 * S(i, j, k) + 1; csr = 2 * S(i, j, k) + 0; nd0 = 0; nc0 = 1; }
          else { csl = 2 * S(i + 1, j, k) + 0; csr = 2 * S(i, j, k) + 1; nd0 = 2; nc0 = 1; }
          const double* wv = &m->rm.w[9 * (2 * S(i, j, k) + nc0) + nd0];   /* w(nd0, s, nc0): stride 3 */
          double dlx[3], drx[3];
          for (int s = 0; s < 3; ++s) {
            dlx[s] = dp[3 * csl + s] - a * wv[3 * s];
            drx[s] = dp[3 * csr + s] - a * wv[3 * s];
          }
          const double* T4 = &m->rm.trans_x1_to_x4[9 * e];
          double dl[2], dr[2];
          for (int r = 0; r < 2; ++r) {
            dl[r] = T4[r] * dlx[0] + T4[r + 3] * dlx[1] + T4[r + 6] * dlx[2];
            dr[r] = T4[r] * drx[0] + T4[r + 3] * drx[1] + T4[r + 6] * drx[2];
          }
          if (fabs(dr[0] - dl[0]) < RM_EPS) dr[0] = dr[0] > dl[0] ? dr[0] + RM_EPS : dr[0] - RM_EPS;
          if (fabs(dl[0] - cl[0]) < RM_EPS) dl[0] = dl[0] > cl[0] ? dl[0] + RM_EPS : dl[0] - RM_EPS;
          if (fabs(dr[0] - cr[0]) < RM_EPS) dr[0] = dr[0] > cr[0] ? dr[0] + RM_EPS : dr[0] - RM_EPS;
          const double mdlr = (dr[1] - dl[1]) / (dr[0] - dl[0]);
          const double bdlr = dl[1] - mdlr * dl[0];
          int dlr_int_clr = 0;
          if (fabs(mdlr) > RM_EPS) {
            ic[0] = -bdlr / mdlr;
            if (ic[0] > cl[0] && ic[0] < cr[0]) dlr_int_clr = 1;
          }
          const double mtcl = tl[1] / (tl[0] - cl[0]), btcl = -mtcl * cl[0];
          const double mbcl = bl[1] / (bl[0] - cl[0]), bbcl = -mbcl * cl[0];
          const double mdcl = dl[1] / (dl[0] - cl[0]);
          int dl_in_left = 0;
          if (dl[0] < cl[0] && mdcl < mbcl && mdcl > mtcl) {
            dl_in_left = 1;
            il[0] = (btcl - bdlr) / (mdlr - mtcl);
            if (il[0] < cl[0]) il[1] = mtcl * il[0] + btcl;
            else { il[0] = (bbcl - bdlr) / (mdlr - mbcl); il[1] = mbcl * il[0] + bbcl; }
          }
          const double mtcr = tr[1] / (tr[0] - cr[0]), btcr = -mtcr * cr[0];
          const double mbcr = br[1] / (br[0] - cr[0]), bbcr = -mbcr * cr[0];
          const double mdcr = dr[1] / (dr[0] - cr[0]);
          int dr_in_right = 0;
          if (dr[0] > cr[0] && mdcr > mbcr && mdcr < mtcr) {
            dr_in_right = 1;
            ir[0] = (btcr - bdlr) / (mdlr - mtcr);
            if (ir[0] > cr[0]) ir[1] = mtcr * ir[0] + btcr;
            else { ir[0] = (bbcr - bdlr) / (mdlr - mbcr); ir[1] = mbcr * ir[0] + bbcr; }
          }
#define SETV(g, v, p) do { tp_[g][v][0] = (p)[0]; tp_[g][v][1] = (p)[1]; } while (0)
          if (dl_in_left) {
            SETV(0, 1, cl); SETV(0, 2, dl); SETV(0, 3, il);
            nds[0] = 4;
            fluxsign[0] = il[1] > 0.0 ? -1.0 : 1.0;
            dl[0] = il[0]; dl[1] = il[1];
          }
          if (dr_in_right) {
            SETV(1, 1, cr); SETV(1, 2, dr); SETV(1, 3, ir);
            nds[1] = 2;
            fluxsign[1] = ir[1] > 0.0 ? -1.0 : 1.0;
            dr[0] = ir[0]; dr[1] = ir[1];
          }
          if (!dlr_int_clr) {
            SETV(2, 1, cl); SETV(2, 2, dl); SETV(2, 3, dr);
            if (dl[1] > 0.0) { nds[2] = 3; fluxsign[2] = -1.0; } else { nds[2] = 1; fluxsign[2] = 1.0; }
            SETV(3, 1, cl); SETV(3, 2, cr); SETV(3, 3, dr);
            if (dl[1] > 0.0) { nds[3] = 3; fluxsign[3] = -1.0; } else { nds[3] = 1; fluxsign[3] = 1.0; }
          } else {
            SETV(2, 1, cl); SETV(2, 2, dl); SETV(2, 3, ic);
            if (dl[1] > 0.0) { nds[2] = 3; fluxsign[2] = -1.0; } else { nds[2] = 1; fluxsign[2] = 1.0; }
            SETV(3, 1, cr); SETV(3, 2, dr); SETV(3, 3, ic);
            if (dl[1] < 0.0) { nds[3] = 3; fluxsign[3] = -1.0; } else { nds[3] = 1; fluxsign[3] = 1.0; }
          }
#undef SETV
          for (int ng = 0; ng < 4; ++ng) {
            if (fluxsign[ng] == 0.0) continue;
            double (*t)[2] = tp_[ng];
            const double area = fluxsign[ng] * 0.5 *
                                fabs((t[2][0] - t[1][0]) * (t[3][1] - t[1][1]) - (t[2][1] - t[1][1]) * (t[3][0] - t[1][0]));
            const int nd = nds[ng] - 1;
            const double* px = &m->rm.p_x4[2 * (nd + 4 * e)];
            const double* T = &m->rm.trans_x4_to_x3[4 * (nd + 4 * e)];
            for (int nv = 1; nv <= 3; ++nv) {
              const double u0 = t[nv][0] - px[0], u1 = t[nv][1] - px[1];
              t[nv][0] = T[0] * u0 + T[2] * u1;
              t[nv][1] = T[1] * u0 + T[3] * u1;
            }
            t[0][0] = (1.0 / 3.0) * (t[1][0] + t[2][0] + t[3][0]);
            t[0][1] = (1.0 / 3.0) * (t[1][1] + t[2][1] + t[3][1]);
            /* transport_integrals, linear: the source cell's field at the midpoint */
            const int ip = i + rm_si[ne][nd], jp = j + rm_sj[ne][nd];
            const size_t sn = S(ip, jp, k);
            const double m0 = mp[sn] + t[0][0] * mx[sn] + t[0][1] * my[sn];
            mflux[e] = mflux[e] + area * m0;
            for (int q = 0; q < nt; ++q) {
              const double t0 = tp[q][sn] + t[0][0] * tx[q][sn] + t[0][1] * ty[q][sn];
              mtflux[q][e] = mtflux[q][e] + area * m0 * t0;
            }
          }
        }
  wrap_edge(m, mflux);
  for (int q = 0; q < nt; ++q) wrap_edge(m, mtflux[q]);

  /* update_fields (remap.F:3972-4145) */
  {
    double* flx = zeros(ns);
    double** mtold = (double**)calloc(nt ? nt : 1, sizeof(double*));
    for (int q = 0; q < nt; ++q) {
      mtold[q] = zeros(ns);
      for (size_t n = 0; n < ns; ++n) mtold[q][n] = mass[n] * trace[q][n];
    }
    div_edge(iim, jjm, m->rm.area_inv_x3, mflux, flx);
    for (size_t n = 0; n < ns; ++n) {
      mass[n] = mass[n] - flx[n];
      if (mass[n] < -RM_EPS) { /* the reference prints a warning and goes on */ }
      else if (mass[n] < RM_EPS) mass[n] = 0.0;
    }
    for (int q = 0; q < nt; ++q) {
      div_edge(iim, jjm, m->rm.area_inv_x3, mtflux[q], flx);
      for (size_t n = 0; n < ns; ++n) trace[q][n] = mass[n] > 0.0 ? (mtold[q][n] - flx[n]) / mass[n] : 0.0;
      free(mtold[q]);
    }
    free(mtold);
    free(flx);
  }
done:
  for (int q = 0; q < nt; ++q) { free(tx[q]); free(ty[q]); free(tp[q]); free(mtflux[q]); }
  free(tx); free(ty); free(tp); free(mtflux);
  free(mmask); free(mx); free(my); free(mp); free(xmav); free(ymav); free(dp);
  return bad;
}

int orc_remap_mass_and_tracers(orc_model* m, double* mass, double* trace, const double* uc, double* edgeflux) {
  const size_t ns = m->ns;
  const int iim = m->iim, jjm = m->jjm, nt = m->nt;
  if (!m->rm_owned) return 2;
  double* u = dup(uc, 4 * ns);
  wrap_vrtx(m, u, 2);   /* remap.F:1391 */
  double** tr = (double**)calloc(nt ? nt : 1, sizeof(double*));
  for (int q = 0; q < nt; ++q) {
    tr[q] = zeros(ns);
    for (int k = 1; k <= 10; ++k)
      for (int j = 1; j <= jjm; ++j)
        for (int i = 1; i <= iim; ++i)
          tr[q][S(i, j, k)] = trace[(size_t)(i - 1) + iim * ((size_t)(j - 1) + jjm * ((size_t)q + nt * (k - 1)))];
  }
  double* mflux = zeros(3 * ns);
  const int bad = rm_remap(m, mass, tr, u, mflux);
  if (!bad) {
    const double dti = 1.0 / m->dt;
    if (edgeflux)
      for (size_t n = 0; n < 3 * ns; ++n) edgeflux[n] = mflux[n] * dti;
    for (int q = 0; q < nt; ++q)
      for (int k = 1; k <= 10; ++k)
        for (int j = 1; j <= jjm; ++j)
          for (int i = 1; i <= iim; ++i)
            trace[(size_t)(i - 1) + iim * ((size_t)(j - 1) + jjm * ((size_t)q + nt * (k - 1)))] = tr[q][S(i, j, k)];
  }
  for (int q = 0; q < nt; ++q) free(tr[q]);
  free(tr); free(mflux); free(u);
  return bad ? 2 : 0;
}

void orc_set_remap_limiter(orc_model* m, int limit_min, int limit_max) {
  m->rm_limit_min = limit_min;
  m->rm_limit_max = limit_max;
}

void orc_set_remap(orc_model* m, const orc_remap* r, int limit_min, int limit_max) {
  const size_t ns = m->ns;
  orc_remap* d = &m->rm;
  if (m->rm_owned) {
    free((void*)d->c); free((void*)d->w); free((void*)d->trans_to_x1); free((void*)d->trans_x1_to_x4);
    free((void*)d->trans_x4_to_x3); free((void*)d->cl_x4); free((void*)d->cr_x4); free((void*)d->tl_x4);
    free((void*)d->tr_x4); free((void*)d->bl_x4); free((void*)d->br_x4); free((void*)d->p_x4); free((void*)d->c_x3);
    free((void*)d->area_inv_x3); free((void*)d->xav); free((void*)d->yav); free((void*)d->xxav); free((void*)d->xyav);
    free((void*)d->yyav); free((void*)d->gwgt_x3); free((void*)d->garea_inv);
  }
  d->radius = r->radius;
  d->c = dup(r->c, 6 * ns); d->w = dup(r->w, 18 * ns); d->trans_to_x1 = dup(r->trans_to_x1, 18 * ns);
  d->trans_x1_to_x4 = dup(r->trans_x1_to_x4, 27 * ns); d->trans_x4_to_x3 = dup(r->trans_x4_to_x3, 48 * ns);
  d->cl_x4 = dup(r->cl_x4, 6 * ns); d->cr_x4 = dup(r->cr_x4, 6 * ns); d->tl_x4 = dup(r->tl_x4, 6 * ns);
  d->tr_x4 = dup(r->tr_x4, 6 * ns); d->bl_x4 = dup(r->bl_x4, 6 * ns); d->br_x4 = dup(r->br_x4, 6 * ns);
  d->p_x4 = dup(r->p_x4, 24 * ns); d->c_x3 = dup(r->c_x3, 12 * ns); d->area_inv_x3 = dup(r->area_inv_x3, ns);
  d->xav = dup(r->xav, ns); d->yav = dup(r->yav, ns); d->xxav = dup(r->xxav, ns); d->xyav = dup(r->xyav, ns);
  d->yyav = dup(r->yyav, ns); d->gwgt_x3 = dup(r->gwgt_x3, 21 * ns); d->garea_inv = dup(r->garea_inv, ns);
  m->rm_owned = 1;
  m->rm_limit_min = limit_min;
  m->rm_limit_max = limit_max;
}

/* ------------------------------------------------- dissipation operator --- */
/* swm/source/swm_dissipation_operator.F:50-203 init_diss_op: per interior
 * cell and corner, the unit vectors from the corner to its three edge
 * midpoints and from the neighbouring corners to the same midpoints, in each
 * corner's X2 frame (t), their lengths (magt) and the distances between the
 * cell centres across the three edges (magn), in metres */
void orc_set_diss(orc_model* m, const double* p, const double* c, const double* w, const double* trans_to_x2,
                  double radius) {
  const int iim = m->iim, jjm = m->jjm;
  const size_t ns = m->ns;
  free(m->ds_t); free(m->ds_magt); free(m->ds_magn);
  m->ds_t = zeros(24 * ns);      /* t(2, ndata, nspace, ncorners, slot) */
  m->ds_magt = zeros(12 * ns);   /* magt(2, ndata, ncorners, slot) */
  m->ds_magn = zeros(6 * ns);    /* magn(ndata, ncorners, slot) */
#define P(i, j, k) (&p[3 * S(i, j, k)])
#define CC(ic, i, j, k) (&c[3 * (2 * S(i, j, k) + (ic))])
#define TX2(ic, i, j, k) (&trans_to_x2[9 * (2 * S(i, j, k) + (ic))])
  for (int k = 1; k <= 10; ++k)
    for (int j = 2; j <= jjm - 1; ++j)
      for (int i = 2; i <= iim - 1; ++i) {
        const size_t s = S(i, j, k);
        const double* d[2][3][2] = {{{P(i - 1, j, k), P(i, j, k)}, {P(i - 1, j - 1, k), P(i - 1, j, k)},
                                     {P(i, j, k), P(i - 1, j - 1, k)}},
                                    {{P(i - 1, j - 1, k), P(i, j, k)}, {P(i, j - 1, k), P(i - 1, j - 1, k)},
                                     {P(i, j, k), P(i, j - 1, k)}}};
        for (int ic = 0; ic < 2; ++ic)
          for (int id = 0; id < 3; ++id) {
            const double t1 = d[ic][id][0][0] - d[ic][id][1][0], t2 = d[ic][id][0][1] - d[ic][id][1][1],
                         t3 = d[ic][id][0][2] - d[ic][id][1][2];
            m->ds_magn[id + 3 * (ic + 2 * s)] = sqrt(t1 * t1 + t2 * t2 + t3 * t3);
          }
        /* the other corner of each edge and its cell */
        const int oi[2][3] = {{i, i - 1, i}, {i, i, i + 1}}, oj[2][3] = {{j + 1, j, j}, {j, j - 1, j}};
        for (int ic = 0; ic < 2; ++ic) {
          const int oc = 1 - ic;
          for (int id = 0; id < 3; ++id) {
            const double* c1 = CC(ic, i, j, k);
            const double* wv = &w[9 * (2 * s + ic) + id];   /* w(id, :, ic): stride 3 */
            const double* c3 = CC(oc, oi[ic][id], oj[ic][id], k);
            const double* T[2] = {TX2(ic, i, j, k), TX2(oc, oi[ic][id], oj[ic][id], k)};
            const double* src[2] = {c1, c3};
            for (int e = 0; e < 2; ++e) {
              double v[3], y[3];
              for (int r = 0; r < 3; ++r) v[r] = src[e][r] - wv[3 * r];
              for (int r = 0; r < 3; ++r) {
                double acc = 0.0;
                for (int q = 0; q < 3; ++q) acc = acc + T[e][r + 3 * q] * v[q];
                y[r] = acc;
              }
              const double rmag = sqrt(y[0] * y[0] + y[1] * y[1] + y[2] * y[2]);
              for (int sp = 0; sp < 2; ++sp) m->ds_t[e + 2 * (id + 3 * (sp + 2 * (ic + 2 * s)))] = y[sp] / rmag;
              m->ds_magt[e + 2 * (id + 3 * (ic + 2 * s))] = rmag;
            }
          }
        }
      }
#undef P
#undef CC
#undef TX2
  for (size_t n = 0; n < 6 * ns; ++n) m->ds_magn[n] = radius * m->ds_magn[n];
  for (size_t n = 0; n < 12 * ns; ++n) m->ds_magt[n] = radius * m->ds_magt[n];
}

/* swm_dissipation_operator.F:212-443 dissipation: norder applications of the
 * vector Laplacian (grad div + k x grad curl from edge divergences and
 * curls) to the corner vectors, times -visc (norder even) or +visc;
 * xin, xout (2,2,iim,jjm,10) */
int orc_dissipation(const orc_model* m, int norder, double visc, const double* xin, double* xout) {
  const int iim = m->iim, jjm = m->jjm;
  const size_t ns = m->ns;
  if (!m->ds_t) return 1;
  double* y = zeros(4 * ns);
  memcpy(xout, xin, 4 * ns * sizeof(double));
#define X(sp, ic, sl) xout[(sp) + 2 * ((ic) + 2 * (sl))]
#define T(a, id, sp, ic) m->ds_t[(a) + 2 * ((id) + 3 * ((sp) + 2 * ((ic) + 2 * s)))]
#define MT(a, id, ic) m->ds_magt[(a) + 2 * ((id) + 3 * ((ic) + 2 * s))]
  for (int it = 0; it < norder; ++it) {
    halo(m->vrtx_dst, m->vrtx_src, m->nvrtx, xout, 2);
    memset(y, 0, 4 * ns * sizeof(double));
    for (int k = 1; k <= 10; ++k)
      for (int j = 2; j <= jjm - 1; ++j)
        for (int i = 2; i <= iim - 1; ++i) {
          const size_t s = S(i, j, k);
          const int oi[2][3] = {{i, i - 1, i}, {i, i, i + 1}}, oj[2][3] = {{j + 1, j, j}, {j, j - 1, j}};
          for (int ic = 0; ic < 2; ++ic) {
            const int oc = 1 - ic;
            double gd[2] = {0.0, 0.0}, gr[2] = {0.0, 0.0};
            for (int id = 0; id < 3; ++id) {
              const size_t so = S(oi[ic][id], oj[ic][id], k);
              const double msum = MT(0, id, ic) + MT(1, id, ic);
              const double dv = (X(0, ic, s) * T(0, id, 0, ic) + X(1, ic, s) * T(0, id, 1, ic) +
                                 X(0, oc, so) * T(1, id, 0, ic) + X(1, oc, so) * T(1, id, 1, ic)) / msum;
              const double rl = (-X(0, ic, s) * T(0, id, 1, ic) + X(1, ic, s) * T(0, id, 0, ic) -
                                 X(0, oc, so) * T(1, id, 1, ic) + X(1, oc, so) * T(1, id, 0, ic)) / msum;
              const double mn = m->ds_magn[id + 3 * (ic + 2 * s)], ca = m->cai[ic + 2 * s];
              for (int sp = 0; sp < 2; ++sp) {
                gd[sp] = gd[sp] - dv * mn * T(0, id, sp, ic) * ca;
                gr[sp] = gr[sp] - rl * mn * T(0, id, sp, ic) * ca;
              }
            }
            y[0 + 2 * (ic + 2 * s)] = gd[0] + -gr[1];
            y[1 + 2 * (ic + 2 * s)] = gd[1] + gr[0];
          }
        }
    memcpy(xout, y, 4 * ns * sizeof(double));
  }
#undef X
#undef T
#undef MT
  const double f = norder % 2 == 0 ? -visc : visc;
  for (size_t n = 0; n < 4 * ns; ++n) xout[n] = f * xout[n];
  free(y);
  return 0;
}

/* --------------------------------------------------------------- model --- */
orc_model* orc_create(const orc_grid* gr) {
  orc_model* m = (orc_model*)calloc(1, sizeof(orc_model));
  m->g = gr->g;
  m->iim = m->jjm = 2 + (1 << gr->g);
  m->nt = gr->ntracer;
  m->ns = (size_t)m->iim * m->jjm * 10;
  m->dt = gr->dt;
  m->grav = gr->grav;
  m->scheme = gr->scheme;
  m->tot_area = gr->tot_area;
  const size_t ns = m->ns;
  m->area = dup(gr->area, ns);
  m->area_inv = dup(gr->area_inv, ns);
  m->cai = dup(gr->corner_area_inv, 2 * ns);
  m->sub_area = dup(gr->sub_area, 6 * ns);
  m->dc = dup(gr->dc, 6 * ns);
  m->normal = dup(gr->normal, 18 * ns);
  m->tangent = dup(gr->tangent, 18 * ns);
  m->vw = dup(gr->v_weights, 12 * ns);
  m->sw = dup(gr->s_weights, 6 * ns);
  m->lw = dup(gr->l_weights, 7 * ns);
  m->f = dup(gr->f, ns);
  m->hs = dup(gr->hs, ns);
  m->mask = dup(gr->center_mask, ns);
  m->nface = gr->nface; m->nvrtx = gr->nvrtx; m->nedge = gr->nedge;
  m->face_dst = dupi(gr->face_dst, gr->nface); m->face_src = dupi(gr->face_src, gr->nface);
  m->vrtx_dst = dupi(gr->vrtx_dst, gr->nvrtx); m->vrtx_src = dupi(gr->vrtx_src, gr->nvrtx);
  m->edge_dst = dupi(gr->edge_dst, gr->nedge); m->edge_src = dupi(gr->edge_src, gr->nedge);
  m->lap_w = (double**)calloc(gr->g, sizeof(double*));
  for (int l = 1; l <= gr->g; ++l) {
    const size_t nl = (size_t)(2 + (1 << l)) * (2 + (1 << l)) * 10;
    m->lap_w[l - 1] = dup(gr->lap_w[l - 1], 7 * nl);
  }
  for (int t = 0; t < 2; ++t) {
    m->h[t] = zeros(ns); m->eta[t] = zeros(ns); m->div[t] = zeros(ns); m->u[t] = zeros(4 * ns);
    m->tr[t] = (double**)calloc(m->nt, sizeof(double*));
    for (int q = 0; q < m->nt; ++q) m->tr[t][q] = zeros(ns);
  }
  for (int t = 0; t < 3; ++t) { m->eta_f[t] = zeros(ns); m->div_f[t] = zeros(ns); m->msh_f[t] = zeros(3 * ns); }
  for (int t = 0; t < 3; ++t) {
    m->h_f[t] = zeros(ns);
    m->tr_f[t] = (double**)calloc(m->nt > 0 ? m->nt : 1, sizeof(double*));
    for (int q = 0; q < m->nt; ++q) m->tr_f[t][q] = zeros(ns);
  }
  m->relative = zeros(ns); m->ke = zeros(ns); m->psi = zeros(ns); m->chi = zeros(ns); m->inv_h = zeros(ns);
  m->xout = zeros(2 * ns);
  m->mass_flux = zeros(3 * ns);
  m->work_old = (double**)calloc(m->nt, sizeof(double*));
  for (int q = 0; q < m->nt; ++q) m->work_old[q] = zeros(ns);
  return m;
}

void orc_destroy(orc_model* m) {
  if (!m) return;
  free(m->ds_t); free(m->ds_magt); free(m->ds_magn);
  if (m->rm_owned) {
    orc_remap* d = &m->rm;
    free((void*)d->c); free((void*)d->w); free((void*)d->trans_to_x1); free((void*)d->trans_x1_to_x4);
    free((void*)d->trans_x4_to_x3); free((void*)d->cl_x4); free((void*)d->cr_x4); free((void*)d->tl_x4);
    free((void*)d->tr_x4); free((void*)d->bl_x4); free((void*)d->br_x4); free((void*)d->p_x4); free((void*)d->c_x3);
    free((void*)d->area_inv_x3); free((void*)d->xav); free((void*)d->yav); free((void*)d->xxav); free((void*)d->xyav);
    free((void*)d->yyav); free((void*)d->gwgt_x3); free((void*)d->garea_inv);
  }
  free(m->area); free(m->area_inv); free(m->cai); free(m->sub_area); free(m->dc); free(m->normal);
  free(m->tangent); free(m->vw); free(m->sw); free(m->lw); free(m->f); free(m->hs); free(m->mask);
  free(m->face_dst); free(m->face_src); free(m->vrtx_dst); free(m->vrtx_src); free(m->edge_dst); free(m->edge_src);
  for (int l = 0; l < m->g; ++l) free(m->lap_w[l]);
  free(m->lap_w);
  for (int t = 0; t < 2; ++t) {
    free(m->h[t]); free(m->eta[t]); free(m->div[t]); free(m->u[t]);
    for (int q = 0; q < m->nt; ++q) free(m->tr[t][q]);
    free(m->tr[t]);
  }
  for (int t = 0; t < 3; ++t) { free(m->eta_f[t]); free(m->div_f[t]); free(m->msh_f[t]); free(m->h_f[t]); }
  for (int t = 0; t < 3; ++t) {
    for (int q = 0; q < m->nt; ++q) free(m->tr_f[t][q]);
    free(m->tr_f[t]);
  }
  free(m->relative); free(m->ke); free(m->psi); free(m->chi); free(m->inv_h); free(m->xout); free(m->mass_flux);
  for (int q = 0; q < m->nt; ++q) free(m->work_old[q]);
  free(m->work_old);
  free(m);
}

/* model_control.F:28-48 get_pointers, :58-89 get_weights */
static void pointers(orc_model* m) {
  const int sc = m->step_count;
  m->n4 = (sc + 1) % 2 + 1; m->n3 = sc % 2 + 1;
  m->n3f = (sc + 2) % 3 + 1; m->n2f = (sc + 1) % 3 + 1; m->n1f = sc % 3 + 1;
}
static void weights(orc_model* m) {
  const int sc = m->step_count;
  if (sc <= 1) { m->w3 = 1.0; m->w2 = 0.0; m->w1 = 0.0; }
  if (sc == 2) { m->w3 = 1.5; m->w2 = -0.5; m->w1 = 0.0; }
  if (sc >= 3) { m->w3 = 1.916666666666666666666; m->w2 = -1.333333333333333333333; m->w1 = 0.416666666666666666666; }
}

/* step.F:438-493 wrap_prognostics */
static void wrap_prognostics(orc_model* m) {
  const int n4 = m->n4 - 1;
  for (size_t n = 0; n < m->ns; ++n) m->inv_h[n] = m->mask[n] == 1.0 ? 1.0 / m->h[n4][n] : 0.0;
  wrap_face(m, m->h[n4]); wrap_face(m, m->eta[n4]); wrap_face(m, m->div[n4]);
  wrap_face(m, m->relative); wrap_face(m, m->ke); wrap_face(m, m->inv_h);
  for (int q = 0; q < m->nt; ++q) wrap_face(m, m->tr[n4][q]);
  wrap_vrtx(m, m->u[0], 2);
  wrap_vrtx(m, m->u[1], 2);
}

/* initialize.F:418-468 with caller-supplied h, u, tracers (hs in the grid) */
void orc_init(orc_model* m, const double* h, const double* u, const double* tracer) {
  const size_t ns = m->ns;
  const int iim = m->iim, jjm = m->jjm;
  m->step_count = 0;
  pointers(m);
  for (int t = 0; t < 2; ++t) {
    memcpy(m->h[t], h, ns * sizeof(double));
    memcpy(m->u[t], u, 4 * ns * sizeof(double));
    wrap_face(m, m->h[t]);
    wrap_vrtx(m, m->u[t], 2);
  }
  wrap_face(m, m->hs);
  /* operators_hl.F:153-266 calculate_rel_div_and_ke */
  double* kec = zeros(2 * ns);
  wrap_vrtx(m, m->u[0], 2);
  for (size_t q = 0; q < 2 * ns; ++q)
    kec[q] = 0.5 * (m->u[0][2 * q] * m->u[0][2 * q] + m->u[0][2 * q + 1] * m->u[0][2 * q + 1]);
  curl_vec(iim, jjm, m->area_inv, m->vw, m->u[0], m->relative);
  div_vec(iim, jjm, m->area_inv, m->vw, m->u[0], m->div[0]);
  center_ke(iim, jjm, m->sub_area, m->area_inv, kec, m->ke);
  free(kec);
  wrap_face(m, m->relative); wrap_face(m, m->div[0]); wrap_face(m, m->ke);
  memcpy(m->div[1], m->div[0], ns * sizeof(double));
  #pragma omp parallel for schedule(static)
  for (size_t n = 0; n < ns; ++n) m->eta[0][n] = m->f[n] + m->relative[n];
  memcpy(m->eta[1], m->eta[0], ns * sizeof(double));
  for (int q = 0; q < m->nt; ++q) {
    #pragma omp parallel for schedule(static)
    for (int k = 1; k <= 10; ++k)
      for (int j = 1; j <= jjm; ++j)
        for (int i = 1; i <= iim; ++i)
          m->tr[0][q][S(i, j, k)] = tracer[(size_t)(i - 1) + iim * ((size_t)(j - 1) + jjm * ((size_t)q + m->nt * (k - 1)))];
    memcpy(m->tr[1][q], m->tr[0][q], ns * sizeof(double));
  }
  wrap_prognostics(m);
}

/* step.F:114-212 step_dt with the FCT scheme (step.F:611-678 get_fct_flux) */
static int step_dt(orc_model* m) {
  const int iim = m->iim, jjm = m->jjm;
  const size_t ns = m->ns;
  const int n3 = m->n3 - 1, n4 = m->n4 - 1;
  double *h_c = zeros(2 * ns), *eta_c = zeros(2 * ns);
  avg_corners(iim, jjm, m->sw, m->h[n3], h_c);
  avg_corners(iim, jjm, m->sw, m->eta[n3], eta_c);
  wrap_vrtx(m, h_c, 1);
  wrap_vrtx(m, eta_c, 1);

  if (m->scheme == ORC_REMAP) {
    /* get_remap_flux (step.F:687-771) */
    if (!m->rm_owned) { free(h_c); free(eta_c); return 2; }
    double* hr = dup(m->h[n3], ns);
    double** trr = (double**)calloc(m->nt > 0 ? m->nt : 1, sizeof(double*));
    for (int q = 0; q < m->nt; ++q) trr[q] = dup(m->tr[n3][q], ns);
    double* mflux = zeros(3 * ns);
    if (rm_remap(m, hr, trr, m->u[n3], mflux)) {
      for (int q = 0; q < m->nt; ++q) free(trr[q]);
      free(trr); free(hr); free(mflux); free(h_c); free(eta_c);
      return 2;
    }
    const double dt = m->dt;
    const double dti_edge = 1.0 / dt;            /* remap.F: real(dbl_kind) */
    const double dti = (double)(float)(1.0 / dt);   /* step.F:705: declared real(int_kind), i.e. real*4 */
    for (size_t n = 0; n < 3 * ns; ++n) m->mass_flux[n] = mflux[n] * dti_edge;
    const int a = m->n3f - 1, b = m->n2f - 1, c = m->n1f - 1;
    for (size_t n = 0; n < ns; ++n) {
      m->h_f[a][n] = (hr[n] - m->h[n3][n]) * dti;
      m->h[n4][n] = m->h[n3][n] + dt * (m->w3 * m->h_f[a][n] + m->w2 * m->h_f[b][n] + m->w1 * m->h_f[c][n]);
    }
    for (int q = 0; q < m->nt; ++q) {
      for (size_t n = 0; n < ns; ++n) {
        m->tr_f[a][q][n] = (hr[n] * trr[q][n] - m->h[n3][n] * m->tr[n3][q][n]) * dti;
        m->tr[n4][q][n] = (m->h[n3][n] * m->tr[n3][q][n] +
                           dt * (m->w3 * m->tr_f[a][q][n] + m->w2 * m->tr_f[b][q][n] + m->w1 * m->tr_f[c][q][n])) /
                          m->h[n4][n];
      }
      free(trr[q]);
    }
    free(trr); free(hr); free(mflux);
  } else if (m->scheme != ORC_FCT) {
    /* get_centered_flux / get_upwind_flux (step.F:495-602) */
    double** tflux = (double**)calloc(m->nt > 0 ? m->nt : 1, sizeof(double*));
    if (m->scheme == ORC_CENTERED) centered_mass_flux(iim, jjm, m->dc, m->normal, h_c, m->u[n3], m->mass_flux);
    else upstream_mass_flux(iim, jjm, m->dc, m->normal, m->u[n3], m->h[n3], m->mass_flux);
    for (int q = 0; q < m->nt; ++q) {
      tflux[q] = zeros(3 * ns);
      if (m->scheme == ORC_CENTERED) centered_tracer_flux(iim, jjm, m->mass_flux, m->tr[n3][q], tflux[q]);
      else upstream_tracer_flux(iim, jjm, m->mass_flux, m->tr[n3][q], tflux[q]);
    }
    wrap_edge(m, m->mass_flux);
    for (int q = 0; q < m->nt; ++q) wrap_edge(m, tflux[q]);
    const double dt = m->dt;
    const int a = m->n3f - 1, b = m->n2f - 1, c = m->n1f - 1;
    double* w = zeros(ns);
    /* step_surface_height (step.F:221-270) */
    div_edge(iim, jjm, m->area_inv, m->mass_flux, w);
    #pragma omp parallel for schedule(static)
    for (size_t n = 0; n < ns; ++n) m->h_f[a][n] = -w[n];
    for (size_t n = 0; n < ns; ++n) {
      m->h[n4][n] = m->h[n3][n] + dt * (m->w3 * m->h_f[a][n] + m->w2 * m->h_f[b][n] + m->w1 * m->h_f[c][n]);
      if (m->h[n4][n] < 1.0e-10) m->h[n4][n] = 1.0e-10;
    }
    /* step_tracer (step.F:278-327) */
    for (int q = 0; q < m->nt; ++q) {
      div_edge(iim, jjm, m->area_inv, tflux[q], w);
      #pragma omp parallel for schedule(static)
      for (size_t n = 0; n < ns; ++n) m->tr_f[a][q][n] = -w[n];
      for (size_t n = 0; n < ns; ++n) {
        double t = m->tr[n3][q][n] * m->h[n3][n] +
                   dt * (m->w3 * m->tr_f[a][q][n] + m->w2 * m->tr_f[b][q][n] + m->w1 * m->tr_f[c][q][n]);
        if (m->h[n4][n] > 1.0e-10) t = t / m->h[n4][n];
        m->tr[n4][q][n] = t;
      }
      free(tflux[q]);
    }
    free(tflux);
    free(w);
  } else {
  /* get_fct_flux */
  double* wm[2] = {dup(m->h[0], ns), dup(m->h[1], ns)};
  double* wlow = zeros(ns);
  double* wflux = zeros(3 * ns);
  mass_fct(m, wm[n3], h_c, m->u[n3], wlow, wflux, wm[n4]);
  memcpy(m->h[n4], wm[n4], ns * sizeof(double));
  memcpy(m->mass_flux, wflux, 3 * ns * sizeof(double));
  wrap_edge(m, m->mass_flux);
  double *wtr = zeros(ns), *wtrh = zeros(ns), *wtn4 = zeros(ns);
  for (int q = 0; q < m->nt; ++q) {
    memcpy(wtr, m->tr[n3][q], ns * sizeof(double));
    for (size_t n = 0; n < ns; ++n)
      wtrh[n] = m->w3 * m->tr[n3][q][n] + m->w2 * m->tr[n4][q][n] + m->w1 * m->work_old[q][n];
    memcpy(m->work_old[q], m->tr[n4][q], ns * sizeof(double));
    tracer_fct(m, wtr, wtrh, wm[n3], wlow, wm[n4], wflux, wtn4);
    memcpy(m->tr[n4][q], wtn4, ns * sizeof(double));
  }
  free(wtr); free(wtrh); free(wtn4); free(wm[0]); free(wm[1]); free(wlow); free(wflux);
  }

  double *vflux = zeros(3 * ns), *dcurl = zeros(3 * ns);
  vorticity_flux(iim, jjm, m->mass_flux, m->h[n3], m->eta[n3], vflux);
  divergence_curl(iim, jjm, m->dc, m->tangent, h_c, m->eta[n3], m->u[n3], m->h[n3], dcurl);
  wrap_edge(m, vflux);
  wrap_edge(m, dcurl);

  /* step_vorticity (step.F:336-370), step_divergence (:379-429) */
  double *w1 = zeros(ns), *w2 = zeros(ns), *w3 = zeros(ns);
  const double dt = m->dt;
  const int a = m->n3f - 1, b = m->n2f - 1, c = m->n1f - 1;
  div_edge(iim, jjm, m->area_inv, vflux, w1);
  #pragma omp parallel for schedule(static)
  for (size_t n = 0; n < ns; ++n) m->eta_f[a][n] = -w1[n];
  for (size_t n = 0; n < ns; ++n)
    m->eta[n4][n] = m->eta[n3][n] + dt * (m->w3 * m->eta_f[a][n] + m->w2 * m->eta_f[b][n] + m->w1 * m->eta_f[c][n]);
  div_edge(iim, jjm, m->area_inv, dcurl, w3);
  #pragma omp parallel for schedule(static)
  for (size_t n = 0; n < ns; ++n) w1[n] = m->ke[n] + m->grav * (m->h[n3][n] + m->hs[n]);
  laplacian(iim, jjm, m->lw, w1, w2);
  #pragma omp parallel for schedule(static)
  for (size_t n = 0; n < ns; ++n) m->div_f[a][n] = -w2[n] + w3[n];
  for (size_t n = 0; n < ns; ++n)
    m->div[n4][n] = m->div[n3][n] + dt * (m->w3 * m->div_f[a][n] + m->w2 * m->div_f[b][n] + m->w1 * m->div_f[c][n]);
  free(w1); free(w2); free(w3); free(vflux); free(dcurl);

  /* elliptic solve for psi and chi (step.F:170-184) */
  double* rhs = zeros(2 * ns);
  for (size_t n = 0; n < ns; ++n) {
    m->relative[n] = m->eta[n4][n] - m->f[n];
    rhs[n] = m->relative[n];
    rhs[ns + n] = m->div[n4][n];
  }
  const int bad = orc_elliptic(m, 2, rhs, m->xout);
  for (size_t n = 0; n < ns; ++n) {
    m->eta[n4][n] = rhs[n] + m->f[n];
    m->psi[n] = m->xout[n];
    m->chi[n] = m->xout[ns + n];
  }
  free(rhs);
  velocity(iim, jjm, m->cai, m->vw, m->psi, m->chi, m->u[n4]);
  wrap_vrtx(m, m->u[n4], 2);
  double* kec = zeros(2 * ns);
  for (size_t q = 0; q < 2 * ns; ++q)
    kec[q] = 0.5 * (m->u[n4][2 * q] * m->u[n4][2 * q] + m->u[n4][2 * q + 1] * m->u[n4][2 * q + 1]);
  center_ke(iim, jjm, m->sub_area, m->area_inv, kec, m->ke);
  free(kec);
  wrap_face(m, m->ke);
  free(h_c); free(eta_c);
  return bad;
}

int orc_step(orc_model* m, int nsteps) {
  for (int s = 0; s < nsteps; ++s) {
    m->step_count += 1;
    pointers(m);
    weights(m);
    const int st = step_dt(m);
    if (st) return st;
    wrap_prognostics(m);
  }
  return 0;
}

int orc_last_vcycles(const orc_model* m) { return m->last_vcycles; }

/* budgets.F:31-247 compute_budgets; global_real_sum (global_reductions.F:
 * 77-98) is sum(X*MASK) in memory order */
void orc_budgets(const orc_model* m, double* out) {
  const size_t ns = m->ns;
  const int n4 = m->n4 - 1;
  const double* h = m->h[n4];
  const double *a = m->area, *mk = m->mask;
  double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (size_t n = 0; n < ns; ++n) {
    s[0] += (h[n] * a[n]) * mk[n];
    s[1] += ((h[n] * m->ke[n]) * a[n]) * mk[n];
    s[2] += ((((0.5 * m->grav) * h[n]) * h[n]) * a[n]) * mk[n];
    s[3] += (((m->grav * h[n]) * m->hs[n]) * a[n]) * mk[n];
    const double pe = mk[n] > 0.0 ? ((m->eta[n4][n] * m->eta[n4][n]) / h[n]) * a[n] : 0.0;
    s[5] += pe * mk[n];
    s[6] += (m->eta[n4][n] * a[n]) * mk[n];
    s[7] += (m->div[n4][n] * a[n]) * mk[n];
  }
  const double inv = 1.0 / m->tot_area;
  out[0] = s[0] * inv;
  out[1] = s[1] * inv;
  out[2] = s[2] * inv;
  out[3] = s[3] * inv;
  out[4] = out[1] + out[2] + out[3];
  out[5] = s[5] * inv;
  out[6] = s[6] * inv;
  out[7] = s[7] * inv;
  for (int q = 0; q < m->nt; ++q) {
    const double* t = m->tr[n4][q];
    double acc = 0.0;
    for (size_t n = 0; n < ns; ++n) acc += (((t[n] * t[n]) * h[n]) * a[n]) * mk[n];
    out[8 + q] = acc * inv;
  }
}

int orc_get(const orc_model* m, const char* name, double* out) {
  const size_t ns = m->ns;
  const int n4 = m->n4 - 1;
  const int iim = m->iim, jjm = m->jjm;
  if (!strcmp(name, "h")) memcpy(out, m->h[n4], ns * 8);
  else if (!strcmp(name, "eta")) memcpy(out, m->eta[n4], ns * 8);
  else if (!strcmp(name, "div")) memcpy(out, m->div[n4], ns * 8);
  else if (!strcmp(name, "relative")) memcpy(out, m->relative, ns * 8);
  else if (!strcmp(name, "ke")) memcpy(out, m->ke, ns * 8);
  else if (!strcmp(name, "psi")) memcpy(out, m->psi, ns * 8);
  else if (!strcmp(name, "chi")) memcpy(out, m->chi, ns * 8);
  else if (!strcmp(name, "u")) memcpy(out, m->u[n4], 4 * ns * 8);
  else if (!strcmp(name, "tracer")) {
    for (int q = 0; q < m->nt; ++q)
      #pragma omp parallel for schedule(static)
      for (int k = 1; k <= 10; ++k)
        for (int j = 1; j <= jjm; ++j)
          for (int i = 1; i <= iim; ++i)
            out[(size_t)(i - 1) + iim * ((size_t)(j - 1) + jjm * ((size_t)q + m->nt * (k - 1)))] = m->tr[n4][q][S(i, j, k)];
  } else return 1;
  return 0;
}

/* ----------------------------------------------- single-operator entries - */
void orc_div_of_an_edge_flux(const orc_model* m, const double* xin, double* xout) { div_edge(m->iim, m->jjm, m->area_inv, xin, xout); }
void orc_laplacian(const orc_model* m, const double* xin, double* xout) { laplacian(m->iim, m->jjm, m->lw, xin, xout); }
void orc_curl_of_a_vector(const orc_model* m, const double* xin, double* xout) { curl_vec(m->iim, m->jjm, m->area_inv, m->vw, xin, xout); }
void orc_div_of_a_vector(const orc_model* m, const double* xin, double* xout) { div_vec(m->iim, m->jjm, m->area_inv, m->vw, xin, xout); }
void orc_center_ke(const orc_model* m, const double* ke_c, double* ke) { center_ke(m->iim, m->jjm, m->sub_area, m->area_inv, ke_c, ke); }
void orc_average_to_corners(const orc_model* m, const double* xin, double* xout) { avg_corners(m->iim, m->jjm, m->sw, xin, xout); }
void orc_velocity(const orc_model* m, const double* psi, const double* chi, double* u) { velocity(m->iim, m->jjm, m->cai, m->vw, psi, chi, u); }
void orc_centered_mass_flux(const orc_model* m, const double* mass_c, const double* u, double* flux) {
  centered_mass_flux(m->iim, m->jjm, m->dc, m->normal, mass_c, u, flux);
}
void orc_upstream_mass_flux(const orc_model* m, const double* u, const double* mass, double* flux) {
  upstream_mass_flux(m->iim, m->jjm, m->dc, m->normal, u, mass, flux);
}
void orc_vorticity_flux(const orc_model* m, const double* mf, const double* mass, const double* eta, double* flux) {
  vorticity_flux(m->iim, m->jjm, mf, mass, eta, flux);
}
void orc_divergence_curl(const orc_model* m, const double* mass_c, const double* eta, const double* u,
                         const double* mass, double* flux) {
  divergence_curl(m->iim, m->jjm, m->dc, m->tangent, mass_c, eta, u, mass, flux);
}
