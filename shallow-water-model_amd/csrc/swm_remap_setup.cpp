// swm_remap_setup.cpp -- the grid quantities of incremental remapping,
// init_remap of utilities/layer_2/remap.F:176-1063, restated in the same
// floating-point operation order (compiled without contraction).
//
// Fortran MATMUL and DOT_PRODUCT sum their products in index order starting
// from zero; the products by the identity matrix X1 are therefore the
// entries of the other factor with a zero of either sign turned into +0.
#include <cmath>
#include <mutex>

#include "swm_setup.h"

namespace swm {

namespace {

constexpr double kEps = 1.0e-10;                      // remap.F eps
constexpr double kP333 = 1.0 / 3.0, kP4 = 0.4, kP6 = 0.6;
constexpr double kP5625m = -9.0 / 16.0, kP52083 = 25.0 / 48.0;

// geodesic_grid_utilities.F:461-484, :493-516, :560-604
inline void cross3(const double* a, const double* b, double* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
inline void unit3(double* a) {
  double len = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
  len = 1.0 / len;
  a[0] = a[0] * len; a[1] = a[1] * len; a[2] = a[2] * len;
}
inline void tangent_plane(const double* p1, const double* p2, double* o) {
  double w[3];
  w[0] = -(p1[1] * (-(p1[1] * p2[0]) + p1[0] * p2[1])) + p1[2] * (p1[2] * p2[0] - p1[0] * p2[2]);
  w[1] = +(p1[0] * (-(p1[1] * p2[0]) + p1[0] * p2[1])) - p1[2] * (-(p1[2] * p2[1]) + p1[1] * p2[2]);
  w[2] = -(p1[0] * (p1[2] * p2[0] - p1[0] * p2[2])) + p1[1] * (-(p1[2] * p2[1]) + p1[1] * p2[2]);
  unit3(w);
  o[0] = w[0]; o[1] = w[1]; o[2] = w[2];
}

// y = A x, A column-major 3x3 (A[r + 3c])
inline void matvec3(const double* A, const double* x, double* y) {
  for (int r = 0; r < 3; ++r) {
    double s = 0.0;
    for (int k = 0; k < 3; ++k) s = s + A[r + 3 * k] * x[k];
    y[r] = s;
  }
}

// remap.F:151-167
inline double tri_area(const double* t) {   // t: (2, 3) vertices
  return 0.5 * std::fabs((t[2] - t[0]) * (t[5] - t[1]) - (t[3] - t[1]) * (t[4] - t[0]));
}

}  // namespace

// ishift_p / jshift_p (remap.F:296-331): the four cells around edge ne of
// cell (i, j), counter-clockwise from (i, j)
const int kRemapShiftI[3][4] = {{0, 0, -1, -1}, {0, -1, -1, 0}, {0, -1, 0, 1}};
const int kRemapShiftJ[3][4] = {{0, 1, 0, -1}, {0, 0, -1, -1}, {0, -1, -1, 0}};

void build_remap_geometry(const Grid& G, RemapGeometry& R) {
  const int iim = G.iim, jjm = G.jjm;
  const size_t ns = (size_t)iim * jjm * 10;
  const double a = G.radius;
  auto S = [&](int i, int j, int k) { return (size_t)(i - 1) + (size_t)iim * ((size_t)(j - 1) + (size_t)jjm * (k - 1)); };
  auto P = [&](int i, int j, int k) { return &G.p[3 * S(i, j, k)]; };
  auto C = [&](int ic, int i, int j, int k) { return &G.c[3 * (2 * S(i, j, k) + ic)]; };
  // w(nd, :, nc, i, j, k)
  auto W = [&](int nd, int nc, int i, int j, int k, int s) { return G.w[9 * (2 * S(i, j, k) + nc) + nd + 3 * s]; };
  auto SUB = [&](int kk, int ic, int i, int j, int k) { return G.sub_area[3 * (2 * S(i, j, k) + ic) + kk]; };

  R.vertex_mask.assign(6 * ns, 0.0);
  R.num_vertices.assign(ns, 0.0);
  R.trans_x1_to_x3.assign(9 * ns, 0.0);
  R.trans_x1_to_x4.assign(27 * ns, 0.0);
  R.trans_x4_to_x3.assign(48 * ns, 0.0);
  for (auto* v : {&R.cl_x4, &R.cr_x4, &R.tl_x4, &R.tr_x4, &R.bl_x4, &R.br_x4}) v->assign(6 * ns, 0.0);
  R.p_x4.assign(24 * ns, 0.0);
  R.c_x3.assign(12 * ns, 0.0);
  for (auto* v : {&R.area_x3, &R.area_inv_x3, &R.xav, &R.yav, &R.xxav, &R.xyav, &R.yyav, &R.xxxav, &R.xxyav,
                  &R.xyyav, &R.yyyav, &R.garea_inv})
    v->assign(ns, 0.0);
  R.gwgt_x3.assign(3 * 7 * ns, 0.0);

  // vertex mask and vertex count (remap.F:336-368)
  for (int k = 1; k <= 10; ++k)
    for (int j = 1; j <= jjm - 1; ++j)
      for (int i = 1; i <= iim - 1; ++i) {
        double* vm = &R.vertex_mask[6 * S(i, j, k)];
        vm[0] = SUB(0, 0, i, j, k);
        vm[1] = SUB(0, 1, i, j, k);
        vm[2] = SUB(1, 0, i + 1, j, k);
        vm[3] = SUB(1, 1, i + 1, j + 1, k);
        vm[4] = SUB(2, 0, i + 1, j + 1, k);
        vm[5] = SUB(2, 1, i, j + 1, k);
      }
  for (double& v : R.vertex_mask)
    if (v > kEps) v = 1.0;
  for (size_t n = 0; n < ns; ++n) {
    const double* vm = &R.vertex_mask[6 * n];
    R.num_vertices[n] = vm[0] + vm[1] + vm[2] + vm[3] + vm[4] + vm[5];
  }

  // X1 -> X3 at cell centres (remap.F:386-440)
  for (int k = 1; k <= 10; ++k)
    for (int j = 1; j <= jjm - 1; ++j)
      for (int i = 1; i <= iim - 1; ++i) {
        double x3[3][3];
        for (int s = 0; s < 3; ++s) x3[2][s] = P(i, j, k)[s];
        unit3(x3[2]);
        const double dot = ((0.0 + x3[2][0] * 0.0) + x3[2][1] * 0.0) + x3[2][2] * 1.0;
        static const double e1[3] = {1.0, 0.0, 0.0}, e3[3] = {0.0, 0.0, 1.0};
        if (std::fabs(dot) < 1.0 - kEps) tangent_plane(x3[2], e3, x3[1]);
        else tangent_plane(x3[2], e1, x3[1]);
        cross3(x3[1], x3[2], x3[0]);
        double* T = &R.trans_x1_to_x3[9 * S(i, j, k)];
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) T[r + 3 * c] = x3[r][c] + 0.0;
      }
  for (int o = 0; o < 9; ++o) apply_halo_strided(G.face, R.trans_x1_to_x3.data(), 9, o);

  // X1 -> X4 at cell edges (remap.F:446-489)
  for (int k = 1; k <= 10; ++k)
    for (int j = 2; j <= jjm - 1; ++j)
      for (int i = 2; i <= iim - 1; ++i)
        for (int ne = 0; ne < 3; ++ne) {
          double x4[3][3];
          const double *cc, *ww;
          double wv[3];
          if (ne == 0) { cc = C(1, i, j + 1, k); for (int s = 0; s < 3; ++s) wv[s] = W(0, 0, i, j, k, s); }
          else if (ne == 1) { cc = C(0, i, j, k); for (int s = 0; s < 3; ++s) wv[s] = W(0, 1, i, j, k, s); }
          else { cc = C(1, i, j, k); for (int s = 0; s < 3; ++s) wv[s] = W(2, 1, i, j, k, s); }
          ww = wv;
          for (int s = 0; s < 3; ++s) { x4[0][s] = cc[s] - ww[s]; x4[2][s] = ww[s]; }
          unit3(x4[0]);
          unit3(x4[2]);
          cross3(x4[2], x4[0], x4[1]);
          double* T = &R.trans_x1_to_x4[9 * (ne + 3 * S(i, j, k))];
          for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) T[r + 3 * c] = x4[r][c] + 0.0;
        }
  for (int o = 0; o < 9; ++o) apply_halo_strided(G.edge, R.trans_x1_to_x4.data(), 9, o);

  // X4 -> X3 for the four cells around each edge (remap.F:495-524)
  for (int k = 1; k <= 10; ++k)
    for (int j = 2; j <= jjm - 1; ++j)
      for (int i = 2; i <= iim - 1; ++i)
        for (int ne = 0; ne < 3; ++ne)
          for (int nd = 0; nd < 4; ++nd) {
            const int ip = i + kRemapShiftI[ne][nd], jp = j + kRemapShiftJ[ne][nd];
            const double* T3 = &R.trans_x1_to_x3[9 * S(ip, jp, k)];
            const double* T4 = &R.trans_x1_to_x4[9 * (ne + 3 * S(i, j, k))];
            double* out = &R.trans_x4_to_x3[4 * (nd + 4 * (ne + 3 * S(i, j, k)))];
            for (int r = 0; r < 2; ++r)
              for (int c = 0; c < 2; ++c) {
                double s = 0.0;
                for (int q = 0; q < 3; ++q) s = s + T3[r + 3 * q] * T4[c + 3 * q];
                out[r + 2 * c] = s;
              }
          }
  for (int o = 0; o < 16; ++o) apply_halo_strided(G.edge, R.trans_x4_to_x3.data(), 16, o);

  // points around each edge in X4 (remap.F:530-589)
  for (int k = 1; k <= 10; ++k)
    for (int j = 2; j <= jjm - 1; ++j)
      for (int i = 2; i <= iim - 1; ++i)
        for (int ne = 0; ne < 3; ++ne) {
          const double *cl, *cr, *tl, *tr, *bl, *br, *pp[4];
          int nd0, nc0;   // the edge's midpoint: w(nd0, :, nc0, i, j)
          if (ne == 0) {
            cl = C(0, i, j, k); cr = C(1, i, j + 1, k); tl = C(1, i - 1, j, k); tr = C(0, i, j + 1, k);
            bl = C(1, i, j, k); br = C(0, i + 1, j + 1, k);
            pp[0] = P(i, j, k); pp[1] = P(i, j + 1, k); pp[2] = P(i - 1, j, k); pp[3] = P(i - 1, j - 1, k);
            nd0 = 0; nc0 = 0;
          } else if (ne == 1) {
            cl = C(1, i, j, k); cr = C(0, i, j, k); tl = C(0, i, j - 1, k); tr = C(1, i - 1, j, k);
            bl = C(0, i + 1, j, k); br = C(1, i, j + 1, k);
            pp[0] = P(i, j, k); pp[1] = P(i - 1, j, k); pp[2] = P(i - 1, j - 1, k); pp[3] = P(i, j - 1, k);
            nd0 = 0; nc0 = 1;
          } else {
            cl = C(0, i + 1, j, k); cr = C(1, i, j, k); tl = C(1, i + 1, j, k); tr = C(0, i, j - 1, k);
            bl = C(1, i + 1, j + 1, k); br = C(0, i, j, k);
            pp[0] = P(i, j, k); pp[1] = P(i - 1, j - 1, k); pp[2] = P(i, j - 1, k); pp[3] = P(i + 1, j, k);
            nd0 = 2; nc0 = 1;
          }
          double wv[3];
          for (int s = 0; s < 3; ++s) wv[s] = W(nd0, nc0, i, j, k, s);
          const double* T = &R.trans_x1_to_x4[9 * (ne + 3 * S(i, j, k))];
          auto rel = [&](const double* x, double* y) {
            double v[3];
            for (int s = 0; s < 3; ++s) v[s] = a * (x[s] - wv[s]);
            matvec3(T, v, y);
          };
          const size_t e = ne + 3 * S(i, j, k);
          double u[3];
          rel(cl, u); R.cl_x4[2 * e] = u[0]; R.cl_x4[2 * e + 1] = 0.0;
          rel(cr, u); R.cr_x4[2 * e] = u[0]; R.cr_x4[2 * e + 1] = 0.0;
          rel(tl, u); R.tl_x4[2 * e] = u[0]; R.tl_x4[2 * e + 1] = u[1];
          rel(tr, u); R.tr_x4[2 * e] = u[0]; R.tr_x4[2 * e + 1] = u[1];
          rel(bl, u); R.bl_x4[2 * e] = u[0]; R.bl_x4[2 * e + 1] = u[1];
          rel(br, u); R.br_x4[2 * e] = u[0]; R.br_x4[2 * e + 1] = u[1];
          for (int nd = 0; nd < 4; ++nd) {
            rel(pp[nd], u);
            R.p_x4[2 * (nd + 4 * e)] = u[0];
            R.p_x4[2 * (nd + 4 * e) + 1] = u[1];
          }
        }

  // X3 corners, areas and moments (remap.F:595-856), poles included
  for (int k = 1; k <= 10; ++k)
    for (int j = 1; j <= jjm - 1; ++j)
      for (int i = 1; i <= iim - 1; ++i) {
        const size_t n = S(i, j, k);
        const double* vm = &R.vertex_mask[6 * n];
        const double* cv[6] = {C(0, i, j, k),         C(1, i, j, k),     C(0, i + 1, j, k),
                               C(1, i + 1, j + 1, k), C(0, i + 1, j + 1, k), C(1, i, j + 1, k)};
        double vt[6][3];
        for (int nv = 0; nv < 6; ++nv)
          for (int s = 0; s < 3; ++s) vt[nv][s] = cv[nv][s] * vm[nv];
        double cmag = 0.0;
        for (int nv = 0; nv < 6; ++nv) cmag = cmag + std::sqrt(vt[nv][0] * vt[nv][0] + vt[nv][1] * vt[nv][1] + vt[nv][2] * vt[nv][2]);
        cmag = cmag / R.num_vertices[n];
        const double* pc = P(i, j, k);
        for (int nv = 0; nv < 6; ++nv)
          for (int s = 0; s < 3; ++s) vt[nv][s] = (pc[s] - vt[nv][s]) * vm[nv];
        double pcmag = 0.0;
        for (int nv = 0; nv < 6; ++nv)
          pcmag = pcmag + std::sqrt(vt[nv][0] * vt[nv][0] + vt[nv][1] * vt[nv][1] + vt[nv][2] * vt[nv][2]);
        pcmag = pcmag / R.num_vertices[n];
        const double dtheta = std::atan(pcmag);
        const double pmag = cmag * std::cos(dtheta);
        double ut[3];
        for (int s = 0; s < 3; ++s) ut[s] = pmag * pc[s];
        const double* T3 = &R.trans_x1_to_x3[9 * n];
        double* cx = &R.c_x3[12 * n];
        for (int nv = 0; nv < 6; ++nv) {
          double d[3], y[3];
          for (int s = 0; s < 3; ++s) d[s] = cv[nv][s] - ut[s];
          matvec3(T3, d, y);
          cx[2 * nv] = a * y[0];
          cx[2 * nv + 1] = a * y[1];
          if (vm[nv] == 0.0) {
            const int nv2 = nv == 0 ? 5 : nv - 1;
            cx[2 * nv] = cx[2 * nv2];
            cx[2 * nv + 1] = cx[2 * nv2 + 1];
          }
        }
        double ar = 0.0, xa = 0.0, ya = 0.0, xxa = 0.0, xya = 0.0, yya = 0.0, xxxa = 0.0, xxya = 0.0, xyya = 0.0,
               yyya = 0.0;
        for (int nv = 0; nv < 6; ++nv) {
          const int nv2 = nv == 5 ? 0 : nv + 1;
          double tv[4][2];   // tv[0] midpoint, tv[1..3] vertices
          tv[1][0] = cx[2 * nv]; tv[1][1] = cx[2 * nv + 1];
          tv[2][0] = cx[2 * nv2]; tv[2][1] = cx[2 * nv2 + 1];
          tv[3][0] = 0.0; tv[3][1] = 0.0;
          const double tri[6] = {tv[1][0], tv[1][1], tv[2][0], tv[2][1], tv[3][0], tv[3][1]};
          const double tarea = tri_area(tri);
          ar = ar + tarea;
          for (int s = 0; s < 2; ++s) tv[0][s] = kP333 * (tv[1][s] + tv[2][s] + tv[3][s]);
          for (int q = 1; q <= 3; ++q)
            for (int s = 0; s < 2; ++s) tv[q][s] = kP6 * tv[0][s] + kP4 * tv[q][s];
          xa = xa + tarea * tv[0][0];
          ya = ya + tarea * tv[0][1];
          // quadrature of x_p x_q (x_r; r = -1: absent) over the triangle;
          // the midpoint term is the left-to-right product p5625m * x * x (* x)
          auto cub = [&](int p, int q, int r) {
            auto f = [&](int v) {
              double x = tv[v][p] * tv[v][q];
              if (r >= 0) x = x * tv[v][r];
              return x;
            };
            double m0 = kP5625m * tv[0][p];
            m0 = m0 * tv[0][q];
            if (r >= 0) m0 = m0 * tv[0][r];
            return m0 + kP52083 * (f(1) + f(2) + f(3));
          };
          xxa = xxa + tarea * cub(0, 0, -1);
          xya = xya + tarea * cub(0, 1, -1);
          yya = yya + tarea * cub(1, 1, -1);
          xxxa = xxxa + tarea * cub(0, 0, 0);
          xxya = xxya + tarea * cub(0, 0, 1);
          xyya = xyya + tarea * cub(0, 1, 1);
          yyya = yyya + tarea * cub(1, 1, 1);
        }
        R.area_x3[n] = ar;
        const double ai = 1.0 / ar;
        R.area_inv_x3[n] = ai;
        R.xav[n] = xa * ai; R.yav[n] = ya * ai;
        R.xxav[n] = xxa * ai; R.xyav[n] = xya * ai; R.yyav[n] = yya * ai;
        R.xxxav[n] = xxxa * ai; R.xxyav[n] = xxya * ai; R.xyyav[n] = xyya * ai; R.yyyav[n] = yyya * ai;
      }
  for (int o = 0; o < 12; ++o) apply_halo_strided(G.face, R.c_x3.data(), 12, o);
  for (auto* v : {&R.area_x3, &R.area_inv_x3, &R.xav, &R.yav, &R.xxav, &R.xyav, &R.yyav, &R.xxxav, &R.xxyav,
                  &R.xyyav, &R.yyyav})
    apply_halo(G.face, v->data(), 1, 0);

  // gradient weights at cell centres (remap.F:913-1060): v_weights to X1,
  // summed per side, then to X3
  auto VW = [&](int nd, int nc, int i, int j, int k, double* out) {   // vwgt_x1(nd, :, nc, i, j, k)
    const size_t cs = 2 * S(i, j, k) + nc;
    const double v[3] = {G.v_weights[6 * cs + nd], G.v_weights[6 * cs + nd + 3], 0.0};
    matvec3(&G.trans_to_x1[9 * cs], v, out);
  };
  std::vector<double> gw1(3 * 7 * ns, 0.0);
  auto sum2 = [&](double* o, int nd1, int nc1, int di1, int dj1, int nd2, int nc2, int di2, int dj2, int i, int j,
                  int k) {
    double x[3], y[3];
    VW(nd1, nc1, i + di1, j + dj1, k, x);
    VW(nd2, nc2, i + di2, j + dj2, k, y);
    for (int s = 0; s < 3; ++s) o[s] = x[s] + y[s];
  };
  for (int k = 1; k <= 10; ++k) {
    for (int j = 2; j <= jjm - 1; ++j)
      for (int i = 2; i <= iim - 1; ++i) {
        double* g = &gw1[21 * S(i, j, k)];
        double v[6][3];
        VW(0, 0, i, j, k, v[0]); VW(0, 1, i, j, k, v[1]); VW(1, 0, i + 1, j, k, v[2]);
        VW(1, 1, i + 1, j + 1, k, v[3]); VW(2, 0, i + 1, j + 1, k, v[4]); VW(2, 1, i, j + 1, k, v[5]);
        for (int s = 0; s < 3; ++s) g[s] = v[0][s] + v[1][s] + v[2][s] + v[3][s] + v[4][s] + v[5][s];
        sum2(g + 3, 1, 0, 0, 0, 1, 1, 0, 1, i, j, k);
        sum2(g + 6, 2, 0, 0, 0, 1, 1, 0, 0, i, j, k);
        sum2(g + 9, 2, 1, 0, 0, 2, 0, 1, 0, i, j, k);
        sum2(g + 12, 0, 0, 1, 0, 2, 1, 1, 1, i, j, k);
        sum2(g + 15, 0, 1, 1, 1, 0, 0, 1, 1, i, j, k);
        sum2(g + 18, 1, 0, 1, 1, 0, 1, 0, 1, i, j, k);
      }
    {   // north pole (1, jjm-1): side 1 and vertex 6 missing
      const int i = 1, j = jjm - 1;
      double* g = &gw1[21 * S(i, j, k)];
      double v[5][3];
      VW(0, 0, i, j, k, v[0]); VW(0, 1, i, j, k, v[1]); VW(1, 0, i + 1, j, k, v[2]);
      VW(1, 1, i + 1, j + 1, k, v[3]); VW(2, 0, i + 1, j + 1, k, v[4]);
      for (int s = 0; s < 3; ++s) g[s] = v[0][s] + v[1][s] + v[2][s] + v[3][s] + v[4][s];
      for (int s = 0; s < 3; ++s) g[3 + s] = 0.0;
      sum2(g + 6, 2, 0, 0, 0, 1, 1, 0, 0, i, j, k);
      sum2(g + 9, 2, 1, 0, 0, 2, 0, 1, 0, i, j, k);
      sum2(g + 12, 0, 0, 1, 0, 2, 1, 1, 1, i, j, k);
      sum2(g + 15, 0, 1, 1, 1, 0, 0, 1, 1, i, j, k);
      sum2(g + 18, 1, 0, 1, 1, 1, 0, 0, 0, i, j, k);
    }
    {   // south pole (iim-1, 1): side 3 and vertex 3 missing
      const int i = iim - 1, j = 1;
      double* g = &gw1[21 * S(i, j, k)];
      double v[5][3];
      VW(0, 0, i, j, k, v[0]); VW(0, 1, i, j, k, v[1]); VW(1, 1, i + 1, j + 1, k, v[2]);
      VW(2, 0, i + 1, j + 1, k, v[3]); VW(2, 1, i, j + 1, k, v[4]);
      for (int s = 0; s < 3; ++s) g[s] = v[0][s] + v[1][s] + v[2][s] + v[3][s] + v[4][s];
      sum2(g + 3, 1, 0, 0, 0, 1, 1, 0, 1, i, j, k);
      sum2(g + 6, 2, 0, 0, 0, 1, 1, 0, 0, i, j, k);
      for (int s = 0; s < 3; ++s) g[9 + s] = 0.0;
      sum2(g + 12, 2, 1, 0, 0, 2, 1, 1, 1, i, j, k);
      sum2(g + 15, 0, 1, 1, 1, 0, 0, 1, 1, i, j, k);
      sum2(g + 18, 1, 0, 1, 1, 0, 1, 0, 1, i, j, k);
    }
    for (int j = 1; j <= jjm - 1; ++j)
      for (int i = 1; i <= iim - 1; ++i) {
        const size_t n = S(i, j, k);
        for (int nv = 0; nv < 7; ++nv) matvec3(&R.trans_x1_to_x3[9 * n], &gw1[21 * n + 3 * nv], &R.gwgt_x3[21 * n + 3 * nv]);
        auto CA = [&](int ic, int ii, int jj) { return G.corner_area[2 * S(ii, jj, k) + ic]; };
        const double t = CA(0, i, j) + CA(1, i, j) + CA(0, i + 1, j) + CA(1, i + 1, j + 1) + CA(0, i + 1, j + 1) +
                         CA(1, i, j + 1);
        R.garea_inv[n] = 1.0 / t;
      }
  }
}

const RemapGeometry& remap_geometry(const Grid& G) {
  static std::mutex mu;
  std::lock_guard<std::mutex> lock(mu);
  if (!G.remap) {
    auto R = std::make_shared<RemapGeometry>();
    build_remap_geometry(G, *R);
    G.remap = R;
  }
  return *G.remap;
}

}  // namespace swm
