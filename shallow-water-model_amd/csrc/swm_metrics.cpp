// swm_metrics.cpp -- masks, index maps, operator metrics, Coriolis parameter and
// the multigrid hierarchy, restated from the reference set-up in the same
// floating-point operation order (compiled without contraction).
#include <cmath>
#include <cstdio>
#include <stdexcept>

#include "swm_setup.h"

namespace swm {

namespace {

inline int ipow2(int n) { return 1 << n; }

// ---- geometry helpers (utilities/layer_2/geodesic_grid_utilities.F) -------
inline void cross3(const double* a, const double* b, double* o) {   // :446-460
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
inline void unit3(double* a) {                                       // :466-480
  double len = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
  len = 1.0 / len;
  a[0] = a[0] * len; a[1] = a[1] * len; a[2] = a[2] * len;
}
inline void tangent_plane(const double* p1, const double* p2, double* o) {  // :500-520
  double w[3];
  w[0] = -(p1[1] * (-(p1[1] * p2[0]) + p1[0] * p2[1])) + p1[2] * (p1[2] * p2[0] - p1[0] * p2[2]);
  w[1] = +(p1[0] * (-(p1[1] * p2[0]) + p1[0] * p2[1])) - p1[2] * (-(p1[2] * p2[1]) + p1[1] * p2[2]);
  w[2] = -(p1[0] * (p1[2] * p2[0] - p1[0] * p2[2])) + p1[1] * (-(p1[2] * p2[1]) + p1[1] * p2[2]);
  unit3(w);
  o[0] = w[0]; o[1] = w[1]; o[2] = w[2];
}
inline double triangle_area(const double* p1, const double* p2, const double* p3) {  // :600-630
  double r[3];
  for (int i = 0; i < 3; ++i) {
    const int j = (i + 1) % 3;
    r[i] = p2[i] * p3[j] - p2[j] * p3[i] - p1[i] * p3[j] + p1[j] * p3[i] + p1[i] * p2[j] - p1[j] * p2[i];
  }
  return 0.5 * std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
}
inline double line_distance(const double* a, const double* b) {     // :640-655
  const double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
  return std::sqrt(dx * dx + dy * dy + dz * dz);
}
inline void lonlat(const double* p, double* s) {                      // :686-739
  s[1] = std::asin(p[2]);
  if (p[0] != 0.0 || p[1] != 0.0) s[0] = std::atan2(p[1], p[0]);
  else s[0] = 0.0;
}
// planar Voronoi corner: the foot of the perpendicular from the origin onto the
// plane of the three centres, located by bisection (:269-377)
void voronoi_corner(const double* p1, const double* p2, const double* p3, double* out) {
  double t1[3], t2[3], t3[3];
  for (int k = 0; k < 3; ++k) { t1[k] = p2[k] - p1[k]; t2[k] = p3[k] - p1[k]; }
  cross3(t1, t2, t3);
  unit3(t3);
  const double d = -(t3[0] * p1[0] + t3[1] * p1[1] + t3[2] * p1[2]);
  double fu = 1.0, fl = 0.0, resid = 1.0, factor = 0.0;
  int iter = 0;
  while (std::fabs(resid) > 1.0e-14) {
    if (++iter == 1000) throw std::runtime_error("voronoi corner: bisection not converging");
    factor = (fl + fu) / 2.0;
    resid = factor * (t3[0] * t3[0] + t3[1] * t3[1] + t3[2] * t3[2]) + d;
    if (resid > 0.0) fu = factor; else fl = factor;
  }
  for (int k = 0; k < 3; ++k) out[k] = factor * t3[k];
}

// index helpers for (…, iim, jjm, 10) arrays, 1-based as in the reference
struct Ix {
  int iim, jjm;
  size_t s(int i, int j, int k) const { return (size_t)(i - 1) + (size_t)iim * ((j - 1) + (size_t)jjm * (k - 1)); }
};

// panel block of a global (im, jm) array (gather_scatter.F:46-171, one rank)
template <class T>
void scatter_panels(const std::vector<T>& glob, int im, int jm, int iim, int jjm, int nv, std::vector<T>& out) {
  const int im5 = im / 5;
  out.assign((size_t)nv * iim * jjm * 10, T(0));
  int nsd = 0;
  for (int panel = 0; panel <= 4; ++panel)
    for (int isd = 0; isd <= 1; ++isd) {
      ++nsd;
      for (int j = 1; j <= jjm; ++j)
        for (int i = 1; i <= iim; ++i) {
          const size_t gi = (size_t)(im5 * panel + (iim - 2) * isd + i - 1) + (size_t)im * (j - 1);
          const size_t pi = (size_t)(i - 1) + (size_t)iim * ((j - 1) + (size_t)jjm * (nsd - 1));
          for (int v = 0; v < nv; ++v) out[(size_t)nv * pi + v] = glob[(size_t)nv * gi + v];
        }
    }
}

// (grid_metrics_generator.F:1144-1223)
std::vector<double> polygon_types(int g) {
  const int iim = 2 + ipow2(g), jjm = iim, jm = jjm, im = 5 * (2 + 2 * ipow2(g)), im5 = im / 5;
  Ix X{iim, jjm};
  std::vector<double> pt((size_t)iim * jjm * 10, 0.0), work((size_t)im * jm, 0.0), w1;
  for (int k = 1; k <= 10; ++k)
    for (int j = 2; j <= jjm - 1; ++j)
      for (int i = 2; i <= iim - 1; ++i) pt[X.s(i, j, k)] = 1.0;
  for (int panel = 0; panel <= 4; ++panel) {
    work[(size_t)(im5 * panel + jm - 1 - 1) + (size_t)im * (jm - 2)] = 1.0;
    work[(size_t)(im5 * (panel + 1) - 1 - 1) + (size_t)im * (jm - 2)] = 1.0;
  }
  scatter_panels(work, im, jm, iim, jjm, 1, w1);
  for (size_t n = 0; n < pt.size(); ++n) pt[n] += w1[n];
  for (int k = 1; k <= 10; ++k) {
    for (int i = 1; i <= iim; ++i) { pt[X.s(i, 1, k)] = 0.0; pt[X.s(i, jjm, k)] = 0.0; }
    for (int j = 1; j <= jjm; ++j) { pt[X.s(1, j, k)] = 0.0; pt[X.s(iim, j, k)] = 0.0; }
  }
  std::fill(work.begin(), work.end(), 0.0);
  work[(size_t)0 + (size_t)im * (jm - 2)] = 3.0;
  work[(size_t)(im - 2)] = 4.0;
  scatter_panels(work, im, jm, iim, jjm, 1, w1);
  for (size_t n = 0; n < pt.size(); ++n) {
    pt[n] += w1[n];
    if (pt[n] < 0.5) pt[n] = -1.0;
  }
  return pt;
}

// global (im, jm) array from the list order (gather_scatter.F:366-408)
std::vector<double> list_to_global(const double* plist, int g, int nv) {
  const int jm = 2 + ipow2(g), im = 5 * (2 + 2 * ipow2(g)), im5 = im / 5;
  std::vector<double> glob((size_t)nv * im * jm, 0.0);
  auto put = [&](int i, int j, int icount) {
    for (int v = 0; v < nv; ++v) glob[(size_t)nv * ((size_t)(i - 1) + (size_t)im * (j - 1)) + v] = plist[(size_t)nv * icount + v];
  };
  put(1, jm - 1, 0);
  put(im - 1, 1, 1);
  int icount = 2;
  for (int j = 2; j <= jm - 1; ++j)
    for (int panel = 0; panel <= 4; ++panel)
      for (int i = 2; i <= im5 - 1; ++i) put(im5 * panel + i, j, icount++);
  return glob;
}

struct LevelMetrics {
  std::vector<double> c, area, area_inv, corner_area, corner_area_inv, sub_area, dc, normal,
      tangent, v_weights, s_weights, l_weights, polygon_type, w, trans_to_x2, trans_to_x1;
  double tot_area = 0.0;
};

// (grid_metrics_generator.F:55-626 generate_grid_metrics_level, l_sphere = .false.)
// p: (3, iim, jjm, 10), already halo-filled
void metrics_level(int g, const std::vector<double>& p, double radius, LevelMetrics& M) {
  const int iim = 2 + ipow2(g), jjm = iim;
  Ix X{iim, jjm};
  const size_t ns = (size_t)iim * jjm * 10;
  M.c.assign(3 * 2 * ns, 0.0);
  M.area.assign(ns, 0.0); M.area_inv.assign(ns, 0.0);
  M.corner_area.assign(2 * ns, 0.0); M.corner_area_inv.assign(2 * ns, 0.0);
  M.sub_area.assign(3 * 2 * ns, 0.0); M.dc.assign(3 * 2 * ns, 0.0);
  M.normal.assign(3 * 3 * 2 * ns, 0.0); M.tangent.assign(3 * 3 * 2 * ns, 0.0);
  M.v_weights.assign(3 * 2 * 2 * ns, 0.0); M.s_weights.assign(3 * 2 * ns, 0.0);
  M.l_weights.assign(7 * ns, 0.0);
  M.w.assign(3 * 3 * 2 * ns, 0.0);
  M.trans_to_x2.assign(3 * 3 * 2 * ns, 0.0);
  M.trans_to_x1.assign(3 * 3 * 2 * ns, 0.0);
  M.polygon_type = polygon_types(g);
  const std::vector<double>& pt = M.polygon_type;

  std::vector<double> mmask(2 * ns, 0.0), cmask(2 * ns, 0.0), gmask(ns, 0.0);
  for (size_t n = 0; n < ns; ++n) gmask[n] = pt[n] > 0.0 ? 1.0 : 0.0;
  for (int k = 1; k <= 10; ++k)
    for (int j = 2; j <= jjm - 1; ++j)
      for (int i = 2; i <= iim - 1; ++i) { cmask[2 * X.s(i, j, k)] = 1.0; cmask[2 * X.s(i, j, k) + 1] = 1.0; }
  for (int k = 1; k <= 10; ++k) {
    for (int j = 2; j <= jjm; ++j)
      for (int i = 2; i <= iim; ++i) { mmask[2 * X.s(i, j, k)] = 1.0; mmask[2 * X.s(i, j, k) + 1] = 1.0; }
    mmask[2 * X.s(iim, 2, k) + 1] = 0.0;
    mmask[2 * X.s(2, jjm, k) + 0] = 0.0;
  }
  for (int k = 1; k <= 10; ++k)
    for (int j = 1; j <= jjm; ++j)
      for (int i = 1; i <= iim; ++i) {
        const double t = pt[X.s(i, j, k)];
        if (t == 2.0) mmask[2 * X.s(i + 1, j + 1, k) + 0] = 0.0;
        if (t == 3.0) {
          mmask[2 * X.s(i, j, k)] = 1.0; mmask[2 * X.s(i, j, k) + 1] = 1.0;
          mmask[2 * X.s(i + 1, j + 1, k) + 0] = 1.0;
        }
        if (t == 4.0) {
          mmask[2 * X.s(i, j, k)] = 1.0; mmask[2 * X.s(i, j, k) + 1] = 1.0;
          mmask[2 * X.s(i + 1, j + 1, k) + 1] = 1.0;
        }
      }

  auto P = [&](int i, int j, int k) { return &p[3 * X.s(i, j, k)]; };
  for (int k = 1; k <= 10; ++k)
    for (int j = 1; j <= jjm; ++j)
      for (int i = 1; i <= iim; ++i)
        for (int ic = 0; ic < 2; ++ic) {
          const size_t cs = 2 * X.s(i, j, k) + ic;   // corner slot
          if (mmask[cs] != 1.0) continue;
          // surrounding centres, counter-clockwise (geodesic_grid_utilities.F:47-147)
          const double* t[3];
          const double typ = pt[X.s(i, j, k)];
          if (typ <= 2.0) {
            if (ic == 0) { t[0] = P(i, j, k); t[1] = P(i - 1, j, k); t[2] = P(i - 1, j - 1, k); }
            else { t[0] = P(i, j, k); t[1] = P(i - 1, j - 1, k); t[2] = P(i, j - 1, k); }
          } else if (typ == 3.0) {
            if (ic == 0) { t[0] = P(i, j, k); t[1] = P(i, j + 1, k); t[2] = P(iim, 1, k); }
            else { t[0] = P(i, j, k); t[1] = P(iim, 1, k); t[2] = P(i, j - 1, k); }
          } else {
            if (ic == 0) { t[0] = P(i, j, k); t[1] = P(i - 1, j, k); t[2] = P(1, jjm, k); }
            else { t[0] = P(i, j, k); t[1] = P(1, jjm, k); t[2] = P(i + 1, j, k); }
          }
          double* c = &M.c[3 * cs];
          voronoi_corner(t[0], t[1], t[2], c);
          // local east / north / up frame
          double h3[3] = {c[0], c[1], c[2]}, h2[3], h1[3];
          unit3(h3);
          const double zhat[3] = {0.0, 0.0, 1.0};
          tangent_plane(h3, zhat, h2);
          cross3(h2, h3, h1);
          const double* X2[3] = {h1, h2, h3};   // rows; the transform to X2 is X2 itself
          // trans_to_X2 = X2 X1^T and trans_to_X1 = X1 X2^T with X1 the
          // identity: the entries of X2, each a sum of one product and two
          // zero products (a zero of either sign ends as +0)
          for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
              M.trans_to_x2[9 * cs + a + 3 * b] = X2[a][b] + 0.0;
              M.trans_to_x1[9 * cs + a + 3 * b] = X2[b][a] + 0.0;
            }
          double r[3][3];
          for (int kk = 0; kk < 3; ++kk) {
            const int l = (kk + 1) % 3;
            for (int s = 0; s < 3; ++s) r[kk][s] = 0.5 * (t[kk][s] + t[l][s]);
            for (int s = 0; s < 3; ++s) M.w[9 * cs + kk + 3 * s] = r[kk][s];   // w(kk, s, corner)
            double gv[3] = {c[0] - r[kk][0], c[1] - r[kk][1], c[2] - r[kk][2]};
            unit3(gv);
            double* tg = &M.tangent[3 * (3 * cs) + 0];
            // tangent(kk, :, corner): components are strided by 3 (ndata first)
            double tl[3];
            for (int a = 0; a < 3; ++a) tl[a] = X2[a][0] * gv[0] + X2[a][1] * gv[1] + X2[a][2] * gv[2];
            unit3(tl);
            double up[3];
            for (int a = 0; a < 3; ++a) up[a] = X2[a][0] * h3[0] + X2[a][1] * h3[1] + X2[a][2] * h3[2];
            double nl[3];
            cross3(tl, up, nl);
            unit3(nl);
            for (int a = 0; a < 3; ++a) {
              tg[kk + 3 * a] = tl[a];
              M.normal[9 * cs + kk + 3 * a] = nl[a];
            }
          }
          for (int kk = 0; kk < 3; ++kk) {
            const int l = (kk + 2) % 3;
            const double f1 = triangle_area(t[kk], r[kk], c);
            const double f2 = triangle_area(t[kk], c, r[l]);
            M.sub_area[3 * cs + kk] = f1 + f2;
            M.dc[3 * cs + kk] = line_distance(r[kk], c);
          }
          M.corner_area[cs] = M.sub_area[3 * cs] + M.sub_area[3 * cs + 1] + M.sub_area[3 * cs + 2];
          M.corner_area_inv[cs] = 1.0 / M.corner_area[cs];
        }

  // cell areas (grid_metrics_generator.F:945-997)
  auto SA = [&](int kk, int ic, int i, int j, int k) { return M.sub_area[3 * (2 * X.s(i, j, k) + ic) + kk]; };
  for (int k = 1; k <= 10; ++k)
    for (int j = 1; j <= jjm; ++j)
      for (int i = 1; i <= iim; ++i)
        if (pt[X.s(i, j, k)] > 0.0) {
          const double a = SA(0, 0, i, j, k) + SA(0, 1, i, j, k) + SA(1, 0, i + 1, j, k) +
                           SA(1, 1, i + 1, j + 1, k) + SA(2, 0, i + 1, j + 1, k) + SA(2, 1, i, j + 1, k);
          M.area[X.s(i, j, k)] = a;
          M.area_inv[X.s(i, j, k)] = 1.0 / a;
        }
  // scale to the sphere of the given radius (left-to-right products)
  const double a = radius, a2 = radius * radius;
  for (double& v : M.dc) v = v * a;
  for (double& v : M.sub_area) v = v * a * a;
  for (double& v : M.corner_area) v = v * a * a;
  for (double& v : M.corner_area_inv) v = v / a2;
  for (double& v : M.area) v = v * a * a;
  for (double& v : M.area_inv) v = v / a2;
  double tot = 0.0;
  for (size_t n = 0; n < ns; ++n) tot = tot + M.area[n] * gmask[n];
  M.tot_area = tot;

  // scalar averaging weights (:634-714)
  for (size_t cs = 0; cs < 2 * ns; ++cs)
    for (int kk = 0; kk < 3; ++kk) M.s_weights[3 * cs + kk] = M.sub_area[3 * cs + kk] * M.corner_area_inv[cs];
  // gradient weights (:722-759)
  for (size_t cs = 0; cs < 2 * ns; ++cs)
    for (int kk = 0; kk < 3; ++kk) {
      const int l = (kk + 2) % 3;
      for (int sp = 0; sp < 2; ++sp)
        M.v_weights[6 * cs + kk + 3 * sp] =
            M.dc[3 * cs + l] * M.normal[9 * cs + l + 3 * sp] - M.dc[3 * cs + kk] * M.normal[9 * cs + kk + 3 * sp];
    }
  // Laplacian weights: divergence of the gradient (:767-937)
  auto VW = [&](int kk, int sp, int ic, int i, int j, int k) { return M.v_weights[6 * (2 * X.s(i, j, k) + ic) + kk + 3 * sp]; };
  auto CAI = [&](int ic, int i, int j, int k) { return M.corner_area_inv[2 * X.s(i, j, k) + ic]; };
  for (int k = 1; k <= 10; ++k)
    for (int l = 1; l <= jjm; ++l)
      for (int kk = 1; kk <= iim; ++kk) {
        double* lw = &M.l_weights[7 * X.s(kk, l, k)];
        if (pt[X.s(kk, l, k)] > 0.0) {
          for (int sp = 0; sp < 2; ++sp) {
            struct Term { int di, dj, ic, id, ie, iff, wd, we, wf; };
            // corner, which v_weight is d/e/f, and which l_weight gets d*d/d*e/d*f
            static const Term T[6] = {
                {0, 0, 0, 0, 1, 2, 6, 0, 1}, {0, 0, 1, 0, 1, 2, 6, 1, 2}, {1, 0, 0, 1, 2, 0, 6, 2, 3},
                {1, 1, 1, 1, 2, 0, 6, 3, 4}, {1, 1, 0, 2, 0, 1, 6, 4, 5}, {0, 1, 1, 2, 0, 1, 6, 5, 0}};
            for (const Term& tm : T) {
              const int i = kk + tm.di, j = l + tm.dj;
              const double ai = CAI(tm.ic, i, j, k);
              const double d = VW(tm.id, sp, tm.ic, i, j, k);
              const double e = VW(tm.ie, sp, tm.ic, i, j, k);
              const double f = VW(tm.iff, sp, tm.ic, i, j, k);
              lw[6] = lw[6] + ai * d * d;
              lw[tm.we] = lw[tm.we] + ai * d * e;
              lw[tm.wf] = lw[tm.wf] + ai * d * f;
            }
          }
        }
        const double ainv = M.area_inv[X.s(kk, l, k)];
        for (int n = 0; n < 7; ++n) lw[n] = -lw[n] * ainv;
      }
  for (size_t n = 0; n < ns; ++n) {
    double* lw = &M.l_weights[7 * n];
    if (pt[n] == 2.0) { lw[5] = lw[5] + lw[4]; lw[4] = 0.0; }
  }
  for (size_t n = 0; n < ns; ++n) {
    double* lw = &M.l_weights[7 * n];
    if (pt[n] == 3.0) { lw[5] = lw[5] + lw[0]; lw[0] = 0.0; }
    if (pt[n] == 4.0) { lw[3] = lw[3] + lw[2]; lw[2] = 0.0; }
  }
}

}  // namespace

void build_grid(Grid& G, int g, const double* p_list, double radius, double omega) {
  G.g = g;
  G.iim = G.jjm = 2 + ipow2(g);
  G.im = 5 * (2 + 2 * ipow2(g));
  G.jm = 2 + ipow2(g);
  G.max_ig = 2 + 10 * ipow2(g) * ipow2(g);
  G.radius = radius;
  G.omega = omega;
  const int iim = G.iim, jjm = G.jjm;
  Ix X{iim, jjm};
  const size_t ns = (size_t)iim * jjm * 10;

  if (p_list) G.p_list.assign(p_list, p_list + 3 * (size_t)G.max_ig);
  else G.p_list = generate_points(g);

  build_halo_tables(g, G.face, G.vrtx, G.edge);

  // index maps (grid_mask.F:60-212, 221-333)
  {
    std::vector<double> idx((size_t)G.max_ig);
    for (int n = 0; n < G.max_ig; ++n) idx[n] = (double)(n + 1);
    std::vector<double> glob = list_to_global(idx.data(), g, 1), pan;
    scatter_panels(glob, G.im, G.jm, iim, jjm, 1, pan);
    apply_halo(G.face, pan.data(), 1, 0);
    G.center_index.resize(ns);
    for (size_t n = 0; n < ns; ++n) G.center_index[n] = (int32_t)std::lround(pan[n]);
    std::vector<double> cidx(2 * ns, 0.0);
    for (size_t n = 0; n < ns; ++n)
      if (G.center_index[n] > 2) {
        cidx[2 * n] = (double)((G.center_index[n] - 3) * 2 + 1);
        cidx[2 * n + 1] = (double)((G.center_index[n] - 3) * 2 + 2);
      }
    apply_halo(G.vrtx, cidx.data(), 1, 0);
    G.corner_index.resize(2 * ns);
    for (size_t n = 0; n < 2 * ns; ++n) G.corner_index[n] = (int32_t)std::lround(cidx[n]);
  }

  // cell centres on the panels, halo-filled (grid_metrics.F:55-120)
  {
    std::vector<double> glob = list_to_global(G.p_list.data(), g, 3);
    scatter_panels(glob, G.im, G.jm, iim, jjm, 3, G.p);
    std::vector<double> work(ns);
    for (int s = 0; s < 3; ++s) {
      for (size_t n = 0; n < ns; ++n) work[n] = G.p[3 * n + s];
      apply_halo(G.face, work.data(), 1, 0);
      for (size_t n = 0; n < ns; ++n) G.p[3 * n + s] = work[n];
    }
  }
  LevelMetrics M;
  metrics_level(g, G.p, radius, M);
  for (int s = 0; s < 3; ++s) apply_halo_strided(G.vrtx, M.c.data(), 3, s);  // corners (grid_metrics.F:151)
  // the corner frames are wrapped (grid_metrics_generator.F:520-524), the
  // edge midpoints are not
  for (int s = 0; s < 9; ++s) {
    apply_halo_strided(G.vrtx, M.trans_to_x1.data(), 9, s);
    apply_halo_strided(G.vrtx, M.trans_to_x2.data(), 9, s);
  }
  G.c.swap(M.c);
  G.w.swap(M.w);
  G.trans_to_x1.swap(M.trans_to_x1);
  G.trans_to_x2.swap(M.trans_to_x2);
  G.area.swap(M.area); G.area_inv.swap(M.area_inv);
  G.corner_area.swap(M.corner_area); G.corner_area_inv.swap(M.corner_area_inv);
  G.sub_area.swap(M.sub_area); G.dc.swap(M.dc);
  G.normal.swap(M.normal); G.tangent.swap(M.tangent);
  G.v_weights.swap(M.v_weights); G.s_weights.swap(M.s_weights); G.l_weights.swap(M.l_weights);
  G.polygon_type.swap(M.polygon_type);
  G.tot_area = M.tot_area;
  G.center_mask.assign(ns, 0.0);
  for (size_t n = 0; n < ns; ++n) G.center_mask[n] = G.polygon_type[n] > 0.0 ? 1.0 : 0.0;
  G.corner_mask.assign(2 * ns, 0.0);
  for (int k = 1; k <= 10; ++k)
    for (int j = 2; j <= jjm - 1; ++j)
      for (int i = 2; i <= iim - 1; ++i) { G.corner_mask[2 * X.s(i, j, k)] = 1.0; G.corner_mask[2 * X.s(i, j, k) + 1] = 1.0; }

  // Coriolis parameter at real cells, then halo (initialize.F:286-321)
  G.f.assign(ns, 0.0);
  for (size_t n = 0; n < ns; ++n)
    if (G.center_mask[n] == 1.0) {
      double s[2];
      lonlat(&G.p[3 * n], s);
      G.f[n] = 2.0 * std::sin(s[1]) * omega;
    }
  apply_halo(G.face, G.f.data(), 1, 0);

  // multigrid hierarchy: coarse levels take every other centre of the finer
  // level and are rediscretised (multigrid_solver.F:272-474)
  G.levels.assign(g, Level());
  std::vector<double> pglob = list_to_global(G.p_list.data(), g, 3);
  for (int lev = g; lev >= 1; --lev) {
    Level& L = G.levels[lev - 1];
    L.g = lev;
    L.iim = L.jjm = 2 + ipow2(lev);
    const int im_l = 5 * (2 + 2 * ipow2(lev)), jm_l = 2 + ipow2(lev);
    HaloTable fv, fe;
    build_halo_tables(lev, L.face, fv, fe);
    std::vector<double> pl;
    scatter_panels(pglob, im_l, jm_l, L.iim, L.jjm, 3, pl);
    const size_t nsl = (size_t)L.iim * L.jjm * 10;
    std::vector<double> work(nsl);
    for (int s = 0; s < 3; ++s) {
      for (size_t n = 0; n < nsl; ++n) work[n] = pl[3 * n + s];
      apply_halo(L.face, work.data(), 1, 0);
      for (size_t n = 0; n < nsl; ++n) pl[3 * n + s] = work[n];
    }
    LevelMetrics LM;
    metrics_level(lev, pl, radius, LM);
    Ix XL{L.iim, L.jjm};
    for (int k = 1; k <= 10; ++k) {
      double& w7n = LM.l_weights[7 * XL.s(1, L.jjm - 1, k) + 6];
      if (w7n == 0.0) w7n = 1.0;
      double& w7s = LM.l_weights[7 * XL.s(L.iim - 1, 1, k) + 6];
      if (w7s == 0.0) w7s = 1.0;
    }
    L.lap_w.swap(LM.l_weights);
    if (lev > 1) {
      const int im_c = 5 * (2 + 2 * ipow2(lev - 1)), jm_c = 2 + ipow2(lev - 1);
      const int im5_c = im_c / 5, im5_f = im_l / 5;
      std::vector<double> coarse(3 * (size_t)im_c * jm_c, 0.0);
      for (int panel = 0; panel <= 4; ++panel)
        for (int j = 2; j <= jm_c - 1; ++j)
          for (int i = 2; i <= im5_c - 1; ++i)
            for (int s = 0; s < 3; ++s)
              coarse[3 * ((size_t)(im5_c * panel + i - 1) + (size_t)im_c * (j - 1)) + s] =
                  pglob[3 * ((size_t)(im5_f * panel + 2 * i - 1 - 1) + (size_t)im_l * (2 * j - 1 - 1)) + s];
      for (int s = 0; s < 3; ++s) {
        coarse[3 * ((size_t)0 + (size_t)im_c * (jm_c - 2)) + s] = pglob[3 * ((size_t)0 + (size_t)im_l * (jm_l - 2)) + s];
        coarse[3 * ((size_t)(im_c - 2)) + s] = pglob[3 * ((size_t)(im_l - 2)) + s];
      }
      pglob.swap(coarse);
    }
  }
}

}  // namespace swm
