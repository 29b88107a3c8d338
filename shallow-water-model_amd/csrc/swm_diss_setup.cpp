// swm_diss_setup.cpp -- the geometry of the dissipation operator,
// init_diss_op of swm/source/swm_dissipation_operator.F:50-203, restated in
// the same floating-point operation order (compiled without contraction).
#include <cmath>

#include "swm_setup.h"

namespace swm {

void build_diss_geometry(const Grid& G, DissGeometry& D) {
  const int iim = G.iim, jjm = G.jjm;
  const size_t ns = (size_t)iim * jjm * 10;
  auto S = [&](int i, int j, int k) { return (size_t)(i - 1) + (size_t)iim * ((size_t)(j - 1) + (size_t)jjm * (k - 1)); };
  auto P = [&](int i, int j, int k) { return &G.p[3 * S(i, j, k)]; };
  D.t.assign(24 * ns, 0.0);
  D.magt.assign(12 * ns, 0.0);
  D.magn.assign(6 * ns, 0.0);
  for (int k = 1; k <= 10; ++k)
    for (int j = 2; j <= jjm - 1; ++j)
      for (int i = 2; i <= iim - 1; ++i) {
        const size_t s = S(i, j, k);
        // distances between the centres across the three edges of each
        // corner (:76-109)
        const double* d[2][3][2] = {{{P(i - 1, j, k), P(i, j, k)}, {P(i - 1, j - 1, k), P(i - 1, j, k)},
                                     {P(i, j, k), P(i - 1, j - 1, k)}},
                                    {{P(i - 1, j - 1, k), P(i, j, k)}, {P(i, j - 1, k), P(i - 1, j - 1, k)},
                                     {P(i, j, k), P(i, j - 1, k)}}};
        for (int ic = 0; ic < 2; ++ic)
          for (int id = 0; id < 3; ++id) {
            double t[3];
            for (int r = 0; r < 3; ++r) t[r] = d[ic][id][0][r] - d[ic][id][1][r];
            D.magn[id + 3 * (ic + 2 * s)] = std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
          }
        // unit vectors from the corner and from the edge's other corner to
        // the edge midpoint, each in its corner's X2 frame (:116-197)
        const int oi[2][3] = {{i, i - 1, i}, {i, i, i + 1}}, oj[2][3] = {{j + 1, j, j}, {j, j - 1, j}};
        for (int ic = 0; ic < 2; ++ic) {
          const int oc = 1 - ic;
          for (int id = 0; id < 3; ++id) {
            const size_t so = S(oi[ic][id], oj[ic][id], k);
            const double* src[2] = {&G.c[3 * (2 * s + ic)], &G.c[3 * (2 * so + oc)]};
            const double* T[2] = {&G.trans_to_x2[9 * (2 * s + ic)], &G.trans_to_x2[9 * (2 * so + oc)]};
            const double* w = &G.w[9 * (2 * s + ic) + id];   // w(id, :, ic): stride 3
            for (int e = 0; e < 2; ++e) {
              double v[3], y[3];
              for (int r = 0; r < 3; ++r) v[r] = src[e][r] - w[3 * r];
              for (int r = 0; r < 3; ++r) {
                double acc = 0.0;
                for (int q = 0; q < 3; ++q) acc = acc + T[e][r + 3 * q] * v[q];
                y[r] = acc;
              }
              const double rmag = std::sqrt(y[0] * y[0] + y[1] * y[1] + y[2] * y[2]);
              for (int sp = 0; sp < 2; ++sp) D.t[e + 2 * (id + 3 * (sp + 2 * (ic + 2 * s)))] = y[sp] / rmag;
              D.magt[e + 2 * (id + 3 * (ic + 2 * s))] = rmag;
            }
          }
        }
      }
  const double a = G.radius;
  for (double& v : D.magn) v = a * v;
  for (double& v : D.magt) v = a * v;
}

}  // namespace swm
