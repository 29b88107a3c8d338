// swm_diss.hip -- the dissipation operator on the device (FP64):
// dissipation of swm/source/swm_dissipation_operator.F:212-443, one
// application of its vector Laplacian per launch.
//
// One thread per slot: the edge divergences and curls of the cell's six
// corner-edge pairs, the gradients at its two corners, k x the curl's
// gradient, their sum, times f (1, or the viscosity with its sign after the
// last application); zero at the other slots.  The corner vectors are four
// planes (space component sp, corner ic) at (2 sp + ic) ns; the geometry is
// 36 planes: t (a, id, sp, ic) at a + 2 (id + 3 (sp + 2 ic)) (24),
// magt(1) + magt(2) (id, ic) at 24 + id + 3 ic, magn (id, ic) at
// 30 + id + 3 ic.

namespace swm {
namespace dev {

constexpr int DS_PLANES = 36;

__global__ void k_diss(int iim, int jjm, int ns, const double* __restrict__ geo, const double* __restrict__ cai,
                       double f, const double* __restrict__ x, double* __restrict__ y) {
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  const int i = c.i, j = c.j, k = c.k;
  const size_t n = (size_t)ns;
  auto X = [&](int sp, int ic, int sl) { return x[(size_t)(2 * sp + ic) * n + sl]; };
  if (!interior(c, iim, jjm)) {
    for (int p = 0; p < 4; ++p) y[(size_t)p * n + s] = f * 0.0;
    return;
  }
  const int oi[2][3] = {{i, i - 1, i}, {i, i, i + 1}}, oj[2][3] = {{j + 1, j, j}, {j, j - 1, j}};
  for (int ic = 0; ic < 2; ++ic) {
    const int oc = 1 - ic;
    double gd[2] = {0.0, 0.0}, gr[2] = {0.0, 0.0};
    for (int id = 0; id < 3; ++id) {
      const int so = (oi[ic][id] - 1) + iim * ((oj[ic][id] - 1) + jjm * (k - 1));
#define T(a, sp) geo[(size_t)((a) + 2 * (id + 3 * ((sp) + 2 * ic))) * n + s]
      const double msum = geo[(size_t)(24 + id + 3 * ic) * n + s];
      const double x1 = X(0, ic, s), x2 = X(1, ic, s), o1 = X(0, oc, so), o2 = X(1, oc, so);
      const double dv = (x1 * T(0, 0) + x2 * T(0, 1) + o1 * T(1, 0) + o2 * T(1, 1)) / msum;
      const double rl = (-x1 * T(0, 1) + x2 * T(0, 0) - o1 * T(1, 1) + o2 * T(1, 0)) / msum;
      const double mn = geo[(size_t)(30 + id + 3 * ic) * n + s], ca = cai[(size_t)ic * n + s];
      for (int sp = 0; sp < 2; ++sp) {
        gd[sp] = gd[sp] - dv * mn * T(0, sp) * ca;
        gr[sp] = gr[sp] - rl * mn * T(0, sp) * ca;
      }
#undef T
    }
    y[(size_t)(0 + ic) * n + s] = f * (gd[0] + -gr[1]);
    y[(size_t)(2 + ic) * n + s] = f * (gd[1] + gr[0]);
  }
}

}  // namespace dev
}  // namespace swm
