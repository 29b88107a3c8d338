// swm_ics.cpp -- the reference's built-in initial state, shallow-water test
// case 5 (zonal flow over an isolated mountain), initial_conditions_swtc5 of
// swm/source/initialize.F:330-513, for the default build (thin_fluid 0).
//
// Host FP64 with the reference's operation order and libm calls, so the
// state is the reference's to the bit on the same grid, halos included: h,
// hs and u are set at real cells and corners and then wrapped
// (initialize.F:416-425), and the tracers derive from the wrapped hs.
#include <cmath>
#include <cstring>

#include "../../include/swm_mi355x.h"
#include "swm_internal.h"
#include "swm_setup.h"

namespace {

constexpr double kPi = 3.14159265358979323846;   // params_physical.F pi

// xyz_to_lonlat (geodesic_grid_utilities.F:686-732): (lon, lat)
inline void to_lonlat(const double* p, double* s) {
  s[1] = std::asin(p[2]);
  s[0] = (p[0] != 0.0 || p[1] != 0.0) ? std::atan2(p[1], p[0]) : 0.0;
}

}  // namespace

extern "C" swm_status swm_swtc5_state(const swm_grid* grid, int ntracer, double grav, double* h, double* hs,
                                      double* u, double* tracer) {
  if (!grid || !h || !hs || !u || !tracer || !(grav > 0.0)) {
    swm::set_error("swm_swtc5_state: bad argument");
    return SWM_ERR_ARG;
  }
  // initialize.F:466-470: the case needs two tracers at least
  if (ntracer < 2) {
    swm::set_error("swm_swtc5_state: test case 5 needs ntracer >= 2");
    return SWM_ERR_ARG;
  }
  const swm::Grid& G = swm::grid_of(grid);
  const size_t pl = (size_t)G.iim * G.jjm, ns = pl * 10;
  const double a = G.radius, omega = G.omega;
  const double f1 = 20.0, f2 = 5960.0, f3 = kPi / 9.0;   // initialize.F:362-364
  std::memset(h, 0, ns * sizeof(double));
  std::memset(hs, 0, ns * sizeof(double));
  std::memset(u, 0, 4 * ns * sizeof(double));
  double s[2];
  // initialize.F:376-410: thickness over the cone-shaped mountain at real
  // cells, zonal wind f1 cos(lat) at real corners
  for (size_t n = 0; n < ns; ++n) {
    if (G.center_mask[n] == 1.0) {
      to_lonlat(&G.p[3 * n], s);
      const double rlat = s[1];
      double rlon = s[0];
      if (rlon < 0.0) rlon = rlon + 2.0 * kPi;
      const double dx = rlon - 1.5 * kPi, dy = rlat - kPi / 6.0;
      const double dist = std::sqrt(dx * dx + dy * dy);
      hs[n] = dist < f3 ? 2000.0 * (1.0 - dist / f3) : 0.0;
      to_lonlat(&G.p[3 * n], s);
      const double sn = std::sin(s[1]);
      h[n] = f2 - (a * omega * f1 + 0.5 * (f1 * f1)) * (sn * sn) / grav;
      h[n] = h[n] - hs[n];
    }
    for (int ic = 0; ic < 2; ++ic) {
      if (G.corner_mask[2 * n + ic] == 1.0) {
        to_lonlat(&G.c[3 * (2 * n + ic)], s);
        u[4 * n + 2 * ic] = f1 * std::cos(s[1]);
        u[4 * n + 2 * ic + 1] = 0.0;
      }
    }
  }
  // initialize.F:416-425: halo fill of h, hs and u
  swm::apply_halo(G.face, h, 1, 0);
  swm::apply_halo(G.face, hs, 1, 0);
  swm::apply_halo_strided(G.vrtx, u, 2, 0);
  swm::apply_halo_strided(G.vrtx, u, 2, 1);
  // initialize.F:471-497: tracer 1 the surface height, tracer 2 a plateau of
  // 1000 over the mountain, the others the Coriolis parameter
  for (int k = 0; k < 10; ++k) {
    for (size_t ij = 0; ij < pl; ++ij) {
      const size_t n = (size_t)k * pl + ij;
      tracer[((size_t)0 + (size_t)ntracer * k) * pl + ij] = hs[n];
      tracer[((size_t)1 + (size_t)ntracer * k) * pl + ij] = hs[n] > 0.0 ? 1000.0 : 0.0;
      for (int q = 2; q < ntracer; ++q) tracer[((size_t)q + (size_t)ntracer * k) * pl + ij] = G.f[n];
    }
  }
  return SWM_OK;
}
