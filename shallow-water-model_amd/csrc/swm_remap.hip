// swm_remap.hip -- incremental remapping of mass and tracers on the device
// (FP64): remap_mass_and_tracers (utilities/layer_2/remap.F:1072-1604) with
// the arguments get_remap_flux passes (swm/source/step.F:687-771) and the
// step's ABM3 update of thickness and tracers from the remapped fields.
//
// One thread per slot in every kernel, the reference's operation order in
// each (compiled without contraction):
//   k_rm_construct  remap_prep, construct_fields: cell-centre gradients of
//                   mass and tracers, centres of mass, the tracers' van Leer
//                   limit (remap.F:1613-2397)
//   k_rm_departure  departure_points_1order at the cell's two corners
//                   (remap.F:2406-2537)
//   k_rm_transport  locate_triangles and the linear transport_integrals of
//                   the cell's three edges (remap.F:2903-3963)
//   k_rm_update     update_fields (remap.F:3972-4145), the edge mass flux
//                   and the ABM3 update of get_remap_flux
// with the halo fills of the reference (face: fields and gradients; vertex:
// departure points; edge: transports) between them on the host side.
//
// Data: cell fields one plane of ns slots, corner fields two planes (corner
// ic at ic * ns), edge fields three planes (edge e at e * ns).  The
// remapping quantities of init_remap are planes of the same kinds:
//   cell   (33): gwgt_x3 (c, nv) at c * 7 + nv (c = 0, 1; nv = 0..6),
//                garea_inv 14, xav 15, yav 16, xxav 17, xyav 18, yyav 19,
//                c_x3 (c, nv) at 20 + 2 * nv + c, area_inv_x3 32
//   edge   (43): trans_X1_to_X4 rows 0-1 at r * 3 + c (0..5),
//                trans_X4_to_X3 (r, c, nd) at 6 + 4 * nd + r + 2 * c,
//                cl 22, cr 23, tl 24-25, tr 26-27, bl 28-29, br 30-31,
//                p_x4 (c, nd) at 32 + 2 * nd + c, edge midpoint w 40-42
//   corner  (9): trans_to_X1 columns 0-1 at r + 3 * c (0..5), corner c 6-8

namespace swm {
namespace dev {

#define SL(i, j, k) ((i) - 1 + iim * ((j) - 1 + jjm * ((k) - 1)))

constexpr double RM_EPS = 1.0e-10;   // remap.F eps
constexpr int RM_CELL = 33, RM_EDGE = 43, RM_CORN = 9;

struct RemapArgs {
  int iim, jjm, ns, nt;
  const double* cellp;
  const double* edgep;
  const double* cornp;
  double a;                        // radius of the sphere
  int limit_min, limit_max;        // l_limit_tracers_min / _max
  int* flag;                       // set to 1 where a thickness is below -eps (remap_prep stops)
};

__device__ __forceinline__ double rm_min(double a, double b) { return b < a ? b : a; }
__device__ __forceinline__ double rm_max(double a, double b) { return b > a ? b : a; }

// remap_prep of one slot: thickness below eps becomes zero, and so do its
// tracers (mask 0)
__device__ __forceinline__ double rm_mass(const double* h3, int s) {
  const double m = h3[s];
  return m < RM_EPS ? 0.0 : m;
}
__device__ __forceinline__ double rm_trace(const double* h3, const double* t3, int s) {
  return h3[s] < RM_EPS ? 0.0 : t3[s];
}

// operators.F:691-835 gradient_at_cell_centers at one slot (x, y); X(sl)
// the field
template <class FX>
__device__ __forceinline__ void rm_gradient_at(const RemapArgs& A, const IJK& c, int s, FX X, double& gx,
                                               double& gy) {
  const int iim = A.iim, jjm = A.jjm, ns = A.ns;
  const int i = c.i, j = c.j, k = c.k;
#define GW(cc, nv) A.cellp[(size_t)((cc) * 7 + (nv)) * ns + s]
  const double ga = A.cellp[(size_t)14 * ns + s];
  gx = 0.0;
  gy = 0.0;
  if (i >= 2 && i <= iim - 1 && j >= 2 && j <= jjm - 1) {
    const double x0 = X(s), x1 = X(SL(i - 1, j, k)), x2 = X(SL(i - 1, j - 1, k)), x3 = X(SL(i, j - 1, k)),
                 x4 = X(SL(i + 1, j, k)), x5 = X(SL(i + 1, j + 1, k)), x6 = X(SL(i, j + 1, k));
    gx = (x0 * GW(0, 0) + x1 * GW(0, 1) + x2 * GW(0, 2) + x3 * GW(0, 3) + x4 * GW(0, 4) + x5 * GW(0, 5) +
          x6 * GW(0, 6)) * ga;
    gy = (x0 * GW(1, 0) + x1 * GW(1, 1) + x2 * GW(1, 2) + x3 * GW(1, 3) + x4 * GW(1, 4) + x5 * GW(1, 5) +
          x6 * GW(1, 6)) * ga;
  } else if (i == 1 && j == jjm - 1) {   // north pole: side 1 missing, (i-1, j-1) in (iim, 1)
    const double x0 = X(s), x2 = X(SL(iim, 1, k)), x3 = X(SL(i, j - 1, k)), x4 = X(SL(i + 1, j, k)),
                 x5 = X(SL(i + 1, j + 1, k)), x6 = X(SL(i, j + 1, k));
    gx = (x0 * GW(0, 0) + x2 * GW(0, 2) + x3 * GW(0, 3) + x4 * GW(0, 4) + x5 * GW(0, 5) + x6 * GW(0, 6)) * ga;
    gy = (x0 * GW(1, 0) + x2 * GW(1, 2) + x3 * GW(1, 3) + x4 * GW(1, 4) + x5 * GW(1, 5) + x6 * GW(1, 6)) * ga;
  } else if (i == iim - 1 && j == 1) {   // south pole: side 3 missing, (i-1, j-1) in (1, jjm)
    const double x0 = X(s), x1 = X(SL(i - 1, j, k)), x2 = X(SL(1, jjm, k)), x4 = X(SL(i + 1, j, k)),
                 x5 = X(SL(i + 1, j + 1, k)), x6 = X(SL(i, j + 1, k));
    gx = (x0 * GW(0, 0) + x1 * GW(0, 1) + x2 * GW(0, 2) + x4 * GW(0, 4) + x5 * GW(0, 5) + x6 * GW(0, 6)) * ga;
    gy = (x0 * GW(1, 0) + x1 * GW(1, 1) + x2 * GW(1, 2) + x4 * GW(1, 4) + x5 * GW(1, 5) + x6 * GW(1, 6)) * ga;
  }
#undef GW
}

// remap_prep and construct_fields: mx, my, mp of the mass and tx, ty, tp of
// each tracer (planes q * ns); tr3 the tracers of level n3
__global__ void k_rm_construct(RemapArgs A, const double* __restrict__ h3, const double* const* __restrict__ tr3,
                               double* __restrict__ mx, double* __restrict__ my, double* __restrict__ mp,
                               double* __restrict__ tx, double* __restrict__ ty, double* __restrict__ tp) {
  const int iim = A.iim, jjm = A.jjm, ns = A.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  const int i = c.i, j = c.j, k = c.k;
  if (h3[s] < -RM_EPS) *A.flag = 1;
  auto M = [&](int sl) { return rm_mass(h3, sl); };
  double gx, gy;
  rm_gradient_at(A, c, s, M, gx, gy);
  mx[s] = gx;
  my[s] = gy;
  const bool own = i <= iim - 1 && j <= jjm - 1;   // the centre-of-mass loops (poles included)
  const double mass = M(s);
  double mpv = 0.0, xmav = 0.0, ymav = 0.0;
  if (own && mass > RM_EPS) {
    const double xav = A.cellp[(size_t)15 * ns + s], yav = A.cellp[(size_t)16 * ns + s];
    const double xxav = A.cellp[(size_t)17 * ns + s], xyav = A.cellp[(size_t)18 * ns + s],
                 yyav = A.cellp[(size_t)19 * ns + s];
    mpv = mass - xav * gx - yav * gy;
    const double wk1 = 1.0 / mass;
    xmav = (gx * xxav + gy * xyav + mpv * xav) * wk1;
    ymav = (gx * xyav + gy * yyav + mpv * yav) * wk1;
  }
  mp[s] = mpv;
  const double mmask = h3[s] < RM_EPS ? 0.0 : 1.0;
  for (int q = 0; q < A.nt; ++q) {
    const double* t3 = tr3[q];
    auto T = [&](int sl) { return rm_trace(h3, t3, sl); };
    double px, py;
    rm_gradient_at(A, c, s, T, px, py);
    const double phi = T(s);
    if (own && mmask > RM_EPS) {
      // limit_gradient (remap.F:2152-2397): range of the neighbours' means
      auto NB = [&](int sl) {
        const double pm = h3[sl] < RM_EPS ? 0.0 : 1.0;
        return pm * T(sl) + (1.0 - pm) * phi;
      };
      double lo = 0.0, hi = 0.0;
      bool have = true;
      if (i >= 2 && i <= iim - 1 && j >= 2 && j <= jjm - 1) {
        const double w[7] = {NB(SL(i - 1, j, k)), NB(SL(i - 1, j - 1, k)), NB(SL(i, j - 1, k)), NB(SL(i + 1, j, k)),
                             NB(SL(i + 1, j + 1, k)), NB(SL(i, j + 1, k)), phi};
        lo = w[0]; hi = w[0];
        for (int n = 1; n < 7; ++n) { lo = rm_min(lo, w[n]); hi = rm_max(hi, w[n]); }
      } else if (i == 1 && j == jjm - 1) {
        const double w[6] = {NB(SL(iim, 1, k)), NB(SL(i, j - 1, k)), NB(SL(i + 1, j, k)), NB(SL(i + 1, j + 1, k)),
                             NB(SL(i, j + 1, k)), phi};
        lo = w[0]; hi = w[0];
        for (int n = 1; n < 6; ++n) { lo = rm_min(lo, w[n]); hi = rm_max(hi, w[n]); }
      } else if (i == iim - 1 && j == 1) {
        const double w[6] = {NB(SL(i - 1, j, k)), NB(SL(1, jjm, k)), NB(SL(i + 1, j, k)), NB(SL(i + 1, j + 1, k)),
                             NB(SL(i, j + 1, k)), phi};
        lo = w[0]; hi = w[0];
        for (int n = 1; n < 6; ++n) { lo = rm_min(lo, w[n]); hi = rm_max(hi, w[n]); }
      } else {
        have = false;   // halo rows: no range (pmn = pmx = 0)
      }
      const double pmn = have ? lo - phi : 0.0, pmx = have ? hi - phi : 0.0;
      const double w0 = -xmav * px - ymav * py;
      double qmn = 0.0, qmx = 0.0;
      for (int nv = 0; nv < 6; ++nv) {
        const double cx = A.cellp[(size_t)(20 + 2 * nv) * ns + s], cy = A.cellp[(size_t)(21 + 2 * nv) * ns + s];
        const double v = w0 + px * cx + py * cy;
        if (nv == 0) { qmn = v; qmx = v; }
        else { qmn = rm_min(qmn, v); qmx = rm_max(qmx, v); }
      }
      const double wk1 = fabs(qmn) > 0.0 ? rm_max(0.0, pmn / qmn) : 1.0;
      const double wk2 = fabs(qmx) > 0.0 ? rm_max(0.0, pmx / qmx) : 1.0;
      double wk3 = rm_min(rm_min(1.0, wk1), wk2);
      if (!A.limit_min) wk3 = rm_min(1.0, wk2);
      if (!A.limit_max) wk3 = rm_min(1.0, wk1);
      if (!A.limit_min && !A.limit_max) wk3 = 1.0;
      px = wk3 * px;
      py = wk3 * py;
    }
    tx[(size_t)q * ns + s] = px;
    ty[(size_t)q * ns + s] = py;
    tp[(size_t)q * ns + s] = (own && mmask > RM_EPS) ? phi - px * xmav - py * ymav : 0.0;
  }
}

// departure_points_1order at the two corners of interior cells; dp: three
// corner fields (x, y, z), zero elsewhere.  The reference's speed test
// squares the first velocity component twice.
__global__ void k_rm_departure(RemapArgs A, double dt, const double* __restrict__ u1, const double* __restrict__ u2,
                               double* __restrict__ dp) {
  const int iim = A.iim, jjm = A.jjm, ns = A.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  const size_t cp = 2 * (size_t)ns;
  for (int nc = 0; nc < 2; ++nc) {
    const size_t cs = (size_t)nc * ns + s;
    double d[3] = {0.0, 0.0, 0.0};
    if (interior(c, iim, jjm)) {
      const double a1 = u1[cs], a2 = u2[cs];
      double cc[3];
      for (int r = 0; r < 3; ++r) cc[r] = A.cornp[(size_t)(6 + r) * cp + cs];
      const double wk1 = sqrt(a1 * a1 + a1 * a1);
      if (wk1 > RM_EPS) {
        for (int r = 0; r < 3; ++r) {
          const double ut = A.cornp[(size_t)r * cp + cs] * a1 + A.cornp[(size_t)(r + 3) * cp + cs] * a2;
          d[r] = A.a * cc[r] - ut * dt;
        }
      } else {
        for (int r = 0; r < 3; ++r) d[r] = A.a * cc[r];
      }
    }
    for (int r = 0; r < 3; ++r) dp[(size_t)r * cp + cs] = d[r];
  }
}

// locate_triangles and transport_integrals (linear) of the three edges of
// interior cells: mass and mass-times-tracer transports (edge fields, zero
// on the edges of other slots)
__global__ void k_rm_transport(RemapArgs A, const double* __restrict__ dp, const double* __restrict__ mx,
                               const double* __restrict__ my, const double* __restrict__ mp,
                               const double* __restrict__ tx, const double* __restrict__ ty,
                               const double* __restrict__ tp, double* __restrict__ mflux,
                               double* __restrict__ mtflux) {
  const int iim = A.iim, jjm = A.jjm, ns = A.ns, nt = A.nt;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  const int i = c.i, j = c.j, k = c.k;
  const size_t ep = 3 * (size_t)ns, cp = 2 * (size_t)ns;
  if (!interior(c, iim, jjm)) {
    for (int ne = 0; ne < 3; ++ne) {
      mflux[(size_t)ne * ns + s] = 0.0;
      for (int q = 0; q < nt; ++q) mtflux[(size_t)q * ep + (size_t)ne * ns + s] = 0.0;
    }
    return;
  }
  const int si[3][4] = {{0, 0, -1, -1}, {0, -1, -1, 0}, {0, -1, 0, 1}};   // ishift_p
  const int sj[3][4] = {{0, 1, 0, -1}, {0, 0, -1, -1}, {0, -1, -1, 0}};   // jshift_p
  for (int ne = 0; ne < 3; ++ne) {
    const size_t e = (size_t)ne * ns + s;
#define EP(p) A.edgep[(size_t)(p) * ep + e]
    const double cl[2] = {EP(22), 0.0}, cr[2] = {EP(23), 0.0};
    const double tl[2] = {EP(24), EP(25)}, tr[2] = {EP(26), EP(27)};
    const double bl[2] = {EP(28), EP(29)}, br[2] = {EP(30), EP(31)};
    // departure points of the edge's ends relative to its midpoint
    size_t csl, csr;
    if (ne == 0) { csl = s; csr = (size_t)ns + SL(i, j + 1, k); }
    else if (ne == 1) { csl = (size_t)ns + s; csr = s; }
    else { csl = SL(i + 1, j, k); csr = (size_t)ns + s; }
    double dlx[3], drx[3];
    for (int r = 0; r < 3; ++r) {
      const double aw = A.a * EP(40 + r);
      dlx[r] = dp[(size_t)r * cp + csl] - aw;
      drx[r] = dp[(size_t)r * cp + csr] - aw;
    }
    double dl[2], dr[2];
    for (int r = 0; r < 2; ++r) {
      dl[r] = EP(3 * r) * dlx[0] + EP(3 * r + 1) * dlx[1] + EP(3 * r + 2) * dlx[2];
      dr[r] = EP(3 * r) * drx[0] + EP(3 * r + 1) * drx[1] + EP(3 * r + 2) * drx[2];
    }
    if (fabs(dr[0] - dl[0]) < RM_EPS) dr[0] = dr[0] > dl[0] ? dr[0] + RM_EPS : dr[0] - RM_EPS;
    if (fabs(dl[0] - cl[0]) < RM_EPS) dl[0] = dl[0] > cl[0] ? dl[0] + RM_EPS : dl[0] - RM_EPS;
    if (fabs(dr[0] - cr[0]) < RM_EPS) dr[0] = dr[0] > cr[0] ? dr[0] + RM_EPS : dr[0] - RM_EPS;
    const double mdlr = (dr[1] - dl[1]) / (dr[0] - dl[0]);
    const double bdlr = dl[1] - mdlr * dl[0];
    double ic[2] = {0.0, 0.0}, il[2] = {0.0, 0.0}, ir[2] = {0.0, 0.0};
    bool dlr_int_clr = false;
    if (fabs(mdlr) > RM_EPS) {
      ic[0] = -bdlr / mdlr;
      if (ic[0] > cl[0] && ic[0] < cr[0]) dlr_int_clr = true;
    }
    const double mtcl = tl[1] / (tl[0] - cl[0]), btcl = -mtcl * cl[0];
    const double mbcl = bl[1] / (bl[0] - cl[0]), bbcl = -mbcl * cl[0];
    const double mdcl = dl[1] / (dl[0] - cl[0]);
    bool dl_in_left = false;
    if (dl[0] < cl[0] && mdcl < mbcl && mdcl > mtcl) {
      dl_in_left = true;
      il[0] = (btcl - bdlr) / (mdlr - mtcl);
      if (il[0] < cl[0]) il[1] = mtcl * il[0] + btcl;
      else { il[0] = (bbcl - bdlr) / (mdlr - mbcl); il[1] = mbcl * il[0] + bbcl; }
    }
    const double mtcr = tr[1] / (tr[0] - cr[0]), btcr = -mtcr * cr[0];
    const double mbcr = br[1] / (br[0] - cr[0]), bbcr = -mbcr * cr[0];
    const double mdcr = dr[1] / (dr[0] - cr[0]);
    bool dr_in_right = false;
    if (dr[0] > cr[0] && mdcr > mbcr && mdcr < mtcr) {
      dr_in_right = true;
      ir[0] = (btcr - bdlr) / (mdlr - mtcr);
      if (ir[0] > cr[0]) ir[1] = mtcr * ir[0] + btcr;
      else { ir[0] = (bbcr - bdlr) / (mdlr - mbcr); ir[1] = mbcr * ir[0] + bbcr; }
    }
    // the four triangle groups: vertices, source cell (nd 1..4), sign
    double tv[4][3][2];
    int nds[4] = {0, 0, 0, 0};
    double sg[4] = {0.0, 0.0, 0.0, 0.0};
    auto set = [&](int g, const double* p1, const double* p2, const double* p3) {
      tv[g][0][0] = p1[0]; tv[g][0][1] = p1[1];
      tv[g][1][0] = p2[0]; tv[g][1][1] = p2[1];
      tv[g][2][0] = p3[0]; tv[g][2][1] = p3[1];
    };
    if (dl_in_left) {
      set(0, cl, dl, il);
      nds[0] = 4;
      sg[0] = il[1] > 0.0 ? -1.0 : 1.0;
      dl[0] = il[0]; dl[1] = il[1];
    }
    if (dr_in_right) {
      set(1, cr, dr, ir);
      nds[1] = 2;
      sg[1] = ir[1] > 0.0 ? -1.0 : 1.0;
      dr[0] = ir[0]; dr[1] = ir[1];
    }
    if (!dlr_int_clr) {
      set(2, cl, dl, dr);
      set(3, cl, cr, dr);
      const bool top = dl[1] > 0.0;
      nds[2] = nds[3] = top ? 3 : 1;
      sg[2] = sg[3] = top ? -1.0 : 1.0;
    } else {
      set(2, cl, dl, ic);
      nds[2] = dl[1] > 0.0 ? 3 : 1;
      sg[2] = dl[1] > 0.0 ? -1.0 : 1.0;
      set(3, cr, dr, ic);
      nds[3] = dl[1] < 0.0 ? 3 : 1;
      sg[3] = dl[1] < 0.0 ? -1.0 : 1.0;
    }
    // per group: signed area, quadrature point in the source cell's X3 and
    // the source slot (groups with no triangle carry area 0 and are skipped)
    double ar[4] = {0.0, 0.0, 0.0, 0.0}, qx[4] = {0.0, 0.0, 0.0, 0.0}, qy[4] = {0.0, 0.0, 0.0, 0.0};
    int src[4] = {-1, -1, -1, -1};
    for (int g = 0; g < 4; ++g) {
      if (sg[g] == 0.0) continue;
      double (*t)[2] = tv[g];
      ar[g] = sg[g] * 0.5 *
              fabs((t[1][0] - t[0][0]) * (t[2][1] - t[0][1]) - (t[1][1] - t[0][1]) * (t[2][0] - t[0][0]));
      const int nd = nds[g] - 1;
      const double p0 = EP(32 + 2 * nd), p1 = EP(33 + 2 * nd);
      const double T00 = EP(6 + 4 * nd), T10 = EP(7 + 4 * nd), T01 = EP(8 + 4 * nd), T11 = EP(9 + 4 * nd);
      for (int v = 0; v < 3; ++v) {
        const double a0 = t[v][0] - p0, a1 = t[v][1] - p1;
        t[v][0] = T00 * a0 + T01 * a1;
        t[v][1] = T10 * a0 + T11 * a1;
      }
      qx[g] = (1.0 / 3.0) * (t[0][0] + t[1][0] + t[2][0]);
      qy[g] = (1.0 / 3.0) * (t[0][1] + t[1][1] + t[2][1]);
      src[g] = SL(i + si[ne][nd], j + sj[ne][nd], k);
    }
    // transport_integrals: the source cell's linear field at the midpoint,
    // summed over the groups in order
    double m0[4] = {0.0, 0.0, 0.0, 0.0};
    double mf = 0.0;
    for (int g = 0; g < 4; ++g) {
      if (src[g] < 0) continue;
      const int sn = src[g];
      m0[g] = mp[sn] + qx[g] * mx[sn] + qy[g] * my[sn];
      mf = mf + ar[g] * m0[g];
    }
    for (int q = 0; q < nt; ++q) {
      double mt = 0.0;
      for (int g = 0; g < 4; ++g) {
        if (src[g] < 0) continue;
        const size_t qo = (size_t)q * ns + src[g];
        const double t0 = tp[qo] + qx[g] * tx[qo] + qy[g] * ty[qo];
        mt = mt + ar[g] * m0[g] * t0;
      }
      mtflux[(size_t)q * ep + e] = mt;
    }
    mflux[e] = mf;
#undef EP
  }
}

// update_fields (remap.F:3972-4145) and get_remap_flux's updates (step.F:
// 713-765): the remapped thickness and tracers, the edge mass flux
// mflux / dt, thickness and tracer tendencies (remapped minus old, times
// dti_step) and their ABM3 steps.  tr3 / tr4: the tracers of levels n3 / n4;
// tfa / tfb / tfc: the tracer tendencies of slots n3_f / n2_f / n1_f.
__global__ void k_rm_update(RemapArgs A, double dt, double dti_edge, double dti_step, double w3, double w2, double w1,
                            const double* __restrict__ h3, const double* const* __restrict__ tr3,
                            const double* __restrict__ mflux, const double* __restrict__ mtflux,
                            double* __restrict__ hfa, const double* __restrict__ hfb, const double* __restrict__ hfc,
                            double* const* __restrict__ tfa, const double* const* __restrict__ tfb,
                            const double* const* __restrict__ tfc, double* __restrict__ h4,
                            double* const* __restrict__ tr4, double* __restrict__ mass_flux) {
  const int iim = A.iim, jjm = A.jjm, ns = A.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  const double ai = A.cellp[(size_t)32 * ns + s];
  const double mass0 = rm_mass(h3, s);
  double mass = mass0 - div_edge_at<double>(iim, jjm, ns, c, mflux, ai);
  if (mass < -RM_EPS) {
    // the reference reports the negative mass and goes on
  } else if (mass < RM_EPS) {
    mass = 0.0;
  }
  const double h3s = h3[s];
  const double hf = (mass - h3s) * dti_step;
  hfa[s] = hf;
  const double h4v = h3s + dt * (w3 * hf + w2 * hfb[s] + w1 * hfc[s]);
  h4[s] = h4v;
  const size_t ep = 3 * (size_t)ns;
  for (int q = 0; q < A.nt; ++q) {
    const double t3 = tr3[q][s];
    const double mtold = mass0 * rm_trace(h3, tr3[q], s);
    const double f = div_edge_at<double>(iim, jjm, ns, c, mtflux + (size_t)q * ep, ai);
    const double tr = mass > 0.0 ? (mtold - f) / mass : 0.0;
    const double tf = (mass * tr - h3s * t3) * dti_step;
    tfa[q][s] = tf;
    tr4[q][s] = (h3s * t3 + dt * (w3 * tf + w2 * tfb[q][s] + w1 * tfc[q][s])) / h4v;
  }
  for (int e = 0; e < 3; ++e) mass_flux[(size_t)e * ns + s] = mflux[(size_t)e * ns + s] * dti_edge;
}

// update_fields alone (remap.F:3972-4145) and the edge flux of
// remap_mass_and_tracers (remap.F:1574-1579): the remapped mass and tracers
// (planes q * ns) and mflux / dt
__global__ void k_rm_fields(RemapArgs A, double dti, const double* __restrict__ h3,
                            const double* const* __restrict__ tr3, const double* __restrict__ mflux,
                            const double* __restrict__ mtflux, double* __restrict__ mass_out,
                            double* __restrict__ trace_out, double* __restrict__ edgeflux) {
  const int iim = A.iim, jjm = A.jjm, ns = A.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  const double ai = A.cellp[(size_t)32 * ns + s];
  const double mass0 = rm_mass(h3, s);
  double mass = mass0 - div_edge_at<double>(iim, jjm, ns, c, mflux, ai);
  if (mass < -RM_EPS) {
  } else if (mass < RM_EPS) {
    mass = 0.0;
  }
  mass_out[s] = mass;
  const size_t ep = 3 * (size_t)ns;
  for (int q = 0; q < A.nt; ++q) {
    const double mtold = mass0 * rm_trace(h3, tr3[q], s);
    const double f = div_edge_at<double>(iim, jjm, ns, c, mtflux + (size_t)q * ep, ai);
    trace_out[(size_t)q * ns + s] = mass > 0.0 ? (mtold - f) / mass : 0.0;
  }
  for (int e = 0; e < 3; ++e) edgeflux[(size_t)e * ns + s] = mflux[(size_t)e * ns + s] * dti;
}

#undef SL

}  // namespace dev
}  // namespace swm
