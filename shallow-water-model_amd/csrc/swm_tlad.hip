// swm_tlad.hip -- tangent-linear (TL) and adjoint (AD) kernels of the FP64
// time step (step.F:114-212) with centred, upstream or flux-corrected
// advection of the mass and the tracers (step.F:495-678, mass_fct.F,
// tracer_fct.F; operators.F as named per kernel), and of the FP64 V-cycle.
//
// Every forward operation y = f(x; base) of the step gets a TL kernel
// (dy = f'(x) dx, gather form, one thread per slot like the forward kernel)
// and an AD kernel (dx += f'(x)^T dy).  AD kernels run one thread per output
// slot, as the forward does, and scatter into the input adjoints with FP64
// atomic adds, so each one is the exact transpose of the forward stencil,
// pole forms and panel edges included.  Halo fills (x[dst] = x[src]) have the
// transpose x[src] += x[dst], x[dst] = 0 (k_gather_adj).  Operations that are
// linear in the state (ABM3 updates, divergence of edge fluxes, Laplacian,
// velocity from streamfunction and velocity potential) reuse the forward
// kernels for TL with a metric bundle whose constant terms (surface height,
// Coriolis parameter) are zero.
#include "swm_kernels.h"

namespace swm {
namespace dev {

#define SL(i, j, k) ((i) - 1 + iim * ((j) - 1 + jjm * ((k) - 1)))

__device__ __forceinline__ void aadd(double* p, double v) { unsafeAtomicAdd(p, v); }

// ------------------------------------------------------------------ halo --
// transpose of k_gather: adjoint of the sources += adjoint of the halo
// slots; then (k_gather_adj_zero) the halo slots' adjoint is cleared.  No
// source is a destination, so the two passes commute with nothing else.
__global__ void k_gather_adj(const int* __restrict__ dst, const int* __restrict__ src, int n, FieldSet fs) {
  const int t = blockIdx.x * bdim() + threadIdx.x;
  if (t >= n) return;
  double* x = fs.p[blockIdx.y];
  const int s = src[t];
  if (s >= 0) aadd(&x[s], x[dst[t]]);
}
__global__ void k_gather_adj_zero(const int* __restrict__ dst, int n, FieldSet fs) {
  const int t = blockIdx.x * bdim() + threadIdx.x;
  if (t >= n) return;
  fs.p[blockIdx.y][dst[t]] = 0.0;
}

// ------------------------------------------------- average to corners -----
// AD of k_avg_corners (operators.F:982-1112): corner values back to the
// three cells with the scalar weights; ghost corners to cell (2,2)
__global__ void k_ad_avg_corners(Geo G, const double* __restrict__ ob, double* __restrict__ xb) {
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  const int i = c.i, j = c.j, k = c.k;
  const double o0 = ob[s], o1 = ob[ns + s];
  if (i >= 2 && i <= iim - 1 && j >= 2 && j <= jjm - 1) {
    const double* sw = G.sw;
    const int a = SL(i, j, k), b = SL(i - 1, j, k), d = SL(i - 1, j - 1, k), e = SL(i, j - 1, k);
    aadd(&xb[a], sw[0 * ns + s] * o0 + sw[3 * ns + s] * o1);
    aadd(&xb[b], sw[1 * ns + s] * o0);
    aadd(&xb[d], sw[2 * ns + s] * o0 + sw[4 * ns + s] * o1);
    aadd(&xb[e], sw[5 * ns + s] * o1);
  } else {
    aadd(&xb[SL(2, 2, k)], o0 + o1);
  }
}

// -------------------------------------------------------- mass flux -------
// The centred edge mass flux of k_mass_fluxes (operators.F:1626-1722):
// F_e = -r0_e + r1_e, r = (u . n) * hc * dc at the half-edges
//   r0_0 (sn,kk=1,ic=1)  r1_0 (s0,0,0)  r0_1 (s0,2,0)  r1_1 (s0,0,1)
//   r0_2 (s0,2,1)        r1_2 (se,1,0)
struct HalfTerm {
  int sl, kk, ic;
};
__device__ __forceinline__ void half_terms(int iim, int jjm, int i, int j, int k, HalfTerm (&t)[2][3]) {
  const int s0 = SL(i, j, k), sn = SL(i, j + 1, k), se = SL(i + 1, j, k);
  t[0][0] = {sn, 1, 1}; t[1][0] = {s0, 0, 0};
  t[0][1] = {s0, 2, 0}; t[1][1] = {s0, 0, 1};
  t[0][2] = {s0, 2, 1}; t[1][2] = {se, 1, 0};
}
#define UPL(u, sp, ic, s_) (u)[((ic) + 2 * (sp)) * ns + (s_)]
#define NPL(n, kk, a, ic, s_) (n)[((kk) + 3 * ((a) + 2 * (ic))) * ns + (s_)]
#define DPL(kk, ic, s_) G.dc[((kk) + 3 * (ic)) * ns + (s_)]

// TL: dF = d(-r0 + r1) with d r = ((du . n) hc + (u . n) dhc) dc; nrm is the
// normal (mass flux) or the tangent (divergence-curl) planes
__device__ __forceinline__ double half_tl(const Geo& G, int ns, const double* __restrict__ nrm,
                                          const double* __restrict__ u, const double* __restrict__ hc,
                                          const double* __restrict__ du, const double* __restrict__ dhc,
                                          const HalfTerm& t) {
  const double n0 = NPL(nrm, t.kk, 0, t.ic, t.sl), n1 = NPL(nrm, t.kk, 1, t.ic, t.sl);
  const double un = UPL(u, 0, t.ic, t.sl) * n0 + UPL(u, 1, t.ic, t.sl) * n1;
  const double dun = UPL(du, 0, t.ic, t.sl) * n0 + UPL(du, 1, t.ic, t.sl) * n1;
  return (dun * hc[t.ic * ns + t.sl] + un * dhc[t.ic * ns + t.sl]) * DPL(t.kk, t.ic, t.sl);
}
// base half-edge value (u . n) hc dc
__device__ __forceinline__ double half_val(const Geo& G, int ns, const double* __restrict__ nrm,
                                           const double* __restrict__ u, const double* __restrict__ hc,
                                           const HalfTerm& t) {
  const double un = UPL(u, 0, t.ic, t.sl) * NPL(nrm, t.kk, 0, t.ic, t.sl) +
                    UPL(u, 1, t.ic, t.sl) * NPL(nrm, t.kk, 1, t.ic, t.sl);
  return un * hc[t.ic * ns + t.sl] * DPL(t.kk, t.ic, t.sl);
}
// AD of one half-edge value with adjoint g
__device__ __forceinline__ void half_ad(const Geo& G, int ns, const double* __restrict__ nrm,
                                        const double* __restrict__ u, const double* __restrict__ hc, double g,
                                        const HalfTerm& t, double* __restrict__ ub, double* __restrict__ hcb) {
  const double n0 = NPL(nrm, t.kk, 0, t.ic, t.sl), n1 = NPL(nrm, t.kk, 1, t.ic, t.sl);
  const double d = DPL(t.kk, t.ic, t.sl);
  const double un = UPL(u, 0, t.ic, t.sl) * n0 + UPL(u, 1, t.ic, t.sl) * n1;
  const double gu = g * hc[t.ic * ns + t.sl] * d;
  aadd(&UPL(ub, 0, t.ic, t.sl), gu * n0);
  aadd(&UPL(ub, 1, t.ic, t.sl), gu * n1);
  aadd(&hcb[t.ic * ns + t.sl], g * un * d);
}

__global__ void k_tl_mass_flux(Geo G, const double* __restrict__ u, const double* __restrict__ hc,
                               const double* __restrict__ du, const double* __restrict__ dhc, double* __restrict__ dF) {
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  if (!interior(c, iim, jjm)) {
    for (int e = 0; e < 3; ++e) dF[e * ns + s] = 0.0;
    return;
  }
  HalfTerm t[2][3];
  half_terms(iim, jjm, c.i, c.j, c.k, t);
  for (int e = 0; e < 3; ++e)
    dF[e * ns + s] = -half_tl(G, ns, G.nrm, u, hc, du, dhc, t[0][e]) + half_tl(G, ns, G.nrm, u, hc, du, dhc, t[1][e]);
}

__global__ void k_ad_mass_flux(Geo G, const double* __restrict__ u, const double* __restrict__ hc,
                               const double* __restrict__ Fb, double* __restrict__ ub, double* __restrict__ hcb) {
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  if (!interior(c, iim, jjm)) return;
  HalfTerm t[2][3];
  half_terms(iim, jjm, c.i, c.j, c.k, t);
  for (int e = 0; e < 3; ++e) {
    const double g = Fb[e * ns + s];
    half_ad(G, ns, G.nrm, u, hc, -g, t[0][e], ub, hcb);
    half_ad(G, ns, G.nrm, u, hc, g, t[1][e], ub, hcb);
  }
}

// ---------------------------------------------------- height update -------
// step.F:221-270 step_surface_height: hf_a = -div(F), h4 = h3 + dt (w3 hf_a
// + w2 hf_b + w1 hf_c); the clip at 1e-10 passes no perturbation where the
// base thickness was clipped (h4b == 1e-10)
__global__ void k_tl_height(Geo G, double dt, double w3, double w2, double w1, const double* __restrict__ dF,
                            const double* __restrict__ dh3, double* __restrict__ dhf_a,
                            const double* __restrict__ dhf_b, const double* __restrict__ dhf_c,
                            const double* __restrict__ h4b, double* __restrict__ dh4) {
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  const double hf = -div_edge_at<double>(iim, jjm, ns, c, dF, G.area_inv[s]);
  dhf_a[s] = hf;
  const double h = dh3[s] + dt * (w3 * hf + w2 * dhf_b[s] + w1 * dhf_c[s]);
  dh4[s] = h4b[s] == 1.0e-10 ? 0.0 : h;
}

// AD of the divergence of edge fluxes at one slot (div_edge_at): the slot's
// value y = ai * (sum of +- F terms); adds g * dy/dF into Fb
__device__ __forceinline__ void div_edge_ad(int iim, int jjm, int ns, const IJK& c, double ai, double g,
                                            double* __restrict__ Fb) {
  const int i = c.i, j = c.j, k = c.k;
  const double v = g * ai;
  if (i == 1 && j == jjm - 1) {
    aadd(&Fb[1 * ns + SL(i, j, k)], -v); aadd(&Fb[2 * ns + SL(i, j, k)], -v); aadd(&Fb[0 * ns + SL(i + 1, j, k)], -v);
    aadd(&Fb[1 * ns + SL(i + 1, j + 1, k)], -v); aadd(&Fb[2 * ns + SL(i, j + 1, k)], -v);
  } else if (i == iim - 1 && j == 1) {
    aadd(&Fb[0 * ns + SL(i, j, k)], -v); aadd(&Fb[1 * ns + SL(i, j, k)], -v); aadd(&Fb[0 * ns + SL(i + 1, j, k)], -v);
    aadd(&Fb[1 * ns + SL(i + 1, j + 1, k)], -v); aadd(&Fb[2 * ns + SL(i, j + 1, k)], -v);
  } else if (i >= 2 && i <= iim - 1 && j >= 2 && j <= jjm - 1) {
    aadd(&Fb[0 * ns + SL(i, j, k)], v); aadd(&Fb[1 * ns + SL(i, j, k)], v); aadd(&Fb[2 * ns + SL(i, j, k)], v);
    aadd(&Fb[0 * ns + SL(i + 1, j, k)], -v); aadd(&Fb[1 * ns + SL(i + 1, j + 1, k)], -v);
    aadd(&Fb[2 * ns + SL(i, j + 1, k)], -v);
  }
}

// AD of k_tl_height: consumes h4b_adj (the new level's adjoint) and the
// adjoint of the slot hf_a (both cleared), adds into h3, hf_b, hf_c, F
__global__ void k_ad_height(Geo G, double dt, double w3, double w2, double w1, double* __restrict__ h4a,
                            double* __restrict__ hfa_a, double* __restrict__ hfb_a, double* __restrict__ hfc_a,
                            const double* __restrict__ h4b, double* __restrict__ h3a, double* __restrict__ Fa) {
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  const double g = h4b[s] == 1.0e-10 ? 0.0 : h4a[s];
  h4a[s] = 0.0;
  h3a[s] += g;
  const double ga = hfa_a[s] + dt * w3 * g;
  hfa_a[s] = 0.0;
  hfb_a[s] += dt * w2 * g;
  hfc_a[s] += dt * w1 * g;
  div_edge_ad(iim, jjm, ns, c, G.area_inv[s], -ga, Fa);
}

// --------------------------------------- vorticity flux, divergence curl --
// k_vort_curl (operators.F:2000-2186): pv = eta / h (0 where h <= 0),
// vf_e = F_e (pv0 + pv_up) / 2, dcurl_e = (-t0_e + t1_e) (pv_up + pv0) / 2 with
// the tangential half-edge values t
__global__ void k_tl_vort_curl(Geo G, const double* __restrict__ F, const double* __restrict__ h3,
                               const double* __restrict__ e3, const double* __restrict__ hc,
                               const double* __restrict__ u, const double* __restrict__ dF,
                               const double* __restrict__ dh3, const double* __restrict__ de3,
                               const double* __restrict__ dhc, const double* __restrict__ du,
                               double* __restrict__ dvf, double* __restrict__ ddc) {
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  if (!interior(c, iim, jjm)) {
    for (int e = 0; e < 3; ++e) { dvf[e * ns + s] = 0.0; ddc[e * ns + s] = 0.0; }
    return;
  }
  const int i = c.i, j = c.j, k = c.k;
  auto PV = [&](int sl, double& p, double& dp) {
    const double hh = h3[sl];
    const double w = hh > 0.0 ? 1.0 / hh : 0.0;
    p = e3[sl] * w;
    dp = hh > 0.0 ? de3[sl] * w - e3[sl] * w * w * dh3[sl] : 0.0;
  };
  const int up[3] = {SL(i - 1, j, k), SL(i - 1, j - 1, k), SL(i, j - 1, k)};
  double p0, dp0;
  PV(s, p0, dp0);
  HalfTerm t[2][3];
  half_terms(iim, jjm, i, j, k, t);
  for (int e = 0; e < 3; ++e) {
    double pu, dpu;
    PV(up[e], pu, dpu);
    const double f = F[e * ns + s], df = dF[e * ns + s];
    dvf[e * ns + s] = 0.5 * (df * (p0 + pu) + f * (dp0 + dpu));
    const double r = -half_val(G, ns, G.tng, u, hc, t[0][e]) + half_val(G, ns, G.tng, u, hc, t[1][e]);
    const double dr = -half_tl(G, ns, G.tng, u, hc, du, dhc, t[0][e]) + half_tl(G, ns, G.tng, u, hc, du, dhc, t[1][e]);
    ddc[e * ns + s] = 0.5 * (dr * (pu + p0) + r * (dpu + dp0));
  }
}

__global__ void k_ad_vort_curl(Geo G, const double* __restrict__ F, const double* __restrict__ h3,
                               const double* __restrict__ e3, const double* __restrict__ hc,
                               const double* __restrict__ u, const double* __restrict__ vfa,
                               const double* __restrict__ dca, double* __restrict__ Fa, double* __restrict__ h3a,
                               double* __restrict__ e3a, double* __restrict__ hca, double* __restrict__ ua) {
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  if (!interior(c, iim, jjm)) return;
  const int i = c.i, j = c.j, k = c.k;
  auto PV = [&](int sl) {
    const double hh = h3[sl];
    return e3[sl] * (hh > 0.0 ? 1.0 / hh : 0.0);
  };
  auto PV_AD = [&](int sl, double g) {   // pv = e / h
    const double hh = h3[sl];
    if (!(hh > 0.0)) return;
    const double w = 1.0 / hh;
    aadd(&e3a[sl], g * w);
    aadd(&h3a[sl], -g * e3[sl] * w * w);
  };
  const int up[3] = {SL(i - 1, j, k), SL(i - 1, j - 1, k), SL(i, j - 1, k)};
  const double p0 = PV(s);
  HalfTerm t[2][3];
  half_terms(iim, jjm, i, j, k, t);
  double gp0 = 0.0;
  for (int e = 0; e < 3; ++e) {
    const double pu = PV(up[e]);
    const double gv = vfa[e * ns + s], gd = dca[e * ns + s];
    const double f = F[e * ns + s];
    const double r = -half_val(G, ns, G.tng, u, hc, t[0][e]) + half_val(G, ns, G.tng, u, hc, t[1][e]);
    aadd(&Fa[e * ns + s], 0.5 * gv * (p0 + pu));
    const double gpv = 0.5 * (gv * f + gd * r);
    gp0 += gpv;
    PV_AD(up[e], gpv);
    const double gr = 0.5 * gd * (pu + p0);
    half_ad(G, ns, G.tng, u, hc, -gr, t[0][e], ua, hca);
    half_ad(G, ns, G.tng, u, hc, gr, t[1][e], ua, hca);
  }
  PV_AD(s, gp0);
}

// ------------------------------------------------------ kinetic energy ----
// k_center_ke (operators.F:1316-1393 of 0.5 |u|^2 at the corners)
#define SAP(kk, ic, sl) G.sub_area[((kk) + 3 * (ic)) * ns + (sl)]
__global__ void k_tl_center_ke(Geo G, const double* __restrict__ u, const double* __restrict__ du,
                               double* __restrict__ dke) {
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  const int i = c.i, j = c.j, k = c.k;
  if (i > iim - 1 || j > jjm - 1) { dke[s] = 0.0; return; }
  auto DK = [&](int ic, int sl) { return u[ic * ns + sl] * du[ic * ns + sl] + u[(ic + 2) * ns + sl] * du[(ic + 2) * ns + sl]; };
  const int se = SL(i + 1, j, k), sne = SL(i + 1, j + 1, k), sn = SL(i, j + 1, k);
  dke[s] = (SAP(0, 0, s) * DK(0, s) + SAP(0, 1, s) * DK(1, s) + SAP(1, 0, se) * DK(0, se) + SAP(1, 1, sne) * DK(1, sne) +
            SAP(2, 0, sne) * DK(0, sne) + SAP(2, 1, sn) * DK(1, sn)) * G.area_inv[s];
}
__global__ void k_ad_center_ke(Geo G, const double* __restrict__ u, const double* __restrict__ kea,
                               double* __restrict__ ua) {
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  const int i = c.i, j = c.j, k = c.k;
  if (i > iim - 1 || j > jjm - 1) return;
  const double g = kea[s] * G.area_inv[s];
  auto KA = [&](int ic, int sl, double w) {
    aadd(&ua[ic * ns + sl], g * w * u[ic * ns + sl]);
    aadd(&ua[(ic + 2) * ns + sl], g * w * u[(ic + 2) * ns + sl]);
  };
  const int se = SL(i + 1, j, k), sne = SL(i + 1, j + 1, k), sn = SL(i, j + 1, k);
  KA(0, s, SAP(0, 0, s)); KA(1, s, SAP(0, 1, s)); KA(0, se, SAP(1, 0, se)); KA(1, sne, SAP(1, 1, sne));
  KA(0, sne, SAP(2, 0, sne)); KA(1, sn, SAP(2, 1, sn));
}
#undef SAP

// ------------------------------------------------------- tendencies -------
// AD of k_tendencies (step.F:336-429): e4 = e3 + dt (w3 ef_a + w2 ef_b +
// w1 ef_c), ef_a = -div(vf); d4 = d3 + dt (w3 df_a + ...), df_a = -lap(X) +
// div(dcurl), X = ke + g (h3 + hs); rhs1 = e4 - f, rhs2 = d4.  Inputs: the
// adjoints of rhs1 (rhs1a) and of d4 (d4a, the new level's adjoint plus
// that of rhs2), both consumed, and of the history slots; outputs added
// into e3, d3, the b and c slots, vf, dcurl and X (Xa, then folded into ke
// and h3 by the caller).
__global__ void k_ad_tendencies(Geo G, double dt, double w3, double w2, double w1, double* __restrict__ rhs1a,
                                double* __restrict__ d4a, double* __restrict__ efa_a, double* __restrict__ efb_a,
                                double* __restrict__ efc_a, double* __restrict__ dfa_a, double* __restrict__ dfb_a,
                                double* __restrict__ dfc_a, double* __restrict__ e3a, double* __restrict__ d3a,
                                double* __restrict__ vfa, double* __restrict__ dca, double* __restrict__ Xa) {
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  const double ai = G.area_inv[s];
  const double ge = rhs1a[s];
  rhs1a[s] = 0.0;
  const double gd = d4a[s];
  d4a[s] = 0.0;
  e3a[s] += ge;
  const double gefa = efa_a[s] + dt * w3 * ge;
  efa_a[s] = 0.0;
  efb_a[s] += dt * w2 * ge;
  efc_a[s] += dt * w1 * ge;
  div_edge_ad(iim, jjm, ns, c, ai, -gefa, vfa);
  d3a[s] += gd;
  const double gdfa = dfa_a[s] + dt * w3 * gd;
  dfa_a[s] = 0.0;
  dfb_a[s] += dt * w2 * gd;
  dfc_a[s] += dt * w1 * gd;
  div_edge_ad(iim, jjm, ns, c, ai, gdfa, dca);
  // -lap(X): the transposed stencil of lap_at at this slot
  const Stencil7<double> st = stencil7<double>(iim, jjm, ns, c, G.lw);
  if (st.real) {
    const double gl = -gdfa;
#pragma unroll
    for (int q = 0; q < 7; ++q)
      if (st.w[q] != 0.0) aadd(&Xa[st.sl[q]], st.w[q] * gl);
  }
}

// X = ke + g (h3 + hs): Xa folded into ke and h3, then cleared
__global__ void k_ad_fold_x(int ns, double grav, double* __restrict__ Xa, double* __restrict__ kea,
                            double* __restrict__ h3a) {
  const int s = blockIdx.x * bdim() + threadIdx.x;
  if (s >= ns) return;
  const double g = Xa[s];
  Xa[s] = 0.0;
  kea[s] += g;
  h3a[s] += grav * g;
}

// ---------------------------------------------------------- velocity ------
// AD of k_velocity (operators.F:1402-1514): u(0,ic) = gc0 - gp1, u(1,ic) =
// gc1 + gp0 with g = cai * sum_k vw(k,sp,ic) x(nb_k); consumes ua (cleared
// on every slot, as the forward writes every slot)
__global__ void k_ad_velocity(Geo G, double* __restrict__ ua, double* __restrict__ psia, double* __restrict__ chia) {
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  double g[4];
  for (int p = 0; p < 4; ++p) { g[p] = ua[p * ns + s]; ua[p * ns + s] = 0.0; }
  if (!interior(c, iim, jjm)) return;
  const int i = c.i, j = c.j, k = c.k;
  const int nb[2][3] = {{s, SL(i - 1, j, k), SL(i - 1, j - 1, k)}, {s, SL(i - 1, j - 1, k), SL(i, j - 1, k)}};
#define VW(kk, sp, ic) G.vw[((kk) + 3 * ((sp) + 2 * (ic))) * ns + s]
  for (int ic = 0; ic < 2; ++ic) {
    const double cai = G.cai[ic * ns + s];
    const double gc0 = g[ic], gp1 = -g[ic], gc1 = g[ic + 2], gp0 = g[ic + 2];
    for (int q = 0; q < 3; ++q) {
      aadd(&chia[nb[ic][q]], cai * (VW(q, 0, ic) * gc0 + VW(q, 1, ic) * gc1));
      aadd(&psia[nb[ic][q]], cai * (VW(q, 0, ic) * gp0 + VW(q, 1, ic) * gp1));
    }
  }
#undef VW
}

// --------------------------------------------------- curl, divergence -----
// AD of k_curl_div (operators.F:47-261), either output may be null
__global__ void k_ad_curl_div(Geo G, const double* __restrict__ curla, const double* __restrict__ dva,
                              double* __restrict__ ua) {
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  const int i = c.i, j = c.j, k = c.k;
  if (i > iim - 1 || j > jjm - 1) return;
#define VW(kk, sp, ic, sl) G.vw[((kk) + 3 * ((sp) + 2 * (ic))) * ns + (sl)]
#define UA(sp, ic, sl) ua[((ic) + 2 * (sp)) * ns + (sl)]
  const int se = SL(i + 1, j, k), sne = SL(i + 1, j + 1, k), sn = SL(i, j + 1, k);
  const double ai = G.area_inv[s];
  const double gc = curla ? curla[s] * ai : 0.0, gv = dva ? dva[s] * ai : 0.0;
  // (kk, ic, slot) of the 6 corner terms
  const int kk6[6] = {0, 0, 1, 1, 2, 2}, ic6[6] = {0, 1, 0, 1, 0, 1};
  const int sl6[6] = {s, s, se, sne, sne, sn};
  for (int q = 0; q < 6; ++q) {
    const int kk = kk6[q], ic = ic6[q], sl = sl6[q];
    // curl: + vw(kk,1,ic) u(0,ic) - vw(kk,0,ic) u(1,ic); div: - vw(kk,0,ic) u(0,ic) - vw(kk,1,ic) u(1,ic)
    aadd(&UA(0, ic, sl), gc * VW(kk, 1, ic, sl) - gv * VW(kk, 0, ic, sl));
    aadd(&UA(1, ic, sl), -gc * VW(kk, 0, ic, sl) - gv * VW(kk, 1, ic, sl));
  }
#undef VW
#undef UA
}

// ------------------------------------------------------------ helpers -----
// y -= am * sums[0] / total (the transpose of the mean removal k_sub_mean,
// whose mean is sum(am * x) / total, is y -= am * sum(y) / total)
__global__ void k_sub_am(int n, const double* __restrict__ am, const double* __restrict__ sums, double inv_total,
                         double* __restrict__ y) {
  const int s = blockIdx.x * bdim() + threadIdx.x;
  if (s >= n) return;
  y[s] = y[s] - am[s] * (sums[0] * inv_total);
}
// y = x * w (w may be the reciprocal area; 0 where mask is 0)
__global__ void k_mul_mask(int n, const double* __restrict__ x, const double* __restrict__ w,
                           const double* __restrict__ mask, double* __restrict__ y) {
  const int s = blockIdx.x * bdim() + threadIdx.x;
  if (s >= n) return;
  y[s] = mask[s] != 0.0 ? x[s] * w[s] : 0.0;
}
// y += x
__global__ void k_axpy(size_t n, double a, const double* __restrict__ x, double* __restrict__ y) {
  const size_t s = (size_t)blockIdx.x * bdim() + threadIdx.x;
  if (s >= n) return;
  y[s] = y[s] + a * x[s];
}
// per-block partial sums of x * y over n values (dot products)
__global__ void __launch_bounds__(256) k_dot_partial(size_t n, const double* __restrict__ x,
                                                     const double* __restrict__ y, double* __restrict__ part) {
  double acc[1] = {0.0};
  for (size_t t = (size_t)blockIdx.x * bdim() + threadIdx.x; t < n; t += (size_t)gridDim.x * bdim())
    acc[0] += x[t] * y[t];
  block_sums<1>(acc, part);
}

// ------------------------------------- mass FCT with a frozen limiter ----
// mass_fct.F:23-539 linearised about the base trajectory with the limiter's
// coefficients and the upstream direction of every edge taken from the base
// step: the low-order (upstream) flux, the antidiffusive flux from the AB3
// history of centred fluxes, the low-order mass and the corrected mass are
// then linear, and the limited flux is the antidiffusive one scaled by a
// fixed coefficient per edge.  The linear parts reuse the forward kernels
// (k_mass_low_beta for the antidiffusive flux and low-order mass, its limiter
// output going to scratch; k_mass_final).

// the frozen limiter coefficient of every edge (base step, before the
// limiting of k_mass_anti): at interior slots min(1, beta(s)) where the base
// antidiffusive flux is >= 0, else min(1, beta(up)); 1 on the other slots,
// which keep the unlimited flux until the edge halo fill
__global__ void k_fct_coef(Geo G, const double* __restrict__ beta, const double* __restrict__ anti,
                           double* __restrict__ coef) {
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  if (!interior(c, iim, jjm)) {
    for (int e = 0; e < 3; ++e) coef[e * ns + s] = 1.0;
    return;
  }
  const int i = c.i, j = c.j, k = c.k;
  const int up[3] = {SL(i - 1, j, k), SL(i - 1, j - 1, k), SL(i, j - 1, k)};
  const double b = dmin(1.0, beta[s]);
  for (int e = 0; e < 3; ++e) coef[e * ns + s] = anti[e * ns + s] >= 0.0 ? b : dmin(1.0, beta[up[e]]);
}

// number of the n coefficients below 1 (active limiter), added to *count:
// one atomic per wave
__global__ void k_count_below_one(size_t n, const double* __restrict__ coef, unsigned long long* __restrict__ count) {
  const size_t t = (size_t)blockIdx.x * bdim() + threadIdx.x;
  const bool below = t < n && coef[t] < 1.0;
  const unsigned long long m = __ballot(below);
  if ((threadIdx.x & 63) == 0 && m) atomicAdd(count, (unsigned long long)__popcll(m));
}

// the limited flux, TL and AD alike (a diagonal map): x *= coef on n values
__global__ void k_fct_scale(size_t n, const double* __restrict__ coef, double* __restrict__ x) {
  const size_t t = (size_t)blockIdx.x * bdim() + threadIdx.x;
  if (t < n) x[t] = x[t] * coef[t];
}

// half-edge normal velocity flux (u . n) dc (the upstream flux drops the
// corner mass, operators.F:1731-1832) and its AD
__device__ __forceinline__ double half_un(const Geo& G, int ns, const double* __restrict__ u, const HalfTerm& t) {
  const double un = UPL(u, 0, t.ic, t.sl) * NPL(G.nrm, t.kk, 0, t.ic, t.sl) +
                    UPL(u, 1, t.ic, t.sl) * NPL(G.nrm, t.kk, 1, t.ic, t.sl);
  return un * DPL(t.kk, t.ic, t.sl);
}
__device__ __forceinline__ void half_un_ad(const Geo& G, int ns, double g, const HalfTerm& t, double* __restrict__ ub) {
  const double d = g * DPL(t.kk, t.ic, t.sl);
  aadd(&UPL(ub, 0, t.ic, t.sl), d * NPL(G.nrm, t.kk, 0, t.ic, t.sl));
  aadd(&UPL(ub, 1, t.ic, t.sl), d * NPL(G.nrm, t.kk, 1, t.ic, t.sl));
}

// TL of the upstream flux of k_mass_fluxes (mass_fct.F:151-214): with the
// base edge velocity flux v = r1 - r0, flx = v h(s) where v >= 0, else
// v h(up); interior slots (0 elsewhere)
__global__ void k_tl_upwind_flux(Geo G, const double* __restrict__ u, const double* __restrict__ h3,
                                 const double* __restrict__ du, const double* __restrict__ dh3,
                                 double* __restrict__ dflx) {
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  if (!interior(c, iim, jjm)) {
    for (int e = 0; e < 3; ++e) dflx[e * ns + s] = 0.0;
    return;
  }
  const int i = c.i, j = c.j, k = c.k;
  const int up[3] = {SL(i - 1, j, k), SL(i - 1, j - 1, k), SL(i, j - 1, k)};
  HalfTerm t[2][3];
  half_terms(iim, jjm, i, j, k, t);
  for (int e = 0; e < 3; ++e) {
    const double v = half_un(G, ns, u, t[1][e]) - half_un(G, ns, u, t[0][e]);
    const double dv = half_un(G, ns, du, t[1][e]) - half_un(G, ns, du, t[0][e]);
    const int sl = v >= 0.0 ? s : up[e];
    dflx[e * ns + s] = dv * h3[sl] + v * dh3[sl];
  }
}

// AD of k_tl_upwind_flux: consumes the flux adjoint (cleared on every slot)
__global__ void k_ad_upwind_flux(Geo G, const double* __restrict__ u, const double* __restrict__ h3,
                                 double* __restrict__ flxa, double* __restrict__ ua, double* __restrict__ h3a) {
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  if (!interior(c, iim, jjm)) {
    for (int e = 0; e < 3; ++e) flxa[e * ns + s] = 0.0;
    return;
  }
  const int i = c.i, j = c.j, k = c.k;
  const int up[3] = {SL(i - 1, j, k), SL(i - 1, j - 1, k), SL(i, j - 1, k)};
  HalfTerm t[2][3];
  half_terms(iim, jjm, i, j, k, t);
  for (int e = 0; e < 3; ++e) {
    const double g = flxa[e * ns + s];
    flxa[e * ns + s] = 0.0;
    const double v = half_un(G, ns, u, t[1][e]) - half_un(G, ns, u, t[0][e]);
    const int sl = v >= 0.0 ? s : up[e];
    aadd(&h3a[sl], g * v);
    const double gv = g * h3[sl];
    half_un_ad(G, ns, gv, t[1][e], ua);
    half_un_ad(G, ns, -gv, t[0][e], ua);
  }
}

// AD of k_mass_final (mflux = flx + anti, h4 = mss_low - dt div(anti) at the
// real cells, 0 elsewhere): consumes the adjoints of mflux and h4
__global__ void k_ad_fct_final(Geo G, double dt, double* __restrict__ Fa, double* __restrict__ h4a,
                               double* __restrict__ flxa, double* __restrict__ antia, double* __restrict__ mla) {
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  for (int e = 0; e < 3; ++e) {
    const double g = Fa[e * ns + s];
    Fa[e * ns + s] = 0.0;
    flxa[e * ns + s] += g;
    aadd(&antia[e * ns + s], g);
  }
  const double g = h4a[s];
  h4a[s] = 0.0;
  const bool real = interior(c, iim, jjm) || (c.i == 1 && c.j == jjm - 1) || (c.i == iim - 1 && c.j == 1);
  if (!real) return;
  mla[s] += g;
  div_edge_ad(iim, jjm, ns, c, G.area_inv[s], -dt * g, antia);
}

// AD of the antidiffusive flux and low-order mass of k_mass_low_beta
// (anti = w3 m3 + w2 m2 + w1 m1 - flx on every slot; mss_low = h3 - dt
// div(flx) at the real cells): consumes the adjoints of anti and mss_low,
// adds into the three flux-history slots, flx and h3
__global__ void k_ad_fct_low(Geo G, double dt, double w3, double w2, double w1, double* __restrict__ antia,
                             double* __restrict__ mla, double* __restrict__ m3a, double* __restrict__ m2a,
                             double* __restrict__ m1a, double* __restrict__ flxa, double* __restrict__ h3a) {
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  for (int e = 0; e < 3; ++e) {
    const double g = antia[e * ns + s];
    antia[e * ns + s] = 0.0;
    m3a[e * ns + s] += w3 * g;
    m2a[e * ns + s] += w2 * g;
    m1a[e * ns + s] += w1 * g;
    aadd(&flxa[e * ns + s], -g);
  }
  const double g = mla[s];
  mla[s] = 0.0;
  const bool real = interior(c, iim, jjm) || (c.i == 1 && c.j == jjm - 1) || (c.i == iim - 1 && c.j == 1);
  if (!real) return;
  aadd(&h3a[s], g);
  div_edge_ad(iim, jjm, ns, c, G.area_inv[s], -dt * g, flxa);
}

// ----------------------------------------------------- tracer transport --
// Tracers do not feed back on the dynamics, but depend on it through the
// total mass flux F and the old and new thickness h3, h4.  Base values (the
// tracer levels t3, t4 and told of the AB3 extrapolation, F, h3, h4 and the
// new tracer t4n) come from the recorded trajectory.

// the frozen tracer limiter coefficient of every edge (base step, on the
// raw antidiffusive flux, before k_tr_anti limits it): the factor k_tr_anti
// applies, min(1, bdn(s), bup(up)) where the flux is >= 0, else
// min(1, bup(s), bdn(up)); 1 on the other slots, which keep the unlimited
// flux until the edge halo fill (tracer_fct.F:539-593)
__global__ void k_tr_coef(Geo G, const double* __restrict__ bdn, const double* __restrict__ bup,
                          const double* __restrict__ anti, double* __restrict__ coef) {
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  if (!interior(c, iim, jjm)) {
    for (int e = 0; e < 3; ++e) coef[e * ns + s] = 1.0;
    return;
  }
  const int i = c.i, j = c.j, k = c.k;
  const int up[3] = {SL(i - 1, j, k), SL(i - 1, j - 1, k), SL(i, j - 1, k)};
  for (int e = 0; e < 3; ++e)
    coef[e * ns + s] = anti[e * ns + s] >= 0.0 ? rmin(rmin(1.0, bdn[s]), bup[up[e]])
                                               : rmin(rmin(1.0, bup[s]), bdn[up[e]]);
}

// TL of the tracer FCT fluxes (tr_fluxes_at, tracer_fct.F:166-243): the
// upstream flux f1 = t3(s) max(0, F) + t3(up) min(0, F) (the upstream cell
// taken where F >= 0 for the flux's own perturbation) and the centred flux
// of the AB3-extrapolated tracer, f2 = (TH(s) + TH(up)) F / 2; interior
// slots (0 elsewhere)
__global__ void k_tl_tr_fluxes(Geo G, double w3, double w2, double w1, const double* __restrict__ F,
                               const double* __restrict__ t3, const double* __restrict__ t4,
                               const double* __restrict__ to, const double* __restrict__ dF,
                               const double* __restrict__ dt3, const double* __restrict__ dt4,
                               const double* __restrict__ dto, double* __restrict__ df1, double* __restrict__ df2) {
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  if (!interior(c, iim, jjm)) {
    for (int e = 0; e < 3; ++e) { df1[e * ns + s] = 0.0; df2[e * ns + s] = 0.0; }
    return;
  }
  const int i = c.i, j = c.j, k = c.k;
  const int up[3] = {SL(i - 1, j, k), SL(i - 1, j - 1, k), SL(i, j - 1, k)};
  auto TH = [&](int sl) { return w3 * t3[sl] + w2 * t4[sl] + w1 * to[sl]; };
  auto DTH = [&](int sl) { return w3 * dt3[sl] + w2 * dt4[sl] + w1 * dto[sl]; };
  const double th0 = TH(s), dth0 = DTH(s);
  for (int e = 0; e < 3; ++e) {
    const double f = F[e * ns + s], df = dF[e * ns + s];
    const int u = up[e];
    df1[e * ns + s] = dt3[s] * dmax(0.0, f) + dt3[u] * dmin(0.0, f) + df * (f >= 0.0 ? t3[s] : t3[u]);
    df2[e * ns + s] = 0.5 * ((dth0 + DTH(u)) * f + (th0 + TH(u)) * df);
  }
}

// AD of k_tl_tr_fluxes: consumes the adjoints of f1 and f2 (cleared on every
// slot), adds into F (own edges) and the three tracer levels
__global__ void k_ad_tr_fluxes(Geo G, double w3, double w2, double w1, const double* __restrict__ F,
                               const double* __restrict__ t3, const double* __restrict__ t4,
                               const double* __restrict__ to, double* __restrict__ f1a, double* __restrict__ f2a,
                               double* __restrict__ Fa, double* __restrict__ t3a, double* __restrict__ t4a,
                               double* __restrict__ toa) {
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  if (!interior(c, iim, jjm)) {
    for (int e = 0; e < 3; ++e) { f1a[e * ns + s] = 0.0; f2a[e * ns + s] = 0.0; }
    return;
  }
  const int i = c.i, j = c.j, k = c.k;
  const int up[3] = {SL(i - 1, j, k), SL(i - 1, j - 1, k), SL(i, j - 1, k)};
  auto TH = [&](int sl) { return w3 * t3[sl] + w2 * t4[sl] + w1 * to[sl]; };
  auto TH_AD = [&](int sl, double g) {
    aadd(&t3a[sl], w3 * g);
    aadd(&t4a[sl], w2 * g);
    aadd(&toa[sl], w1 * g);
  };
  const double th0 = TH(s);
  double g3 = 0.0, gth = 0.0;
  for (int e = 0; e < 3; ++e) {
    const double g1 = f1a[e * ns + s], g2 = f2a[e * ns + s];
    f1a[e * ns + s] = 0.0;
    f2a[e * ns + s] = 0.0;
    const double f = F[e * ns + s];
    const int u = up[e];
    g3 += g1 * dmax(0.0, f);
    aadd(&t3a[u], g1 * dmin(0.0, f));
    const double gh = 0.5 * g2 * f;
    gth += gh;
    TH_AD(u, gh);
    Fa[e * ns + s] += g1 * (f >= 0.0 ? t3[s] : t3[u]) + 0.5 * (th0 + TH(u)) * g2;
  }
  aadd(&t3a[s], g3);
  TH_AD(s, gth);
}

// TL of k_tr_low's low-order tracer mass (tracer_fct.F:251-321): ml = h3 t3
// - dt div(f1) at the real cells (edge_update's forms), 0 elsewhere
__global__ void k_tl_tr_low(Geo G, double dt, const double* __restrict__ h3, const double* __restrict__ t3,
                            const double* __restrict__ dh3, const double* __restrict__ dt3,
                            const double* __restrict__ df1, double* __restrict__ dml) {
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  bool set;
  dml[s] = edge_update<double>(iim, jjm, ns, c, dh3[s] * t3[s] + h3[s] * dt3[s], dt * G.area_inv[s], df1, set);
}

// AD of k_tl_tr_low: consumes the adjoint of ml, adds into h3, t3 and f1
__global__ void k_ad_tr_low(Geo G, double dt, const double* __restrict__ h3, const double* __restrict__ t3,
                            double* __restrict__ mla, double* __restrict__ h3a, double* __restrict__ t3a,
                            double* __restrict__ f1a) {
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  const double g = mla[s];
  mla[s] = 0.0;
  const bool real = interior(c, iim, jjm) || (c.i == 1 && c.j == jjm - 1) || (c.i == iim - 1 && c.j == 1);
  if (!real) return;
  aadd(&h3a[s], g * t3[s]);
  aadd(&t3a[s], g * h3[s]);
  div_edge_ad(iim, jjm, ns, c, G.area_inv[s], -dt * g, f1a);
}

// TL of the limited antidiffusive flux with the frozen coefficients:
// anti = coef (f2 - f1) on every value (k_tr_limiter forms f2 - f1 on all
// slots, k_tr_anti limits the interior ones)
__global__ void k_tl_tr_anti(size_t n, const double* __restrict__ coef, const double* __restrict__ df1,
                             const double* __restrict__ df2, double* __restrict__ dan) {
  const size_t t = (size_t)blockIdx.x * bdim() + threadIdx.x;
  if (t < n) dan[t] = coef[t] * (df2[t] - df1[t]);
}
// its AD: consumes the adjoint of anti, adds into f1 and f2
__global__ void k_ad_tr_anti(size_t n, const double* __restrict__ coef, double* __restrict__ ana,
                             double* __restrict__ f1a, double* __restrict__ f2a) {
  const size_t t = (size_t)blockIdx.x * bdim() + threadIdx.x;
  if (t >= n) return;
  const double g = coef[t] * ana[t];
  ana[t] = 0.0;
  f2a[t] += g;
  f1a[t] -= g;
}

// TL of k_tr_final (tracer_fct.F:601-665, step.F:661): mf = ml - dt
// div(anti) at the real cells (0 elsewhere), the old level kept as told,
// the new tracer mf / h4 (0 where h4 = 0): dt4 = (dmf - t4n dh4) / h4
__global__ void k_tl_tr_final(Geo G, double dt, const double* __restrict__ h4, const double* __restrict__ t4n,
                              const double* __restrict__ dml, const double* __restrict__ dan,
                              const double* __restrict__ dh4, double* __restrict__ dt4, double* __restrict__ dto) {
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  bool set;
  const double dmf = edge_update<double>(iim, jjm, ns, c, dml[s], dt * G.area_inv[s], dan, set);
  dto[s] = dt4[s];
  const double hh = h4[s];
  dt4[s] = hh != 0.0 ? (dmf - t4n[s] * dh4[s]) / hh : 0.0;
}

// AD of k_tl_tr_final: the new level's adjoint (t4a) is consumed and
// replaced by that of told (the copy's transpose: told held the old level),
// told's is cleared; adds into ml, anti and h4
__global__ void k_ad_tr_final(Geo G, double dt, const double* __restrict__ h4, const double* __restrict__ t4n,
                              double* __restrict__ t4a, double* __restrict__ toa, double* __restrict__ mla,
                              double* __restrict__ ana, double* __restrict__ h4a) {
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  const double g = t4a[s];
  t4a[s] = toa[s];
  toa[s] = 0.0;
  const double hh = h4[s];
  if (hh == 0.0) return;
  const double gm = g / hh;
  aadd(&h4a[s], -t4n[s] * gm);
  const bool real = interior(c, iim, jjm) || (c.i == 1 && c.j == jjm - 1) || (c.i == iim - 1 && c.j == 1);
  if (!real) return;
  aadd(&mla[s], gm);
  div_edge_ad(iim, jjm, ns, c, G.area_inv[s], -dt * gm, ana);
}

// TL of the centred (operators.F:1923-1991) or upstream (:1841-1914) tracer
// edge flux of k_tr_flux with the base mass flux F and tracer t; interior
// slots (0 elsewhere)
template <bool CENTERED>
__global__ void k_tl_tr_flux(Geo G, const double* __restrict__ F, const double* __restrict__ t,
                             const double* __restrict__ dF, const double* __restrict__ dtr, double* __restrict__ dout) {
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  if (!interior(c, iim, jjm)) {
    for (int e = 0; e < 3; ++e) dout[e * ns + s] = 0.0;
    return;
  }
  const int i = c.i, j = c.j, k = c.k;
  const int up[3] = {SL(i - 1, j, k), SL(i - 1, j - 1, k), SL(i, j - 1, k)};
  for (int e = 0; e < 3; ++e) {
    const double f = F[e * ns + s], df = dF[e * ns + s];
    const int u = up[e];
    if (CENTERED)
      dout[e * ns + s] = 0.5 * (df * (t[s] + t[u]) + f * (dtr[s] + dtr[u]));
    else
      dout[e * ns + s] = dmax(0.0, f) * dtr[s] + dmin(0.0, f) * dtr[u] + df * (f >= 0.0 ? t[s] : t[u]);
  }
}
template <bool CENTERED>
__global__ void k_ad_tr_flux(Geo G, const double* __restrict__ F, const double* __restrict__ t,
                             double* __restrict__ outa, double* __restrict__ Fa, double* __restrict__ ta) {
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  if (!interior(c, iim, jjm)) {
    for (int e = 0; e < 3; ++e) outa[e * ns + s] = 0.0;
    return;
  }
  const int i = c.i, j = c.j, k = c.k;
  const int up[3] = {SL(i - 1, j, k), SL(i - 1, j - 1, k), SL(i, j - 1, k)};
  double g0 = 0.0;
  for (int e = 0; e < 3; ++e) {
    const double g = outa[e * ns + s];
    outa[e * ns + s] = 0.0;
    const double f = F[e * ns + s];
    const int u = up[e];
    if (CENTERED) {
      Fa[e * ns + s] += 0.5 * g * (t[s] + t[u]);
      g0 += 0.5 * f * g;
      aadd(&ta[u], 0.5 * f * g);
    } else {
      Fa[e * ns + s] += g * (f >= 0.0 ? t[s] : t[u]);
      g0 += dmax(0.0, f) * g;
      aadd(&ta[u], dmin(0.0, f) * g);
    }
  }
  aadd(&ta[s], g0);
}

// TL of k_tracer_update (step.F:278-327): tf = -div(tflux) into the AB3
// history slot a, t = t3 h3 + dt (w3 tf_a + w2 tf_b + w1 tf_c), divided by
// h4 where the base h4 exceeds 1e-10: dt4 = (dt - t4n dh4) / h4
__global__ void k_tl_tracer_update(Geo G, double dt, double w3, double w2, double w1,
                                   const double* __restrict__ dtflux, const double* __restrict__ t3,
                                   const double* __restrict__ h3, const double* __restrict__ h4,
                                   const double* __restrict__ t4n, const double* __restrict__ dt3,
                                   const double* __restrict__ dh3, const double* __restrict__ dh4,
                                   double* __restrict__ dtf_a, const double* __restrict__ dtf_b,
                                   const double* __restrict__ dtf_c, double* __restrict__ dt4) {
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  const double dtf = -div_edge_at<double>(iim, jjm, ns, c, dtflux, G.area_inv[s]);
  dtf_a[s] = dtf;
  const double d = dt3[s] * h3[s] + t3[s] * dh3[s] + dt * (w3 * dtf + w2 * dtf_b[s] + w1 * dtf_c[s]);
  const double hn = h4[s];
  dt4[s] = hn > 1.0e-10 ? (d - t4n[s] * dh4[s]) / hn : d;
}
// its AD: consumes the new level's adjoint and that of history slot a
__global__ void k_ad_tracer_update(Geo G, double dt, double w3, double w2, double w1, const double* __restrict__ t3,
                                   const double* __restrict__ h3, const double* __restrict__ h4,
                                   const double* __restrict__ t4n, double* __restrict__ t4a,
                                   double* __restrict__ tfa_a, double* __restrict__ tfb_a,
                                   double* __restrict__ tfc_a, double* __restrict__ t3a, double* __restrict__ h3a,
                                   double* __restrict__ h4a, double* __restrict__ tfluxa) {
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  const double g = t4a[s];
  t4a[s] = 0.0;
  const double hn = h4[s];
  double gd = g;
  if (hn > 1.0e-10) {
    gd = g / hn;
    aadd(&h4a[s], -t4n[s] * gd);
  }
  aadd(&t3a[s], gd * h3[s]);
  aadd(&h3a[s], gd * t3[s]);
  const double gtf = tfa_a[s] + dt * w3 * gd;
  tfa_a[s] = 0.0;
  tfb_a[s] += dt * w2 * gd;
  tfc_a[s] += dt * w1 * gd;
  div_edge_ad(iim, jjm, ns, c, G.area_inv[s], -gtf, tfluxa);
}

// ------------------------------------------------------- multigrid AD --
// Transposes of the FP64 V-cycle kernels (k_relax, k_inject, k_interp of
// swm_kernels.hip, multigrid_solver.F:593-1054).  Each kernel of the cycle is
// linear in its fields with the level's weights as constant coefficients, so
// the cycle run n times from a warm start is a linear map of (warm start,
// right-hand side), and these kernels, applied in reverse order, are its
// exact transpose.  One thread per slot of the forward kernel's output; the
// adjoint of that output is read and cleared (it has been consumed), and the
// adjoints of the inputs are added: at the thread's own slot directly, at
// neighbour slots (read through the level's halo map as the forward reads
// them) with atomic adds.

// transpose of k_relax<double, NV, ZERO>: ob the adjoint of the sweep's
// output (cleared), ab += its transpose onto the input iterate, bb += onto
// the right-hand side
template <int NV, bool ZERO>
__global__ void __launch_bounds__(256) k_relax_ad(Lev L, double* __restrict__ ob, double* __restrict__ ab,
                                                  double* __restrict__ bb) {
  const int iim = L.iim, jjm = L.jjm, ns = L.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  const int i = c.i, j = c.j, k = c.k;
  const bool in = i >= 2 && i <= iim - 1 && j >= 2 && j <= jjm - 1;
  const bool np = i == 1 && j == jjm - 1, sp = i == iim - 1 && j == 1;
  double g[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    g[v] = ob[(size_t)v * ns + s];
    ob[(size_t)v * ns + s] = 0.0;
  }
  if (!(in || np || sp)) return;   // the forward writes 0 there
  const bool deep = i >= 3 && i <= iim - 2 && j >= 3 && j <= jjm - 2;
  const int* __restrict__ m = L.src;
  int nb[6];
  if (in) {
    nb[0] = SL(i - 1, j, k); nb[1] = SL(i - 1, j - 1, k); nb[2] = SL(i, j - 1, k);
    nb[3] = SL(i + 1, j, k); nb[4] = SL(i + 1, j + 1, k); nb[5] = SL(i, j + 1, k);
  } else if (np) {
    nb[0] = -1; nb[1] = SL(iim, 1, k); nb[2] = SL(i, j - 1, k);
    nb[3] = SL(i + 1, j, k); nb[4] = SL(i + 1, j + 1, k); nb[5] = SL(i, j + 1, k);
  } else {
    nb[0] = SL(i - 1, j, k); nb[1] = SL(1, jjm, k); nb[2] = -1;
    nb[3] = SL(i + 1, j, k); nb[4] = SL(i + 1, j + 1, k); nb[5] = SL(i, j + 1, k);
  }
  int me = s;
  if (!ZERO && !deep) {
#pragma unroll
    for (int q = 0; q < 6; ++q) nb[q] = nb[q] < 0 ? -1 : m[nb[q]];
    me = m[s];
  }
  const double wt = L.lwt[s];
  const double omega = 0.8;
  double w[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) w[q] = L.lw[(size_t)q * ns + s];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    if (g[v] == 0.0) continue;
    // out = (1 - omega) a(me) + omega wt (b - sum_q w_q a(nb_q))
    bb[(size_t)v * ns + s] += omega * wt * g[v];
    if (ZERO) continue;
    double* __restrict__ av = ab + (size_t)v * ns;
    aadd(&av[me], (1.0 - omega) * g[v]);
    const double h = -omega * wt * g[v];
#pragma unroll
    for (int q = 0; q < 6; ++q)
      if (nb[q] >= 0) aadd(&av[nb[q]], w[q] * h);
  }
}

// transpose of k_inject<double, NV> (with its fused zero-start sweep z0 when
// czb is given): cbb, czb the adjoints of the coarse right-hand side and of
// that sweep (cleared at every coarse slot); fbb += at the fine point's own
// slot, fab += onto the seven fine values of the residual
template <int NV>
__global__ void k_inject_ad(Lev F, Lev C, double* __restrict__ fab, double* __restrict__ fbb,
                            double* __restrict__ cbb, double* __restrict__ czb) {
  int iim = C.iim, jjm = C.jjm;
  const int s1 = xcd_slot();
  if (s1 >= C.ns) return;
  const IJK c = ijk_of(s1, iim, jjm);
  const bool in = c.i >= 2 && c.i <= iim - 1 && c.j >= 2 && c.j <= jjm - 1;
  const bool np = c.k == 1 && c.i == 1 && c.j == jjm - 1;
  const bool sp = c.k == 10 && c.i == iim - 1 && c.j == 1;
  double r[NV];
  const double zc = czb ? 0.8 * C.lwt[s1] : 0.0;   // zero_sweep: 0.8 wt r
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    r[v] = cbb[(size_t)v * C.ns + s1];
    cbb[(size_t)v * C.ns + s1] = 0.0;
    if (czb) {
      r[v] += zc * czb[(size_t)v * C.ns + s1];
      czb[(size_t)v * C.ns + s1] = 0.0;
    }
  }
  // outside the injected slots the coarse values are constants of memory
  if (!(in || np || sp)) return;
  const bool deep = c.i >= 3 && c.i <= iim - 2 && c.j >= 3 && c.j <= jjm - 2;
  const int k = c.k;
  int i2, j2;
  if (in) { i2 = 2 * c.i - 1; j2 = 2 * c.j - 1; }
  else if (np) { i2 = 1; j2 = F.jjm - 1; }
  else { i2 = F.iim - 1; j2 = 1; }
  iim = F.iim; jjm = F.jjm;
  const int ns = F.ns;
  const int s = SL(i2, j2, k);
  const int* __restrict__ m = F.src;
  int q[7];
  q[0] = s;
  if (in) {
    q[1] = SL(i2 - 1, j2, k); q[2] = SL(i2 - 1, j2 - 1, k); q[3] = SL(i2, j2 - 1, k);
  } else if (np) {
    q[1] = s; q[2] = SL(iim, 1, k); q[3] = SL(i2, j2 - 1, k);
  } else {
    q[1] = SL(i2 - 1, j2, k); q[2] = SL(1, jjm, k); q[3] = s;
  }
  q[4] = SL(i2 + 1, j2, k); q[5] = SL(i2 + 1, j2 + 1, k); q[6] = SL(i2, j2 + 1, k);
  if (!deep) {
#pragma unroll
    for (int n = 0; n < 7; ++n) q[n] = m[q[n]];
  }
  double w[7];
#pragma unroll
  for (int n = 0; n < 7; ++n) w[n] = F.lw[(size_t)n * ns + s];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    if (r[v] == 0.0) continue;
    // r = b(s) - (w7 x(q0) + sum_n w_n x(q_n)), pole forms without one term
    fbb[(size_t)v * ns + s] += r[v];
    double* __restrict__ a = fab + (size_t)v * ns;
    aadd(&a[q[0]], -w[6] * r[v]);
    if (!np) aadd(&a[q[1]], -w[0] * r[v]);
    aadd(&a[q[2]], -w[1] * r[v]);
    if (!sp) aadd(&a[q[3]], -w[2] * r[v]);
    aadd(&a[q[4]], -w[3] * r[v]);
    aadd(&a[q[5]], -w[4] * r[v]);
    aadd(&a[q[6]], -w[5] * r[v]);
  }
}

// transpose of k_interp<double, NV> (x2 += P x1): x1b += P^T x2b; the fine
// adjoint is unchanged (the identity part of the update)
template <int NV>
__global__ void k_interp_ad(Lev C, Lev F, double* __restrict__ x1b, const double* __restrict__ x2b) {
  const int iim = F.iim, jjm = F.jjm;
  const int s = xcd_slot();
  if (s >= F.ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  const int i = c.i, j = c.j, k = c.k;
  const int ci = C.iim, cj = C.jjm;
  auto CS = [&](int ii, int jj) {
    const int a = (ii + 1) / 2, b = (jj + 1) / 2;
    const int sl = (a - 1) + ci * ((b - 1) + cj * (k - 1));
    return (a == 1 || a == ci || b == 1 || b == cj) ? C.src[sl] : sl;
  };
  const bool io = (i & 1), jo = (j & 1);
  const bool pn = i == 1 && j == jjm - 1, ps = i == iim - 1 && j == 1;
  const bool iA = i >= 3 && i <= iim - 1, iB = i >= 2 && i <= iim - 2;
  const bool jA = j >= 3 && j <= jjm - 1, jB = j >= 2 && j <= jjm - 2;
  const bool cA = io && jo && iA && jA;
  const bool cB = !io && jo && iB && jA;
  const bool cC = !io && !jo && iB && jB;
  const bool cD = io && !jo && iA && jB;
  const bool two = cB || cC || cD;
  if (!(pn || ps || cA || two)) return;
  const int di = (cB || cC) ? 1 : 0, dj = (cC || cD) ? 1 : 0;
  const int c0 = CS(i + di, j + dj);
  const int c1 = two ? CS(i - di, j - dj) : c0;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const double g = x2b[(size_t)v * F.ns + s];
    if (g == 0.0) continue;
    double* __restrict__ xc = x1b + (size_t)v * C.ns;
    if (two) {
      aadd(&xc[c0], 0.5 * g);
      aadd(&xc[c1], 0.5 * g);
    } else {
      aadd(&xc[c0], g);
    }
  }
}

// transpose of the coarse-level cycle (k_coarse_cycle from a zero start),
// alpha = M beta at the n slots of its finest level: bb += M^T ab, with
// mt = M^T stored by rows (row j: the derivatives of alpha with respect to
// beta(j)); one wave per row, field blockIdx.y
__global__ void k_coarse_ad(const double* __restrict__ mt, int n, const double* __restrict__ ab,
                            double* __restrict__ bb) {
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * (bdim() >> 6) + (threadIdx.x >> 6);
  if (row >= n) return;
  const double* __restrict__ r = mt + (size_t)row * n;
  const double* __restrict__ a = ab + (size_t)blockIdx.y * n;
  double acc = 0.0;
  for (int s = lane; s < n; s += 64) acc += r[s] * a[s];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  if (lane == 0) bb[(size_t)blockIdx.y * n + row] += acc;
}

// x = unit vector e_j on n values (the coarse cycle's probes)
__global__ void k_unit(double* __restrict__ x, int n, int j) {
  const int t = blockIdx.x * bdim() + threadIdx.x;
  if (t < n) x[t] = t == j ? 1.0 : 0.0;
}

#undef UPL
#undef NPL
#undef DPL
#undef SL

}  // namespace dev
}  // namespace swm
