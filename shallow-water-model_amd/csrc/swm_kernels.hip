// swm_kernels.hip -- gfx950 kernels of the shallow-water dynamical core (FP64).
//
// Data layout: every field is a set of planes; one plane is the reference's
// (iim, jjm, 10) panel array of one component, i fastest (256-byte aligned
// allocations, one thread per slot, consecutive threads on consecutive i, so
// every plane access is coalesced).  Multi-component fields use the plane
// orders
//   corner scalar   (ic)            plane ic
//   corner vector   (space, ic)     plane ic + 2*space
//   edge field      (e)             plane e
//   dc, sub_area, s_weights (kk, ic)        plane kk + 3*ic
//   normal, tangent, v_weights (kk, a, ic)  plane kk + 3*(a + 2*ic)
//   l_weights (n)                           plane n
// so that one halo table indexes (component, slot) as plane*ns + slot for all
// of them.  All stencils are HBM-bandwidth bound (a few flops per byte), so
// the kernels are written for coalescing and few passes, not for MFMA.
//
// Each kernel restates the reference routine named in its comment with the
// same floating-point operation order (compiled without contraction), so
// the results agree with the reference to the last bit or to round-off.
#include "swm_kernels.h"

namespace swm {
namespace dev {

#define SL(i, j, k) ((i) - 1 + iim * ((j) - 1 + jjm * ((k) - 1)))

struct IJK {
  int i, j, k;
};
// q = s / d and r = s % d for 0 <= s < 2^24 and 4 <= d < 2^12 in a dozen
// instructions instead of the integer division's ~25: float(s) is exact and
// the product with the approximate reciprocal of d is within 3/4 of s / d,
// so its truncation is the quotient or one off, which the remainder's sign
// and size correct.
__device__ __forceinline__ void divmod_small(int s, int d, int& q, int& r) {
  const float rd = __builtin_amdgcn_rcpf((float)d);
  q = (int)((float)s * rd);
  r = s - q * d;
  if (r < 0) { q -= 1; r += d; }
  if (r >= d) { q += 1; r -= d; }
}
__device__ __forceinline__ IJK ijk_of(int s, int iim, int jjm) {
  IJK r;
  int q, t;
  divmod_small(s, iim, q, t);
  r.i = t + 1;
  int k1, j1;
  divmod_small(q, jjm, k1, j1);
  r.j = j1 + 1;
  r.k = k1 + 1;
  return r;
}
// load and store of element i (0 <= i, i * sizeof(T) < 2^32) of a buffer
// whose base is the same for the whole wave: the byte offset is formed in 32
// bits, so the access takes the scalar base plus one 32-bit offset register
// instead of a 64-bit address computed per lane
template <class T>
__device__ __forceinline__ T ldg(const T* __restrict__ p, int i) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) + (uint32_t)i * (uint32_t)sizeof(T));
}
template <class T>
__device__ __forceinline__ void stg(T* __restrict__ p, int i, T x) {
  *reinterpret_cast<T*>(reinterpret_cast<char*>(p) + (uint32_t)i * (uint32_t)sizeof(T)) = x;
}
// two consecutive elements i, i + 1 in one 16-byte (FP64) or 8-byte (FP32)
// access; the element index need only be even for the natural alignment
// (odd ones give 8- or 4-byte aligned accesses, which global memory takes)
struct __attribute__((aligned(8))) D2 {
  double x, y;
};
struct __attribute__((aligned(4))) F2 {
  float x, y;
};
__device__ __forceinline__ D2 ld2(const double* __restrict__ p, int i) {
  return *reinterpret_cast<const D2*>(reinterpret_cast<const char*>(p) + (uint32_t)i * 8u);
}
__device__ __forceinline__ F2 ld2(const float* __restrict__ p, int i) {
  return *reinterpret_cast<const F2*>(reinterpret_cast<const char*>(p) + (uint32_t)i * 4u);
}
__device__ __forceinline__ void st2(double* __restrict__ p, int i, double a, double b) {
  *reinterpret_cast<D2*>(reinterpret_cast<char*>(p) + (uint32_t)i * 8u) = D2{a, b};
}
__device__ __forceinline__ void st2(float* __restrict__ p, int i, float a, float b) {
  *reinterpret_cast<F2*>(reinterpret_cast<char*>(p) + (uint32_t)i * 4u) = F2{a, b};
}
// (p[i], p[i + 1]) of a plane of n elements, for -1 <= i <= n - 1: the
// element outside [0, n) is a copy of the one inside
template <class T>
__device__ __forceinline__ auto ld2c(const T* __restrict__ p, int i, int n) {
  const int b = i < 0 ? 0 : (i > n - 2 ? n - 2 : i);
  auto t = ld2(p, b);
  if (b > i) t.y = t.x;
  if (b < i) t.x = t.y;
  return t;
}
__device__ __forceinline__ bool interior(const IJK& c, int iim, int jjm) {
  return c.i >= 2 && c.i <= iim - 1 && c.j >= 2 && c.j <= jjm - 1;
}
// ------------------------------------------- producer-side halo writes ---
// Every halo table entry copies a slot one cell inside a panel edge (or a
// pole slot) to a halo slot, so only slots within one cell of the edge have
// push words; the others read nothing.  The words (with the first three
// entries) are one load at the start of a kernel, beside its other loads, so
// that the copies at its end wait for no further load.
struct PushW {
  int4 w;
};
__device__ __forceinline__ PushW push_load(const Push& P, const IJK& c, int iim, int jjm, int s) {
  PushW r;
  const bool rim = c.i <= 2 || c.i >= iim - 1 || c.j <= 2 || c.j >= jjm - 1;
  r.w = (P.hp != nullptr && rim) ? ldg(P.hp, s) : int4{0, 0, 0, 0};
  return r;
}
// the copies of the slot's values v0, v1, v2 (components 0..NC-1 of field
// x, planes ns apart) into the halo slots that hold them.  The values come
// as scalars: a selected element of a local array would be read through a
// selected address, which puts the array in memory.
template <class R, int NC>
__device__ __forceinline__ void push_one(R* __restrict__ x, int t, R v0, R v1, R v2) {
  const int comp = (int)((unsigned)t >> 28);
  R val = v0;
  if constexpr (NC > 1) val = comp == 1 ? v1 : val;
  if constexpr (NC > 2) val = comp == 2 ? v2 : val;
  stg(x, t & 0x0fffffff, val);
}
template <class R, int NC>
__device__ __forceinline__ void push(const Push& P, const PushW& W, R* __restrict__ x, R v0, R v1, R v2) {
  const int n = W.w.x & 15;
  if (n > 0) push_one<R, NC>(x, W.w.y, v0, v1, v2);
  if (n > 1) push_one<R, NC>(x, W.w.z, v0, v1, v2);
  if (n > 2) push_one<R, NC>(x, W.w.w, v0, v1, v2);
  const int e0 = (int)((unsigned)W.w.x >> 10);
  for (int q = 3; q < n; ++q) push_one<R, NC>(x, ldg(P.e, e0 + q - 3), v0, v1, v2);   // corner and pole slots only
}
// the slot's values of field x (components 0..NC-1, planes ns apart):
// stored, and with a push word (w != 0) stored only where the slot is not a
// halo copy (zero where the fill has no source) and copied into the halo
// slots that hold them.  Slots without a word take the plain stores.
template <class R, int NC>
__device__ __forceinline__ void put_all(const Push& P, const PushW& W, R* __restrict__ x, int ns, int s, R v0,
                                        R v1 = R(0), R v2 = R(0)) {
  const int w = W.w.x;
  if (w == 0) {
    stg(x, s, v0);
    if constexpr (NC > 1) stg(x + (size_t)ns, s, v1);
    if constexpr (NC > 2) stg(x + (size_t)2 * ns, s, v2);
    return;
  }
  const int m0 = (w >> 4) & 3, m1 = (w >> 6) & 3, m2 = (w >> 8) & 3;
  if (m0 != 1) stg(x, s, m0 == 2 ? R(0) : v0);
  if constexpr (NC > 1)
    if (m1 != 1) stg(x + (size_t)ns, s, m1 == 2 ? R(0) : v1);
  if constexpr (NC > 2)
    if (m2 != 1) stg(x + (size_t)2 * ns, s, m2 == 2 ? R(0) : v2);
  push<R, NC>(P, W, x, v0, v1, v2);
}
template <class R>
__device__ __forceinline__ void put1(const Push& P, const PushW& W, R* __restrict__ x, int s, R v) {
  put_all<R, 1>(P, W, x, 0, s, v);
}
// a kernel that leaves the slot's own values as they are: only the zeros of
// the fill where it has no source
template <class R, int NC>
__device__ __forceinline__ void put_zeros(const PushW& W, R* __restrict__ x, int ns, int s) {
  const int w = W.w.x;
  if (w == 0) return;
#pragma unroll
  for (int c = 0; c < NC; ++c)
    if (((w >> (4 + 2 * c)) & 3) == 2) stg(x + (size_t)c * ns, s, R(0));
}

__device__ __forceinline__ double dmin(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }


// ------------------------------------------------------ XCD-aware slots ---
// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  Stencil kernels therefore renumber their blocks so that XCD x works on
// one contiguous range of blocks: neighbouring rows (j-1, j+1) are then
// mostly fetched into the same L2 instead of into three different ones.
// Bijective on [0, gridDim.x) for any block count.
__device__ __forceinline__ int xcd_block() {
  const int nb = gridDim.x, b = blockIdx.x;
  const int per = (nb + 7) >> 3;          // blocks of the fullest XCDs
  const int full = nb - 8 * (per - 1);    // XCDs that get 'per' blocks
  const int x = b & 7, q = b >> 3;
  const int start = x <= full ? x * per : full * per + (x - full) * (per - 1);
  return start + q;
}
// the workgroup's size with one scalar load: hidden_group_size_x, the 16
// bits at byte 12 of the implicit kernel arguments (code object v5), which
// hold for the whole dispatch.  blockDim.x reads the same field with a
// vector load at the start of every wave, choosing the remainder field for
// a partial last workgroup; every launch here uses whole workgroups.
__device__ __forceinline__ int bdim() {
  const __attribute__((address_space(4))) unsigned* p =
      (const __attribute__((address_space(4))) unsigned*)__builtin_amdgcn_implicitarg_ptr();
  return (int)(p[3] & 0xffffu);
}
__device__ __forceinline__ int xcd_slot() { return xcd_block() * bdim() + threadIdx.x; }

// the workgroup size bdim() reads from the implicit kernel arguments
// (code-object v5 layout, which the build pins) against blockDim.x: 1 per
// block where they agree
__global__ void k_block_size_probe(int* out) {
  if (threadIdx.x == 0) out[blockIdx.x] = bdim() == (int)blockDim.x ? 1 : 0;
}

template <class R>
__device__ __forceinline__ R rmin(R a, R b) { return a < b ? a : b; }
template <class R>
__device__ __forceinline__ R rmax(R a, R b) { return a > b ? a : b; }

// the mixed path stores thickness-like fields less h0 and the vorticity less
// the Coriolis parameter; these give the absolute values (identity in FP64)
template <class R>
__device__ __forceinline__ R habs(const GeoT<R>& G, R x) {
  if constexpr (sizeof(R) == 4) return x + G.h0;
  else return x;
}
template <class R>
__device__ __forceinline__ R eabs(const GeoT<R>& G, const R* __restrict__ eta, int s) {
  if constexpr (sizeof(R) == 4) return eta[s] + G.fo[s];
  else return eta[s];
}

// ------------------------------------------------------------------ halo --
// x[dst] = (src < 0 ? 0 : x[src]) for every table entry, on up to 8 fields
// (blockIdx.y); no source is a destination, so it runs in place.
template <class R>
__global__ void k_gather(const int* __restrict__ dst, const int* __restrict__ src, int n, FieldSetT<R> fs) {
  const int t = blockIdx.x * bdim() + threadIdx.x;
  if (t >= n) return;
  R* x = fs.p[blockIdx.y];
  at_member(fs.mb, x);
  const int s = src[t];
  x[dst[t]] = s < 0 ? R(0) : x[s];
}

// ---------------------------------------------------- corner averages ----
// operators.F:982-1112 average_to_cell_corners (flag = .true.) of the
// thickness (absolute values); the step's corner vorticity (step.F:132-141)
// only feeds an unused work array (step.F:391-393) and is not formed
template <class R>
__global__ void k_avg_corners(GeoT<R> G, const R* __restrict__ sw, const R* __restrict__ x1, R* __restrict__ o1) {
  at_member(G.mb, x1, o1);
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  const int i = c.i, j = c.j, k = c.k;
  const PushW w = push_load(G.pv, c, iim, jjm, s);
  R v[2];
  if (i >= 2 && i <= iim - 1 && j >= 2 && j <= jjm - 1) {
    const int a = SL(i, j, k), b = SL(i - 1, j, k), d = SL(i - 1, j - 1, k), e = SL(i, j - 1, k);
    const R w10 = sw[0 * ns + s], w20 = sw[1 * ns + s], w30 = sw[2 * ns + s];
    const R w11 = sw[3 * ns + s], w21 = sw[4 * ns + s], w31 = sw[5 * ns + s];
    const R xa = habs(G, x1[a]), xb = habs(G, x1[b]), xd = habs(G, x1[d]), xe = habs(G, x1[e]);
    v[0] = +w10 * xa + w20 * xb + w30 * xd;
    v[1] = +w11 * xa + w21 * xd + w31 * xe;
  } else {
    const R f1 = habs(G, x1[SL(2, 2, k)]);
    v[0] = f1;
    v[1] = f1;
  }
  put_all<R, 2>(G.pv, w, o1, ns, s, v[0], v[1]);
}

// half-edge fluxes of cell (i,j) as combined by centered_mass_flux
// (operators.F:1626-1722); with mc == nullptr the corner-mass factor is
// dropped (upstream_mass_flux, operators.F:1731-1832).  nrm is either the
// normal or the tangent planes.
template <class R>
struct Half {
  R r[2][3];
};
template <class R>
__device__ __forceinline__ Half<R> half_edges(int iim, int jjm, int ns, int i, int j, int k, const R* __restrict__ dc,
                                              const R* __restrict__ nrm, const R* __restrict__ mc,
                                              const R* __restrict__ u) {
  // u planes: (ic + 2*sp); nrm planes: kk + 3*(a + 2*ic); dc planes kk + 3*ic
#define U(sp, ic, s_) u[((ic) + 2 * (sp)) * ns + (s_)]
#define N(kk, a, ic, s_) nrm[((kk) + 3 * ((a) + 2 * (ic))) * ns + (s_)]
#define D(kk, ic, s_) dc[((kk) + 3 * (ic)) * ns + (s_)]
#define HN(s_, kk, ic) (U(0, ic, s_) * N(kk, 0, ic, s_) + U(1, ic, s_) * N(kk, 1, ic, s_))
  Half<R> h;
  const int s0 = SL(i, j, k), sn = SL(i, j + 1, k), se = SL(i + 1, j, k);
  if (mc) {
    h.r[0][0] = HN(sn, 1, 1) * mc[ns + sn]; h.r[0][0] = h.r[0][0] * D(1, 1, sn);
    h.r[1][0] = HN(s0, 0, 0) * mc[s0];      h.r[1][0] = h.r[1][0] * D(0, 0, s0);
    h.r[0][1] = HN(s0, 2, 0) * mc[s0];      h.r[0][1] = h.r[0][1] * D(2, 0, s0);
    h.r[1][1] = HN(s0, 0, 1) * mc[ns + s0]; h.r[1][1] = h.r[1][1] * D(0, 1, s0);
    h.r[0][2] = HN(s0, 2, 1) * mc[ns + s0]; h.r[0][2] = h.r[0][2] * D(2, 1, s0);
    h.r[1][2] = HN(se, 1, 0) * mc[se];      h.r[1][2] = h.r[1][2] * D(1, 0, se);
  } else {
    h.r[0][0] = HN(sn, 1, 1); h.r[0][0] = h.r[0][0] * D(1, 1, sn);
    h.r[1][0] = HN(s0, 0, 0); h.r[1][0] = h.r[1][0] * D(0, 0, s0);
    h.r[0][1] = HN(s0, 2, 0); h.r[0][1] = h.r[0][1] * D(2, 0, s0);
    h.r[1][1] = HN(s0, 0, 1); h.r[1][1] = h.r[1][1] * D(0, 1, s0);
    h.r[0][2] = HN(s0, 2, 1); h.r[0][2] = h.r[0][2] * D(2, 1, s0);
    h.r[1][2] = HN(se, 1, 0); h.r[1][2] = h.r[1][2] * D(1, 0, se);
  }
#undef U
#undef N
#undef D
#undef HN
  return h;
}

// half_edges for the mixed path from premultiplied planes nd = n * dc (or
// t * dc): r = (u . nd) * mc, or (u . nd) without the corner mass
template <class R>
__device__ __forceinline__ Half<R> half_edges_pre(int iim, int jjm, int ns, int i, int j, int k,
                                                  const R* __restrict__ nd, const R* __restrict__ mc,
                                                  const R* __restrict__ u) {
#define U(sp, ic, s_) u[((ic) + 2 * (sp)) * ns + (s_)]
#define N(kk, a, ic, s_) nd[((kk) + 3 * ((a) + 2 * (ic))) * ns + (s_)]
#define HN(s_, kk, ic) (U(0, ic, s_) * N(kk, 0, ic, s_) + U(1, ic, s_) * N(kk, 1, ic, s_))
  Half<R> h;
  const int s0 = SL(i, j, k), sn = SL(i, j + 1, k), se = SL(i + 1, j, k);
  h.r[0][0] = HN(sn, 1, 1); h.r[1][0] = HN(s0, 0, 0); h.r[0][1] = HN(s0, 2, 0);
  h.r[1][1] = HN(s0, 0, 1); h.r[0][2] = HN(s0, 2, 1); h.r[1][2] = HN(se, 1, 0);
  if (mc) {
    h.r[0][0] = h.r[0][0] * mc[ns + sn]; h.r[1][0] = h.r[1][0] * mc[s0]; h.r[0][1] = h.r[0][1] * mc[s0];
    h.r[1][1] = h.r[1][1] * mc[ns + s0]; h.r[0][2] = h.r[0][2] * mc[ns + s0]; h.r[1][2] = h.r[1][2] * mc[se];
  }
#undef U
#undef N
#undef HN
  return h;
}

// the tangential half-edge fluxes (u . t) dc * mc of half_edges_pre from
// the normal planes: the wall's unit tangent is its unit normal turned by a
// right angle in the corner's east/north frame, t = (-n_N, n_E) (the metric
// planes agree with it to one FP64 rounding on every slot), so
// u . t dc = -u_E (n_N dc) + u_N (n_E dc)
template <class R>
__device__ __forceinline__ Half<R> half_edges_tan_pre(int iim, int jjm, int ns, int i, int j, int k,
                                                      const R* __restrict__ nd, const R* __restrict__ mc,
                                                      const R* __restrict__ u) {
#define U(sp, ic, s_) u[((ic) + 2 * (sp)) * ns + (s_)]
#define N(kk, a, ic, s_) nd[((kk) + 3 * ((a) + 2 * (ic))) * ns + (s_)]
#define HT(s_, kk, ic) (U(0, ic, s_) * (-N(kk, 1, ic, s_)) + U(1, ic, s_) * N(kk, 0, ic, s_))
  Half<R> h;
  const int s0 = SL(i, j, k), sn = SL(i, j + 1, k), se = SL(i + 1, j, k);
  h.r[0][0] = HT(sn, 1, 1); h.r[1][0] = HT(s0, 0, 0); h.r[0][1] = HT(s0, 2, 0);
  h.r[1][1] = HT(s0, 0, 1); h.r[0][2] = HT(s0, 2, 1); h.r[1][2] = HT(se, 1, 0);
  h.r[0][0] = h.r[0][0] * mc[ns + sn]; h.r[1][0] = h.r[1][0] * mc[s0]; h.r[0][1] = h.r[0][1] * mc[s0];
  h.r[1][1] = h.r[1][1] * mc[ns + s0]; h.r[0][2] = h.r[0][2] * mc[ns + s0]; h.r[1][2] = h.r[1][2] * mc[se];
#undef U
#undef N
#undef HT
  return h;
}

// premultiplied metric planes of the mixed path: out(p) = nrm(p) * dc(kk, ic)
// for the plane p = kk + 3 (a + 2 ic), formed in FP64
__global__ void k_premul_dc(int ns, const double* __restrict__ nrm, const double* __restrict__ dc,
                            float* __restrict__ out) {
  const int s = blockIdx.x * bdim() + threadIdx.x;
  if (s >= ns) return;
  for (int ic = 0; ic < 2; ++ic)
    for (int a = 0; a < 2; ++a)
      for (int kk = 0; kk < 3; ++kk) {
        const int p = kk + 3 * (a + 2 * ic);
        out[(size_t)p * ns + s] = (float)(nrm[(size_t)p * ns + s] * dc[(size_t)(kk + 3 * ic) * ns + s]);
      }
}

// upstream (first-order) and centred (high-order tendency) edge mass fluxes
// (mass_fct.F:151-214 via operators.F:1626-1832)
template <class R>
__global__ void k_mass_fluxes(GeoT<R> G, const R* __restrict__ u, const R* __restrict__ h3, const R* __restrict__ hc,
                              R* __restrict__ flx1, R* __restrict__ mshf, const R* __restrict__ eta3,
                              R* __restrict__ dcurl, R* __restrict__ ke) {
  if constexpr (sizeof(R) == 8) {
    // FP64: the loads where the forms use them (measured: every load first
    // is 40 us slower at g=9 for the FP64 planes)
    at_member(G.mb, u, h3, hc, flx1, mshf);
    const bool dcr = dcurl != nullptr;
    if (dcr) at_member(G.mb, eta3, dcurl);
    const int iim = G.iim, jjm = G.jjm, ns = G.ns;
    const int s = xcd_slot();
    if (s >= ns) return;
    const IJK c = ijk_of(s, iim, jjm);
    const PushW w = push_load(G.pe, c, iim, jjm, s);
    if (ke) {
      // the kinetic energy of this (old) level's velocity, k_center_ke's
      // arithmetic (step.F:192-203, operators.F:1316-1393), from the velocity
      // this kernel reads anyway; the previous step left it out
      at_member(G.mb, ke);
      const PushW wk = push_load(G.pf, c, iim, jjm, s);
      const int i = c.i, j = c.j, k = c.k;
      R kv = R(0);
      if (i <= iim - 1 && j <= jjm - 1) {
        auto KC = [&](int ic, int sl) {
          const R a = u[ic * ns + sl], b = u[(ic + 2) * ns + sl];
          return R(0.5) * (a * a + b * b);
        };
  #define SA(kk, ic, sl) G.sub_area[((kk) + 3 * (ic)) * ns + (sl)]
        const int se = SL(i + 1, j, k), sne = SL(i + 1, j + 1, k), sn = SL(i, j + 1, k);
        kv = (+SA(0, 0, s) * KC(0, s) + SA(0, 1, s) * KC(1, s) + SA(1, 0, se) * KC(0, se) +
              SA(1, 1, sne) * KC(1, sne) + SA(2, 0, sne) * KC(0, sne) + SA(2, 1, sn) * KC(1, sn)) * G.area_inv[s];
  #undef SA
      }
      put1(G.pf, wk, ke, s, kv);
    }
    if (!interior(c, iim, jjm)) {
      put_all<R, 3>(G.pe, w, flx1, ns, s, R(0), R(0), R(0));
      put_all<R, 3>(G.pe, w, mshf, ns, s, R(0), R(0), R(0));
      if (dcr) put_all<R, 3>(G.pe, w, dcurl, ns, s, R(0), R(0), R(0));
      return;
    }
    const int i = c.i, j = c.j, k = c.k;
    Half<R> hc_, hu;
    if constexpr (sizeof(R) == 4) {
      hc_ = half_edges_pre<R>(iim, jjm, ns, i, j, k, G.ndc, hc, u);
      hu = half_edges_pre<R>(iim, jjm, ns, i, j, k, G.ndc, nullptr, u);
    } else {
      hc_ = half_edges<R>(iim, jjm, ns, i, j, k, G.dc, G.nrm, hc, u);
      hu = half_edges<R>(iim, jjm, ns, i, j, k, G.dc, G.nrm, nullptr, u);
    }
    const int up[3] = {SL(i - 1, j, k), SL(i - 1, j - 1, k), SL(i, j - 1, k)};
    const R hs0 = habs(G, h3[s]);
    R vm[3], vf[3];
  #pragma unroll
    for (int e = 0; e < 3; ++e) {
      vm[e] = -hc_.r[0][e] + hc_.r[1][e];
      const R v = hu.r[1][e] - hu.r[0][e];
      const R mx = v > R(0) ? v : R(0), mn = v < R(0) ? v : R(0);
      vf[e] = mx * hs0 + mn * habs(G, h3[up[e]]);
    }
    put_all<R, 3>(G.pe, w, mshf, ns, s, vm[0], vm[1], vm[2]);
    put_all<R, 3>(G.pe, w, flx1, ns, s, vf[0], vf[1], vf[2]);
    if (dcr) {
      // dcurl: centered_divergence_curl (operators.F:2083-2186, k_vort_curl's
      // arithmetic) from the same old-level u, corner thickness, h and eta
      Half<R> ht;
      if constexpr (sizeof(R) == 4) ht = half_edges_tan_pre<R>(iim, jjm, ns, i, j, k, G.ndc, hc, u);
      else ht = half_edges<R>(iim, jjm, ns, i, j, k, G.dc, G.tng, hc, u);
      auto PV = [&](int sl) {
        const R hh = habs(G, h3[sl]);
        const R wi = hh > R(0) ? R(1) / hh : R(0);
        return eabs(G, eta3, sl) * wi;
      };
      const R pv0 = PV(s);
      R vd[3];
  #pragma unroll
      for (int e = 0; e < 3; ++e) vd[e] = (-ht.r[0][e] + ht.r[1][e]) * (PV(up[e]) + pv0) * R(0.5);
      put_all<R, 3>(G.pe, w, dcurl, ns, s, vd[0], vd[1], vd[2]);
    }
  } else {
    at_member(G.mb, u, h3, hc, flx1, mshf);
    const bool dcr = dcurl != nullptr;
    if (dcr) at_member(G.mb, eta3, dcurl);
    if (ke) at_member(G.mb, ke);
    const int iim = G.iim, jjm = G.jjm, ns = G.ns;
    const int s = xcd_slot();
    if (s >= ns) return;
    const IJK c = ijk_of(s, iim, jjm);
    const int i = c.i, j = c.j, k = c.k;
    const PushW w = push_load(G.pe, c, iim, jjm, s);
    const PushW wk = ke ? push_load(G.pf, c, iim, jjm, s) : PushW{};
    const bool in = interior(c, iim, jjm);
    const bool kreal = i <= iim - 1 && j <= jjm - 1;
    // every value first, so that the loads overlap; neighbour indices are
    // clamped into the array (values a slot's forms do not use are not used)
    auto cl = [&](int sl) { return min(max(sl, 0), ns - 1); };
    const int s0 = s, sn = cl(SL(i, j + 1, k)), se = cl(SL(i + 1, j, k)), sne = cl(SL(i + 1, j + 1, k));
    const int up[3] = {cl(SL(i - 1, j, k)), cl(SL(i - 1, j - 1, k)), cl(SL(i, j - 1, k))};
    // velocity (planes ic + 2 sp): both corners at s0, corner 1 at sn, corner
    // 0 at se, and for the kinetic energy both at sne
    auto U = [&](int sp, int ic, int sl) { return ldg(u + (size_t)(ic + 2 * sp) * ns, sl); };
    const R u00 = U(0, 0, s0), u10 = U(1, 0, s0), u01 = U(0, 1, s0), u11 = U(1, 1, s0);
    const R un0 = U(0, 1, sn), un1 = U(1, 1, sn), ue0 = U(0, 0, se), ue1 = U(1, 0, se);
    R une[2][2] = {{R(0), R(0)}, {R(0), R(0)}};   // [ic][sp]
    R sa[6] = {R(0), R(0), R(0), R(0), R(0), R(0)};
    R ai = R(0);
    if (ke) {
  #pragma unroll
      for (int ic = 0; ic < 2; ++ic) { une[ic][0] = U(0, ic, sne); une[ic][1] = U(1, ic, sne); }
      auto SA = [&](int kk, int ic, int sl) { return ldg(G.sub_area + (size_t)(kk + 3 * ic) * ns, sl); };
      sa[0] = SA(0, 0, s0); sa[1] = SA(0, 1, s0); sa[2] = SA(1, 0, se);
      sa[3] = SA(1, 1, sne); sa[4] = SA(2, 0, sne); sa[5] = SA(2, 1, sn);
      ai = ldg(G.area_inv, s0);
    }
    // corner thickness: corner 1 at sn and s0, corner 0 at s0 and se
    const R mcn = ldg(hc + (size_t)ns, sn), mc0 = ldg(hc, s0), mc1 = ldg(hc + (size_t)ns, s0), mce = ldg(hc, se);
    const R hs = ldg(h3, s0);
    R hup[3], es = R(0), eup[3] = {R(0), R(0), R(0)};
  #pragma unroll
    for (int e = 0; e < 3; ++e) hup[e] = ldg(h3, up[e]);
    if (dcr) {
      es = eabs(G, eta3, s0);
  #pragma unroll
      for (int e = 0; e < 3; ++e) eup[e] = eabs(G, eta3, up[e]);
    }
    // the metric values of the six half edges: (point, kk, ic) =
    // (sn,1,1) (s0,0,0) (s0,2,0) (s0,0,1) (s0,2,1) (se,1,0)
    const int hp[6] = {sn, s0, s0, s0, s0, se}, hk[6] = {1, 0, 2, 0, 2, 1}, hi[6] = {1, 0, 0, 1, 1, 0};
    R nE[6], nN[6], tE[6], tN[6], dd[6];
  #pragma unroll
    for (int q = 0; q < 6; ++q) {
      if constexpr (sizeof(R) == 4) {
        // premultiplied n dc; the tangent turned from the normal
        nE[q] = ldg(G.ndc + (size_t)(hk[q] + 3 * (0 + 2 * hi[q])) * ns, hp[q]);
        nN[q] = ldg(G.ndc + (size_t)(hk[q] + 3 * (1 + 2 * hi[q])) * ns, hp[q]);
        tE[q] = R(0); tN[q] = R(0); dd[q] = R(0);
      } else {
        nE[q] = ldg(G.nrm + (size_t)(hk[q] + 3 * (0 + 2 * hi[q])) * ns, hp[q]);
        nN[q] = ldg(G.nrm + (size_t)(hk[q] + 3 * (1 + 2 * hi[q])) * ns, hp[q]);
        dd[q] = ldg(G.dc + (size_t)(hk[q] + 3 * hi[q]) * ns, hp[q]);
        if (dcr) {
          tE[q] = ldg(G.tng + (size_t)(hk[q] + 3 * (0 + 2 * hi[q])) * ns, hp[q]);
          tN[q] = ldg(G.tng + (size_t)(hk[q] + 3 * (1 + 2 * hi[q])) * ns, hp[q]);
        } else {
          tE[q] = R(0); tN[q] = R(0);
        }
      }
    }
    // the velocity and corner thickness of each half edge
    const R ua[6] = {un0, u00, u00, u01, u01, ue0}, ub[6] = {un1, u10, u10, u11, u11, ue1};
    const R mcq[6] = {mcn, mc0, mc0, mc1, mc1, mce};
    // half-edge normal fluxes with (hc) and without (hu) the corner thickness,
    // and tangential ones with it (ht): r[0][e], r[1][e] = half edges 2e', as
    // half_edges / half_edges_pre / half_edges_tan_pre order them
    R hcq[6], huq[6], htq[6];
  #pragma unroll
    for (int q = 0; q < 6; ++q) {
      const R hn = ua[q] * nE[q] + ub[q] * nN[q];
      if constexpr (sizeof(R) == 4) {
        hcq[q] = hn * mcq[q];
        huq[q] = hn;
        htq[q] = (ua[q] * (-nN[q]) + ub[q] * nE[q]) * mcq[q];
      } else {
        hcq[q] = hn * mcq[q];
        hcq[q] = hcq[q] * dd[q];
        huq[q] = hn;
        huq[q] = huq[q] * dd[q];
        htq[q] = (ua[q] * tE[q] + ub[q] * tN[q]) * mcq[q];
        htq[q] = htq[q] * dd[q];
      }
    }
    // r[0][e] is half edge 2e (0: sn, 2: s0 k=2 ic=0, 4: s0 k=2 ic=1),
    // r[1][e] half edge 2e+1 (1: s0 k=0 ic=0, 3: s0 k=0 ic=1, 5: se)
    R vm[3] = {R(0), R(0), R(0)}, vf[3] = {R(0), R(0), R(0)}, vd[3] = {R(0), R(0), R(0)};
    if (in) {
      const R hs0 = habs(G, hs);
  #pragma unroll
      for (int e = 0; e < 3; ++e) {
        vm[e] = -hcq[2 * e] + hcq[2 * e + 1];
        const R v = huq[2 * e + 1] - huq[2 * e];
        const R mx = v > R(0) ? v : R(0), mn = v < R(0) ? v : R(0);
        vf[e] = mx * hs0 + mn * habs(G, hup[e]);
      }
      if (dcr) {
        // centered_divergence_curl (operators.F:2083-2186), k_vort_curl's PV
        auto PV = [&](R hv, R ev) {
          const R hh = habs(G, hv);
          const R wi = hh > R(0) ? R(1) / hh : R(0);
          return ev * wi;
        };
        const R pv0 = PV(hs, es);
  #pragma unroll
        for (int e = 0; e < 3; ++e) vd[e] = (-htq[2 * e] + htq[2 * e + 1]) * (PV(hup[e], eup[e]) + pv0) * R(0.5);
      }
    }
    if (ke) {
      // the kinetic energy of this (old) level's velocity, k_center_ke's
      // arithmetic (step.F:192-203, operators.F:1316-1393), from the velocity
      // this kernel reads anyway; the previous step left it out
      R kv = R(0);
      if (kreal) {
        auto KC = [&](R a, R b) { return R(0.5) * (a * a + b * b); };
        kv = (+sa[0] * KC(u00, u10) + sa[1] * KC(u01, u11) + sa[2] * KC(ue0, ue1) + sa[3] * KC(une[1][0], une[1][1]) +
              sa[4] * KC(une[0][0], une[0][1]) + sa[5] * KC(un0, un1)) * ai;
      }
      put1(G.pf, wk, ke, s, kv);
    }
    put_all<R, 3>(G.pe, w, mshf, ns, s, vm[0], vm[1], vm[2]);
    put_all<R, 3>(G.pe, w, flx1, ns, s, vf[0], vf[1], vf[2]);
    if (dcr) put_all<R, 3>(G.pe, w, dcurl, ns, s, vd[0], vd[1], vd[2]);
  }
}

// update of a cell field by the divergence of edge fluxes, pole forms
// included (mass_fct.F:350-445, :478-532; tracer_fct.F:251-321, :601-665)
template <class R>
__device__ __forceinline__ R edge_update(int iim, int jjm, int ns, const IJK& c, R base, R dt_ai,
                                         const R* __restrict__ F, bool& set) {
  const int i = c.i, j = c.j, k = c.k;
  set = true;
  if (i == 1 && j == jjm - 1)
    return base + dt_ai * (F[1 * ns + SL(i, j, k)] + F[2 * ns + SL(i, j, k)] + F[0 * ns + SL(i + 1, j, k)] +
                           F[1 * ns + SL(i + 1, j + 1, k)] + F[2 * ns + SL(i, j + 1, k)]);
  if (i == iim - 1 && j == 1)
    return base + dt_ai * (F[0 * ns + SL(i, j, k)] + F[1 * ns + SL(i, j, k)] + F[0 * ns + SL(i + 1, j, k)] +
                           F[1 * ns + SL(i + 1, j + 1, k)] + F[2 * ns + SL(i, j + 1, k)]);
  if (i >= 2 && i <= iim - 1 && j >= 2 && j <= jjm - 1)
    return base - dt_ai * ((F[0 * ns + SL(i, j, k)] + F[1 * ns + SL(i, j, k)] + F[2 * ns + SL(i, j, k)]) -
                           (F[0 * ns + SL(i + 1, j, k)] + F[1 * ns + SL(i + 1, j + 1, k)] + F[2 * ns + SL(i, j + 1, k)]));
  set = false;
  return R(0);
}

// antidiffusive flux (all slots), low-order mass and the limiter beta_down
// (mass_fct.F:205-214 and :342-402; the beta loop starts at i,j = 1 and so
// overwrites its own pole values, which is reproduced).  The low-order mass
// is stored like h (less h0 in the mixed path).
template <class R>
__global__ void k_mass_low_beta(GeoT<R> G, R dt, R w3, R w2, R w1, const R* __restrict__ flx1, const R* __restrict__ m3,
                                const R* __restrict__ m2, const R* __restrict__ m1, const R* __restrict__ h3,
                                R* __restrict__ anti, R* __restrict__ mss_low, R* __restrict__ beta) {
  at_member(G.mb, flx1, m3, m2, m1, h3, anti, mss_low, beta);
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  const int i = c.i, j = c.j, k = c.k;
  const PushW wb = push_load(G.pf, c, iim, jjm, s);
  // every value first (the loads overlap): the three edges at the slot and
  // the one edge each at (i+1, j), (i+1, j+1), (i, j+1), those clamped into
  // the array where they fall outside it (their values are then not used)
  const int nbs[3] = {min(SL(i + 1, j, k), ns - 1), min(SL(i + 1, j + 1, k), ns - 1), min(SL(i, j + 1, k), ns - 1)};
  R fo[3], ho3[3], ho2[3], ho1[3], fn[3], hn3[3], hn2[3], hn1[3];
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    const size_t o = (size_t)e * ns;
    fo[e] = ldg(flx1 + o, s); ho3[e] = ldg(m3 + o, s); ho2[e] = ldg(m2 + o, s); ho1[e] = ldg(m1 + o, s);
    fn[e] = ldg(flx1 + o, nbs[e]); hn3[e] = ldg(m3 + o, nbs[e]); hn2[e] = ldg(m2 + o, nbs[e]); hn1[e] = ldg(m1 + o, nbs[e]);
  }
  const R hs3 = ldg(h3, s), ai = ldg(G.area_inv, s), ar = ldg(G.area, s);
  auto AB = [&](R x3, R x2, R x1, R f) {
    R f2 = R(0);
    f2 = f2 + w3 * x3;
    f2 = f2 + w2 * x2;
    f2 = f2 + w1 * x1;
    return f2 - f;
  };
  R ao[3];
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    ao[e] = AB(ho3[e], ho2[e], ho1[e], fo[e]);
    stg(anti + (size_t)e * ns, s, ao[e]);
  }
  // edge_update's forms, on the loaded fluxes
  const R dt_ai = dt * ai;
  R ml;
  if (i == 1 && j == jjm - 1)
    ml = hs3 + dt_ai * (fo[1] + fo[2] + fn[0] + fn[1] + fn[2]);
  else if (i == iim - 1 && j == 1)
    ml = hs3 + dt_ai * (fo[0] + fo[1] + fn[0] + fn[1] + fn[2]);
  else if (i >= 2 && i <= iim - 1 && j >= 2 && j <= jjm - 1)
    ml = hs3 - dt_ai * ((fo[0] + fo[1] + fo[2]) - (fn[0] + fn[1] + fn[2]));
  else
    ml = R(0);
  stg(mss_low, s, ml);
  R bt = R(0);
  if (i <= iim - 1 && j <= jjm - 1) {
    const R a_out = (rmax(R(0), ao[0]) + rmax(R(0), ao[1]) + rmax(R(0), ao[2])) -
                    (rmin(R(0), AB(hn3[0], hn2[0], hn1[0], fn[0])) + rmin(R(0), AB(hn3[1], hn2[1], hn1[1], fn[1])) +
                     rmin(R(0), AB(hn3[2], hn2[2], hn1[2], fn[2])));
    bt = ar * (habs(G, ml) - R(1.0e-6)) / (dt * (a_out + R(1.0e-18)));
  }
  put1(G.pf, wb, beta, s, bt);
}

// mass_fct.F:410-470 set_corrected_mss_anti (in place; non-interior slots
// keep the uncorrected value until the edge halo fill)
template <class R>
__global__ void k_mass_anti(GeoT<R> G, const R* __restrict__ beta, R* __restrict__ anti) {
  at_member(G.mb, beta, anti);
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  const PushW w = push_load(G.pe, c, iim, jjm, s);
  if (!interior(c, iim, jjm)) {
    put_zeros<R, 3>(w, anti, ns, s);
    return;
  }
  const int i = c.i, j = c.j, k = c.k;
  const R b = rmin(R(1), beta[s]);
  const int up[3] = {SL(i - 1, j, k), SL(i - 1, j - 1, k), SL(i, j - 1, k)};
  R v[3];
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    const R a = anti[e * ns + s];
    v[e] = b * rmax(R(0), a) + rmin(R(1), beta[up[e]]) * rmin(R(0), a);
  }
  put_all<R, 3>(G.pe, w, anti, ns, s, v[0], v[1], v[2]);
}

// tracer FCT, first- and high-order edge fluxes (tracer_fct.F:166-243) of
// one slot from the total mass flux F of its edges; the high-order tracer is
// the AB3 extrapolation of step.F:657-660
template <class R>
__device__ __forceinline__ void tr_fluxes_at(int iim, int jjm, int ns, const IJK& c, int s, R w3, R w2, R w1,
                                             const R* __restrict__ t3, const R* __restrict__ t4,
                                             const R* __restrict__ told, const R (&F)[3], R* __restrict__ f1,
                                             R* __restrict__ f2) {
  if (!interior(c, iim, jjm)) {
    for (int e = 0; e < 3; ++e) { f1[e * ns + s] = R(0); f2[e * ns + s] = R(0); }
    return;
  }
  const int i = c.i, j = c.j, k = c.k;
  auto TH = [&](int sl) { return w3 * t3[sl] + w2 * t4[sl] + w1 * told[sl]; };
  const int up[3] = {SL(i - 1, j, k), SL(i - 1, j - 1, k), SL(i, j - 1, k)};
  const R ths = TH(s);
  for (int e = 0; e < 3; ++e) {
    f1[e * ns + s] = t3[s] * rmax(R(0), F[e]) + t3[up[e]] * rmin(R(0), F[e]);
    f2[e * ns + s] = R(0.5) * (ths + TH(up[e])) * F[e];
  }
}

// the first tracer's fluxes, fused into k_mass_final (t3 == nullptr: none)
template <class R>
struct TrFluxArgs {
  const R *t3, *t4, *told;
  R *f1, *f2;
  R w3, w2, w1;
};

// the vorticity edge flux, fused into k_mass_final (vf == nullptr: none)
template <class R>
struct VortFluxArgs {
  const R *h3, *eta3;
  R* vf;
};

// flux-corrected mass and the total mass flux (mass_fct.F:121-128, :478-532);
// with Vf also the vorticity edge flux of the new mass flux,
// centered_vorticity_flux (operators.F:2000-2074, k_vort_curl's arithmetic)
// With beta (not nullptr) the kernel also performs set_corrected_mss_anti
// (mass_fct.F:410-470, k_mass_anti's arithmetic) on the antidiffusive fluxes
// it reads, which are then the uncorrected ones of k_mass_low_beta: a
// slot's own edges with beta at the slot and its upstream neighbours, the
// forward edge e at n_e = (i+1, j), (i+1, j+1), (i, j+1) with beta at n_e
// and at the slot, whose upstream neighbour n_e's edge e reaches.  A forward
// edge in the halo holds a copy of an edge that the slot n_e's image owns
// (the set-up checks this on the fill tables), so the same form with the
// halo copy of beta gives the corrected copy; the real poles' own edges copy
// edges of the slots psrc, whose upstream neighbour is the pole.
template <class R>
__global__ void k_mass_final(GeoT<R> G, R dt, const R* __restrict__ flx1, const R* __restrict__ anti,
                             const R* __restrict__ mss_low, R* __restrict__ mflux, R* __restrict__ h4,
                             TrFluxArgs<R> T, VortFluxArgs<R> Vf, const R* __restrict__ beta) {
  at_member(G.mb, flx1, anti, mss_low, mflux, h4);
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  const int i = c.i, j = c.j, k = c.k;
  auto cl = [&](int sl) { return min(max(sl, 0), ns - 1); };
  // every value first, so that the loads overlap (indices clamped into the
  // array; values a form does not use are not used): the slot's three edges
  // of both fluxes, the one edge each of anti at (i+1, j), (i+1, j+1),
  // (i, j+1) (edge_update), and the first tracer at the slot and at the
  // three upstream neighbours (tr_fluxes_at)
  const int en[3] = {cl(SL(i + 1, j, k)), cl(SL(i + 1, j + 1, k)), cl(SL(i, j + 1, k))};
  const PushW wf = push_load(G.pf, c, iim, jjm, s);
  const bool vfl = Vf.vf != nullptr;
  const PushW we = (T.t3 != nullptr || vfl) ? push_load(G.pe, c, iim, jjm, s) : PushW{};
  R fl[3], an[6];
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    const size_t o = (size_t)e * ns;
    fl[e] = ldg(flx1 + o, s);
    an[e] = ldg(anti + o, s);
    an[3 + e] = ldg(anti + o, en[e]);
  }
  const R base = ldg(mss_low, s), ai = ldg(G.area_inv, s);
  const bool corr = beta != nullptr;
  const bool npole = i == 1 && j == jjm - 1 && k == 1, spole = i == iim - 1 && j == 1 && k == 10;
  R bs = R(0), bup[3] = {R(0), R(0), R(0)}, bfw[3] = {R(0), R(0), R(0)}, bp[2] = {R(0), R(0)};
  if (corr) {
    at_member(G.mb, beta);
    bs = rmin(R(1), ldg(beta, s));
#pragma unroll
    for (int e = 0; e < 3; ++e) bfw[e] = rmin(R(1), ldg(beta, en[e]));
    const int upc[3] = {cl(SL(i - 1, j, k)), cl(SL(i - 1, j - 1, k)), cl(SL(i, j - 1, k))};
#pragma unroll
    for (int e = 0; e < 3; ++e) bup[e] = rmin(R(1), ldg(beta, upc[e]));
    if (npole) { bp[0] = rmin(R(1), ldg(beta, G.psrc[0])); bp[1] = rmin(R(1), ldg(beta, G.psrc[1])); }
    if (spole) { bp[0] = rmin(R(1), ldg(beta, G.psrc[2])); bp[1] = rmin(R(1), ldg(beta, G.psrc[3])); }
  }
  const bool tr = T.t3 != nullptr;
  const R* t3 = T.t3;
  const R* t4 = T.t4;
  const R* told = T.told;
  R* f1 = T.f1;
  R* f2 = T.f2;
  if (tr) at_member(G.mb, t3, t4, told, f1, f2);
  const bool in = interior(c, iim, jjm);
  const int up[3] = {cl(SL(i - 1, j, k)), cl(SL(i - 1, j - 1, k)), cl(SL(i, j - 1, k))};
  R a3[4], a4[4], ao[4];
  if (tr) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int sl = q == 0 ? s : up[q - 1];
      a3[q] = ldg(t3, sl);
      a4[q] = ldg(t4, sl);
      ao[q] = ldg(told, sl);
    }
  }
  // the potential vorticity eta / h of the old level at the slot and its
  // three upstream neighbours (k_vort_curl's PV)
  R pv[4];
  if (vfl) {
    const R* h3 = Vf.h3;
    const R* eta3 = Vf.eta3;
    at_member(G.mb, h3, eta3);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int sl = q == 0 ? s : up[q - 1];
      const R hh = habs(G, ldg(h3, sl));
      const R wi = hh > R(0) ? R(1) / hh : R(0);
      pv[q] = eabs(G, eta3, sl) * wi;
    }
  }
  if (corr) {
    // k_mass_anti's form: b_owner * max(0, a) + b_upstream * min(0, a)
    auto cor = [](R bo, R bu, R a) { return bo * rmax(R(0), a) + bu * rmin(R(0), a); };
    if (in) {
#pragma unroll
      for (int e = 0; e < 3; ++e) an[e] = cor(bs, bup[e], an[e]);
    } else if (npole) {
      an[1] = cor(bp[0], bs, an[1]);
      an[2] = cor(bp[1], bs, an[2]);
    } else if (spole) {
      an[0] = cor(bp[0], bs, an[0]);
      an[1] = cor(bp[1], bs, an[1]);
    }
#pragma unroll
    for (int e = 0; e < 3; ++e) an[3 + e] = cor(bfw[e], bs, an[3 + e]);
  }
  R F[3];
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    F[e] = fl[e] + an[e];
    stg(mflux + (size_t)e * ns, s, F[e]);
  }
  // edge_update's forms, on the loaded antidiffusive fluxes
  const R dt_ai = dt * ai;
  R hn;
  if (i == 1 && j == jjm - 1)
    hn = base + dt_ai * (an[1] + an[2] + an[3] + an[4] + an[5]);
  else if (i == iim - 1 && j == 1)
    hn = base + dt_ai * (an[0] + an[1] + an[3] + an[4] + an[5]);
  else if (i >= 2 && i <= iim - 1 && j >= 2 && j <= jjm - 1)
    hn = base - dt_ai * ((an[0] + an[1] + an[2]) - (an[3] + an[4] + an[5]));
  else
    hn = R(0);
  put1(G.pf, wf, h4, s, hn);
  // the first tracer's edge fluxes take the total flux of the slot's own
  // edges only, which this slot has just formed (k_tr_fluxes fused;
  // tr_fluxes_at's arithmetic)
  if (tr) {
    const R w3 = T.w3, w2 = T.w2, w1 = T.w1;
    R v1[3] = {R(0), R(0), R(0)}, v2[3] = {R(0), R(0), R(0)};
    if (in) {
      auto TH = [&](int q) { return w3 * a3[q] + w2 * a4[q] + w1 * ao[q]; };
      const R ths = TH(0);
#pragma unroll
      for (int e = 0; e < 3; ++e) {
        v1[e] = a3[0] * rmax(R(0), F[e]) + a3[1 + e] * rmin(R(0), F[e]);
        v2[e] = R(0.5) * (ths + TH(1 + e)) * F[e];
      }
    }
    put_all<R, 3>(G.pe, we, f1, ns, s, v1[0], v1[1], v1[2]);
    put_all<R, 3>(G.pe, we, f2, ns, s, v2[0], v2[1], v2[2]);
  }
  if (vfl) {
    R* vf = Vf.vf;
    at_member(G.mb, vf);
    R v[3] = {R(0), R(0), R(0)};
    if (in) {
#pragma unroll
      for (int e = 0; e < 3; ++e) v[e] = R(0.5) * (F[e]) * (pv[0] + pv[1 + e]);
    }
    put_all<R, 3>(G.pe, we, vf, ns, s, v[0], v[1], v[2]);
  }
}

// tracer FCT fluxes of a tracer over all slots (k_mass_final forms the
// first tracer's itself); tracer_fct.F:166-243, the
// high-order tracer is the AB3 extrapolation of step.F:657-660
template <class R>
__global__ void k_tr_fluxes(GeoT<R> G, R w3, R w2, R w1, const R* __restrict__ t3, const R* __restrict__ t4,
                            const R* __restrict__ told, const R* __restrict__ mflux, R* __restrict__ f1,
                            R* __restrict__ f2) {
  at_member(G.mb, t3, t4, told, mflux, f1, f2);
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  const R F[3] = {mflux[s], mflux[ns + s], mflux[2 * ns + s]};
  tr_fluxes_at<R>(iim, jjm, ns, c, s, w3, w2, w1, t3, t4, told, F, f1, f2);
}

// tracer_fct.F:251-321 set_msstrc_low, and the low-order tracer ml / h4 (0
// where h4 <= 0), which the limiter takes at the cell and its neighbours
template <class R>
__global__ void k_tr_low(GeoT<R> G, R dt, const R* __restrict__ h3, const R* __restrict__ t3, const R* __restrict__ f1,
                         const R* __restrict__ h4, R* __restrict__ ml, R* __restrict__ tlow) {
  at_member(G.mb, h3, t3, f1, h4, ml, tlow);
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  bool set;
  const R m = edge_update<R>(iim, jjm, ns, c, habs(G, h3[s]) * t3[s], dt * G.area_inv[s], f1, set);
  ml[s] = m;
  const R hh = habs(G, h4[s]);
  tlow[s] = hh > R(0) ? m / hh : R(0);
}

// tracer_fct.F:329-531: local extrema over the cell and its neighbours of the
// tracer and its low-order update, and the limiter weights beta_dn / beta_up
template <class R>
__global__ void k_tr_limiter(GeoT<R> G, R dt, const R* __restrict__ t3, const R* __restrict__ ml,
                             const R* __restrict__ tlow, const R* __restrict__ mss_low, const R* __restrict__ f1,
                             const R* __restrict__ f2, R* __restrict__ anti, R* __restrict__ bdn, R* __restrict__ bup) {
  at_member(G.mb, t3, ml, tlow, mss_low, f1, f2, anti, bdn, bup);
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  const int i = c.i, j = c.j, k = c.k;
  auto A = [&](int e, int sl) { return f2[e * ns + sl] - f1[e * ns + sl]; };
  for (int e = 0; e < 3; ++e) anti[e * ns + s] = A(e, s);
  auto TL = [&](int sl) { return tlow[sl]; };
  const int p = iim - 1, q = jjm - 1;
  const bool np = (i == 1 && j == q), sp = (i == p && j == 1);
  const bool in = interior(c, iim, jjm);
  if (!(np || sp || in)) { bdn[s] = R(0); bup[s] = R(0); return; }
  R mn, mx, a_out, a_in;
  if (in) {
    const int nb[7] = {s, SL(i - 1, j, k), SL(i - 1, j - 1, k), SL(i, j - 1, k), SL(i + 1, j, k),
                       SL(i + 1, j + 1, k), SL(i, j + 1, k)};
    mn = t3[s]; mx = t3[s];
    const R tl0 = TL(s);
    mn = rmin(mn, tl0); mx = rmax(mx, tl0);
    for (int n = 1; n < 7; ++n) { mn = rmin(mn, t3[nb[n]]); mx = rmax(mx, t3[nb[n]]); }
    for (int n = 1; n < 7; ++n) { const R v = TL(nb[n]); mn = rmin(mn, v); mx = rmax(mx, v); }
    const R o1 = A(0, s), o2 = A(1, s), o3 = A(2, s);
    const R n1 = A(0, SL(i + 1, j, k)), n2 = A(1, SL(i + 1, j + 1, k)), n3 = A(2, SL(i, j + 1, k));
    a_out = (rmax(R(0), o1) + rmax(R(0), o2) + rmax(R(0), o3)) - (rmin(R(0), n1) + rmin(R(0), n2) + rmin(R(0), n3));
    a_in = -((rmin(R(0), o1) + rmin(R(0), o2) + rmin(R(0), o3)) - (rmax(R(0), n1) + rmax(R(0), n2) + rmax(R(0), n3)));
  } else {
    int nb[5];
    R ee[5];
    if (np) {
      const int t[5] = {SL(p, 1, k), SL(1, q - 1, k), SL(2, q, k), SL(2, q + 1, k), SL(1, q + 1, k)};
      for (int n = 0; n < 5; ++n) nb[n] = t[n];
      ee[0] = A(1, s); ee[1] = A(2, s); ee[2] = A(0, SL(i + 1, j, k)); ee[3] = A(1, SL(i + 1, j + 1, k));
      ee[4] = A(2, SL(i, j + 1, k));
    } else {
      const int t[5] = {SL(p - 1, 1, k), SL(1, q, k), SL(p + 1, 1, k), SL(p + 1, 2, k), SL(p, 2, k)};
      for (int n = 0; n < 5; ++n) nb[n] = t[n];
      ee[0] = A(0, s); ee[1] = A(1, s); ee[2] = A(0, SL(i + 1, j, k)); ee[3] = A(1, SL(i + 1, j + 1, k));
      ee[4] = A(2, SL(i, j + 1, k));
    }
    mn = t3[s]; mx = t3[s];
    const R tl0 = TL(s);
    mn = rmin(mn, tl0); mx = rmax(mx, tl0);
    for (int n = 0; n < 5; ++n) { mn = rmin(mn, t3[nb[n]]); mx = rmax(mx, t3[nb[n]]); }
    for (int n = 0; n < 5; ++n) { const R v = TL(nb[n]); mn = rmin(mn, v); mx = rmax(mx, v); }
    a_out = -(rmin(R(0), ee[0]) + rmin(R(0), ee[1]) + rmin(R(0), ee[2]) + rmin(R(0), ee[3]) + rmin(R(0), ee[4]));
    a_in = (rmax(R(0), ee[0]) + rmax(R(0), ee[1]) + rmax(R(0), ee[2]) + rmax(R(0), ee[3]) + rmax(R(0), ee[4]));
  }
  const R ar = G.area[s] / dt;
  const R mlow = habs(G, mss_low[s]);
  bdn[s] = ar * (ml[s] - mlow * mn) / (a_out + R(1.0e-18));
  bup[s] = ar * (mlow * mx - ml[s]) / (a_in + R(1.0e-18));
}

// tracer_fct.F:539-593 set_corrected_trc_anti (in place)
template <class R>
__global__ void k_tr_anti(GeoT<R> G, const R* __restrict__ bdn, const R* __restrict__ bup, R* __restrict__ anti) {
  at_member(G.mb, bdn, bup, anti);
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  if (!interior(c, iim, jjm)) return;
  const int i = c.i, j = c.j, k = c.k;
  const int up[3] = {SL(i - 1, j, k), SL(i - 1, j - 1, k), SL(i, j - 1, k)};
  for (int e = 0; e < 3; ++e) {
    const R a = anti[e * ns + s];
    anti[e * ns + s] = rmin(rmin(R(1), bdn[s]), bup[up[e]]) * rmax(R(0), a) +
                       rmin(rmin(R(1), bup[s]), bdn[up[e]]) * rmin(R(0), a);
  }
}

// tracer_fct.F:601-665 set_msstrc_fct and the division by the new mass; the
// previous tracer level is kept for the next AB3 extrapolation (step.F:661).
// With bdn (not nullptr) the kernel also performs set_corrected_trc_anti
// (tracer_fct.F:539-593, k_tr_anti's arithmetic) on the antidiffusive
// fluxes it reads (then k_tr_limiter's uncorrected ones), with the roles of
// owner and upstream cell as k_mass_final takes them for the mass.
template <class R>
__global__ void k_tr_final(GeoT<R> G, R dt, const R* __restrict__ ml, const R* __restrict__ anti,
                           const R* __restrict__ h4, R* __restrict__ t4, R* __restrict__ told,
                           const R* __restrict__ bdn, const R* __restrict__ bup) {
  at_member(G.mb, ml, anti, h4, t4, told);
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  if (bdn == nullptr) {
    bool set;
    const R mf = edge_update<R>(iim, jjm, ns, c, ml[s], dt * G.area_inv[s], anti, set);
    told[s] = t4[s];
    const R hh = habs(G, h4[s]);
    t4[s] = hh != R(0) ? mf / hh : R(0);
    return;
  }
  at_member(G.mb, bdn, bup);
  const int i = c.i, j = c.j, k = c.k;
  auto cl = [&](int sl) { return min(max(sl, 0), ns - 1); };
  const int en[3] = {cl(SL(i + 1, j, k)), cl(SL(i + 1, j + 1, k)), cl(SL(i, j + 1, k))};
  const int upc[3] = {cl(SL(i - 1, j, k)), cl(SL(i - 1, j - 1, k)), cl(SL(i, j - 1, k))};
  const bool in = interior(c, iim, jjm);
  const bool np = i == 1 && j == jjm - 1, sp = i == iim - 1 && j == 1;
  const bool npole = np && k == 1, spole = sp && k == 10;
  // every value first: the slot's three edges and edge e at en[e], beta
  // down / up at the slot, its upstream and forward neighbours (and the
  // real poles' source slots)
  R a[6];
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    a[e] = ldg(anti + (size_t)e * ns, s);
    a[3 + e] = ldg(anti + (size_t)e * ns, en[e]);
  }
  const R ds = ldg(bdn, s), us = ldg(bup, s);
  R dup[3], uup[3], dfw[3], ufw[3];
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    dup[e] = ldg(bdn, upc[e]); uup[e] = ldg(bup, upc[e]);
    dfw[e] = ldg(bdn, en[e]); ufw[e] = ldg(bup, en[e]);
  }
  R dps[2] = {R(0), R(0)}, ups[2] = {R(0), R(0)};
  if (npole || spole) {
    const int o = npole ? 0 : 2;
#pragma unroll
    for (int q = 0; q < 2; ++q) { dps[q] = ldg(bdn, G.psrc[o + q]); ups[q] = ldg(bup, G.psrc[o + q]); }
  }
  const R base = ldg(ml, s), dt_ai = dt * ldg(G.area_inv, s);
  const R t4o = ldg(t4, s), hh = habs(G, ldg(h4, s));
  // k_tr_anti's form: owner (down, up), upstream (down, up)
  auto cor = [](R dn_o, R up_o, R dn_u, R up_u, R x) {
    return rmin(rmin(R(1), dn_o), up_u) * rmax(R(0), x) + rmin(rmin(R(1), up_o), dn_u) * rmin(R(0), x);
  };
  if (in) {
#pragma unroll
    for (int e = 0; e < 3; ++e) a[e] = cor(ds, us, dup[e], uup[e], a[e]);
  } else if (npole) {
    a[1] = cor(dps[0], ups[0], ds, us, a[1]);
    a[2] = cor(dps[1], ups[1], ds, us, a[2]);
  } else if (spole) {
    a[0] = cor(dps[0], ups[0], ds, us, a[0]);
    a[1] = cor(dps[1], ups[1], ds, us, a[1]);
  }
#pragma unroll
  for (int e = 0; e < 3; ++e) a[3 + e] = cor(dfw[e], ufw[e], ds, us, a[3 + e]);
  // edge_update's forms (tracer_fct.F:601-665)
  R mf;
  if (np) mf = base + dt_ai * (a[1] + a[2] + a[3] + a[4] + a[5]);
  else if (sp) mf = base + dt_ai * (a[0] + a[1] + a[3] + a[4] + a[5]);
  else if (in) mf = base - dt_ai * ((a[0] + a[1] + a[2]) - (a[3] + a[4] + a[5]));
  else mf = R(0);
  stg(told, s, t4o);
  stg(t4, s, hh != R(0) ? mf / hh : R(0));
}

// centred (operators.F:1923-1991) or upstream (:1841-1914) tracer edge flux
// of the mass flux, interior slots (0 elsewhere)
template <class R, bool CENTERED>
__global__ void k_tr_flux(GeoT<R> G, const R* __restrict__ mflux, const R* __restrict__ t, R* __restrict__ out) {
  at_member(G.mb, mflux, t, out);
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  if (!interior(c, iim, jjm)) {
    for (int e = 0; e < 3; ++e) out[e * ns + s] = R(0);
    return;
  }
  const int i = c.i, j = c.j, k = c.k;
  const int up[3] = {SL(i - 1, j, k), SL(i - 1, j - 1, k), SL(i, j - 1, k)};
  for (int e = 0; e < 3; ++e) {
    const R F = mflux[e * ns + s];
    if (CENTERED) {
      out[e * ns + s] = R(0.5) * (F) * (t[s] + t[up[e]]);
    } else {
      const R mx = F > R(0) ? F : R(0), mn = F < R(0) ? F : R(0);   // maxval / minval of (F, 0)
      out[e * ns + s] = mx * t[s] + mn * t[up[e]];
    }
  }
}

// step.F:221-270 step_surface_height: ABM3 on minus the divergence of the
// mass flux, then h >= 1e-10 on every slot (stored less h0 in the mixed path)
template <class R>
__global__ void k_height_update(GeoT<R> G, R dt, R w3, R w2, R w1, const R* __restrict__ mflux,
                                const R* __restrict__ h3, R* __restrict__ hf_a, const R* __restrict__ hf_b,
                                const R* __restrict__ hf_c, R* __restrict__ h4) {
  at_member(G.mb, mflux, h3, hf_a, hf_b, hf_c, h4);
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  const R w = div_edge_at<R>(iim, jjm, ns, c, mflux, G.area_inv[s]);
  const R hf = -w;
  hf_a[s] = hf;
  R h = h3[s] + dt * (w3 * hf + w2 * hf_b[s] + w1 * hf_c[s]);
  if (habs(G, h) < R(1.0e-10)) h = R(1.0e-10) - habs(G, R(0));
  h4[s] = h;
}

// step.F:278-327 step_tracer: ABM3 on mass times tracer, divided by the new
// thickness where it exceeds 1e-10
template <class R>
__global__ void k_tracer_update(GeoT<R> G, R dt, R w3, R w2, R w1, const R* __restrict__ tflux,
                                const R* __restrict__ t3, const R* __restrict__ h3, const R* __restrict__ h4,
                                R* __restrict__ tf_a, const R* __restrict__ tf_b, const R* __restrict__ tf_c,
                                R* __restrict__ t4) {
  at_member(G.mb, tflux, t3, h3, h4, tf_a, tf_b, tf_c, t4);
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  const R w = div_edge_at<R>(iim, jjm, ns, c, tflux, G.area_inv[s]);
  const R tf = -w;
  tf_a[s] = tf;
  R t = t3[s] * habs(G, h3[s]) + dt * (w3 * tf + w2 * tf_b[s] + w1 * tf_c[s]);
  const R hn = habs(G, h4[s]);
  if (hn > R(1.0e-10)) t = t / hn;
  t4[s] = t;
}

// operators.F:2000-2186 centered_vorticity_flux and centered_divergence_curl
template <class R>
__global__ void k_vort_curl(GeoT<R> G, const R* __restrict__ mflux, const R* __restrict__ h3,
                            const R* __restrict__ eta3, const R* __restrict__ hc, const R* __restrict__ u,
                            R* __restrict__ vf, R* __restrict__ dcurl) {
  at_member(G.mb, mflux, h3, eta3, hc, u, vf, dcurl);
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  const PushW w = push_load(G.pe, c, iim, jjm, s);
  if (!interior(c, iim, jjm)) {
    put_all<R, 3>(G.pe, w, vf, ns, s, R(0), R(0), R(0));
    put_all<R, 3>(G.pe, w, dcurl, ns, s, R(0), R(0), R(0));
    return;
  }
  const int i = c.i, j = c.j, k = c.k;
  auto PV = [&](int sl) {
    const R hh = habs(G, h3[sl]);
    const R w = hh > R(0) ? R(1) / hh : R(0);
    return eabs(G, eta3, sl) * w;
  };
  const int up[3] = {SL(i - 1, j, k), SL(i - 1, j - 1, k), SL(i, j - 1, k)};
  const R pv0 = PV(s);
  Half<R> ht;
  if constexpr (sizeof(R) == 4) ht = half_edges_pre<R>(iim, jjm, ns, i, j, k, G.tdc, hc, u);
  else ht = half_edges<R>(iim, jjm, ns, i, j, k, G.dc, G.tng, hc, u);
  R v1[3], v2[3];
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    const R pvu = PV(up[e]);
    v1[e] = R(0.5) * (mflux[e * ns + s]) * (pv0 + pvu);
    v2[e] = (-ht.r[0][e] + ht.r[1][e]) * (pvu + pv0) * R(0.5);
  }
  put_all<R, 3>(G.pe, w, vf, ns, s, v1[0], v1[1], v1[2]);
  put_all<R, 3>(G.pe, w, dcurl, ns, s, v2[0], v2[1], v2[2]);
}

// operators.F:270-392 div_of_an_edge_flux at one slot (0 where not defined)
template <class R>
__device__ __forceinline__ R div_edge_at(int iim, int jjm, int ns, const IJK& c, const R* __restrict__ F, R ai) {
  const int i = c.i, j = c.j, k = c.k;
  if (i == 1 && j == jjm - 1)
    return (-F[1 * ns + SL(i, j, k)] - F[2 * ns + SL(i, j, k)] - F[0 * ns + SL(i + 1, j, k)] -
            F[1 * ns + SL(i + 1, j + 1, k)] - F[2 * ns + SL(i, j + 1, k)]) * ai;
  if (i == iim - 1 && j == 1)
    return (-F[0 * ns + SL(i, j, k)] - F[1 * ns + SL(i, j, k)] - F[0 * ns + SL(i + 1, j, k)] -
            F[1 * ns + SL(i + 1, j + 1, k)] - F[2 * ns + SL(i, j + 1, k)]) * ai;
  if (i >= 2 && i <= iim - 1 && j >= 2 && j <= jjm - 1)
    return (F[0 * ns + SL(i, j, k)] + F[1 * ns + SL(i, j, k)] + F[2 * ns + SL(i, j, k)] - F[0 * ns + SL(i + 1, j, k)] -
            F[1 * ns + SL(i + 1, j + 1, k)] - F[2 * ns + SL(i, j + 1, k)]) * ai;
  return R(0);
}

// operators.F:402-518 laplacian_of_a_scalar at one slot; X(sl) gives the field
template <class R, class FX>
__device__ __forceinline__ R lap_at(int iim, int jjm, int ns, const IJK& c, const R* __restrict__ lw, FX X) {
  const int i = c.i, j = c.j, k = c.k;
  const int s = SL(i, j, k);
#define LW(n) lw[((n) - 1) * ns + s]
  if (i == 1 && j == jjm - 1)
    return +LW(7) * X(s) + LW(2) * X(SL(iim, 1, k)) + LW(3) * X(SL(i, j - 1, k)) + LW(4) * X(SL(i + 1, j, k)) +
           LW(5) * X(SL(i + 1, j + 1, k)) + LW(6) * X(SL(i, j + 1, k));
  if (i == iim - 1 && j == 1)
    return +LW(7) * X(s) + LW(1) * X(SL(i - 1, j, k)) + LW(2) * X(SL(1, jjm, k)) + LW(4) * X(SL(i + 1, j, k)) +
           LW(5) * X(SL(i + 1, j + 1, k)) + LW(6) * X(SL(i, j + 1, k));
  if (i >= 2 && i <= iim - 1 && j >= 2 && j <= jjm - 1)
    return +LW(7) * X(s) + LW(1) * X(SL(i - 1, j, k)) + LW(2) * X(SL(i - 1, j - 1, k)) + LW(3) * X(SL(i, j - 1, k)) +
           LW(4) * X(SL(i + 1, j, k)) + LW(5) * X(SL(i + 1, j + 1, k)) + LW(6) * X(SL(i, j + 1, k));
#undef LW
  return R(0);
}

// lap_at's stencil at one slot with the weights loaded once for several
// fields: 7 terms in lap_at's summation order (centre, then neighbours 1..6).
// A term a pole lacks gets weight 0 on the centre slot, which leaves every
// partial sum unchanged, so apply() returns lap_at's value to the bit; off
// the real cells the operator is 0 and nothing is loaded.
template <class R>
struct Stencil7 {
  bool real;
  int sl[7];
  R w[7];
  template <class FX>
  __device__ __forceinline__ R apply(FX X) const {
    if (!real) return R(0);
    R v = w[0] * X(sl[0]);
    v = v + w[1] * X(sl[1]);
    v = v + w[2] * X(sl[2]);
    v = v + w[3] * X(sl[3]);
    v = v + w[4] * X(sl[4]);
    v = v + w[5] * X(sl[5]);
    v = v + w[6] * X(sl[6]);
    return v;
  }
};

template <class R>
__device__ __forceinline__ Stencil7<R> stencil7(int iim, int jjm, int ns, const IJK& c, const R* __restrict__ lw) {
  const int i = c.i, j = c.j, k = c.k, s = SL(i, j, k);
  const bool in = i >= 2 && i <= iim - 1 && j >= 2 && j <= jjm - 1;
  const bool np = i == 1 && j == jjm - 1, sp = i == iim - 1 && j == 1;
  Stencil7<R> t;
  t.real = in || np || sp;
#pragma unroll
  for (int q = 0; q < 7; ++q) { t.sl[q] = s; t.w[q] = R(0); }
  if (!t.real) return t;
  t.w[0] = lw[6 * ns + s];
  if (!np) { t.sl[1] = SL(i - 1, j, k); t.w[1] = lw[s]; }
  t.sl[2] = np ? SL(iim, 1, k) : sp ? SL(1, jjm, k) : SL(i - 1, j - 1, k);
  t.w[2] = lw[ns + s];
  if (!sp) { t.sl[3] = SL(i, j - 1, k); t.w[3] = lw[2 * ns + s]; }
  t.sl[4] = SL(i + 1, j, k); t.w[4] = lw[3 * ns + s];
  t.sl[5] = SL(i + 1, j + 1, k); t.w[5] = lw[4 * ns + s];
  t.sl[6] = SL(i, j + 1, k); t.w[6] = lw[5 * ns + s];
  return t;
}

// The same Laplacian written on differences from the centre value, for the
// mixed path: the weights sum to zero, so sum_k w_k (x_k - x_0) is the same
// operator, and rounding the weights to FP32 then errs relative to the
// differences, not to the (much larger) field values.
template <class R, class FX>
__device__ __forceinline__ R lap_diff_at(int iim, int jjm, int ns, const IJK& c, const R* __restrict__ lw, FX X) {
  const int i = c.i, j = c.j, k = c.k;
  const int s = SL(i, j, k);
  const R x0 = X(s);
  auto D = [&](int sl) { return X(sl) - x0; };
#define LW(n) lw[((n) - 1) * ns + s]
  if (i == 1 && j == jjm - 1)
    return LW(2) * D(SL(iim, 1, k)) + LW(3) * D(SL(i, j - 1, k)) + LW(4) * D(SL(i + 1, j, k)) +
           LW(5) * D(SL(i + 1, j + 1, k)) + LW(6) * D(SL(i, j + 1, k));
  if (i == iim - 1 && j == 1)
    return LW(1) * D(SL(i - 1, j, k)) + LW(2) * D(SL(1, jjm, k)) + LW(4) * D(SL(i + 1, j, k)) +
           LW(5) * D(SL(i + 1, j + 1, k)) + LW(6) * D(SL(i, j + 1, k));
  if (i >= 2 && i <= iim - 1 && j >= 2 && j <= jjm - 1)
    return LW(1) * D(SL(i - 1, j, k)) + LW(2) * D(SL(i - 1, j - 1, k)) + LW(3) * D(SL(i, j - 1, k)) +
           LW(4) * D(SL(i + 1, j, k)) + LW(5) * D(SL(i + 1, j + 1, k)) + LW(6) * D(SL(i, j + 1, k));
#undef LW
  return R(0);
}

// step.F:336-429 step_vorticity / step_divergence (ABM3) and the elliptic
// right-hand sides of step.F:170-173.  Mixed path: eta is stored as the
// relative vorticity (eta - f), which is then also the right-hand side, and
// the Laplacian of ke + g (h + hs) is formed on differences (h stored less
// h0, a constant the Laplacian does not see).  One slot; k_tendencies runs
// it over the grid.
template <class R>
__device__ __forceinline__ void tendencies_at(const GeoT<R>& G, int s, R dt, R grav, R w3, R w2, R w1,
                                              const R* __restrict__ vf, const R* __restrict__ dcurl,
                                              const R* __restrict__ ke, const R* __restrict__ h3,
                                              const R* __restrict__ eta3, const R* __restrict__ div3,
                                              R* __restrict__ etaf_a, const R* __restrict__ etaf_b,
                                              const R* __restrict__ etaf_c, R* __restrict__ divf_a,
                                              const R* __restrict__ divf_b, const R* __restrict__ divf_c,
                                              R* __restrict__ eta4, R* __restrict__ div4, R* __restrict__ relative,
                                              R& rhs1, R& rhs2) {
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const IJK c = ijk_of(s, iim, jjm);
  const PushW w = push_load(G.pf, c, iim, jjm, s);
  const R ai = G.area_inv[s];
  const R wv = div_edge_at<R>(iim, jjm, ns, c, vf, ai);
  const R ef = -wv;
  etaf_a[s] = ef;
  const R e4 = eta3[s] + dt * (w3 * ef + w2 * etaf_b[s] + w1 * etaf_c[s]);
  const R wc = div_edge_at<R>(iim, jjm, ns, c, dcurl, ai);
  const R* hs = G.hs;
  auto X = [&](int sl) { return ke[sl] + grav * (h3[sl] + hs[sl]); };
  R wl;
  if constexpr (sizeof(R) == 4) wl = lap_diff_at<R>(iim, jjm, ns, c, G.lw, X);
  else wl = lap_at<R>(iim, jjm, ns, c, G.lw, X);
  const R df = -wl + wc;
  divf_a[s] = df;
  const R d4 = div3[s] + dt * (w3 * df + w2 * divf_b[s] + w1 * divf_c[s]);
  eta4[s] = e4;
  R rel;
  if constexpr (sizeof(R) == 4) rel = e4;
  else rel = e4 - G.f[s];
  // the new divergence and relative vorticity with their halo copies (the
  // new vorticity is formed again after the solve, which stores its copies)
  put1(G.pf, w, div4, s, d4);
  put1(G.pf, w, relative, s, rel);
  rhs1 = rel;
  rhs2 = d4;
}

// ------------------------------------------------------- reductions -------
// per-block partial sums of x*w (w may be null) for nf fields (blockIdx.y),
// accumulated in FP64
template <class R, class W = R>
__global__ void k_sum_partial(const R* __restrict__ x0, const R* __restrict__ x1, const W* __restrict__ w, int n,
                              double* __restrict__ part, size_t mb) {
  at_member(mb, x0, x1, part);
  __shared__ double sh[256];
  const R* x = blockIdx.y == 0 ? x0 : x1;
  double acc = 0.0;
  for (int t = blockIdx.x * bdim() + threadIdx.x; t < n; t += gridDim.x * bdim())
    acc += w ? (double)x[t] * (double)w[t] : (double)x[t];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int o = bdim() / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.y * gridDim.x + blockIdx.x] = sh[0];
}
// per-block partial sums of x^2 * w for field blockIdx.y (x0, x1), FP64
template <class R>
__global__ void k_sumsq_mixed(const R* __restrict__ x0, const R* __restrict__ x1, const double* __restrict__ w, int n,
                              double* __restrict__ part, size_t mb) {
  at_member(mb, x0, x1, part);
  __shared__ double sh[256];
  const R* x = blockIdx.y == 0 ? x0 : x1;
  double acc = 0.0;
  for (int t = blockIdx.x * bdim() + threadIdx.x; t < n; t += gridDim.x * bdim()) {
    const double v = (double)x[t];
    acc += v * v * w[t];
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int o = bdim() / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.y * gridDim.x + blockIdx.x] = sh[0];
}

// compensated sum of the partials of field blockIdx.x (256 threads, fixed
// order: strided Neumaier sums per thread, then a pairwise tree of
// (sum, error) pairs combined with Knuth's two-sum); deterministic
__device__ __forceinline__ void two_sum_add(double& s, double& c, double v) {
  const double t = s + v;
  const double bp = t - s;
  c += (s - (t - bp)) + (v - bp);
  s = t;
}
constexpr int SUM_THREADS = 1024;
// (blocks nf1.. sum the columns of a second set part2 / nb2 into out2)
__global__ void __launch_bounds__(SUM_THREADS) k_sum_final(const double* __restrict__ part, int nb,
                                                           double* __restrict__ out, size_t mb, int nf1,
                                                           const double* __restrict__ part2, int nb2,
                                                           double* __restrict__ out2) {
  int f = blockIdx.x;
  if (part2 && f >= nf1) {
    part = part2;
    nb = nb2;
    out = out2;
    f -= nf1;
  }
  at_member(mb, part, out);
  __shared__ double ss[SUM_THREADS], cc[SUM_THREADS];
  const int t = threadIdx.x;
  double s = 0.0, c = 0.0;
  // the partials in batches of eight independent loads, added in the same
  // order as one at a time
  const double* __restrict__ pf = part + (size_t)f * nb;
  int b = t;
  for (; b + 7 * SUM_THREADS < nb; b += 8 * SUM_THREADS) {
    double v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = pf[b + q * SUM_THREADS];
#pragma unroll
    for (int q = 0; q < 8; ++q) two_sum_add(s, c, v[q]);
  }
  for (; b < nb; b += SUM_THREADS) two_sum_add(s, c, pf[b]);
  ss[t] = s;
  cc[t] = c;
  __syncthreads();
  for (int o = SUM_THREADS / 2; o > 0; o >>= 1) {
    if (t < o) {
      double s1 = ss[t], c1 = cc[t] + cc[t + o];
      two_sum_add(s1, c1, ss[t + o]);
      ss[t] = s1;
      cc[t] = c1;
    }
    __syncthreads();
  }
  if (t == 0) out[f] = ss[0] + cc[0];
}

// rhs -= mean (elliptic_solver.F:80-92); sums[v] / total gives the mean
template <class R>
__global__ void k_sub_mean(int n, const double* __restrict__ sums, double total, R* __restrict__ r1,
                           R* __restrict__ r2, size_t mb) {
  at_member(mb, sums, r1, r2);
  const int t = blockIdx.x * bdim() + threadIdx.x;
  if (t >= n) return;
  const double m1 = sums[0] / total, m2 = sums[1] / total;
  r1[t] = r1[t] - (R)m1;
  r2[t] = r2[t] - (R)m2;
}

// block-wide sums of NV per-thread values, written to part[v * gridDim.x + block]:
// a butterfly over the 64 lanes of each wave, then the waves' sums (at most
// four, blocks of 64..256 threads) in wave order; one barrier
template <int NV>
__device__ __forceinline__ void block_sums(const double (&acc)[NV], double* __restrict__ part) {
  __shared__ double sh[NV][4];
  double a[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) a[v] = acc[v];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
    for (int v = 0; v < NV; ++v) a[v] = a[v] + __shfl_xor(a[v], o, 64);
  }
  const int w = (int)threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int v = 0; v < NV; ++v) sh[v][w] = a[v];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = bdim() >> 6;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      double s = sh[v][0];
      for (int q = 1; q < nw; ++q) s = s + sh[v][q];
      part[v * gridDim.x + blockIdx.x] = s;
    }
  }
}

// k_tendencies over all slots; SUMS (mixed path): also the per-block sums of
// am * rhs of both fields, which the solve's mean removal takes
template <class R, bool SUMS = false>
__global__ void k_tendencies(GeoT<R> G, R dt, R grav, R w3, R w2, R w1, const R* __restrict__ vf,
                             const R* __restrict__ dcurl, const R* __restrict__ ke, const R* __restrict__ h3,
                             const R* __restrict__ eta3, const R* __restrict__ div3, R* __restrict__ etaf_a,
                             const R* __restrict__ etaf_b, const R* __restrict__ etaf_c, R* __restrict__ divf_a,
                             const R* __restrict__ divf_b, const R* __restrict__ divf_c, R* __restrict__ eta4,
                             R* __restrict__ div4, R* __restrict__ relative, R* __restrict__ rhs1,
                             R* __restrict__ rhs2, const double* __restrict__ am, double* __restrict__ part) {
  at_member(G.mb, vf, dcurl, ke, h3, eta3, div3, etaf_a, etaf_b, etaf_c, divf_a, divf_b, divf_c, eta4, div4, relative, rhs1, rhs2);
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if constexpr (SUMS) {
    // with the per-block sums of am * rhs (the solve's mean removal) every
    // thread reaches the block reduction
    at_member(G.mb, part);
    double acc[2] = {0.0, 0.0};
    if (s < ns) {
      R r1, r2;
      tendencies_at<R>(G, s, dt, grav, w3, w2, w1, vf, dcurl, ke, h3, eta3, div3, etaf_a, etaf_b, etaf_c, divf_a,
                       divf_b, divf_c, eta4, div4, relative, r1, r2);
      rhs1[s] = r1;
      rhs2[s] = r2;
      acc[0] = (double)r1 * am[s];
      acc[1] = (double)r2 * am[s];
    }
    block_sums<2>(acc, part);
    return;
  }
  if (s >= ns) return;
  R r1, r2;
  tendencies_at<R>(G, s, dt, grav, w3, w2, w1, vf, dcurl, ke, h3, eta3, div3, etaf_a, etaf_b, etaf_c, divf_a, divf_b,
                   divf_c, eta4, div4, relative, r1, r2);
  rhs1[s] = r1;
  rhs2[s] = r2;
}

// FP64 path: squared Poisson residual at real cells of NV fields (planes ns
// apart), one slot per thread and the weights read once, per-block sums
// (elliptic_solver.F:203-239 check_convergence, :341-358 l2_norm)
template <int NV>
__global__ void __launch_bounds__(256) k_resid2(Geo G, const double* __restrict__ x, const double* __restrict__ rhs,
                                                double* __restrict__ part) {
  at_member(G.mb, x, rhs, part);
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  double acc[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) acc[v] = 0.0;
  if (s < ns) {
    const IJK c = ijk_of(s, iim, jjm);
    const double m = G.mask[s];
    const Stencil7<double> st = stencil7<double>(iim, jjm, ns, c, G.lw);
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const double* xv = x + (size_t)v * ns;
      auto X = [&](int sl) { return xv[sl]; };
      const double w = st.apply(X);
      const double r = (w - 0.0 * xv[s]) - rhs[(size_t)v * ns + s];
      acc[v] = (r * r) * m;
    }
  }
  block_sums<NV>(acc, part);
}

// the mixed solve's FP32 right-hand side of field v less its area-weighted
// mean (elliptic_solver.F:80-92), bsum[v] / total, as k_sub_mean would store
// it; bsum == nullptr: b already has its mean removed
__device__ __forceinline__ float rhs_less_mean(float b, const double* __restrict__ bsum, int v, double total) {
  return bsum ? b - (float)(bsum[v] / total) : b;
}

// The mixed path's FP64 defects form the Laplacian from its symmetric edge
// coefficients: area * L is symmetric, so the coefficient of the edge from s
// to a forward neighbour is the one that neighbour holds for the edge back,
// L(x)_s = area_inv(s) sum_e sym_e (x_e - x_s) over the six edges (3 held at
// s, 3 at the forward neighbours; the weights sum to zero, so the centre term
// is the differences').  Three coefficients and one masked area_inv per slot
// replace seven weights and the mask; the operator is the reference's to
// rounding (1e-14 of the terms).  The poles keep lap_at's form.
struct LapSym {
  bool real, pole;   // real: a cell of the mask (pole: the pole slot)
  int nb[6];
  double S[6], ai;
};
__device__ __forceinline__ LapSym lap_sym(const Geo& G, int iim, int jjm, int ns, const IJK& c) {
  LapSym t;
  const int i = c.i, j = c.j, k = c.k, s = SL(i, j, k);
  t.ai = ldg(G.aim, s);
  t.real = t.ai != 0.0;
  t.pole = (i == 1 && j == jjm - 1) || (i == iim - 1 && j == 1);
  if (!t.real) {
#pragma unroll
    for (int q = 0; q < 6; ++q) t.nb[q] = s;   // in-range indices; the result is 0 there
    return t;
  }
  t.nb[0] = SL(i - 1, j, k); t.nb[1] = SL(i - 1, j - 1, k); t.nb[2] = SL(i, j - 1, k);
  t.nb[3] = SL(i + 1, j, k); t.nb[4] = SL(i + 1, j + 1, k); t.nb[5] = SL(i, j + 1, k);
  if (t.pole) {
    // the pole's own weights l_weights(1..6) in the same difference form
    // (they sum to minus l_weights(7)); the neighbour its form leaves out
    // gets weight 0 and the slot itself
    const bool np = i == 1;
    if (np) { t.nb[0] = s; t.nb[1] = SL(iim, 1, k); }
    else { t.nb[1] = SL(1, jjm, k); t.nb[2] = s; }
#pragma unroll
    for (int q = 0; q < 6; ++q) t.S[q] = ldg(G.lw + (size_t)q * ns, s);
    if (np) t.S[0] = 0.0;
    else t.S[2] = 0.0;
    t.ai = 1.0;
    return t;
  }
  const double *s0 = G.sym, *s1 = G.sym + ns, *s2 = G.sym + 2 * (size_t)ns;
  t.S[0] = ldg(s0, s); t.S[1] = ldg(s1, s); t.S[2] = ldg(s2, s);
  t.S[3] = ldg(s0, t.nb[3]); t.S[4] = ldg(s1, t.nb[4]); t.S[5] = ldg(s2, t.nb[5]);
  return t;
}
template <class FX>
__device__ __forceinline__ double lap_sym_apply(const LapSym& t, const Geo& G, int iim, int jjm, int ns, const IJK& c,
                                                FX X) {
  if (!t.real) return 0.0;
  const double x0 = X(SL(c.i, c.j, c.k));
  double v = t.S[0] * (X(t.nb[0]) - x0);
#pragma unroll
  for (int q = 1; q < 6; ++q) v = v + t.S[q] * (X(t.nb[q]) - x0);
  return v * t.ai;
}

// the weighted Jacobi sweep from zero, k_relax<R, NV, true>, at a slot it
// computes: 0.2 * 0 + 0.8 * (wt * b).  k_relax subtracts the neighbours'
// terms w * 0 from b first; that changes b at most in the sign of a zero,
// and 0.2 * 0 + 0.8 * (+-0) is +0 either way, so the two agree bit for bit.
template <class R>
__device__ __forceinline__ R zero_sweep(R wt, R b) {
  const R omega = R(0.8);
  const R tt = wt * b;
  return (R(1) - omega) * R(0) + omega * tt;
}

// the first smoothing sweep of the next correction cycle, from zero, on the
// defect d just formed: k_relax<float, NV, true> at the finest level gives
// zero_sweep at the real slots and 0 elsewhere
__device__ __forceinline__ float first_sweep(bool real, const float* __restrict__ lwt, int s, float d) {
  if (!real) return 0.0f;
  return zero_sweep<float>(ldg(lwt, s), d);
}

// Mixed path, FP64 defect of the Poisson problem for NV fields: d = rhs -
// L(x') at real cells (0 elsewhere), x' = x + e with CORR (the FP32
// correction e, halos filled; x' is stored to xn on every slot, halos
// included, so the iterate stays halo-consistent) or x' = x.  d is stored in
// FP32 as the right-hand side of the next correction cycle (r == nullptr:
// not stored, where no cycle is expected to follow); per-block sums of d^2
// for the convergence test (elliptic_solver.F:203-239).
//
// Two slots per thread, s even and s + 1 (one row: iim is even).  Each row
// of the stencil is read with 16-byte loads that carry the neighbours i - 1
// and i + 2 as well (rows j - 1 and j + 1 in two loads each), so a value is
// fetched through the vector L1 about twice instead of seven times: the
// seven separate loads per point limited the one-slot form to 3.6 TB/s.
// Every term is the one-slot form's, in lap_sym_apply's order, so the
// results are the same bit for bit.  A real pole is always the even slot
// (columns 1 and iim - 1 are odd): it takes lap_sym's pole form through the
// same terms, with its own weights, its far neighbour for (i-1, j-1), the
// neighbour the form leaves out read as the slot itself at weight 0, and
// area factor 1.
template <int NV, bool CORR>
__global__ void __launch_bounds__(256) k_defect(Geo G, const double* __restrict__ x, const float* __restrict__ e,
                                                double* __restrict__ xn, const float* __restrict__ b,
                                                float* __restrict__ r, double* __restrict__ part,
                                                const double* __restrict__ gauge, double inv_area,
                                                const double* __restrict__ bsum, double total,
                                                const float* __restrict__ lwt, float* __restrict__ e0) {
  at_member(G.mb, x, e, xn, b, r, part);
  if (gauge) at_member(G.mb, gauge);
  if (bsum) at_member(G.mb, bsum);
  if (e0) at_member(G.mb, e0);
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int t = xcd_slot();
  const bool act = 2 * t < ns;
  // threads past the end work on the last pair and store nothing
  const int s = act ? 2 * t : ns - 2;
  const int su = s >= iim ? s - iim : s;
  const int sd = s + iim < ns ? s + iim : s;
  const int u1 = su + 1 < ns ? su + 1 : su, d2 = sd + 2 < ns ? sd + 2 : sd, s2 = s + 2 < ns ? s + 2 : s;
  const double* s0p = G.sym;
  const double* s1p = G.sym + ns;
  const double* s2p = G.sym + 2 * (size_t)ns;
  const IJK cA = ijk_of(s, iim, jjm);
  const D2 ai = ld2(G.aim, s);
  const bool npA = cA.i == 1 && cA.j == jjm - 1, spA = cA.i == iim - 1 && cA.j == 1;
  const bool poleA = (npA || spA) && ai.x != 0.0;
  // the weights: slot A sym(0..2) at s, sym0 at s + 1, sym1 at sd + 1,
  // sym2 at sd; slot B the same one slot on
  D2 S0 = ld2(s0p, s), S1 = ld2(s1p, s), S2 = ld2(s2p, s);
  const double S0B3 = ldg(s0p, s2);
  D2 S1d = ld2c(s1p, sd + 1, ns);
  D2 S2d = ld2(s2p, sd);
  const double S0B0 = S0.y;
  double aiA = ai.x;
  int far = s;
  if (poleA) {
    S0.x = npA ? 0.0 : ldg(G.lw, s);
    S1.x = ldg(G.lw + ns, s);
    S2.x = spA ? 0.0 : ldg(G.lw + 2 * (size_t)ns, s);
    S0.y = ldg(G.lw + 3 * (size_t)ns, s);
    S1d.x = ldg(G.lw + 4 * (size_t)ns, s);
    S2d.x = ldg(G.lw + 5 * (size_t)ns, s);
    aiA = 1.0;
    far = npA ? SL(iim, 1, cA.k) : SL(1, jjm, cA.k);
  }
  double wA[NV], wB[NV], xA[NV], xB[NV];
  F2 bq[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const double* xv = x + (size_t)v * ns;
    const float* ev = e + (size_t)v * ns;
    // row j: (i-1, i), (i+1, i+2); row j-1: (i-1, i), i+1; row j+1: (i, i+1), i+2
    const D2 r0 = ld2c(xv, s - 1, ns), r1 = ld2c(xv, s + 1, ns), u0 = ld2c(xv, su - 1, ns), d0 = ld2(xv, sd);
    const double ux = ldg(xv, u1), dx = ldg(xv, d2);
    F2 er0{0.f, 0.f}, er1{0.f, 0.f}, eu0{0.f, 0.f}, ed0{0.f, 0.f};
    float eux = 0.f, edx = 0.f;
    if (CORR) {
      er0 = ld2c(ev, s - 1, ns); er1 = ld2c(ev, s + 1, ns); eu0 = ld2c(ev, su - 1, ns); ed0 = ld2(ev, sd);
      eux = ldg(ev, u1); edx = ldg(ev, d2);
    }
    bq[v] = ld2(b + (size_t)v * ns, s);
    auto X = [&](double xq, float eq) { return CORR ? xq + (double)eq : xq; };
    const double c0 = X(r0.y, er0.y), c1 = X(r1.x, er1.x), p2 = X(r1.y, er1.y);
    const double uc0 = X(u0.y, eu0.y), up1 = X(ux, eux);
    const double dc0 = X(d0.x, ed0.x), dp1 = X(d0.y, ed0.y), dp2 = X(dx, edx);
    double m1 = X(r0.x, er0.x), um1 = X(u0.x, eu0.x), uA = uc0;
    if (poleA) {
      um1 = X(ldg(xv, far), CORR ? ldg(ev, far) : 0.0f);
      if (npA) m1 = c0;
      else uA = c0;
    }
    // lap_sym_apply's terms in its order: i-1, (i-1, j-1), j-1, i+1,
    // (i+1, j+1), j+1
    double w = S0.x * (m1 - c0);
    w = w + S1.x * (um1 - c0);
    w = w + S2.x * (uA - c0);
    w = w + S0.y * (c1 - c0);
    w = w + S1d.x * (dp1 - c0);
    w = w + S2d.x * (dc0 - c0);
    wA[v] = w * aiA;
    w = S0B0 * (c0 - c1);
    w = w + S1.y * (uc0 - c1);
    w = w + S2.y * (up1 - c1);
    w = w + S0B3 * (p2 - c1);
    w = w + S1d.y * (dp2 - c1);
    w = w + S2d.y * (dp1 - c1);
    wB[v] = w * ai.y;
    xA[v] = c0;
    xB[v] = c1;
  }
  // the first sweep's slots are k_relax's (interior and pole slots)
  const IJK cB{cA.i + 1, cA.j, cA.k};
  const bool rA = interior(cA, iim, jjm) || npA || spA;
  const bool rB = interior(cB, iim, jjm);
  const F2 wt = e0 ? ld2(lwt, s) : F2{0.0f, 0.0f};
  double acc[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const double dA = ai.x != 0.0 ? (double)rhs_less_mean(bq[v].x, bsum, v, total) - wA[v] : 0.0;
    const double dB = ai.y != 0.0 ? (double)rhs_less_mean(bq[v].y, bsum, v, total) - wB[v] : 0.0;
    acc[v] = 0.0;
    if (act) {
      if (CORR) {
        const double gs = gauge ? gauge[v] * inv_area : 0.0;
        st2(xn + (size_t)v * ns, s, gauge ? xA[v] - gs : xA[v], gauge ? xB[v] - gs : xB[v]);
      }
      if (r) st2(r + (size_t)v * ns, s, (float)dA, (float)dB);
      if (e0)
        st2(e0 + (size_t)v * ns, s, rA ? zero_sweep<float>(wt.x, (float)dA) : 0.0f,
            rB ? zero_sweep<float>(wt.y, (float)dB) : 0.0f);
      acc[v] = (ai.x != 0.0 ? dA * dA : 0.0) + (ai.y != 0.0 ? dB * dB : 0.0);
    }
  }
  block_sums<NV>(acc, part);
}

// Mixed path, first FP64 defect of a solve from the linear start
// x0 = 2 x - xp (x the last solution, xp the one before; xp = x gives the
// plain warm start exactly): x0 is stored to xn on every slot and
// d = rhs - L(x0) as in k_defect.  The solution's constant part is free (L
// of a constant is zero), so extrapolating it would let it drift; the
// per-block sums of area * (x - xp) (columns NV.. of part) give the mean
// change, which the first correction takes off again (k_defect's gauge).
// With gauge (k_defect) the stored iterate loses gauge[v] * inv_area; the
// defect, of the unshifted iterate, is the same up to rounding.
template <int NV>
__global__ void __launch_bounds__(256) k_defect_extrap(Geo G, const double* __restrict__ am,
                                                       const double* __restrict__ x, const double* __restrict__ xp,
                                                       double* __restrict__ xn, const float* __restrict__ b,
                                                       float* __restrict__ r, double* __restrict__ part,
                                                       const double* __restrict__ bsum, double total,
                                                       const float* __restrict__ lwt, float* __restrict__ e0) {
  at_member(G.mb, x, xp, xn, b, r, part);
  if (bsum) at_member(G.mb, bsum);
  if (e0) at_member(G.mb, e0);
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  double acc[2 * NV];
#pragma unroll
  for (int v = 0; v < 2 * NV; ++v) acc[v] = 0.0;
  if (s < ns) {
    const IJK c = ijk_of(s, iim, jjm);
    const LapSym st = lap_sym(G, iim, jjm, ns, c);
    const double m = st.real ? 1.0 : 0.0;
    const bool rslot = interior(c, iim, jjm) || (c.i == 1 && c.j == jjm - 1) || (c.i == iim - 1 && c.j == 1);
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const double* xv = x + (size_t)v * ns;
      const double* pv = xp + (size_t)v * ns;
      auto X = [&](int sl) { return 2.0 * ldg(xv, sl) - ldg(pv, sl); };
      const double w = lap_sym_apply(st, G, iim, jjm, ns, c, X);
      const double d = m != 0.0 ? (double)rhs_less_mean(ldg(b + (size_t)v * ns, s), bsum, v, total) - w : 0.0;
      stg(xn + (size_t)v * ns, s, X(s));
      stg(r + (size_t)v * ns, s, (float)d);
      if (e0) stg(e0 + (size_t)v * ns, s, first_sweep(rslot, lwt, s, (float)d));
      acc[v] = (d * d) * m;
      acc[NV + v] = ldg(am, s) * (ldg(xv, s) - ldg(pv, s));
    }
  }
  block_sums<2 * NV>(acc, part);
}

// Mixed path, the quadratic start of a solve, pointwise: x0 = 3 (x - xp) +
// xpp (x the last solution, xp and xpp the two before) stored to xn on every
// slot, and the per-block sums of area * (x0 - x) of both fields in columns
// 2 and 3 of part (the mean change, which the first correction takes off);
// k_defect<2, false> then forms the defect of x0 with one read per stencil
// point instead of three.
__global__ void __launch_bounds__(256) k_extrap_start(int ns, const double* __restrict__ am,
                                                      const double* __restrict__ x, const double* __restrict__ xp,
                                                      const double* __restrict__ xpp, double* __restrict__ xn,
                                                      double* __restrict__ part, size_t mb) {
  at_member(mb, x, xp, xpp, xn, part);
  const int s = xcd_slot();
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  if (s < ns) {
    const double a = ldg(am, s);
#pragma unroll
    for (int v = 0; v < 2; ++v) {
      const size_t o = (size_t)v * ns;
      const double xv = ldg(x + o, s);
      const double x0 = 3.0 * (xv - ldg(xp + o, s)) + ldg(xpp + o, s);
      stg(xn + o, s, x0);
      acc[2 + v] = a * (x0 - xv);
    }
  }
  block_sums<4>(acc, part);
}

// the convergence test of elliptic_solver.F:203-239 (absolute form) on the
// device, for a velocity kernel queued before the host has read the norms:
// go = 1 when both fields' mean squared residuals, sums / n, meet tol (the
// host's test, same division and comparison, so the same decision; a NaN
// fails it), else 0.  sums: the two final sums (one member).
__global__ void k_conv_go(const double* __restrict__ sums, double n, double tol, double* __restrict__ go) {
  if (threadIdx.x == 0) go[0] = (sums[0] / n <= tol && sums[1] / n <= tol) ? 1.0 : 0.0;
}

// fixed-schedule solves: per member (blockIdx.z), the largest mean squared
// residual of the two fields so far (track[0], track[1]) and the number of
// solves that ended above tol (track[2]); a NaN counts as above
__global__ void k_track_resid(const double* __restrict__ sums, double inv_n, double tol, double* __restrict__ track,
                              size_t mb) {
  at_member(mb, sums, track);
  if (threadIdx.x != 0) return;
  const double a = sums[0] * inv_n, b = sums[1] * inv_n;
  if (!(a <= track[0])) track[0] = a;
  if (!(b <= track[1])) track[1] = b;
  if (!(a <= tol) || !(b <= tol)) track[2] = track[2] + 1.0;
}

// x += scale * (*g) on n values
__global__ void k_add_scaled(size_t n, const double* __restrict__ g, double scale, double* __restrict__ x,
                             size_t mb) {
  at_member(mb, g, x);
  const size_t t = (size_t)blockIdx.x * bdim() + threadIdx.x;
  if (t < n) x[t] = x[t] + scale * g[0];
}

// x += e (FP64 solution, FP32 correction), nvar planes of ns
__global__ void k_correct(size_t n, const float* __restrict__ e, double* __restrict__ x) {
  const size_t t = (size_t)blockIdx.x * bdim() + threadIdx.x;
  if (t < n) x[t] = x[t] + (double)e[t];
}

// ------------------------------------------------------------- multigrid --
// Halo reads go through the level's closed-form map (multigrid_solver.F:
// 1105-1158 wrap_simple): src[slot] is the slot whose value the halo holds.

// weighted Jacobi sweep (multigrid_solver.F:593-682) of NV fields (planes
// ns apart) with one thread per slot, so the weights are read once; halo
// slots become 0.  ZERO: the input is identically zero (the coarse levels'
// first sweep after multigrid_solver.F:520-524 clears them), so it is not
// read; the arithmetic is the same, so the result is the same to the bit.
// Slots at least two away from the panel edge have no halo neighbour and
// read their neighbours without the map.
template <class R, int NV, bool ZERO>
__global__ void __launch_bounds__(256) k_relax(LevT<R> L, const R* __restrict__ beta, const R* __restrict__ a,
                                               R* __restrict__ out) {
  at_member(L.mb, beta, a, out);
  const int iim = L.iim, jjm = L.jjm, ns = L.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  const int i = c.i, j = c.j, k = c.k;
  const bool in = i >= 2 && i <= iim - 1 && j >= 2 && j <= jjm - 1;
  const bool np = i == 1 && j == jjm - 1, sp = i == iim - 1 && j == 1;
  if (!(in || np || sp)) {
#pragma unroll
    for (int v = 0; v < NV; ++v) stg(out + (size_t)v * ns, s, R(0));
    return;
  }
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
  int me = s;   // the slot's own value is read through the map as well
  if (!ZERO && !deep) {
#pragma unroll
    for (int q = 0; q < 6; ++q) nb[q] = nb[q] < 0 ? -1 : ldg(m, nb[q]);
    me = ldg(m, s);
  }
  R w[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) w[q] = ldg(L.lw + (size_t)q * ns, s);
  const R wt = ldg(L.lwt, s);
  const R omega = R(0.8);
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const R* __restrict__ av = a + (size_t)v * ns;
    auto A = [&](int q) { return ZERO ? R(0) : ldg(av, nb[q]); };
    const R bv = ldg(beta + (size_t)v * ns, s);
    R tt;
    if (in)
      tt = wt * (bv - w[0] * A(0) - w[1] * A(1) - w[2] * A(2) - w[3] * A(3) - w[4] * A(4) - w[5] * A(5));
    else if (np)
      tt = wt * (bv - w[1] * A(1) - w[2] * A(2) - w[3] * A(3) - w[4] * A(4) - w[5] * A(5));
    else
      tt = wt * (bv - w[0] * A(0) - w[1] * A(1) - w[3] * A(3) - w[4] * A(4) - w[5] * A(5));
    const R self = ZERO ? R(0) : ldg(av, me);
    stg(out + (size_t)v * ns, s, (R(1) - omega) * self + omega * tt);
  }
}

// Mixed path, levels above the coarse three: a non-zero weighted Jacobi sweep of k_relax with
// the half-precision weights of LevT, in difference form (the weights sum to
// zero): out = x0 + hcw (S beta - sum_k hw_k (x_k - x0)), the same operator
// as k_relax's to the weights' rounding.  It only smooths inside the FP64
// defect correction, so that rounding sets no accuracy limit.
template <int NV>
__global__ void __launch_bounds__(256) k_relax_h(LevT<float> L, const float* __restrict__ beta,
                                                 const float* __restrict__ a, float* __restrict__ out) {
  at_member(L.mb, beta, a, out);
  const int iim = L.iim, jjm = L.jjm, ns = L.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  const int i = c.i, j = c.j, k = c.k;
  const bool in = i >= 2 && i <= iim - 1 && j >= 2 && j <= jjm - 1;
  const bool np = i == 1 && j == jjm - 1, sp = i == iim - 1 && j == 1;
  if (!(in || np || sp)) {
#pragma unroll
    for (int v = 0; v < NV; ++v) stg(out + (size_t)v * ns, s, 0.0f);
    return;
  }
  const bool deep = i >= 3 && i <= iim - 2 && j >= 3 && j <= jjm - 2;
  const int* __restrict__ m = L.src;
  int nb[6];
  if (in) {
    nb[0] = SL(i - 1, j, k); nb[1] = SL(i - 1, j - 1, k); nb[2] = SL(i, j - 1, k);
  } else if (np) {
    nb[0] = s; nb[1] = SL(iim, 1, k); nb[2] = SL(i, j - 1, k);
  } else {
    nb[0] = SL(i - 1, j, k); nb[1] = SL(1, jjm, k); nb[2] = s;
  }
  nb[3] = SL(i + 1, j, k); nb[4] = SL(i + 1, j + 1, k); nb[5] = SL(i, j + 1, k);
  int me = s;
  if (!deep) {
#pragma unroll
    for (int q = 0; q < 6; ++q) nb[q] = ldg(m, nb[q]);
    me = ldg(m, s);
  }
  float w[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) w[q] = (float)ldg(L.hw + (size_t)q * ns, s);
  // the term a pole lacks: weight 0 (its neighbour index is the slot itself)
  if (np) w[0] = 0.0f;
  if (sp) w[2] = 0.0f;
  const float cw = (float)ldg(L.hcw, s);
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const float* __restrict__ av = a + (size_t)v * ns;
    const float x0 = ldg(av, me);
    float acc = L.hs * ldg(beta + (size_t)v * ns, s);
#pragma unroll
    for (int q = 0; q < 6; ++q) acc -= w[q] * (ldg(av, nb[q]) - x0);
    stg(out + (size_t)v * ns, s, x0 + cw * acc);
  }
}

// With z0 (the coarse level's scratch) an injection also performs the coarse
// level's first smoothing sweep, from zero, which needs the right-hand side
// just injected at its own slot only: z0 = zero_sweep there, at the slots
// k_relax<R, NV, true> computes.  This covers those outside the injection:
// the pole form's slots of the panels that do not hold the pole, from the
// right-hand side in memory as k_relax reads it, and zero at the halo slots.
template <class R, int NV>
__device__ __forceinline__ void coarse_zero_sweep_rest(const LevT<R>& C, const IJK& c, int s1,
                                                       const R* __restrict__ x1, R* __restrict__ z0) {
  const bool pole_form = (c.i == 1 && c.j == C.jjm - 1) || (c.i == C.iim - 1 && c.j == 1);
#pragma unroll
  for (int v = 0; v < NV; ++v)
    z0[(size_t)v * C.ns + s1] = pole_form ? zero_sweep<R>(C.lwt[s1], x1[(size_t)v * C.ns + s1]) : R(0);
}

// the order of a level's real cells in the dense coarse solve (interior
// slots panel by panel, then the north and the south pole) and their count
__device__ __forceinline__ int dense_index(const IJK& c, int iim, int jjm, bool np, bool sp) {
  const int e = iim - 2;
  if (np) return 10 * e * e;
  if (sp) return 10 * e * e + 1;
  return (c.k - 1) * e * e + (c.j - 2) * e + (c.i - 2);
}
__device__ __forceinline__ int dense_count(int iim) { return 10 * (iim - 2) * (iim - 2) + 2; }

// Mixed path, levels above the coarse three: k_inject with the half-precision weights in
// difference form, r = beta - (1/S) sum_k hw_k (x_k - x0) at the coincident
// fine points.  xc (the level of the dense coarse solve): the right-hand
// side also in the dense solve's cell order, NV planes of dense_count
template <int NV>
__global__ void k_inject_h(LevT<float> F, LevT<float> C, const float* __restrict__ xa, const float* __restrict__ xb,
                           float* __restrict__ x1, float* __restrict__ z0, float* __restrict__ xc) {
  at_member(F.mb, xa, xb, x1);
  if (z0) at_member(F.mb, z0);
  if (xc) at_member(F.mb, xc);
  int iim = C.iim, jjm = C.jjm;
  const int s1 = xcd_slot();
  if (s1 >= C.ns) return;
  const IJK c = ijk_of(s1, iim, jjm);
  const bool in = c.i >= 2 && c.i <= iim - 1 && c.j >= 2 && c.j <= jjm - 1;
  const bool np = c.k == 1 && c.i == 1 && c.j == jjm - 1;
  const bool sp = c.k == 10 && c.i == iim - 1 && c.j == 1;
  if (!(in || np || sp)) {
    if (z0) coarse_zero_sweep_rest<float, NV>(C, c, s1, x1, z0);
    return;
  }
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
    for (int n = 0; n < 7; ++n) q[n] = ldg(m, q[n]);
  }
  float w[6];
#pragma unroll
  for (int n = 0; n < 6; ++n) w[n] = (float)ldg(F.hw + (size_t)n * ns, s);
  if (np) w[0] = 0.0f;
  if (sp) w[2] = 0.0f;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const float* __restrict__ a = xa + (size_t)v * ns;
    const float x0 = ldg(a, q[0]);
    float acc = 0.0f;
#pragma unroll
    for (int n = 0; n < 6; ++n) acc += w[n] * (ldg(a, q[n + 1]) - x0);
    const float r = ldg(xb + (size_t)v * ns, s) - F.his * acc;
    stg(x1 + (size_t)v * C.ns, s1, r);
    if (z0) stg(z0 + (size_t)v * C.ns, s1, zero_sweep<float>(ldg(C.lwt, s1), r));
    if (xc) stg(xc + (size_t)v * dense_count(C.iim), dense_index(c, C.iim, C.jjm, np, sp), r);
  }
}

// residual at the odd-odd fine points becomes the coarse right-hand side
// (multigrid_solver.F:684-886 inject_residual); one thread per coarse slot
// and NV fields (planes ns apart), so the weights are read once.  Fine
// neighbours of interior coarse points other than the outermost ring are
// interior themselves and are read without the halo map.
template <class R, int NV>
__global__ void k_inject(LevT<R> F, LevT<R> C, const R* __restrict__ xa, const R* __restrict__ xb, R* __restrict__ x1,
                         R* __restrict__ z0, R* __restrict__ xc) {
  at_member(F.mb, xa, xb, x1);
  if (z0) at_member(F.mb, z0);
  if (xc) at_member(F.mb, xc);
  int iim = C.iim, jjm = C.jjm;
  const int s1 = xcd_slot();
  if (s1 >= C.ns) return;
  const IJK c = ijk_of(s1, iim, jjm);
  const bool in = c.i >= 2 && c.i <= iim - 1 && c.j >= 2 && c.j <= jjm - 1;
  const bool np = c.k == 1 && c.i == 1 && c.j == jjm - 1;
  const bool sp = c.k == 10 && c.i == iim - 1 && c.j == 1;
  if (!(in || np || sp)) {
    if (z0) coarse_zero_sweep_rest<R, NV>(C, c, s1, x1, z0);
    return;
  }
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
  // the seven fine slots, mapped (pole forms leave out one neighbour)
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
    for (int n = 0; n < 7; ++n) q[n] = ldg(m, q[n]);
  }
  R w[7];
#pragma unroll
  for (int n = 0; n < 7; ++n) w[n] = ldg(F.lw + (size_t)n * ns, s);
  // w[n] is l_weights(n+1); q[0] is the centre, q[1..6] the neighbours 1..6
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const R* __restrict__ a = xa + (size_t)v * F.ns;
    const R bv = ldg(xb + (size_t)v * F.ns, s);
    auto X = [&](int n) { return ldg(a, q[n]); };
    R r;
    if (in)
      r = bv - ((w[6] - R(0)) * X(0) + w[0] * X(1) + w[1] * X(2) + w[2] * X(3) + w[3] * X(4) +
                w[4] * X(5) + w[5] * X(6));
    else if (np)
      r = bv - ((w[6] - R(0)) * X(0) + w[1] * X(2) + w[2] * X(3) + w[3] * X(4) + w[4] * X(5) +
                w[5] * X(6));
    else
      r = bv - ((w[6] - R(0)) * X(0) + w[0] * X(1) + w[1] * X(2) + w[3] * X(4) + w[4] * X(5) +
                w[5] * X(6));
    stg(x1 + (size_t)v * C.ns, s1, r);
    if (z0) stg(z0 + (size_t)v * C.ns, s1, zero_sweep<R>(ldg(C.lwt, s1), r));
    if (xc) stg(xc + (size_t)v * dense_count(C.iim), dense_index(c, C.iim, C.jjm, np, sp), r);
  }
}

// coarse correction added at coincident points, two-point averages between
// (multigrid_solver.F:888-1054 interpolate); one thread per fine slot and NV
// fields.  Coarse reads go through the map only on the coarse halo ring.
template <class R, int NV>
__global__ void k_interp(LevT<R> C, LevT<R> F, const R* __restrict__ x1, R* __restrict__ x2) {
  at_member(F.mb, x1, x2);
  const int iim = F.iim, jjm = F.jjm;
  const int s = xcd_slot();
  if (s >= F.ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  const int i = c.i, j = c.j, k = c.k;
  // temp(2a-1, 2b-1) = coarse (a, b) for a, b = 1..iim1-1 (halo reads mapped)
  const int ci = C.iim, cj = C.jjm;
  auto CS = [&](int ii, int jj) {
    const int a = (ii + 1) / 2, b = (jj + 1) / 2;
    const int sl = (a - 1) + ci * ((b - 1) + cj * (k - 1));
    return (a == 1 || a == ci || b == 1 || b == cj) ? ldg(C.src, sl) : sl;
  };
  // the cases (pole forms, coincident, between along i, diagonal, between
  // along j) exclude one another; they are selected, not branched on, so
  // that neighbouring lanes of the other parity do not diverge: c0 at
  // (i + di, j + dj), c1 at (i - di, j - dj) for the two-point averages
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
    const R* __restrict__ xc = x1 + (size_t)v * C.ns;
    const R y0 = ldg(xc, c0), y1 = ldg(xc, c1);
    const R add = two ? R(0.5) * (y0 + y1) : y0;
    R* __restrict__ xf = x2 + (size_t)v * F.ns;
    stg(xf, s, ldg((const R*)xf, s) + add);
  }
}

template <class R>
__global__ void k_fill(R* __restrict__ x, size_t n, R v, size_t mb) {
  at_member(mb, x);
  const size_t t = (size_t)blockIdx.x * bdim() + threadIdx.x;
  if (t < n) x[t] = v;
}
__global__ void k_copy(double* __restrict__ y, const double* __restrict__ x, size_t n, size_t mb) {
  at_member(mb, y, x);
  const size_t t = (size_t)blockIdx.x * bdim() + threadIdx.x;
  if (t < n) y[t] = x[t];
}

// ----------------------------------------------------- velocity and KE ----
// operators.F:1402-1514 calculate_velocity: grad(chi) + k x grad(psi), from
// the FP64 streamfunction and velocity potential.  Mixed path: the gradient
// is formed on differences from the corner's first cell (its weights sum to
// zero), in FP64 from the FP32 weights.
//
// go != nullptr: launched before the host has tested the solve; it runs only
// where k_conv_go found the solve converged (go[0] == 1), else exits.
template <class R>
__global__ void k_velocity(GeoT<R> G, const double* __restrict__ psi, const double* __restrict__ chi,
                           R* __restrict__ u, const R* __restrict__ rhs1, const double* __restrict__ bsum,
                           double total, R* __restrict__ eta4, const double* __restrict__ go) {
  if (go) {
    at_member(G.mb, go);
    if (go[0] == 0.0) return;
  }
  at_member(G.mb, psi, chi, u);
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  if (eta4) {
    // mixed path: the new level's vorticity is the solve's right-hand side
    // less its mean (k_post_elliptic fused here)
    at_member(G.mb, rhs1, eta4);
    if (bsum) at_member(G.mb, bsum);
    if constexpr (sizeof(R) == 4) {
      const PushW wf = push_load(G.pf, c, iim, jjm, s);
      const R e4 = rhs_less_mean(rhs1[s], bsum, 0, total);
      put1(G.pf, wf, eta4, s, e4);
    }
  }
  const PushW w = push_load(G.pv, c, iim, jjm, s);
  if (!interior(c, iim, jjm)) {
    // the corner table's components are the corners; each space component
    // is one field of it
    put_all<R, 2>(G.pv, w, u, ns, s, R(0), R(0));
    put_all<R, 2>(G.pv, w, u + (size_t)2 * ns, ns, s, R(0), R(0));
    return;
  }
  const int i = c.i, j = c.j, k = c.k;
  const int nb[2][3] = {{s, SL(i - 1, j, k), SL(i - 1, j - 1, k)}, {s, SL(i - 1, j - 1, k), SL(i, j - 1, k)}};
#define VW(kk, sp, ic) G.vw[((kk) + 3 * ((sp) + 2 * (ic))) * ns + s]
  R uv[4];   // planes ic + 2 * space
  for (int ic = 0; ic < 2; ++ic) {
    if constexpr (sizeof(R) == 8) {
      const double cai = G.cai[ic * ns + s];
      double gc[2], gp[2];
      for (int sp = 0; sp < 2; ++sp) {
        gc[sp] = (+VW(0, sp, ic) * chi[nb[ic][0]] + VW(1, sp, ic) * chi[nb[ic][1]] + VW(2, sp, ic) * chi[nb[ic][2]]) * cai;
        gp[sp] = (+VW(0, sp, ic) * psi[nb[ic][0]] + VW(1, sp, ic) * psi[nb[ic][1]] + VW(2, sp, ic) * psi[nb[ic][2]]) * cai;
      }
      uv[ic + 0] = gc[0] + (-gp[1]);
      uv[ic + 2] = gc[1] + gp[0];
    } else {
      const double cai = (double)G.cai[ic * ns + s];
      const double c1 = chi[nb[ic][1]] - chi[nb[ic][0]], c2 = chi[nb[ic][2]] - chi[nb[ic][0]];
      const double p1 = psi[nb[ic][1]] - psi[nb[ic][0]], p2 = psi[nb[ic][2]] - psi[nb[ic][0]];
      double gc[2], gp[2];
      for (int sp = 0; sp < 2; ++sp) {
        const double v1 = (double)VW(1, sp, ic), v2 = (double)VW(2, sp, ic);
        gc[sp] = (v1 * c1 + v2 * c2) * cai;
        gp[sp] = (v1 * p1 + v2 * p2) * cai;
      }
      uv[ic + 0] = (R)(gc[0] - gp[1]);
      uv[ic + 2] = (R)(gc[1] + gp[0]);
    }
  }
#undef VW
  put_all<R, 2>(G.pv, w, u, ns, s, uv[0], uv[1]);
  put_all<R, 2>(G.pv, w, u + (size_t)2 * ns, ns, s, uv[2], uv[3]);
}

// operators.F:1316-1393 with the corner kinetic energy 0.5 |u|^2 (step.F:192-203)
template <class R>
__global__ void k_center_ke(GeoT<R> G, const R* __restrict__ u, R* __restrict__ ke) {
  at_member(G.mb, u, ke);
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  const int i = c.i, j = c.j, k = c.k;
  const PushW w = push_load(G.pf, c, iim, jjm, s);
  if (i > iim - 1 || j > jjm - 1) { put1(G.pf, w, ke, s, R(0)); return; }
  auto KC = [&](int ic, int sl) {
    const R a = u[ic * ns + sl], b = u[(ic + 2) * ns + sl];
    return R(0.5) * (a * a + b * b);
  };
#define SA(kk, ic, sl) G.sub_area[((kk) + 3 * (ic)) * ns + (sl)]
  const int se = SL(i + 1, j, k), sne = SL(i + 1, j + 1, k), sn = SL(i, j + 1, k);
  const R v = (+SA(0, 0, s) * KC(0, s) + SA(0, 1, s) * KC(1, s) + SA(1, 0, se) * KC(0, se) +
               SA(1, 1, sne) * KC(1, sne) + SA(2, 0, sne) * KC(0, sne) + SA(2, 1, sn) * KC(1, sn)) * G.area_inv[s];
#undef SA
  put1(G.pf, w, ke, s, v);
}

// operators.F:47-261 curl and divergence of a corner vector (initial state)
__global__ void k_curl_div(Geo G, const double* __restrict__ u, double* __restrict__ curl, double* __restrict__ dv) {
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = xcd_slot();
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  const int i = c.i, j = c.j, k = c.k;
  if (i > iim - 1 || j > jjm - 1) { if (curl) curl[s] = 0.0; if (dv) dv[s] = 0.0; return; }
#define VW(kk, sp, ic, sl) G.vw[((kk) + 3 * ((sp) + 2 * (ic))) * ns + (sl)]
#define UU(sp, ic, sl) u[((ic) + 2 * (sp)) * ns + (sl)]
  const int se = SL(i + 1, j, k), sne = SL(i + 1, j + 1, k), sn = SL(i, j + 1, k);
  const double ai = G.area_inv[s];
  if (curl)
    curl[s] = (+VW(0, 1, 0, s) * UU(0, 0, s) - VW(0, 0, 0, s) * UU(1, 0, s) + VW(0, 1, 1, s) * UU(0, 1, s) -
               VW(0, 0, 1, s) * UU(1, 1, s) + VW(1, 1, 0, se) * UU(0, 0, se) - VW(1, 0, 0, se) * UU(1, 0, se) +
               VW(1, 1, 1, sne) * UU(0, 1, sne) - VW(1, 0, 1, sne) * UU(1, 1, sne) + VW(2, 1, 0, sne) * UU(0, 0, sne) -
               VW(2, 0, 0, sne) * UU(1, 0, sne) + VW(2, 1, 1, sn) * UU(0, 1, sn) - VW(2, 0, 1, sn) * UU(1, 1, sn)) *
              ai;
  if (dv)
    dv[s] = (-VW(0, 0, 0, s) * UU(0, 0, s) - VW(0, 1, 0, s) * UU(1, 0, s) - VW(0, 0, 1, s) * UU(0, 1, s) -
             VW(0, 1, 1, s) * UU(1, 1, s) - VW(1, 0, 0, se) * UU(0, 0, se) - VW(1, 1, 0, se) * UU(1, 0, se) -
             VW(1, 0, 1, sne) * UU(0, 1, sne) - VW(1, 1, 1, sne) * UU(1, 1, sne) - VW(2, 0, 0, sne) * UU(0, 0, sne) -
             VW(2, 1, 0, sne) * UU(1, 0, sne) - VW(2, 0, 1, sn) * UU(0, 1, sn) - VW(2, 1, 1, sn) * UU(1, 1, sn)) *
            ai;
#undef VW
#undef UU
}

// eta = relative + f (step.F:177-184; mixed path: the stored relative
// vorticity is the solve's right-hand side itself); inv_h (step.F:448-452)
template <class R>
__global__ void k_post_elliptic(GeoT<R> G, const R* __restrict__ rhs1, R* __restrict__ eta4) {
  at_member(G.mb, rhs1, eta4);
  const int s = xcd_slot();
  if (s >= G.ns) return;
  R v;
  if constexpr (sizeof(R) == 4) v = rhs1[s];
  else v = rhs1[s] + G.f[s];
  const PushW w = G.pf.hp ? push_load(G.pf, ijk_of(s, G.iim, G.jjm), G.iim, G.jjm, s) : PushW{};
  put1(G.pf, w, eta4, s, v);
}
template <class R>
__global__ void k_inv_h(GeoT<R> G, const R* __restrict__ h4, R* __restrict__ inv_h) {
  at_member(G.mb, h4, inv_h);
  const int s = xcd_slot();
  if (s >= G.ns) return;
  inv_h[s] = G.mask[s] == R(1) ? R(1) / habs(G, h4[s]) : R(0);
}

// ------------------------------------------------------------- budgets ---
// budgets.F:31-247 compute_budgets integrands summed over real cells, one
// quantity per blockIdx.y: 0 h, 1 h ke, 2 g h^2 / 2, 3 g h hs, 4 (unused:
// the total energy is formed from 1..3), 5 eta^2 / h, 6 eta, 7 div, 8 + q the
// mass-weighted square of tracer q; each times the cell area.  Per-block
// partial sums in FP64 (stored values converted to absolute FP64 first).
constexpr int BUDGET_MAXTR = 16;
template <class R>
struct BudgetIn {
  const R *h, *ke, *eta, *div;
  const R* tr[BUDGET_MAXTR];
  const double *area, *mask, *hs, *f;   // f: mixed path (eta stored less f), else null
  double h0, grav;
  int ns;
};
template <class R>
__global__ void __launch_bounds__(256) k_budgets(BudgetIn<R> B, double* __restrict__ part) {
  __shared__ double sh[256];
  const int q = blockIdx.y;
  double acc = 0.0;
  for (int s = blockIdx.x * bdim() + threadIdx.x; s < B.ns; s += gridDim.x * bdim()) {
    const double m = B.mask[s];
    if (m == 0.0) continue;
    const double h = (double)B.h[s] + B.h0, a = B.area[s];
    const double eta = (double)B.eta[s] + (B.f ? B.f[s] : 0.0);
    double w;
    switch (q) {
      case 0: w = h * a; break;
      case 1: w = (h * (double)B.ke[s]) * a; break;
      case 2: w = (((0.5 * B.grav) * h) * h) * a; break;
      case 3: w = ((B.grav * h) * B.hs[s]) * a; break;
      case 5: w = ((eta * eta) / h) * a; break;
      case 6: w = eta * a; break;
      case 7: w = (double)B.div[s] * a; break;
      case 4: w = 0.0; break;
      default: {
        const double t = (double)B.tr[q - 8][s];
        w = ((t * t) * h) * a;
      }
    }
    acc += w * m;
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int o = bdim() / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[q * gridDim.x + blockIdx.x] = sh[0];
}

// FP64 state -> FP32 storage of the mixed path: y = x - off (off may be null)
__global__ void k_to_f32(size_t n, const double* __restrict__ x, const double* __restrict__ off, double c,
                         float* __restrict__ y) {
  const size_t t = (size_t)blockIdx.x * bdim() + threadIdx.x;
  if (t < n) y[t] = (float)(x[t] - (off ? off[t] : 0.0) - c);
}
// and back: y = x + off + c
__global__ void k_to_f64(size_t n, const float* __restrict__ x, const double* __restrict__ off, double c,
                         double* __restrict__ y) {
  const size_t t = (size_t)blockIdx.x * bdim() + threadIdx.x;
  if (t < n) y[t] = (double)x[t] + (off ? off[t] : 0.0) + c;
}

}  // namespace dev
}  // namespace swm

namespace swm {
namespace dev {

// ------------------------------------------------- synthetic initial state --
// Real 4-pi normalised spherical-harmonic sums at the cell centres p (3
// planes): out[0] = sum c1 l^-1.5 Y / (-l(l+1)) (streamfunction shape),
// out[1] = sum c1 l^-1.5 Y (vorticity shape), out[2], out[3] the same with c2
// (velocity potential / divergence), out[4] = sum c3 l^-1.5 Y (thickness).
// coef holds, for l = 1..L and m = -l..l, the three coefficient sets at
// index 3*(l*l - 1 + l + m) + set.
__global__ void k_harmonics(int ns, const double* __restrict__ p, int lmax, const double* __restrict__ coef,
                            double* __restrict__ o0, double* __restrict__ o1, double* __restrict__ o2,
                            double* __restrict__ o3, double* __restrict__ o4) {
  const int s = blockIdx.x * bdim() + threadIdx.x;
  if (s >= ns) return;
  const double x = p[s], y = p[ns + s], z = p[2 * ns + s];
  const double mu = fmin(1.0, fmax(-1.0, z));
  const double cphi = sqrt(fmax(0.0, 1.0 - mu * mu));
  const double lon = atan2(y, x);
  double a0 = 0, a1 = 0, a2 = 0, a3 = 0, a4 = 0;
  double pmm = 1.0;
  for (int m = 0; m <= lmax; ++m) {
    if (m == 1) pmm = sqrt(3.0) * cphi;
    else if (m > 1) pmm = sqrt((2.0 * m + 1.0) / (2.0 * m)) * cphi * pmm;
    double sm, cm;
    sincos(m * lon, &sm, &cm);
    double pl2 = 0.0, pl1 = 0.0;
    for (int l = m; l <= lmax; ++l) {
      double pl;
      if (l == m) pl = pmm;
      else if (l == m + 1) pl = sqrt(2.0 * m + 3.0) * mu * pmm;
      else {
        const double a = sqrt((2.0 * l - 1.0) * (2.0 * l + 1.0) / ((double)(l - m) * (l + m)));
        const double b = sqrt((2.0 * l + 1.0) * (l + m - 1.0) * (l - m - 1.0) / ((double)(l - m) * (l + m) * (2.0 * l - 3.0)));
        pl = a * mu * pl1 - b * pl2;
      }
      pl2 = pl1; pl1 = pl;
      if (l < 1) continue;
      const double w = pow((double)l, -1.5);
      const double inv = -1.0 / (l * (l + 1.0));
      const int ip = 3 * (l * l - 1 + l + m), in_ = 3 * (l * l - 1 + l - m);
      double t1, t2, t3;
      if (m == 0) {
        t1 = coef[ip] * pl; t2 = coef[ip + 1] * pl; t3 = coef[ip + 2] * pl;
      } else {
        t1 = pl * (coef[ip] * cm + coef[in_] * sm);
        t2 = pl * (coef[ip + 1] * cm + coef[in_ + 1] * sm);
        t3 = pl * (coef[ip + 2] * cm + coef[in_ + 2] * sm);
      }
      a0 += w * inv * t1; a1 += w * t1; a2 += w * inv * t2; a3 += w * t2; a4 += w * t3;
    }
  }
  o0[s] = a0; o1[s] = a1; o2[s] = a2; o3[s] = a3; o4[s] = a4;
}

// per-block partial sums of squares (masked) and maxima of |u| at corners
__global__ void k_sumsq_partial(const double* __restrict__ x, const double* __restrict__ w, int n,
                                double* __restrict__ part) {
  __shared__ double sh[256];
  double acc = 0.0;
  for (int t = blockIdx.x * bdim() + threadIdx.x; t < n; t += gridDim.x * bdim()) acc += x[t] * x[t] * w[t];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int o = bdim() / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}
__global__ void k_speedmax_partial(Geo G, const double* __restrict__ u, double* __restrict__ part) {
  __shared__ double sh[256];
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  double acc = 0.0;
  for (int s = blockIdx.x * bdim() + threadIdx.x; s < ns; s += gridDim.x * bdim()) {
    const IJK c = ijk_of(s, iim, jjm);
    if (!interior(c, iim, jjm)) continue;
    for (int ic = 0; ic < 2; ++ic) {
      const double a = u[ic * ns + s], b = u[(ic + 2) * ns + s];
      acc = fmax(acc, sqrt(a * a + b * b));
    }
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int o = bdim() / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// thickness h0 + scale*hp at real cells (0 elsewhere); chi scaled in place;
// u scaled; tracer = sum of Gaussian bumps (all slots)
__global__ void k_synth_finish(Geo G, double h0, double hscale, const double* __restrict__ hp, double* __restrict__ h,
                               double cscale, double* __restrict__ chi) {
  const int s = blockIdx.x * bdim() + threadIdx.x;
  if (s >= G.ns) return;
  h[s] = G.mask[s] == 1.0 ? h0 + hscale * hp[s] : 0.0;
  chi[s] = chi[s] * cscale;
}
__global__ void k_scale(double* __restrict__ x, size_t n, double a) {
  const size_t t = (size_t)blockIdx.x * bdim() + threadIdx.x;
  if (t < n) x[t] = x[t] * a;
}
__global__ void k_bumps(int ns, const double* __restrict__ p, int nb, const double* __restrict__ centres,
                        double radius, double width, double* __restrict__ t) {
  const int s = blockIdx.x * bdim() + threadIdx.x;
  if (s >= ns) return;
  const double x = p[s], y = p[ns + s], z = p[2 * ns + s];
  double acc = 0.0;
  for (int b = 0; b < nb; ++b) {
    const double d = radius * acos(fmin(1.0, fmax(-1.0, centres[3 * b] * x + centres[3 * b + 1] * y + centres[3 * b + 2] * z)));
    const double r = d / width;
    acc += exp(-(r * r));
  }
  t[s] = acc;
}

}  // namespace dev
}  // namespace swm

namespace swm {
namespace dev {

// single reference operators on every slot (for the operator entry points)
__global__ void k_op_div_edge(Geo G, const double* __restrict__ F, double* __restrict__ out) {
  const int s = xcd_slot();
  if (s >= G.ns) return;
  const IJK c = ijk_of(s, G.iim, G.jjm);
  out[s] = div_edge_at(G.iim, G.jjm, G.ns, c, F, G.area_inv[s]);
}
__global__ void k_op_lap(Geo G, const double* __restrict__ x, double* __restrict__ out) {
  const int s = xcd_slot();
  if (s >= G.ns) return;
  const IJK c = ijk_of(s, G.iim, G.jjm);
  auto X = [&](int sl) { return x[sl]; };
  out[s] = lap_at(G.iim, G.jjm, G.ns, c, G.lw, X);
}
// operators.F:1316-1393 with corner values given
__global__ void k_op_center_ke(Geo G, const double* __restrict__ kec, double* __restrict__ ke) {
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = blockIdx.x * bdim() + threadIdx.x;
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  const int i = c.i, j = c.j, k = c.k;
  if (i > iim - 1 || j > jjm - 1) { ke[s] = 0.0; return; }
  const int se = (i + 1) - 1 + iim * ((j) - 1 + jjm * (k - 1)), sne = (i + 1) - 1 + iim * ((j + 1) - 1 + jjm * (k - 1)),
            sn = (i) - 1 + iim * ((j + 1) - 1 + jjm * (k - 1));
#define SA(kk, ic, sl) G.sub_area[((kk) + 3 * (ic)) * ns + (sl)]
  ke[s] = (+SA(0, 0, s) * kec[s] + SA(0, 1, s) * kec[ns + s] + SA(1, 0, se) * kec[se] + SA(1, 1, sne) * kec[ns + sne] +
           SA(2, 0, sne) * kec[sne] + SA(2, 1, sn) * kec[ns + sn]) * G.area_inv[s];
#undef SA
}

}  // namespace dev
}  // namespace swm

namespace swm {
namespace dev {

// ------------------------------------------- coarse end of the V-cycle ----
// Levels 1..nlev (nlev <= 3) of one V-cycle (multigrid_solver.F:476-591) in
// one workgroup: the levels' alpha, the coarse right-hand sides and the
// smoother's second buffer live in LDS; weights and halo maps are read from
// global memory (L2) into registers once per visit of a level, so each
// Jacobi sweep is LDS traffic and one barrier.  Same arithmetic, in the same
// order, as k_relax / k_inject / k_interp.
constexpr int CO_THREADS = 1024;
constexpr int CO_MAXI = 2;   // items (field, slot) per thread: nvar * ns(level 3) <= 2048
constexpr int CO_MAXLEV = 3;

template <class R>
struct CoarseArgs {
  LevT<R> lev[CO_MAXLEV];   // levels 1..nlev
  int nlev, nvar;
  const R* beta;   // right-hand side at level nlev (nvar planes of ns)
  R* alpha;        // solution at level nlev (in: warm start unless zero)
  int zero;        // alpha at level nlev is zero on entry
  size_t mb;       // member stride of beta / alpha (one workgroup per member, blockIdx.z)
};

__device__ __forceinline__ int co_ns(int l) { return (2 + (1 << l)) * (2 + (1 << l)) * 10; }

// relax nsw sweeps at level L: x (LDS, result) and t (LDS, scratch); b is the
// right-hand side (LDS or global)
template <class R>
__device__ __forceinline__ void co_relax(const LevT<R>& L, int nvar, int nsw, R* x, R* t, const R* b) {
  const int iim = L.iim, jjm = L.jjm, ns = L.ns;
  const int nitem = nvar * ns;
  int kind[CO_MAXI], self[CO_MAXI], out_i[CO_MAXI], nb[CO_MAXI][6];
  R w[CO_MAXI][6], wt[CO_MAXI], bv[CO_MAXI];
#pragma unroll
  for (int r = 0; r < CO_MAXI; ++r) {
    const int it = threadIdx.x + r * CO_THREADS;
    kind[r] = -1;
    if (it >= nitem) continue;
    const int v = it / ns, s = it - v * ns;
    const IJK c = ijk_of(s, iim, jjm);
    const int i = c.i, j = c.j, k = c.k;
    const bool in = i >= 2 && i <= iim - 1 && j >= 2 && j <= jjm - 1;
    const bool np = i == 1 && j == jjm - 1, sp = i == iim - 1 && j == 1;
    out_i[r] = v * ns + s;
    kind[r] = in ? 0 : (np ? 1 : (sp ? 2 : 3));
    if (kind[r] == 3) continue;
    int q6[6];
    if (in) {
      q6[0] = SL(i - 1, j, k); q6[1] = SL(i - 1, j - 1, k); q6[2] = SL(i, j - 1, k);
    } else if (np) {
      q6[0] = s; q6[1] = SL(iim, 1, k); q6[2] = SL(i, j - 1, k);
    } else {
      q6[0] = SL(i - 1, j, k); q6[1] = SL(1, jjm, k); q6[2] = s;
    }
    q6[3] = SL(i + 1, j, k); q6[4] = SL(i + 1, j + 1, k); q6[5] = SL(i, j + 1, k);
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      nb[r][q] = v * ns + L.src[q6[q]];
      w[r][q] = L.lw[q * ns + s];
    }
    self[r] = v * ns + L.src[s];
    wt[r] = L.lwt[s];
    bv[r] = b[v * ns + s];
  }
  const R omega = R(0.8);
  R* a = x;
  R* o = t;
  for (int sw = 0; sw < nsw; ++sw) {
#pragma unroll
    for (int r = 0; r < CO_MAXI; ++r) {
      if (kind[r] < 0) continue;
      if (kind[r] == 3) { o[out_i[r]] = R(0); continue; }
      R tt;
      if (kind[r] == 0)
        tt = wt[r] * (bv[r] - w[r][0] * a[nb[r][0]] - w[r][1] * a[nb[r][1]] - w[r][2] * a[nb[r][2]] -
                      w[r][3] * a[nb[r][3]] - w[r][4] * a[nb[r][4]] - w[r][5] * a[nb[r][5]]);
      else if (kind[r] == 1)
        tt = wt[r] * (bv[r] - w[r][1] * a[nb[r][1]] - w[r][2] * a[nb[r][2]] - w[r][3] * a[nb[r][3]] -
                      w[r][4] * a[nb[r][4]] - w[r][5] * a[nb[r][5]]);
      else
        tt = wt[r] * (bv[r] - w[r][0] * a[nb[r][0]] - w[r][1] * a[nb[r][1]] - w[r][3] * a[nb[r][3]] -
                      w[r][4] * a[nb[r][4]] - w[r][5] * a[nb[r][5]]);
      o[out_i[r]] = (R(1) - omega) * a[self[r]] + omega * tt;
    }
    __syncthreads();
    R* sw_ = a; a = o; o = sw_;
  }
  // nsw is even: the result is in x
}

// residual at odd-odd fine points of level F -> right-hand side of level Cc
template <class R>
__device__ __forceinline__ void co_inject(const LevT<R>& F, const LevT<R>& Cc, int nvar, const R* xa, const R* xb, R* x1) {
  const int nitem = nvar * Cc.ns;
  for (int it = threadIdx.x; it < nitem; it += CO_THREADS) {
    const int v = it / Cc.ns, s1 = it - v * Cc.ns;
    int iim = Cc.iim, jjm = Cc.jjm;
    const IJK c = ijk_of(s1, iim, jjm);
    const bool in = c.i >= 2 && c.i <= iim - 1 && c.j >= 2 && c.j <= jjm - 1;
    const bool np = c.k == 1 && c.i == 1 && c.j == jjm - 1;
    const bool sp = c.k == 10 && c.i == iim - 1 && c.j == 1;
    if (!(in || np || sp)) continue;
    const int k = c.k;
    int i2, j2;
    if (in) { i2 = 2 * c.i - 1; j2 = 2 * c.j - 1; }
    else if (np) { i2 = 1; j2 = F.jjm - 1; }
    else { i2 = F.iim - 1; j2 = 1; }
    iim = F.iim; jjm = F.jjm;
    const int ns = F.ns;
    const int s = SL(i2, j2, k);
    const int* __restrict__ m = F.src;
    const R* xav = xa + v * ns;
    auto A = [&](int sl) { return xav[m[sl]]; };
#define LW(n) F.lw[((n) - 1) * ns + s]
    R r;
    if (in)
      r = xb[v * ns + s] - ((LW(7) - R(0)) * A(s) + LW(1) * A(SL(i2 - 1, j2, k)) + LW(2) * A(SL(i2 - 1, j2 - 1, k)) +
                            LW(3) * A(SL(i2, j2 - 1, k)) + LW(4) * A(SL(i2 + 1, j2, k)) +
                            LW(5) * A(SL(i2 + 1, j2 + 1, k)) + LW(6) * A(SL(i2, j2 + 1, k)));
    else if (np)
      r = xb[v * ns + s] - ((LW(7) - R(0)) * A(s) + LW(2) * A(SL(iim, 1, k)) + LW(3) * A(SL(i2, j2 - 1, k)) +
                            LW(4) * A(SL(i2 + 1, j2, k)) + LW(5) * A(SL(i2 + 1, j2 + 1, k)) +
                            LW(6) * A(SL(i2, j2 + 1, k)));
    else
      r = xb[v * ns + s] - ((LW(7) - R(0)) * A(s) + LW(1) * A(SL(i2 - 1, j2, k)) + LW(2) * A(SL(1, jjm, k)) +
                            LW(4) * A(SL(i2 + 1, j2, k)) + LW(5) * A(SL(i2 + 1, j2 + 1, k)) +
                            LW(6) * A(SL(i2, j2 + 1, k)));
#undef LW
    x1[v * Cc.ns + s1] = r;
  }
  __syncthreads();
}

// coarse correction of level Cc added into level F
template <class R>
__device__ __forceinline__ void co_interp(const LevT<R>& Cc, const LevT<R>& F, int nvar, const R* x1, R* x2) {
  const int nitem = nvar * F.ns;
  for (int it = threadIdx.x; it < nitem; it += CO_THREADS) {
    const int v = it / F.ns, s = it - v * F.ns;
    const int iim = F.iim, jjm = F.jjm;
    const IJK c = ijk_of(s, iim, jjm);
    const int i = c.i, j = c.j, k = c.k;
    const int p2 = iim - 1, q2 = jjm - 1;
    const int ci = Cc.iim, cj = Cc.jjm;
    const R* x1v = x1 + v * Cc.ns;
    auto T = [&](int ii, int jj) {
      const int a = (ii + 1) / 2, b = (jj + 1) / 2;
      return x1v[Cc.src[(a - 1) + ci * ((b - 1) + cj * (k - 1))]];
    };
    // the cases selected, not branched on (as k_interp)
    const bool io = (i & 1), jo = (j & 1);
    const bool pn = i == 1 && j == q2, ps = i == p2 && j == 1;
    const bool iA = i >= 3 && i <= iim - 1, iB = i >= 2 && i <= iim - 2;
    const bool jA = j >= 3 && j <= jjm - 1, jB = j >= 2 && j <= jjm - 2;
    const bool cA = io && jo && iA && jA;
    const bool cB = !io && jo && iB && jA;
    const bool cC = !io && !jo && iB && jB;
    const bool cD = io && !jo && iA && jB;
    const bool two = cB || cC || cD;
    if (pn || ps || cA || two) {
      const int di = (cB || cC) ? 1 : 0, dj = (cC || cD) ? 1 : 0;
      const R y0 = T(i + di, j + dj);
      const R y1 = two ? T(i - di, j - dj) : y0;
      const R add = two ? R(0.5) * (y0 + y1) : y0;
      x2[v * F.ns + s] = x2[v * F.ns + s] + add;
    }
  }
  __syncthreads();
}

// level l of the bundle with constant indices (no dynamic indexing of the
// kernel argument, which would go through scratch memory)
template <class R>
__device__ __forceinline__ LevT<R> co_lev(const CoarseArgs<R>& P, int l) {
  return l == 1 ? P.lev[0] : (l == 2 ? P.lev[1] : P.lev[2]);
}

template <class R>
__global__ void __launch_bounds__(CO_THREADS) k_coarse_cycle(CoarseArgs<R> P) {
  // the member's right-hand side and solution (local copies: writing to the
  // argument bundle itself would move it to scratch memory)
  const R* p_beta = P.beta;
  R* p_alpha = P.alpha;
  at_member(P.mb, p_beta, p_alpha);
  // LDS: alpha of levels 1..nlev, rhs of levels 1..nlev-1, one scratch level
  __shared__ R lds[2 * (160 + 360 + 1000) + 2 * (160 + 360) + 2 * 1000];
  const int nv = P.nvar, top = P.nlev;
  // offsets: alpha(l) at nv * (ns(1)+..+ns(l-1)), rhs(l) after all alphas
  auto cum = [&](int l) {   // ns(1) + .. + ns(l)
    int c = 0;
    for (int m = 1; m <= l; ++m) c += co_ns(m);
    return c;
  };
  auto AL = [&](int l) { return lds + nv * cum(l - 1); };
  auto BE = [&](int l) { return lds + nv * (cum(top) + cum(l - 1)); };
  R* tmp = lds + nv * (cum(top) + cum(top - 1));
  // coarse levels start from zero (multigrid_solver.F:520-524); the top
  // level from the caller's alpha unless that is zero as well
  {
    const int nz = nv * cum(top - 1), n = nv * co_ns(top);
    for (int t = threadIdx.x; t < nz; t += CO_THREADS) lds[t] = R(0);
    R* at = AL(top);
    for (int t = threadIdx.x; t < n; t += CO_THREADS) at[t] = P.zero ? R(0) : p_alpha[t];
  }
  __syncthreads();
  for (int l = top; l >= 2; --l) {
    const R* b = l == top ? p_beta : BE(l);
    co_relax(co_lev(P, l), nv, 2, AL(l), tmp, b);
    co_inject(co_lev(P, l), co_lev(P, l - 1), nv, AL(l), b, BE(l - 1));
  }
  co_relax(co_lev(P, 1), nv, 24, AL(1), tmp, top == 1 ? p_beta : BE(1));
  for (int l = 1; l <= top - 1; ++l) {
    co_interp(co_lev(P, l), co_lev(P, l + 1), nv, AL(l), AL(l + 1));
    co_relax(co_lev(P, l + 1), nv, 2, AL(l + 1), tmp, l + 1 == top ? p_beta : BE(l + 1));
  }
  const int n = nv * co_ns(top);
  const R* at = AL(top);
  for (int t = threadIdx.x; t < n; t += CO_THREADS) p_alpha[t] = at[t];
}

constexpr int CD_MAXN = 10 * 64 + 2;     // real cells of level 3
constexpr int CD_MAXN4 = 10 * 256 + 2;   // real cells of level 4
// Mixed path: the coarse correction of the coarsest levels as one dense
// solve.  alpha = M beta at the real cells of the level (cell[0..n)), with M
// the FP32 n x n block of the inverse of the level's Laplacian bordered by
// ones (A alpha + lambda 1 = beta, sum alpha = 0: the solution without the
// constant the singular operator leaves free, for beta less its component
// along 1 outside A's range).  One wave per row, NV fields, at most MAXN
// cells; the lanes take the columns in a fixed order and a butterfly sums
// them, so the result does not depend on the launch.  Halo slots of alpha
// are not written (the interpolation reads the level through its map).
template <int NV, int MAXN>
__global__ void __launch_bounds__(512) k_coarse_direct(const float* __restrict__ M, const int* __restrict__ cell, int n,
                                                       int ns, const float* __restrict__ bc,
                                                       float* __restrict__ alpha, size_t mb) {
  at_member(mb, bc, alpha);
  // beta at the real cells in the cell order (bc, NV planes of n, written
  // by the injection), staged once per workgroup in LDS, so the rows' dot
  // products read M and LDS only
  __shared__ float sb[NV][MAXN];
  {
    // all of a thread's loads issued together (workgroups of 256 or more)
    const int nt = bdim();
    constexpr int NL = (MAXN + 255) / 256;
    float t[NV][NL];
#pragma unroll
    for (int q = 0; q < NL; ++q) {
      const int c = (int)threadIdx.x + nt * q;
#pragma unroll
      for (int v = 0; v < NV; ++v) t[v][q] = c < n ? ldg(bc + (size_t)v * n, c) : 0.0f;
    }
#pragma unroll
    for (int q = 0; q < NL; ++q) {
      const int c = (int)threadIdx.x + nt * q;
      if (c < n) {
#pragma unroll
        for (int v = 0; v < NV; ++v) sb[v][c] = t[v][q];
      }
    }
  }
  __syncthreads();
  const int r = (int)(blockIdx.x * (bdim() >> 6) + (threadIdx.x >> 6));
  const int lane = (int)(threadIdx.x & 63);
  if (r >= n) return;   // whole waves leave together
  const float* __restrict__ row = M + (size_t)r * n;
  float acc[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) acc[v] = 0.0f;
  // a lane's loads of the row issued together
  constexpr int NQ = (MAXN + 63) / 64, BQ = NQ;
#pragma unroll
  for (int q0 = 0; q0 < NQ; q0 += BQ) {
    float mr[BQ];
#pragma unroll
    for (int q = 0; q < BQ; ++q) {
      const int c = lane + 64 * (q0 + q);
      mr[q] = (q0 + q < NQ && c < n) ? row[c] : 0.0f;
    }
#pragma unroll
    for (int q = 0; q < BQ; ++q) {
      const int c = lane + 64 * (q0 + q);
      if (q0 + q < NQ && c < n) {
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[v] = acc[v] + mr[q] * sb[v][c];
      }
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = acc[v] + __shfl_xor(acc[v], o, 64);
  }
  if (lane == 0) {
#pragma unroll
    for (int v = 0; v < NV; ++v) alpha[(size_t)v * ns + ldg(cell, r)] = acc[v];
  }
}

// Set-up of a dense coarse solve: Gauss-Jordan inversion in place of the
// m x m row-major FP64 matrix A with partial pivoting, one pivot column c per
// pair of launches.  k_gj_pivot (one workgroup of GJ_THREADS) picks the row
// of the largest |A(r, c)|, r >= c (the first of equal ones), swaps it into
// row c and copies row c and column c aside; k_gj_update then forms every
// element from those copies.  The host undoes the row swaps as column swaps
// in reverse order afterwards.
constexpr int GJ_THREADS = 1024;
__global__ void __launch_bounds__(GJ_THREADS) k_gj_pivot(double* __restrict__ A, int m, int c, double* __restrict__ rowb,
                                                         double* __restrict__ colb, int* __restrict__ perm) {
  __shared__ double bv[GJ_THREADS];
  __shared__ int bi[GJ_THREADS];
  const int t = threadIdx.x;
  double best = -1.0;
  int bidx = m;
  for (int r = c + t; r < m; r += GJ_THREADS) {
    const double v = fabs(A[(size_t)r * m + c]);
    if (v > best) { best = v; bidx = r; }
  }
  bv[t] = best;
  bi[t] = bidx;
  __syncthreads();
  for (int h = GJ_THREADS / 2; h > 0; h >>= 1) {
    if (t < h) {
      const double v = bv[t + h];
      const int ix = bi[t + h];
      if (v > bv[t] || (v == bv[t] && ix < bi[t])) { bv[t] = v; bi[t] = ix; }
    }
    __syncthreads();
  }
  const int p = bi[0];
  if (p != c)
    for (int q = t; q < m; q += GJ_THREADS) {
      const double a = A[(size_t)p * m + q];
      A[(size_t)p * m + q] = A[(size_t)c * m + q];
      A[(size_t)c * m + q] = a;
    }
  __syncthreads();
  for (int q = t; q < m; q += GJ_THREADS) {
    rowb[q] = A[(size_t)c * m + q];
    colb[q] = A[(size_t)q * m + c];
  }
  if (t == 0) perm[c] = p;
}
__global__ void k_gj_update(double* __restrict__ A, int m, int c, const double* __restrict__ rowb,
                            const double* __restrict__ colb) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)m * m) return;
  const int r = (int)(e / m), q = (int)(e % m);
  const double pinv = 1.0 / rowb[c];
  if (r == c) A[e] = q == c ? pinv : rowb[q] * pinv;
  else if (q == c) A[e] = -colb[r] * pinv;
  else A[e] = A[e] - colb[r] * (rowb[q] * pinv);
}

}  // namespace dev
}  // namespace swm

#undef SL

namespace swm {
namespace dev {

// ------------------------------------------------ single operators --------
// Kernels of the operator entry points that the time step does not use in
// this form.  Same slot rules as the reference loops: results on
// i = 2..nx-1, j = 2..ny-1 of every panel (npad = 1), zero elsewhere.
#define SLP(i, j, k) ((i) - 1 + iim * ((j) - 1 + jjm * ((k) - 1)))

// operators.F:592-682 gradient at the two corners; out planes ic + 2*space
__global__ void k_op_gradient(Geo G, const double* __restrict__ x, double* __restrict__ out) {
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = blockIdx.x * bdim() + threadIdx.x;
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  if (!interior(c, iim, jjm)) {
    for (int p = 0; p < 4; ++p) out[p * ns + s] = 0.0;
    return;
  }
  const int i = c.i, j = c.j, k = c.k;
  const int nb[2][3] = {{s, SLP(i - 1, j, k), SLP(i - 1, j - 1, k)}, {s, SLP(i - 1, j - 1, k), SLP(i, j - 1, k)}};
#define VW(kk, sp, ic) G.vw[((kk) + 3 * ((sp) + 2 * (ic))) * ns + s]
  for (int ic = 0; ic < 2; ++ic)
    for (int sp = 0; sp < 2; ++sp)
      out[(ic + 2 * sp) * ns + s] =
          (+VW(0, sp, ic) * x[nb[ic][0]] + VW(1, sp, ic) * x[nb[ic][1]] + VW(2, sp, ic) * x[nb[ic][2]]) * G.cai[ic * ns + s];
#undef VW
}

// operators.F:528-582 k_cross_v on every slot; planes ic + 2*space
__global__ void k_op_cross(int ns, const double* __restrict__ x, double* __restrict__ out) {
  const int s = blockIdx.x * bdim() + threadIdx.x;
  if (s >= ns) return;
  for (int ic = 0; ic < 2; ++ic) {
    out[ic * ns + s] = -x[(ic + 2) * ns + s];
    out[(ic + 2) * ns + s] = +x[ic * ns + s];
  }
}

// operators.F:982-1112 average_to_cell_corners; mode 0: flag = .true.
// (s_weights), 1: flag = .false. (weights 0.3333333), 2: min_value = .true.
__global__ void k_op_avg_corners(Geo G, int mode, const double* __restrict__ x, double* __restrict__ out) {
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = blockIdx.x * bdim() + threadIdx.x;
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  const int i = c.i, j = c.j, k = c.k;
  if (!interior(c, iim, jjm)) {
    const double f = x[SLP(2, 2, k)];
    out[s] = f;
    out[ns + s] = f;
    return;
  }
  const int a = s, b = SLP(i - 1, j, k), d = SLP(i - 1, j - 1, k), e = SLP(i, j - 1, k);
  const double* sw = G.sw;
  if (mode == 0) {
    out[s] = +sw[0 * ns + s] * x[a] + sw[1 * ns + s] * x[b] + sw[2 * ns + s] * x[d];
    out[ns + s] = +sw[3 * ns + s] * x[a] + sw[4 * ns + s] * x[d] + sw[5 * ns + s] * x[e];
  } else if (mode == 1) {
    const double f = 0.3333333;
    out[s] = +f * x[a] + f * x[b] + f * x[d];
    out[ns + s] = +f * x[a] + f * x[d] + f * x[e];
  } else {
    out[s] = dmin(dmin(sw[0 * ns + s] * x[a], sw[1 * ns + s] * x[b]), sw[2 * ns + s] * x[d]);
    out[ns + s] = dmin(dmin(sw[3 * ns + s] * x[a], sw[4 * ns + s] * x[d]), sw[5 * ns + s] * x[e]);
  }
}

__device__ __forceinline__ double pv_of(const double* __restrict__ h, const double* __restrict__ eta, int sl) {
  const double w = h[sl] > 0.0 ? 1.0 / h[sl] : 0.0;
  return eta[sl] * w;
}

// operators.F:2000-2074 centered_vorticity_flux
__global__ void k_op_vort_flux(Geo G, const double* __restrict__ mflux, const double* __restrict__ h,
                               const double* __restrict__ eta, double* __restrict__ out) {
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = blockIdx.x * bdim() + threadIdx.x;
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  if (!interior(c, iim, jjm)) {
    for (int e = 0; e < 3; ++e) out[e * ns + s] = 0.0;
    return;
  }
  const int i = c.i, j = c.j, k = c.k;
  const int up[3] = {SLP(i - 1, j, k), SLP(i - 1, j - 1, k), SLP(i, j - 1, k)};
  const double pv0 = pv_of(h, eta, s);
  for (int e = 0; e < 3; ++e) out[e * ns + s] = 0.5 * (mflux[e * ns + s]) * (pv0 + pv_of(h, eta, up[e]));
}

// operators.F:2083-2186 centered_divergence_curl
__global__ void k_op_div_curl(Geo G, const double* __restrict__ hc, const double* __restrict__ eta,
                              const double* __restrict__ u, const double* __restrict__ h, double* __restrict__ out) {
  const int iim = G.iim, jjm = G.jjm, ns = G.ns;
  const int s = blockIdx.x * bdim() + threadIdx.x;
  if (s >= ns) return;
  const IJK c = ijk_of(s, iim, jjm);
  if (!interior(c, iim, jjm)) {
    for (int e = 0; e < 3; ++e) out[e * ns + s] = 0.0;
    return;
  }
  const int i = c.i, j = c.j, k = c.k;
  const int up[3] = {SLP(i - 1, j, k), SLP(i - 1, j - 1, k), SLP(i, j - 1, k)};
  const double pv0 = pv_of(h, eta, s);
  const Half<double> ht = half_edges<double>(iim, jjm, ns, i, j, k, G.dc, G.tng, hc, u);
  for (int e = 0; e < 3; ++e) out[e * ns + s] = (-ht.r[0][e] + ht.r[1][e]) * (pv_of(h, eta, up[e]) + pv0) * 0.5;
}

#undef SLP

}  // namespace dev
}  // namespace swm
