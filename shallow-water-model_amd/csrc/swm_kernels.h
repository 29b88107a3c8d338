// swm_kernels.h -- argument bundles of the device kernels.
//
// The step and multigrid kernels are templates on the storage type R of the
// state and metric planes: R = double is the FP64 path (the reference's
// arithmetic, operation for operation), R = float the mixed path (FP32
// storage and stencils; see the kernels for the offset forms it uses).
#pragma once
#include <hip/hip_runtime.h>

namespace swm {
namespace dev {

// Producer-side halo writes of one halo table (face, vrtx or edge).  The
// table copies element (c', s') to halo element (c, g); here it is inverted
// so that the kernel that forms the value at slot s' also stores its copies.
// hp holds four words per slot, read with one load:
//   x   bits 0-3    number of entries whose source is this slot (at most 15)
//       bits 4-9    two bits per component of the slot as a destination:
//                   0 not one (the kernel stores its own value), 1 copied
//                   from another slot (left to that slot's thread), 2 zero
//       bits 10-31  where its entries beyond the third start in e
//   y, z, w         its first three entries
// An entry is (source component << 28) | destination plane index c * ns + g.
// hp == nullptr: no producer-side writes (the halo fills run instead).
struct Push {
  const int4* hp;
  const int* e;
};

// grid extents and metric planes of the model level.  Mixed path only: h0
// is subtracted from every stored thickness-like field (h, the low-order
// mass) and fo (the Coriolis parameter) from the stored vorticity, so that
// FP32 keeps the small increments of nearly constant fields.
//
// Ensembles: the members of a context share the grid, so the metric planes
// here are shared, while each member's state and scratch form one block, the
// blocks mb bytes apart.  A launch's z index is the member; the kernels add
// z * mb to every state pointer they are given (mb = 0 with one member).
template <class R>
struct GeoT {
  int iim, jjm, ns;
  const R *area, *area_inv, *cai, *sub_area, *dc, *nrm, *tng, *vw, *sw, *lw, *f, *hs, *mask;
  R h0;
  const R* fo;
  size_t mb;
  // mixed path: normal and tangent planes times the half-edge lengths,
  // (kk, a, ic) planes as nrm / tng (formed in FP64, stored in R)
  const R *ndc, *tdc;
  // mixed path (FP64 defects): the Laplacian's symmetric edge coefficients
  // sym(e) = area * l_weights(e+1) of the edge towards up_e (an edge field,
  // halos filled) and area_inv at real cells, 0 elsewhere
  const double *sym, *aim;
  // producer-side halo writes of the step kernels (all null: none)
  Push pf, pv, pe;
  // the slots whose edges the real poles' own edges copy (north pole edges
  // 1 and 2, south pole edges 0 and 1): the final mass kernel corrects
  // those edges itself when it takes over the antidiffusive correction
  int psrc[4];
};
using Geo = GeoT<double>;

// one multigrid level: Laplacian weights, reciprocal central weight and the
// closed-form halo map (src[slot] = slot whose value the halo slot holds)
template <class R>
struct LevT {
  int iim, jjm, ns;
  const R* lw;
  const R* lwt;
  const int* src;
  size_t mb;   // member stride of the state (see GeoT)
  // mixed path, levels above the coarse three: half-precision weights for
  // the smoother and restriction in difference form, hw(k) = S
  // l_weights(k) (k = 1..6) and hcw = 0.8 / (S l_weights(7)); scale S and
  // 1 / S (hw == nullptr: none)
  const _Float16* hw;
  const _Float16* hcw;
  float hs, his;
};
using Lev = LevT<double>;

template <class R>
struct FieldSetT {
  R* p[8];
  size_t mb;   // member stride (see GeoT)
};

// state pointers of member blockIdx.z: p + z * mb bytes
template <class... P>
__device__ __forceinline__ void at_member(size_t mb, P&... p) {
  const size_t o = (size_t)blockIdx.z * mb;
  ((p = (P)((char*)(p) + o)), ...);
}
using FieldSet = FieldSetT<double>;

}  // namespace dev
}  // namespace swm
