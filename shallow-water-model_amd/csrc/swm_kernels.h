// swm_kernels.h -- argument bundles of the device kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace swm {
namespace dev {

// grid extents and metric planes of the model level
struct Geo {
  int iim, jjm, ns;
  const double *area, *area_inv, *cai, *sub_area, *dc, *nrm, *tng, *vw, *sw, *lw, *f, *hs, *mask;
};

// one multigrid level: Laplacian weights, reciprocal central weight and the
// closed-form halo map (src[slot] = slot whose value the halo slot holds)
struct Lev {
  int iim, jjm, ns;
  const double* lw;
  const double* lwt;
  const int* src;
};

struct FieldSet {
  double* p[8];
};

}  // namespace dev
}  // namespace swm
