// swm_internal.h -- declarations shared by the host set-up and the device code.
#pragma once
#include <cstdint>
#include <string>

#include "../../include/swm_mi355x.h"

#include "swm_setup.h"

struct swm_grid;

namespace swm {
void set_error(const std::string& s);
const Grid& grid_of(const swm_grid* g);

// global cell numbering of the slots (grid_center_index, 1-based) and the
// real-cell mask, as the restart files need them
struct CellIndex {
  int iim, jjm, max_ig;
  const int32_t* index;
  const double* mask;
};
inline CellIndex cell_index(const Grid& G) {
  return CellIndex{G.iim, G.jjm, G.max_ig, G.center_index.data(), G.center_mask.data()};
}
swm_status restart_write(const CellIndex& G, const char* path, int nt, long long nsecond, const double* h,
                         const double* u, const double* tracer);
swm_status restart_read(const CellIndex& G, const char* path, int nt, long long* nsecond, double* h, double* u,
                        double* tracer);
swm_status restart_name(long long nsecond, char* buf, int len);
}  // namespace swm
