// swm_setup.h -- host-side (FP64) set-up of the icosahedral grid consumed by the
// device step: grid points, panel/pole index maps, halo gather tables, masks,
// operator metrics and the multigrid hierarchy.
//
// Arrays use the reference's Fortran (column-major) index order so that they
// can be compared with the reference slot by slot; the device upload turns the
// multi-component ones into separate planes.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace swm {

// Halo fill of one field kind as a gather: x[dst[k]] = (src[k] < 0 ? 0 : x[src[k]]).
// Indices are flat Fortran offsets into (ncomp, iim, jjm, 10) with the component
// fastest (ncomp = 1 for cell data, 2 for corner data, 3 for edge data).
// No source is ever a destination, so the gather may run in place.
struct HaloTable {
  int ncomp = 1;
  std::vector<int32_t> dst, src;
};

struct Level {   // one multigrid level (the finest level equals the model grid)
  int g = 0, iim = 0, jjm = 0;
  std::vector<double> lap_w;        // (7, iim, jjm, 10)
  HaloTable face;                   // general face halo table of this level
};

struct Grid {
  int g = 0, iim = 0, jjm = 0, nsdm = 10, im = 0, jm = 0, max_ig = 0;
  double radius = 6.37122e6, omega = 7.292e-5;
  // grid
  std::vector<double> p_list;          // (3, max_ig) cell centres, global order
  std::vector<double> p;               // (3, iim, jjm, 10)
  std::vector<double> c;               // (3, 2, iim, jjm, 10) Voronoi corners
  // metrics (all in metres after scaling by the radius)
  std::vector<double> area, area_inv;                // (iim, jjm, 10)
  std::vector<double> corner_area, corner_area_inv;  // (2, iim, jjm, 10)
  std::vector<double> sub_area, dc;                  // (3, 2, iim, jjm, 10)
  std::vector<double> normal, tangent;               // (3, 3, 2, iim, jjm, 10)
  std::vector<double> v_weights;                     // (3, 2, 2, iim, jjm, 10)
  std::vector<double> s_weights;                     // (3, 2, iim, jjm, 10)
  std::vector<double> l_weights;                     // (7, iim, jjm, 10)
  std::vector<double> f;                             // (iim, jjm, 10) Coriolis
  double tot_area = 0.0;
  // masks and index maps
  std::vector<double> center_mask, polygon_type;     // (iim, jjm, 10)
  std::vector<double> corner_mask;                   // (2, iim, jjm, 10)
  std::vector<int32_t> center_index;                 // (iim, jjm, 10) 1-based, 0 = none
  std::vector<int32_t> corner_index;                 // (2, iim, jjm, 10)
  // halo gather tables
  HaloTable face, vrtx, edge;
  // multigrid hierarchy, levels 1..g at index 0..g-1
  std::vector<Level> levels;
};

// Generates the unit-sphere cell centres of the recursively bisected
// icosahedron at level g (10*4^g+2 points) in the reference's global order.
std::vector<double> generate_points(int g);

// Builds the whole grid from cell centres (p_list == nullptr: generate them).
void build_grid(Grid& G, int g, const double* p_list, double radius, double omega);

// Halo tables of the general face/vertex/edge wrap at level g.
void build_halo_tables(int g, HaloTable& face, HaloTable& vrtx, HaloTable& edge);

// Closed-form face halo of the multigrid levels (10 panels).
HaloTable build_simple_face_table(int g);

// Applies a halo table to a host array of nlayer layers (for tests and set-up).
void apply_halo(const HaloTable& t, double* x, int nlayer, int64_t layer_stride);

// Same for an array whose table entries are interleaved with a faster index of
// extent `stride` (e.g. the 3 coordinates of the corner positions).
void apply_halo_strided(const HaloTable& t, double* x, int stride, int offset);

}  // namespace swm
