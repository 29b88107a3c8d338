// swm_setup.h -- host-side (FP64) set-up of the icosahedral grid consumed by the
// device step: grid points, panel/pole index maps, halo gather tables, masks,
// operator metrics and the multigrid hierarchy.
//
// Arrays use the reference's Fortran (column-major) index order so that they
// can be compared with the reference slot by slot; the device upload turns the
// multi-component ones into separate planes.
#pragma once
#include <cstdint>
#include <memory>
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

// Grid quantities of incremental remapping (init_remap, remap.F:176-1063),
// in the reference layout with the slot (iim, jjm, 10) last: X3 is the local
// east/north/up frame at cell centres, X4 the tangent/outward-normal/up frame
// at cell edges (origin at the edge midpoint), distances in metres.
struct RemapGeometry {
  std::vector<double> vertex_mask;      // (6, slot) 1 for real vertices
  std::vector<double> num_vertices;     // (slot)
  std::vector<double> trans_x1_to_x3;   // (3, 3, slot)
  std::vector<double> trans_x1_to_x4;   // (3, 3, 3 edges, slot)
  std::vector<double> trans_x4_to_x3;   // (2, 2, 4 cells, 3 edges, slot)
  std::vector<double> cl_x4, cr_x4, tl_x4, tr_x4, bl_x4, br_x4;   // (2, 3 edges, slot)
  std::vector<double> p_x4;             // (2, 4 cells, 3 edges, slot)
  std::vector<double> c_x3;             // (2, 6, slot) corners relative to the centre
  std::vector<double> area_x3, area_inv_x3, xav, yav, xxav, xyav, yyav, xxxav, xxyav, xyyav, yyyav;  // (slot)
  std::vector<double> gwgt_x3;          // (3, 0:6, slot) cell-centre gradient weights
  std::vector<double> garea_inv;        // (slot)
};

// Geometry of the dissipation operator (init_diss_op,
// swm/source/swm_dissipation_operator.F:50-203) in the reference layout:
// t (2, ndata, nspace, ncorners, slot) unit vectors in the corners' X2
// frames, magt (2, ndata, ncorners, slot) their lengths and magn (ndata,
// ncorners, slot) the centre-to-centre distances across the edges, in metres.
struct DissGeometry {
  std::vector<double> t, magt, magn;
};
void build_diss_geometry(const struct Grid& G, DissGeometry& D);

// the four cells around edge ne (0-based) of cell (i, j): offsets of cell
// nd = 0..3 (remap.F ishift_p / jshift_p)
extern const int kRemapShiftI[3][4];
extern const int kRemapShiftJ[3][4];

struct Grid {
  int g = 0, iim = 0, jjm = 0, nsdm = 10, im = 0, jm = 0, max_ig = 0;
  double radius = 6.37122e6, omega = 7.292e-5;
  // grid
  std::vector<double> p_list;          // (3, max_ig) cell centres, global order
  std::vector<double> p;               // (3, iim, jjm, 10)
  std::vector<double> c;               // (3, 2, iim, jjm, 10) Voronoi corners
  std::vector<double> w;               // (3, 3, 2, iim, jjm, 10) edge midpoints (ndata, xyz, corner)
  std::vector<double> trans_to_x2;     // (3, 3, 2, iim, jjm, 10) X1 -> local east/north/up at corners
  std::vector<double> trans_to_x1;     // (3, 3, 2, iim, jjm, 10) its inverse (transpose)
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
  // remapping quantities, built on first use (remap_geometry)
  mutable std::shared_ptr<const RemapGeometry> remap;
};

// init_remap for the grid.
void build_remap_geometry(const Grid& G, RemapGeometry& R);
// The grid's remapping quantities, built on the first call (thread-safe).
const RemapGeometry& remap_geometry(const Grid& G);

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
