// swm_capi_grid.cpp -- C ABI of the host-side grid set-up.
#include <cstring>
#include <exception>
#include <string>
#include <utility>
#include <vector>

#include "../../include/swm_mi355x.h"
#include "swm_internal.h"
#include "swm_setup.h"

struct swm_grid {
  swm::Grid G;
};

namespace {
thread_local std::string g_err;
}

namespace swm {
void set_error(const std::string& s) { g_err = s; }
const Grid& grid_of(const swm_grid* g) { return g->G; }
}  // namespace swm

extern "C" {

const char* swm_last_error(void) { return g_err.c_str(); }

swm_status swm_grid_points(int g, double* p_list) {
  if (g < 1 || g > 10 || !p_list) { swm::set_error("swm_grid_points: bad argument"); return SWM_ERR_ARG; }
  try {
    std::vector<double> p = swm::generate_points(g);
    std::memcpy(p_list, p.data(), p.size() * sizeof(double));
  } catch (const std::exception& e) {
    swm::set_error(e.what());
    return SWM_ERR_ARG;
  }
  return SWM_OK;
}

swm_status swm_grid_create(int g, const double* p_list, double radius, double omega, swm_grid** out) {
  if (g < 1 || g > 10 || !out || !(radius > 0.0)) { swm::set_error("swm_grid_create: bad argument"); return SWM_ERR_ARG; }
  swm_grid* gr = new swm_grid;
  try {
    swm::build_grid(gr->G, g, p_list, radius, omega);
  } catch (const std::exception& e) {
    delete gr;
    swm::set_error(std::string("swm_grid_create: ") + e.what());
    return SWM_ERR_ARG;
  }
  *out = gr;
  return SWM_OK;
}

void swm_grid_destroy(swm_grid* grid) { delete grid; }

swm_status swm_grid_dims(const swm_grid* grid, int* g, int* iim, int* jjm, int* max_ig) {
  if (!grid) return SWM_ERR_ARG;
  if (g) *g = grid->G.g;
  if (iim) *iim = grid->G.iim;
  if (jjm) *jjm = grid->G.jjm;
  if (max_ig) *max_ig = grid->G.max_ig;
  return SWM_OK;
}

namespace {
struct ArrayRef {
  const void* data = nullptr;
  int64_t n = -1;
  int esize = 8;
};

ArrayRef lookup(const swm_grid* grid, const char* name) {
  const swm::Grid& G = grid->G;
  ArrayRef r;
  auto D = [&](const std::vector<double>& v) { r.data = v.data(); r.n = (int64_t)v.size(); r.esize = 8; };
  auto I = [&](const std::vector<int32_t>& v) { r.data = v.data(); r.n = (int64_t)v.size(); r.esize = 4; };
  const std::string s(name);
  if (s == "p") D(G.p);
  else if (s == "c") D(G.c);
  else if (s == "w") D(G.w);
  else if (s == "trans_to_X1") D(G.trans_to_x1);
  else if (s == "trans_to_X2") D(G.trans_to_x2);
  else if (s == "area") D(G.area);
  else if (s == "area_inv") D(G.area_inv);
  else if (s == "corner_area") D(G.corner_area);
  else if (s == "corner_area_inv") D(G.corner_area_inv);
  else if (s == "sub_area") D(G.sub_area);
  else if (s == "dc") D(G.dc);
  else if (s == "normal") D(G.normal);
  else if (s == "tangent") D(G.tangent);
  else if (s == "v_weights") D(G.v_weights);
  else if (s == "s_weights") D(G.s_weights);
  else if (s == "l_weights") D(G.l_weights);
  else if (s == "f") D(G.f);
  else if (s == "grid_center_mask") D(G.center_mask);
  else if (s == "grid_corner_mask") D(G.corner_mask);
  else if (s == "polygon_type") D(G.polygon_type);
  else if (s == "p_list") D(G.p_list);
  else if (s == "tot_area") { r.data = &G.tot_area; r.n = 1; }
  else if (s == "grid_center_index") I(G.center_index);
  else if (s == "grid_corner_index") I(G.corner_index);
  else if (s == "face_dst") I(G.face.dst);
  else if (s == "face_src") I(G.face.src);
  else if (s == "vrtx_dst") I(G.vrtx.dst);
  else if (s == "vrtx_src") I(G.vrtx.src);
  else if (s == "edge_dst") I(G.edge.dst);
  else if (s == "edge_src") I(G.edge.src);
  else if (s.rfind("remap_", 0) == 0) {
    const swm::RemapGeometry& R = swm::remap_geometry(G);
    const std::string t = s.substr(6);
    const std::pair<const char*, const std::vector<double>*> names[] = {
        {"vertex_mask", &R.vertex_mask}, {"num_vertices", &R.num_vertices}, {"trans_X1_to_X3", &R.trans_x1_to_x3},
        {"trans_X1_to_X4", &R.trans_x1_to_x4}, {"trans_X4_to_X3", &R.trans_x4_to_x3}, {"cl_x4", &R.cl_x4},
        {"cr_x4", &R.cr_x4}, {"tl_x4", &R.tl_x4}, {"tr_x4", &R.tr_x4}, {"bl_x4", &R.bl_x4}, {"br_x4", &R.br_x4},
        {"p_x4", &R.p_x4}, {"c_x3", &R.c_x3}, {"area_x3", &R.area_x3}, {"area_inv_x3", &R.area_inv_x3},
        {"xav", &R.xav}, {"yav", &R.yav}, {"xxav", &R.xxav}, {"xyav", &R.xyav}, {"yyav", &R.yyav},
        {"xxxav", &R.xxxav}, {"xxyav", &R.xxyav}, {"xyyav", &R.xyyav}, {"yyyav", &R.yyyav},
        {"gwgt_x3", &R.gwgt_x3}, {"garea_inv", &R.garea_inv}};
    for (const auto& nv : names)
      if (t == nv.first) D(*nv.second);
  } else if (s.rfind("lap_w", 0) == 0) {
    const int l = std::atoi(s.c_str() + 5);
    if (l >= 1 && l <= G.g) D(G.levels[l - 1].lap_w);
  }
  return r;
}
}  // namespace

int64_t swm_grid_array_len(const swm_grid* grid, const char* name) {
  if (!grid || !name) return -1;
  return lookup(grid, name).n;
}

swm_status swm_grid_array(const swm_grid* grid, const char* name, void* out) {
  if (!grid || !name || !out) { swm::set_error("swm_grid_array: bad argument"); return SWM_ERR_ARG; }
  ArrayRef r = lookup(grid, name);
  if (r.n < 0) { swm::set_error(std::string("swm_grid_array: unknown array ") + name); return SWM_ERR_ARG; }
  std::memcpy(out, r.data, (size_t)r.n * r.esize);
  return SWM_OK;
}

}  // extern "C"
