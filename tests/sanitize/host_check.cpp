// host_check.cpp -- runs the host half of the library (grid set-up, halo
// tables, metrics, multigrid weights, restart codec, test case 5, namelist
// reader) and the oracle (the CPU FP64 restatement of the reference step)
// under AddressSanitizer and UndefinedBehaviorSanitizer; built and run by
// tests/test_sanitize_host.py.  Exits non-zero on any failure.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/swm_mi355x.h"
#include "../../oracle/swm_oracle.h"

static void check(bool ok, const char* what) {
  if (!ok) {
    std::fprintf(stderr, "host_check: %s failed: %s\n", what, swm_last_error());
    std::exit(1);
  }
}

template <class T>
static std::vector<T> arr(const swm_grid* g, const char* name) {
  const int64_t n = swm_grid_array_len(g, name);
  check(n >= 0, name);
  std::vector<T> v((size_t)n);
  check(swm_grid_array(g, name, v.data()) == SWM_OK, name);
  return v;
}

int main(int argc, char** argv) {
  const std::string tmp = argc > 1 ? argv[1] : "/tmp";
  const char* dnames[] = {"p", "c", "area", "area_inv", "corner_area", "corner_area_inv", "sub_area", "dc", "normal",
                          "tangent", "v_weights", "s_weights", "l_weights", "f", "grid_center_mask",
                          "grid_corner_mask", "polygon_type", "tot_area"};
  const char* inames[] = {"grid_center_index", "grid_corner_index", "face_dst", "face_src", "vrtx_dst", "vrtx_src",
                          "edge_dst", "edge_src"};
  for (int g = 1; g <= 5; ++g) {
    swm_grid* grid = nullptr;
    check(swm_grid_create(g, nullptr, 6.37122e6, 7.292e-5, &grid) == SWM_OK, "swm_grid_create");
    for (const char* n : dnames) (void)arr<double>(grid, n);
    for (const char* n : inames) (void)arr<int32_t>(grid, n);
    for (int l = 1; l <= g; ++l) (void)arr<double>(grid, ("lap_w" + std::to_string(l)).c_str());
    // the remapping quantities (init_remap), built on first use
    for (const char* n : {"w", "trans_to_X1", "trans_to_X2", "remap_c_x3", "remap_gwgt_x3", "remap_trans_X4_to_X3",
                          "remap_p_x4", "remap_area_inv_x3", "remap_garea_inv"})
      (void)arr<double>(grid, n);
    check(swm_grid_array_len(grid, "no_such_array") == -1, "unknown array");
    swm_grid_destroy(grid);
  }
  check(swm_grid_create(0, nullptr, 1.0, 0.0, nullptr) != SWM_OK, "bad level refused");

  // level 3: the oracle and the host codecs
  const int g = 3, nt = 2;
  swm_grid* grid = nullptr;
  check(swm_grid_create(g, nullptr, 6.37122e6, 7.292e-5, &grid) == SWM_OK, "grid g=3");
  int gg, iim, jjm, max_ig;
  check(swm_grid_dims(grid, &gg, &iim, &jjm, &max_ig) == SWM_OK, "dims");
  const size_t ns = (size_t)iim * jjm * 10;
  std::vector<double> h(ns), hs(ns), u(4 * ns), tr((size_t)nt * ns);
  check(swm_swtc5_state(grid, nt, 9.80616, h.data(), hs.data(), u.data(), tr.data()) == SWM_OK, "swtc5");
  const std::string rpath = tmp + "/host_check_restart";
  check(swm_restart_file_write(grid, rpath.c_str(), nt, 1800, h.data(), u.data(), tr.data()) == SWM_OK,
        "restart write");
  std::vector<double> h2(ns), u2(4 * ns), t2((size_t)nt * ns);
  long long ts = 0;
  check(swm_restart_file_read(grid, rpath.c_str(), nt, &ts, h2.data(), u2.data(), t2.data()) == SWM_OK,
        "restart read");
  check(ts == 1800, "restart time stamp");
  char name[64];
  check(swm_restart_name(86400 + 3661, name, sizeof name) == SWM_OK, "restart name");
  const std::string npath = tmp + "/host_check_namelist";
  FILE* f = std::fopen(npath.c_str(), "w");
  check(f != nullptr, "namelist file");
  std::fputs("   5.00   0.00 200.00\n T\n F\n F\n T\n F\n 10 10 1\n", f);
  std::fclose(f);
  swm_namelist nl;
  (void)swm_read_namelist(npath.c_str(), &nl);

  auto area = arr<double>(grid, "area"), area_inv = arr<double>(grid, "area_inv"),
       cai = arr<double>(grid, "corner_area_inv"), sub = arr<double>(grid, "sub_area"), dc = arr<double>(grid, "dc"),
       nrm = arr<double>(grid, "normal"), tng = arr<double>(grid, "tangent"), vw = arr<double>(grid, "v_weights"),
       sw = arr<double>(grid, "s_weights"), lw = arr<double>(grid, "l_weights"), fc = arr<double>(grid, "f"),
       mask = arr<double>(grid, "grid_center_mask");
  auto fd = arr<int32_t>(grid, "face_dst"), fs = arr<int32_t>(grid, "face_src"), vd = arr<int32_t>(grid, "vrtx_dst"),
       vs = arr<int32_t>(grid, "vrtx_src"), ed = arr<int32_t>(grid, "edge_dst"), es = arr<int32_t>(grid, "edge_src");
  std::vector<std::vector<double>> lap(g);
  std::vector<const double*> lp(g);
  for (int l = 1; l <= g; ++l) {
    lap[l - 1] = arr<double>(grid, ("lap_w" + std::to_string(l)).c_str());
    lp[l - 1] = lap[l - 1].data();
  }
  double tot = 0.0;
  check(swm_grid_array(grid, "tot_area", &tot) == SWM_OK, "tot_area");
  for (int scheme = 0; scheme < 3; ++scheme) {
    orc_grid og{};
    og.g = g; og.ntracer = nt; og.dt = 1800.0; og.grav = 9.80616; og.scheme = scheme;
    og.area = area.data(); og.area_inv = area_inv.data(); og.corner_area_inv = cai.data(); og.sub_area = sub.data();
    og.dc = dc.data(); og.normal = nrm.data(); og.tangent = tng.data(); og.v_weights = vw.data();
    og.s_weights = sw.data(); og.l_weights = lw.data(); og.f = fc.data(); og.hs = hs.data();
    og.center_mask = mask.data();
    og.nface = (int)fd.size(); og.nvrtx = (int)vd.size(); og.nedge = (int)ed.size();
    og.face_dst = fd.data(); og.face_src = fs.data(); og.vrtx_dst = vd.data(); og.vrtx_src = vs.data();
    og.edge_dst = ed.data(); og.edge_src = es.data();
    og.lap_w = lp.data(); og.tot_area = tot;
    orc_model* m = orc_create(&og);
    check(m != nullptr, "orc_create");
    orc_set_sum_mode(m, scheme == 1);
    orc_init(m, h.data(), u.data(), tr.data());
    check(orc_step(m, 3) == 0, "orc_step");
    std::vector<double> out(4 * ns), b(8 + nt), e(3 * ns), s1(ns), s2(ns);
    check(orc_get(m, "u", out.data()) == 0 && orc_get(m, "tracer", out.data()) == 0, "orc_get");
    orc_budgets(m, b.data());
    orc_get(m, "h", s1.data());
    orc_laplacian(m, s1.data(), s2.data());
    orc_div_of_an_edge_flux(m, out.data(), s2.data());
    orc_average_to_corners(m, s1.data(), e.data());
    orc_wrap(m, 0, 1, s1.data());
    orc_wrap(m, 2, 1, e.data());
    std::vector<double> rhs(2 * ns), x(2 * ns, 0.0);
    orc_get(m, "relative", rhs.data());
    orc_get(m, "div", rhs.data() + ns);
    (void)orc_elliptic(m, 2, rhs.data(), x.data());
    // the two FCT routines under their own arguments
    std::vector<double> hc(2 * ns), hist(9 * ns, 1.0), low(ns), flux(3 * ns), hn(ns), tf(ns);
    orc_average_to_corners(m, h.data(), hc.data());
    const int ti[3] = {2, 3, 1};
    const double tw[3] = {23.0 / 12.0, -16.0 / 12.0, 5.0 / 12.0};
    orc_set_mass_fct(m, ti, tw, h.data(), hc.data(), u.data(), hist.data(), low.data(), flux.data(), hn.data());
    orc_set_tracer_fct(m, tr.data(), tr.data(), h.data(), low.data(), hn.data(), flux.data(), tf.data());
    for (size_t k = 0; k < ns; ++k) check(std::isfinite(tf[k]), "finite tracer");
    orc_destroy(m);
  }
  swm_grid_destroy(grid);
  std::remove(rpath.c_str());
  std::remove(npath.c_str());
  std::printf("host_check ok\n");
  return 0;
}
