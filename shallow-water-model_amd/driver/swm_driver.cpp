// swm_driver -- the reference's stand-alone driver (swm/source/driver.F:
// 40-166) on the device library: run settings from data/namelist
// (model_control.F:98-139), the grid from data/grid_centers_NNNNNNN
// (grid_metrics_generator.F:1073-1136; generated when absent), the
// shallow-water test case 5 initial state (initialize.F:330-513) or, with
// l_restart, restarts/swm_restart (initialize.F:222-231), then the time loop
// with the time manager's flags (time_manager_module.F:82-113): step, restart
// files every n_restart steps, min/max fields (step.F:780-826) and budgets
// (budgets.F:31-247, stdout and the real*4 records of fort.31) every
// n_diagnostics steps, until the model time passes tau_end.
//
//   swm_driver [-g G] [-precision fp64|mixed] [-ntracer N] [-device D]
//
// run in a reference run directory (data/, restarts/).  The quick-plot
// netCDF file of swm_qp_module.F (qp_output/hswm_d000000p000.g2.nc, the
// directory made when absent) holds the initial state and every n_output-th
// step, as initialize.F:153-159, :275-276 and driver.F:133-135 write it.
// l_remap selects incremental remapping (FP64).  Not supported, as outside
// the library: the OpenDX files of l_dx (dx_write_utility.F), reported once
// and skipped.
#include <dirent.h>
#include <sys/stat.h>

#include <cerrno>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/swm_mi355x.h"

namespace {

[[noreturn]] void die(const std::string& what) {
  std::fprintf(stderr, "swm_driver: %s\n", what.c_str());
  std::exit(1);
}

void check(swm_status st, const char* where) {
  if (st != SWM_OK) die(std::string(where) + ": " + swm_last_error());
}

// Fortran Ew.d output (mantissa in [0.1, 1), two exponent digits)
std::string fortran_e(double v, int w, int d) {
  char buf[64];
  if (v == 0.0 || !std::isfinite(v)) {
    if (!std::isfinite(v)) std::snprintf(buf, sizeof buf, "%*s", w, std::string(w, '*').c_str());
    else std::snprintf(buf, sizeof buf, "%*s", w, ("0." + std::string(d, '0') + "E+00").c_str());
    return buf;
  }
  char sci[64];
  std::snprintf(sci, sizeof sci, "%.*E", d - 1, std::fabs(v));   // d.ddddE+xx
  std::string digits;
  for (const char* p = sci; *p && *p != 'E'; ++p)
    if (*p != '.') digits += *p;
  const int e = std::atoi(std::strchr(sci, 'E') + 1) + 1;
  std::snprintf(buf, sizeof buf, "%s0.%sE%c%02d", v < 0 ? "-" : "", digits.c_str(), e < 0 ? '-' : '+', std::abs(e));
  std::string s(buf);
  if ((int)s.size() < w) s = std::string(w - s.size(), ' ') + s;
  return s;
}

int level_of_cells(long n) {
  for (int g = 1; g <= 10; ++g)
    if (10L * (1L << (2 * g)) + 2 == n) return g;
  return -1;
}

std::vector<double> read_grid_file(const std::string& path, long n) {
  std::ifstream in(path);
  std::vector<double> p(3 * (size_t)n);
  for (long k = 0; k < 3 * n; ++k)
    if (!(in >> p[(size_t)k])) die("short grid file " + path);
  return p;
}

struct MinMax {
  double lo = INFINITY, hi = -INFINITY;
  void add(double x) {
    lo = std::min(lo, x);
    hi = std::max(hi, x);
  }
};

}  // namespace

int main(int argc, char** argv) {
  int g = -1, ntracer = 5, device = 0, precision = SWM_FP64;
  for (int k = 1; k < argc; ++k) {
    const std::string a = argv[k];
    auto next = [&]() -> std::string {
      if (k + 1 >= argc) die("missing value after " + a);
      return argv[++k];
    };
    if (a == "-g") g = std::atoi(next().c_str());
    else if (a == "-ntracer") ntracer = std::atoi(next().c_str());
    else if (a == "-device") device = std::atoi(next().c_str());
    else if (a == "-precision") {
      const std::string v = next();
      if (v == "fp64") precision = SWM_FP64;
      else if (v == "mixed") precision = SWM_MIXED;
      else die("precision must be fp64 or mixed");
    } else die("unknown argument " + a);
  }

  swm_namelist nl;
  check(swm_read_namelist("data/namelist", &nl), "reading data/namelist");
  if (nl.l_upwind + nl.l_centered + nl.l_fct + nl.l_remap != 1)
    die("exactly one of l_upwind, l_centered, l_fct, l_remap must be set");
  if (nl.l_remap && precision != SWM_FP64) die("l_remap: incremental remapping runs in FP64");
  const double tau_end = 24.0 * nl.tau_end_day + nl.tau_end_hour - 0.001;   // model_control.F:124
  std::printf("\n  input : tau_end        = %.15g\n  input : tau_end_day    = %.15g\n"
              "  input : tau_end_hour   = %.15g\n  input : time step      = %.15g\n\n",
              tau_end, nl.tau_end_day, nl.tau_end_hour, nl.dt);

  // grid: data/grid_centers_NNNNNNN as the reference reads it, else generated
  std::string gfile;
  if (DIR* dd = opendir("data")) {
    while (dirent* e = readdir(dd)) {
      const std::string nm = e->d_name;
      if (nm.rfind("grid_centers_", 0) == 0) {
        const int gl = level_of_cells(std::atol(nm.c_str() + 13));
        if (gl > 0 && (g < 0 || gl == g)) { gfile = "data/" + nm; g = gl; }
      }
    }
    closedir(dd);
  }
  if (g < 1) die("no level: pass -g or provide data/grid_centers_NNNNNNN");
  const long max_ig = 10L * (1L << (2 * g)) + 2;
  std::vector<double> pts;
  if (!gfile.empty()) {
    pts = read_grid_file(gfile, max_ig);
    std::printf("  grid: %s\n", gfile.c_str());
  } else {
    std::printf("  grid: generated, level %d\n", g);
  }
  const double radius = 6.37122e6, omega = 7.292e-5, grav = 9.80616;   // params_physical.F:45-52
  swm_grid* grid = nullptr;
  check(swm_grid_create(g, pts.empty() ? nullptr : pts.data(), radius, omega, &grid), "swm_grid_create");
  int iim = 0, jjm = 0, mig = 0;
  check(swm_grid_dims(grid, nullptr, &iim, &jjm, &mig), "swm_grid_dims");
  const size_t ns = (size_t)iim * jjm * 10;
  std::vector<double> cmask(ns), kmask(2 * ns);
  check(swm_grid_array(grid, "grid_center_mask", cmask.data()), "grid_center_mask");
  check(swm_grid_array(grid, "grid_corner_mask", kmask.data()), "grid_corner_mask");

  swm_config cfg{};
  cfg.ntracer = ntracer;
  cfg.dt = nl.dt;
  cfg.grav = grav;
  cfg.precision = precision;
  cfg.device = device;
  cfg.advection = nl.l_fct ? SWM_ADV_FCT : nl.l_centered ? SWM_ADV_CENTERED : nl.l_remap ? SWM_ADV_REMAP : SWM_ADV_UPWIND;
  swm_ctx* ctx = nullptr;
  check(swm_create(grid, &cfg, &ctx), "swm_create");

  // initialize.F:153-159: the qp file, named for model time 0 (the clock is
  // set from a restart file only afterwards)
  if (mkdir("qp_output", 0755) != 0 && errno != EEXIST) die("cannot create qp_output/");
  swm_qp* qp = nullptr;
  check(swm_qp_create(grid, "SWTC#5", 0, nullptr, &qp), "initialize_swm_qp");

  // initialize.F:222-231: test case 5, then the restart file over it
  std::printf("\n  generating initial conditions for SW Test Case 5\n\n");
  std::vector<double> h(ns), hs(ns), u(4 * ns), tr((size_t)ntracer * ns);
  check(swm_swtc5_state(grid, ntracer, grav, h.data(), hs.data(), u.data(), tr.data()), "swm_swtc5_state");
  check(swm_init_state(ctx, h.data(), hs.data(), u.data(), tr.data()), "swm_init_state");
  if (nl.l_restart) {
    check(swm_read_restart(ctx, "restarts/swm_restart"), "read_swm_restart");
    std::printf("\n finished reading initial conditions\n\n");
  }
  auto write_qp = [&]() {   // pack_swm_qp, write_swm_qp
    check(swm_qp_pack(qp, ctx), "pack_swm_qp");
    check(swm_qp_write(qp, swm_time_seconds(ctx)), "write_swm_qp");
  };

  FILE* f31 = std::fopen("fort.31", "wb");
  if (!f31) die("cannot open fort.31");
  std::vector<double> b0;
  auto budgets = [&]() {
    std::vector<double> b((size_t)8 + ntracer);
    check(swm_budgets(ctx, b.data(), (int)b.size()), "compute_budgets");
    const double eps = 1.0e-15;   // budgets.F:84
    double te_diff = 0.0, lte_diff = 0.0, penst_diff = 0.0, lpenst_diff = 0.0, m1 = 1.0, m2 = 1.0;
    if (b0.empty()) {
      b0 = b;
    } else {   // budgets.F:197-204
      te_diff = std::fabs(b[4] - b0[4]) / (b0[4] + eps);
      lte_diff = std::log10(te_diff + eps);
      penst_diff = std::fabs(b[5] - b0[5]) / (b0[5] + eps);
      lpenst_diff = std::log10(penst_diff + eps);
      m1 = b[8] / b0[8];
      m2 = ntracer > 1 ? b[9] / b0[9] : 0.0;
    }
    const float r4[14] = {(float)b[0], (float)b[1], (float)b[2], (float)b[3], (float)b[4], (float)te_diff,
                          (float)lte_diff, (float)b[5], (float)penst_diff, (float)lpenst_diff, (float)b[6],
                          (float)b[7], 0.0f, 0.0f};   // the reference's lm1 / lm2 are zero
    const int32_t len = (int32_t)sizeof r4;
    std::fwrite(&len, 4, 1, f31);
    std::fwrite(r4, sizeof r4, 1, f31);
    std::fwrite(&len, 4, 1, f31);
    std::fflush(f31);
    std::printf("\n\n Total Mass : %s\n", fortran_e(b[0], 17, 8).c_str());
    std::printf(" Log10 of Frac Energy Change : %s\n", fortran_e(lte_diff, 17, 8).c_str());
    std::printf(" Log10 of Frac Pot. Enstrophy Change : %s\n\n\n", fortran_e(lpenst_diff, 17, 8).c_str());
    std::printf(" Fractional mass-weighted (tracer 1)^2  : %s\n", fortran_e(m1, 17, 8).c_str());
    std::printf(" Fractional mass-weighted (tracer 2)^2  : %s\n\n\n", fortran_e(m2, 17, 8).c_str());
  };
  auto dump_min_max = [&](double tau) {   // step.F:780-826
    std::printf("\n\n GLOBAL MIN AND MAX OF SELECT FIELDS at time (hours) :%20.5f\n", tau);
    const char* label[4] = {"divergence     ", "rel vorticity  ", "kinetic energy ", "mass thickness "};
    const char* field[4] = {"div", "relative", "ke", "h"};
    std::vector<double> x(4 * ns);
    for (int k = 0; k < 4; ++k) {
      check(swm_get_field(ctx, field[k], x.data()), "swm_get_field");
      MinMax mm;
      for (size_t n = 0; n < ns; ++n)
        if (cmask[n] == 1.0) mm.add(x[n]);
      std::printf(" min/max %s: %s%s\n", label[k], fortran_e(mm.lo, 20, 10).c_str(), fortran_e(mm.hi, 20, 10).c_str());
    }
    check(swm_get_field(ctx, "u", x.data()), "swm_get_field");
    for (int c = 0; c < 2; ++c) {
      MinMax mm;
      for (size_t k = 0; k < 2 * ns; ++k)
        if (kmask[k] == 1.0) mm.add(x[2 * k + c]);
      std::printf(" min/max %s: %s%s\n", c == 0 ? "eastward wind  " : "northward wind ", fortran_e(mm.lo, 20, 10).c_str(),
                  fortran_e(mm.hi, 20, 10).c_str());
    }
    std::printf("\n\n");
  };
  budgets();   // initialize.F:236
  write_qp();  // initialize.F:275-276

  bool warned_dx = false, halt = false;
  int step_count = 0;
  const auto t0 = std::chrono::steady_clock::now();
  while (!halt) {
    // time_manager_module.F:82-113 (the library advances the clock in swm_step)
    ++step_count;
    check(swm_step(ctx, 1), "step_dt");
    const long long nsec = swm_time_seconds(ctx);
    const double tau = (double)((float)nsec / 3600.0f);   // a default-real quotient in the reference
    halt = tau > tau_end;
    const bool l_restart = nl.n_restart > 0 && step_count % nl.n_restart == 0;
    const bool l_output = nl.n_output > 0 && step_count % nl.n_output == 0;
    const bool l_diag = nl.n_diagnostics > 0 && step_count % nl.n_diagnostics == 0;
    std::printf(" %4d %6lld %24.15E\n", step_count, nsec, tau);
    if (l_diag) dump_min_max(tau);
    if (l_restart) {
      char name[64];
      check(swm_restart_name(nsec, name, (int)sizeof name), "swm_restart_name");
      check(swm_write_restart(ctx, name), "write_swm_restart");
      std::printf("\n finished writing restart file %s\n\n", name);
    }
    if (l_output) {
      write_qp();
      if (!warned_dx) {
        std::printf("\n  l_output: OpenDX output is not part of the device library; skipped\n\n");
        warned_dx = true;
      }
    }
    if (l_diag) budgets();
  }
  const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::fclose(f31);
  std::printf("\n swm_driver: %d steps at level %d (%ld cells), %.3f s, %.4e cell-steps/s\n", step_count, g, max_ig,
              wall, (double)max_ig * step_count / wall);
  std::printf(" qp output: %d records in %s\n", swm_qp_records(qp), swm_qp_path(qp));
  swm_qp_destroy(qp);
  swm_destroy(ctx);
  swm_grid_destroy(grid);
  return 0;
}
