// swm_qp.cpp -- the model's quick-plot ("qp") output: the netCDF file of
// swm/source/swm_qp_module.F, written as utilities/layer_1/netcdf_geodesic.F
// (ginit_nc, gwrt_nc) and utilities/layer_1/netcdf_general_tools.F
// (add_global_atts, define_time, set_time) make it.
//
// There is no netCDF library in this build: the file is written in the
// netCDF classic format (CDF-1, what nf_create with NF_CLOBBER makes),
// big-endian, with the reference's definitions in the reference's order, so
// that a netCDF reader sees the same dimensions, variables, attributes and
// values:
//
//   dimensions  time (record), char_len = 10     netcdf_general_tools.F:244-245
//               grid_cells = grid_size = cells,
//               grid_corners = 6, grid_rank = 1  netcdf_geodesic.F:419-425
//   global atts calendar, institution, history, run, grid, version,
//               extravaratts, total_grid_size    netcdf_general_tools.F:505-531
//   variables   time, char_time(time, char_len),
//               start_period, end_period,
//               period_length                    netcdf_general_tools.F:251-277
//               grid_center_lat/lon(grid_cells),
//               grid_dims(grid_rank),
//               grid_corner_lat/lon(grid_cells,
//               grid_corners)                    netcdf_geodesic.F:431-458
//               the 11 qp fields (time, grid_cells) as float, each with
//               long_name, title, units, alt_units, type, positions,
//               missing_value, mult_factor, contour_min, contour_int
//                                                netcdf_geodesic.F:339-360
//               interp_indx(grid_size)           netcdf_geodesic.F:142
//
// The data offsets are those the netCDF library gives these definitions with
// its default alignment (data right after the header, every variable's slab
// padded to 4 bytes, the record variables interleaved per record in
// definition order) and fill mode on: grid_dims, which the reference defines
// but never writes, holds the int fill value, and the two bytes that pad
// char_time to 12 bytes per record are NUL.
//
// A cell is written when grid_center_mask > 0.5, in the order panel, j, i
// (the Fortran PACK of netcdf_geodesic.F:650, :707); its six corners are
// get_surrounding_corners (utilities/layer_2/geodesic_grid_utilities.F:
// 156-260) of the corner latitudes and longitudes, a pentagon's sixth corner
// repeating its fifth (swm_qp_module.F:167-168).  The fields are real_kind
// (FP32): each value is the FP64 model value rounded to nearest
// (qp_utilities.F qpsum, replace mode, as pack_swm_qp uses it).
//
// As in the reference, every write opens the file, adds one record and
// closes it again (gwrt_nc, netcdf_geodesic.F:177-241), so a run that stops
// leaves a readable file.
#include <sys/stat.h>

#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <exception>
#include <string>
#include <vector>

#include "../../include/swm_mi355x.h"
#include "swm_internal.h"
#include "swm_setup.h"

namespace {

// netCDF classic format constants (the format's published specification)
enum : int32_t { kNcChar = 2, kNcInt = 4, kNcFloat = 5, kNcDouble = 6 };
enum : uint32_t { kTagDimension = 0x0A, kTagVariable = 0x0B, kTagAttribute = 0x0C };
constexpr int32_t kFillInt = -2147483647;

constexpr int kNqp = 11;        // swm_qp_module.F nqp
constexpr int kCorners = 6;     // 3 * ncorners, swm_qp_module.F:146
constexpr int kStrLen = 10;     // netcdf_general_tools.F STRLEN
constexpr float kMissing = 1e36f;   // netcdf_general_tools.F:42 MISSING

// set_attribute_data (swm_qp_module.F:379-509) over the defaults of
// init_qp_type (qp_utilities.F): name, units, description; every field is
// logged, category "xxx", multfac 1, contmin = contint = 1e36, alt_units
// "no qp alternate units"
struct QpField {
  const char* name;
  const char* units;
  const char* descr;
};
constexpr QpField kFields[kNqp] = {
    {"thickness", "meters", "thickness of fliud layer"},
    {"height", "meters", "surface height of fliud layer"},
    {"relative", "1/s", "relative vorticity of fluid"},
    {"divergence", "1/s", "divergence of fluid"},
    {"absolute", "1/s", "absolute vorticity of fliud layer"},
    {"kinetic_energy", "1/s", "kinetic energy of fluid"},
    {"tracer_1", "tracer per unit thickness", "tracer field #1"},
    {"tracer_2", "1/s", "tracer field #2"},
    {"tracer_3", "tracer per unit thickness", "tracer field #3"},
    {"tracer_4", "tracer per unit thickness", "tracer field #4"},
    {"tracer_5", "tracer per unit thickness", "tracer field #5"},
};

// ---- big-endian encoding ---------------------------------------------------
void be32(std::vector<uint8_t>& b, uint32_t v) {
  for (int s = 24; s >= 0; s -= 8) b.push_back((uint8_t)(v >> s));
}
void be64(std::vector<uint8_t>& b, uint64_t v) {
  for (int s = 56; s >= 0; s -= 8) b.push_back((uint8_t)(v >> s));
}
void bef(std::vector<uint8_t>& b, float x) {
  uint32_t v;
  std::memcpy(&v, &x, 4);
  be32(b, v);
}
void bed(std::vector<uint8_t>& b, double x) {
  uint64_t v;
  std::memcpy(&v, &x, 8);
  be64(b, v);
}
void pad4(std::vector<uint8_t>& b) {
  while (b.size() % 4) b.push_back(0);
}
void name(std::vector<uint8_t>& b, const std::string& s) {
  be32(b, (uint32_t)s.size());
  b.insert(b.end(), s.begin(), s.end());
  pad4(b);
}

struct Att {
  std::string name;
  int32_t type = kNcChar;
  uint32_t nelems = 0;
  std::vector<uint8_t> values;   // encoded, unpadded
};
Att text(const std::string& n, const std::string& s) {
  Att a{n, kNcChar, (uint32_t)s.size(), std::vector<uint8_t>(s.begin(), s.end())};
  return a;
}
Att real(const std::string& n, float x) {
  Att a{n, kNcFloat, 1, {}};
  bef(a.values, x);
  return a;
}
Att integer(const std::string& n, int32_t x) {
  Att a{n, kNcInt, 1, {}};
  be32(a.values, (uint32_t)x);
  return a;
}

struct Dim {
  std::string name;
  uint32_t len;   // 0: the record dimension
};
struct Var {
  std::string name;
  std::vector<int> dims;   // C order, slowest first
  std::vector<Att> atts;
  int32_t type;
  int64_t vsize = 0, begin = 0;
  bool record = false;
};

int type_size(int32_t t) { return t == kNcChar ? 1 : t == kNcDouble ? 8 : 4; }

void att_list(std::vector<uint8_t>& b, const std::vector<Att>& atts) {
  if (atts.empty()) {
    be32(b, 0);
    be32(b, 0);
    return;
  }
  be32(b, kTagAttribute);
  be32(b, (uint32_t)atts.size());
  for (const Att& a : atts) {
    name(b, a.name);
    be32(b, (uint32_t)a.type);
    be32(b, a.nelems);
    b.insert(b.end(), a.values.begin(), a.values.end());
    pad4(b);
  }
}

struct Schema {
  std::vector<Dim> dims;
  std::vector<Att> gatts;
  std::vector<Var> vars;
  int64_t header = 0, begin_rec = 0, recsize = 0;

  std::vector<uint8_t> encode(uint32_t numrecs) const {
    std::vector<uint8_t> b = {'C', 'D', 'F', 1};
    be32(b, numrecs);
    be32(b, kTagDimension);
    be32(b, (uint32_t)dims.size());
    for (const Dim& d : dims) {
      name(b, d.name);
      be32(b, d.len);
    }
    att_list(b, gatts);
    be32(b, kTagVariable);
    be32(b, (uint32_t)vars.size());
    for (const Var& v : vars) {
      name(b, v.name);
      be32(b, (uint32_t)v.dims.size());
      for (int d : v.dims) be32(b, (uint32_t)d);
      att_list(b, v.atts);
      be32(b, (uint32_t)v.type);
      be32(b, (uint32_t)v.vsize);
      be32(b, (uint32_t)v.begin);   // CDF-1: 32-bit offsets
    }
    return b;
  }

  // vsize and begin of every variable; false when an offset does not fit
  // the classic format's 32 bits
  bool layout() {
    int nrec = 0;
    for (Var& v : vars) {
      int64_t n = type_size(v.type);
      v.record = false;
      for (int d : v.dims) {
        if (dims[d].len == 0) v.record = true;
        else n *= dims[d].len;
      }
      v.vsize = (n + 3) / 4 * 4;
      nrec += v.record;
    }
    header = (int64_t)encode(0).size();
    int64_t off = header;
    for (Var& v : vars)
      if (!v.record) { v.begin = off; off += v.vsize; }
    begin_rec = off;
    recsize = 0;
    for (Var& v : vars)
      if (v.record) { v.begin = begin_rec + recsize; recsize += v.vsize; }
    // one record variable alone is not padded (not the case of this file)
    if (nrec == 1)
      for (Var& v : vars)
        if (v.record) recsize = v.vsize;
    for (const Var& v : vars)
      if (v.begin > INT32_MAX || v.vsize > (int64_t)UINT32_MAX - 3) return false;
    return true;
  }
};

// current_time (netcdf_general_tools.F:448-470): "mm/dd/yyyy at hh:mm zone GMT "
// with the local time and its zone as date_and_time gives them ("+hhmm")
std::string current_time() {
  std::time_t t = std::time(nullptr);
  std::tm lt;
  localtime_r(&t, &lt);
  const long off = lt.tm_gmtoff / 60;
  char buf[64];
  std::snprintf(buf, sizeof buf, "%02d/%02d/%04d at %02d:%02d %c%02ld%02ld GMT ", lt.tm_mon + 1, lt.tm_mday,
                lt.tm_year + 1900, lt.tm_hour, lt.tm_min, off < 0 ? '-' : '+', std::labs(off) / 60,
                std::labs(off) % 60);
  std::string s(buf);
  s.resize(30, ' ');
  return s;
}

// xyz_to_lonlat (geodesic_grid_utilities.F:686-732)
void lonlat(const double* p, double* s) {
  s[1] = std::asin(p[2]);
  s[0] = (p[0] != 0.0 || p[1] != 0.0) ? std::atan2(p[1], p[0]) : 0.0;
}

// Fortran I<w> of v into w characters (asterisks when it does not fit)
std::string fortran_int(long long v, int w) {
  char buf[32];
  std::snprintf(buf, sizeof buf, "%*lld", w, v);
  std::string s(buf);
  if ((int)s.size() > w) s.assign(w, '*');
  return s;
}

bool write_at(FILE* f, int64_t off, const void* p, size_t n) {
  return fseeko(f, (off_t)off, SEEK_SET) == 0 && std::fwrite(p, 1, n, f) == n;
}

}  // namespace

struct swm_qp {
  std::string path;
  Schema schema;
  int64_t ns = 0;                      // slots of a field, iim * jjm * 10
  std::vector<int64_t> slots;          // the written cells' slots, in file order
  std::vector<float> data;             // (kNqp, cells): the packed fields
  int32_t timerec = 0;                 // records written
  long long nsecondinit = 0;
};

namespace {

swm_status create(const swm::Grid& G, const char* tag, long long nsecond, const char* path, const double* clat,
                  const double* clon, const double* crlat, const double* crlon, swm_qp** out) {
  const int iim = G.iim, jjm = G.jjm;
  const int64_t ns = (int64_t)iim * jjm * 10;
  swm_qp* q = new swm_qp;
  q->ns = ns;
  for (int64_t n = 0; n < ns; ++n)
    if (G.center_mask[n] > 0.5) q->slots.push_back(n);
  const int64_t npts = (int64_t)q->slots.size();

  // coordinates (swm_qp_module.F:155-172)
  std::vector<double> lat(npts), lon(npts), corner_lat(kCorners * npts), corner_lon(kCorners * npts);
  std::vector<int32_t> indx(npts);
  for (int64_t c = 0; c < npts; ++c) {
    const int64_t n = q->slots[c];
    lat[c] = clat[n];
    lon[c] = clon[n];
    indx[c] = G.center_index[n];
    const int i = (int)(n % iim), j = (int)(n / iim % jjm);
    if (i + 1 >= iim || j + 1 >= jjm) {
      delete q;
      swm::set_error("swm_qp_create: a real cell on the last row or column of its panel");
      return SWM_ERR_STATE;
    }
    const int64_t panel = n / ((int64_t)iim * jjm) * iim * jjm;
    auto at = [&](int ic, int di, int dj) { return 2 * (panel + (i + di) + (int64_t)iim * (j + dj)) + ic; };
    // get_surrounding_corners (geodesic_grid_utilities.F:207-256), corner
    // index ic 0/1, slot offsets (di, dj)
    int64_t k[kCorners];
    const double type = G.polygon_type[n];
    k[0] = at(0, 0, 0);
    k[1] = at(1, 0, 0);
    if (type == 4.0) {            // south pole
      k[2] = at(1, 1, 1); k[3] = at(0, 1, 1); k[4] = at(1, 0, 1);
    } else {
      k[2] = at(0, 1, 0); k[3] = at(1, 1, 1);
      k[4] = type == 2.0 ? at(1, 0, 1) : at(0, 1, 1);
    }
    k[5] = type == 1.0 ? at(1, 0, 1) : -1;   // pentagons: the fifth corner again
    for (int m = 0; m < kCorners; ++m) {
      const int64_t src = k[m] >= 0 ? k[m] : k[4];
      corner_lat[kCorners * c + m] = crlat[src];
      corner_lon[kCorners * c + m] = crlon[src];
    }
  }

  // definitions (ginit_nc, netcdf_geodesic.F:58-171)
  Schema& S = q->schema;
  S.dims = {{"time", 0}, {"char_len", (uint32_t)kStrLen}, {"grid_cells", (uint32_t)npts},
            {"grid_size", (uint32_t)npts}, {"grid_corners", (uint32_t)kCorners}, {"grid_rank", 1}};
  enum { dTime, dChar, dCells, dSize, dCorners, dRank };
  std::string run = tag ? tag : "";
  if (run.size() > 40) run.resize(40);   // character*40 simulation_tag
  while (!run.empty() && run.back() == ' ') run.pop_back();
  S.gatts = {text("calendar", "noleap"),
             text("institution", "Colorado State University"),
             text("history", "Created: " + current_time()),
             text("run", run),
             text("grid", "geodesic"),
             text("version", "0.6"),
             text("extravaratts", "yes"),
             integer("total_grid_size", G.max_ig)};
  const std::string days = "days since 1-1-1";
  S.vars.push_back({"time", {dTime}, {text("quantity", "time"), text("units", days), text("calendar", "noleap")},
                    kNcDouble});
  S.vars.push_back({"char_time", {dTime, dChar}, {text("format", "mm/dd/yyyy")}, kNcChar});
  S.vars.push_back({"start_period", {dTime}, {text("long_name", "start of averaged period"), text("units", days)},
                    kNcDouble});
  S.vars.push_back({"end_period", {dTime}, {text("long_name", "end of averaged period"), text("units", days)},
                    kNcDouble});
  S.vars.push_back({"period_length", {dTime}, {text("long_name", "length of averaged period"), text("units", "days")},
                    kNcDouble});
  S.vars.push_back({"grid_center_lat", {dCells}, {text("units", "radians")}, kNcDouble});
  S.vars.push_back({"grid_center_lon", {dCells}, {text("units", "radians")}, kNcDouble});
  S.vars.push_back({"grid_dims", {dRank}, {}, kNcInt});
  S.vars.push_back({"grid_corner_lat", {dCells, dCorners}, {text("units", "radians")}, kNcDouble});
  S.vars.push_back({"grid_corner_lon", {dCells, dCorners}, {text("units", "radians")}, kNcDouble});
  for (const QpField& f : kFields)
    S.vars.push_back({f.name,
                      {dTime, dCells},
                      {text("long_name", f.descr), text("title", f.descr), text("units", f.units),
                       text("alt_units", "no qp alternate units"), text("type", "xxx"), text("positions", "center"),
                       real("missing_value", kMissing), real("mult_factor", 1.0f), real("contour_min", 1e36f),
                       real("contour_int", 1e36f)},
                      kNcFloat});
  S.vars.push_back({"interp_indx", {dSize}, {}, kNcInt});
  if (!S.layout()) {
    delete q;
    swm::set_error("swm_qp_create: the grid is too large for the netCDF classic format's 32-bit offsets");
    return SWM_ERR_DIMS;
  }

  // header and the non-record data (nf_enddef, init_geo_coord; interp_indx
  // is the reference's first gwrt_nc, written here already)
  std::vector<uint8_t> b = S.encode(0);
  auto doubles = [&](const std::vector<double>& v) { for (double x : v) bed(b, x); };
  for (const Var& v : S.vars) {
    if (v.record) continue;
    if ((int64_t)b.size() != v.begin) { delete q; swm::set_error("swm_qp_create: internal layout error"); return SWM_ERR_STATE; }
    if (v.name == "grid_center_lat") doubles(lat);
    else if (v.name == "grid_center_lon") doubles(lon);
    else if (v.name == "grid_dims") be32(b, (uint32_t)kFillInt);
    else if (v.name == "grid_corner_lat") doubles(corner_lat);
    else if (v.name == "grid_corner_lon") doubles(corner_lon);
    else if (v.name == "interp_indx") for (int32_t x : indx) be32(b, (uint32_t)x);
    pad4(b);
  }
  q->path = path;
  FILE* f = std::fopen(path, "wb");
  if (!f) {
    swm::set_error(std::string("swm_qp_create: cannot create ") + path + ": " + std::strerror(errno));
    delete q;
    return SWM_ERR_ARG;
  }
  const bool ok = std::fwrite(b.data(), 1, b.size(), f) == b.size();
  if (std::fclose(f) != 0 || !ok) {
    swm::set_error(std::string("swm_qp_create: write error on ") + path);
    delete q;
    return SWM_ERR_STATE;
  }
  q->data.assign((size_t)kNqp * npts, 0.0f);
  q->nsecondinit = nsecond;
  *out = q;
  return SWM_OK;
}

// convert_time (netcdf_general_tools.F:354-444): "mm/dd/yyyy" of a model time
// in the no-leap calendar, year 0 first
std::string char_time(long long secs) {
  static const int start[12] = {0, 31, 59, 90, 120, 151, 181, 212, 243, 273, 304, 334};
  const long long year = secs / (365LL * 86400LL);
  int day = (int)(secs % (365LL * 86400LL) / 86400LL) + 1;
  int month = 12;
  while (month > 1 && day <= start[month - 1]) --month;
  day -= start[month - 1];
  // (I2.2, I4.4) read back as A2, A4
  char mm[3];
  std::snprintf(mm, sizeof mm, "%02d", month);
  std::string cday = fortran_int(day, 2), cyear = fortran_int(year, 4);
  for (char& ch : cday) if (ch == ' ') ch = '0';
  for (char& ch : cyear) if (ch == ' ') ch = '0';
  return std::string(mm) + "/" + cday + "/" + cyear;
}

}  // namespace

extern "C" {

swm_status swm_qp_name(long long nsecond, int task, char* buf, int len) {
  if (!buf || nsecond < 0 || task < 0) { swm::set_error("swm_qp_name: bad argument"); return SWM_ERR_ARG; }
  // swm_qp_module.F:130-136, :176: '(i7)' of 1000000 + day and '(i4)' of
  // 1000 + task, each with its first character replaced
  std::string t = fortran_int(1000000LL + nsecond / 86400, 7), p = fortran_int(1000LL + task, 4);
  t[0] = 'd';
  p[0] = 'p';
  const std::string s = "qp_output/hswm_" + t + p + ".g2.nc";
  if ((int)s.size() + 1 > len) { swm::set_error("swm_qp_name: buffer too short"); return SWM_ERR_ARG; }
  std::memcpy(buf, s.c_str(), s.size() + 1);
  return SWM_OK;
}

swm_status swm_qp_create_coords(const swm_grid* grid, const char* simulation_tag, long long nsecond,
                                const char* path, const double* center_lat, const double* center_lon,
                                const double* corner_lat, const double* corner_lon, swm_qp** out) {
  if (!grid || !out || !center_lat || !center_lon || !corner_lat || !corner_lon || nsecond < 0) {
    swm::set_error("swm_qp_create: bad argument");
    return SWM_ERR_ARG;
  }
  char name[64];
  if (!path) {
    swm_status st = swm_qp_name(nsecond, 0, name, sizeof name);
    if (st) return st;
    path = name;
  }
  try {
    return create(swm::grid_of(grid), simulation_tag, nsecond, path, center_lat, center_lon, corner_lat, corner_lon,
                  out);
  } catch (const std::exception& e) {
    swm::set_error(std::string("swm_qp_create: ") + e.what());
    return SWM_ERR_STATE;
  }
}

swm_status swm_qp_create(const swm_grid* grid, const char* simulation_tag, long long nsecond, const char* path,
                         swm_qp** out) {
  if (!grid) { swm::set_error("swm_qp_create: bad argument"); return SWM_ERR_ARG; }
  // grid_center_lat/lon and grid_corner_lat/lon from the (wrapped) cell
  // centres and corners (utilities/layer_3/grid_metrics.F:194-206)
  const swm::Grid& G = swm::grid_of(grid);
  const size_t ns = (size_t)G.iim * G.jjm * 10;
  std::vector<double> clat(ns), clon(ns), crlat(2 * ns), crlon(2 * ns);
  for (size_t n = 0; n < ns; ++n) {
    double s[2];
    lonlat(&G.p[3 * n], s);
    clon[n] = s[0];
    clat[n] = s[1];
    for (int ic = 0; ic < 2; ++ic) {
      lonlat(&G.c[3 * (2 * n + ic)], s);
      crlon[2 * n + ic] = s[0];
      crlat[2 * n + ic] = s[1];
    }
  }
  return swm_qp_create_coords(grid, simulation_tag, nsecond, path, clat.data(), clon.data(), crlat.data(),
                              crlon.data(), out);
}

swm_status swm_qp_pack_field(swm_qp* q, const char* field, const double* x) {
  if (!q || !field || !x) { swm::set_error("swm_qp_pack_field: bad argument"); return SWM_ERR_ARG; }
  int k = 0;
  while (k < kNqp && std::strcmp(kFields[k].name, field) != 0) ++k;
  if (k == kNqp) { swm::set_error(std::string("swm_qp_pack_field: no qp field ") + field); return SWM_ERR_ARG; }
  const int64_t npts = (int64_t)q->slots.size();
  float* d = q->data.data() + (size_t)k * npts;
  for (int64_t c = 0; c < npts; ++c) d[c] = (float)x[q->slots[c]];
  return SWM_OK;
}

swm_status swm_qp_write(swm_qp* q, long long nsecond) {
  if (!q || nsecond < 0) { swm::set_error("swm_qp_write: bad argument"); return SWM_ERR_ARG; }
  const Schema& S = q->schema;
  const int64_t npts = (int64_t)q->slots.size();
  const int32_t rec = q->timerec + 1;   // next_time_step
  // set_time with the start of the period (netcdf_general_tools.F:288-352)
  const long long mid = (q->nsecondinit + nsecond) / 2;
  std::vector<uint8_t> r;
  r.reserve((size_t)S.recsize);
  for (const Var& v : S.vars) {
    if (!v.record) continue;
    if (v.name == "time") bed(r, (double)mid / 86400.0);
    else if (v.name == "char_time") { const std::string s = char_time(mid); r.insert(r.end(), s.begin(), s.end()); }
    else if (v.name == "start_period") bed(r, (double)q->nsecondinit / 86400.0);
    else if (v.name == "end_period") bed(r, (double)nsecond / 86400.0);
    else if (v.name == "period_length") bed(r, (double)(nsecond - q->nsecondinit) / 86400.0);
    else {
      int k = 0;
      while (k < kNqp && v.name != kFields[k].name) ++k;
      const float* d = q->data.data() + (size_t)k * npts;
      for (int64_t c = 0; c < npts; ++c) bef(r, d[c]);
    }
    pad4(r);
  }
  if ((int64_t)r.size() != S.recsize) { swm::set_error("swm_qp_write: internal layout error"); return SWM_ERR_STATE; }
  FILE* f = std::fopen(q->path.c_str(), "r+b");
  if (!f) {
    swm::set_error("swm_qp_write: cannot open " + q->path + ": " + std::strerror(errno));
    return SWM_ERR_ARG;
  }
  std::vector<uint8_t> nrec;
  be32(nrec, (uint32_t)rec);
  bool ok = write_at(f, S.begin_rec + (int64_t)(rec - 1) * S.recsize, r.data(), r.size()) &&
            write_at(f, 4, nrec.data(), 4);
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) { swm::set_error("swm_qp_write: write error on " + q->path); return SWM_ERR_STATE; }
  q->timerec = rec;
  // write_swm_qp (swm_qp_module.F:280-299): the period restarts and the
  // fields return to zero
  q->nsecondinit = nsecond;
  std::fill(q->data.begin(), q->data.end(), 0.0f);
  return SWM_OK;
}

swm_status swm_qp_pack(swm_qp* q, swm_ctx* ctx) {
  if (!q || !ctx) { swm::set_error("swm_qp_pack: bad argument"); return SWM_ERR_ARG; }
  int iim = 0, jjm = 0, nt = 0;
  swm_status st = swm_ctx_layout(ctx, &iim, &jjm, &nt);
  if (st) return st;
  const int64_t ns = (int64_t)iim * jjm * 10;
  if (ns != q->ns) { swm::set_error("swm_qp_pack: the context is on another grid"); return SWM_ERR_DIMS; }
  // pack_swm_qp (swm_qp_module.F:309-370): the current time level
  std::vector<double> x((size_t)ns), y((size_t)ns);
  auto put = [&](const char* field, const char* qp) -> swm_status {
    swm_status s = swm_get_field(ctx, field, x.data());
    return s ? s : swm_qp_pack_field(q, qp, x.data());
  };
  if ((st = put("h", "thickness"))) return st;
  if ((st = swm_get_field(ctx, "hs", y.data()))) return st;
  for (int64_t n = 0; n < ns; ++n) x[n] += y[n];   // h + hs in FP64
  if ((st = swm_qp_pack_field(q, "height", x.data()))) return st;
  if ((st = put("relative", "relative")) || (st = put("div", "divergence")) || (st = put("eta", "absolute")) ||
      (st = put("ke", "kinetic_energy")))
    return st;
  const int nqt = nt < 5 ? nt : 5;   // tracer_1..tracer_5 for the tracers the model carries
  if (nqt > 0) {
    std::vector<double> t((size_t)ns * nt);
    if ((st = swm_get_field(ctx, "tracer", t.data()))) return st;
    const size_t pl = (size_t)iim * jjm;   // tracer layout (iim, jjm, ntracer, 10)
    for (int m = 0; m < nqt; ++m) {
      for (int k = 0; k < 10; ++k) std::memcpy(&x[k * pl], &t[((size_t)m + (size_t)nt * k) * pl], pl * sizeof(double));
      const std::string qp = "tracer_" + std::to_string(m + 1);
      if ((st = swm_qp_pack_field(q, qp.c_str(), x.data()))) return st;
    }
  }
  return SWM_OK;
}

int swm_qp_records(const swm_qp* q) { return q ? q->timerec : -1; }

const char* swm_qp_path(const swm_qp* q) { return q ? q->path.c_str() : nullptr; }

void swm_qp_destroy(swm_qp* q) { delete q; }

}  // extern "C"
