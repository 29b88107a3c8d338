// swm_restart.cpp -- restart files of the shallow-water model.
//
// The file is the direct-access format of utilities/layer_2/io_utilities.F
// as swm/source/swm_restart.F uses it: fixed-length records of lrecl bytes,
// record n at byte (n - 1) * lrecl, every field little-endian.
//
//   record 1          int32 nhead, int32 lrecl          (io_utilities.F:318, :563)
//   record 2          int32 nvar, int32 real_bytes,
//                     int32 ncells, logical(4) restart  (io_utilities.F:345-346, :564-565)
//   records 3..nvar+2 char[40] short_name, int32 nvar_dims,
//                     int32 var_dims[4], int32 start_index  (io_utilities.F:356-360, :568-571)
//   record nhead      int64 time stamp (s), restart files only (io_utilities.F:630, :672)
//   record npt+nhead  the values of global cell npt, real_bytes each, variables
//                     packed from their start_index (io_utilities.F:707-713, :862-869)
//
// nhead = nvar + 2, plus 1 for a restart file; lrecl = max(values per cell *
// real_bytes, 64) (io_utilities.F:547-551).  The model's restart holds
// H (1 value), TRACER (ntracer) and U (2 x 2, space index fastest) per cell
// (swm_restart.F:147-149).  A cell's record is written from the slot whose
// mask is 1 (io_utilities.F:862, write_record with grid_center_mask), and
// read into every slot through grid_center_index, halos included
// (swm_restart.F:35-45).  Bytes the reference leaves undefined (record tails,
// unused var_dims) are written as zeros.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/swm_mi355x.h"
#include "swm_internal.h"
#include "swm_setup.h"

namespace {

constexpr int kNameLen = 40;   // io_utilities.F head_clen
constexpr int kMaxDims = 4;    // io_utilities.F max_var_dims

struct Header {
  std::string name;
  int32_t ndims = 0;
  int32_t dims[kMaxDims] = {0, 0, 0, 0};
  int32_t start = 0;
};

template <class T>
T get(const std::vector<unsigned char>& r, size_t off) {
  T v;
  std::memcpy(&v, r.data() + off, sizeof(T));
  return v;
}

struct File {
  FILE* f = nullptr;
  ~File() {
    if (f) std::fclose(f);
  }
};

bool read_at(FILE* f, long long off, unsigned char* p, size_t n) {
  return fseeko(f, (off_t)off, SEEK_SET) == 0 && std::fread(p, 1, n, f) == n;
}

}  // namespace

namespace swm {

swm_status restart_write(const CellIndex& G, const char* path, int nt, long long nsecond, const double* h,
                         const double* u, const double* tracer) {
  const int nvar = 3, nhead = nvar + 3;
  const int nval = 1 + nt + 4;
  const int lrecl = std::max(nval * 8, (2 + kMaxDims) * 4 + kNameLen);
  const size_t pl = (size_t)G.iim * G.jjm, ns = pl * 10;
  // records of the real cells, in global cell order
  std::vector<int32_t> owner((size_t)G.max_ig + 1, -1);
  for (size_t n = 0; n < ns; ++n) {
    const int32_t npt = G.index[n];
    if ((int32_t)(G.mask[n] * npt) == 0) continue;
    if (npt < 1 || npt > G.max_ig) { set_error("swm_restart_file_write: grid cell index out of range"); return SWM_ERR_STATE; }
    owner[npt] = (int32_t)n;   // the last slot wins, as with the reference's record overwrite
  }
  for (int k = 1; k <= G.max_ig; ++k)
    if (owner[k] < 0) { set_error("swm_restart_file_write: a global cell has no real slot"); return SWM_ERR_STATE; }
  File F;
  F.f = std::fopen(path, "wb");
  if (!F.f) { set_error(std::string("swm_restart_file_write: cannot open ") + path); return SWM_ERR_ARG; }
  std::vector<unsigned char> buf((size_t)(nhead + G.max_ig) * lrecl, 0);
  auto rec = [&](int n) { return (size_t)(n - 1) * lrecl; };
  std::memcpy(buf.data() + rec(1), &nhead, 4);
  std::memcpy(buf.data() + rec(1) + 4, &lrecl, 4);
  const int32_t r2[4] = {nvar, 8, G.max_ig, 1};
  std::memcpy(buf.data() + rec(2), r2, sizeof r2);
  Header hd[3];
  hd[0].name = "H"; hd[0].ndims = 1; hd[0].dims[0] = 1; hd[0].start = 1;
  hd[1].name = "TRACER"; hd[1].ndims = 1; hd[1].dims[0] = nt; hd[1].start = 2;
  hd[2].name = "U"; hd[2].ndims = 2; hd[2].dims[0] = 2; hd[2].dims[1] = 2; hd[2].start = 2 + nt;
  for (int v = 0; v < nvar; ++v) {
    unsigned char* p = buf.data() + rec(3 + v);
    std::memset(p, ' ', kNameLen);
    std::memcpy(p, hd[v].name.data(), hd[v].name.size());
    std::memcpy(p + kNameLen, &hd[v].ndims, 4);
    std::memcpy(p + kNameLen + 4, hd[v].dims, 4 * kMaxDims);
    std::memcpy(p + kNameLen + 4 + 4 * kMaxDims, &hd[v].start, 4);
  }
  const int64_t ts = nsecond;
  std::memcpy(buf.data() + rec(nhead), &ts, 8);
  std::vector<double> vals(nval);
  for (int k = 1; k <= G.max_ig; ++k) {
    const size_t n = (size_t)owner[k];
    const size_t kp = n / pl, ij = n % pl;
    vals[0] = h[n];
    for (int q = 0; q < nt; ++q) vals[1 + q] = tracer[((size_t)q + (size_t)nt * kp) * pl + ij];
    for (int c = 0; c < 4; ++c) vals[1 + nt + c] = u[4 * n + c];
    std::memcpy(buf.data() + rec(nhead + k), vals.data(), (size_t)nval * 8);
  }
  if (std::fwrite(buf.data(), 1, buf.size(), F.f) != buf.size()) {
    set_error(std::string("swm_restart_file_write: write failed on ") + path);
    return SWM_ERR_ARG;
  }
  if (std::fclose(F.f) != 0) {
    F.f = nullptr;
    set_error(std::string("swm_restart_file_write: close failed on ") + path);
    return SWM_ERR_ARG;
  }
  F.f = nullptr;
  return SWM_OK;
}

swm_status restart_read(const CellIndex& G, const char* path, int nt, long long* nsecond, double* h, double* u,
                        double* tracer) {
  const char* who = "swm_restart_file_read: ";
  File F;
  F.f = std::fopen(path, "rb");
  if (!F.f) { set_error(std::string(who) + "cannot open " + path); return SWM_ERR_ARG; }
  std::vector<unsigned char> r(16);
  if (!read_at(F.f, 0, r.data(), 8)) { set_error(std::string(who) + "short file"); return SWM_ERR_ARG; }
  const int32_t nhead = get<int32_t>(r, 0), lrecl = get<int32_t>(r, 4);
  if (lrecl < 64 || nhead < 3 || nhead > 1000) { set_error(std::string(who) + "not a direct-access model file"); return SWM_ERR_ARG; }
  r.assign((size_t)lrecl, 0);
  if (!read_at(F.f, lrecl, r.data(), 16)) { set_error(std::string(who) + "short file"); return SWM_ERR_ARG; }
  const int32_t nvar = get<int32_t>(r, 0), real_bytes = get<int32_t>(r, 4), ncells = get<int32_t>(r, 8);
  const bool restart = get<int32_t>(r, 12) != 0;
  if (nvar < 1 || nhead != nvar + 2 + (restart ? 1 : 0) || (real_bytes != 4 && real_bytes != 8)) {
    set_error(std::string(who) + "inconsistent file header");
    return SWM_ERR_ARG;
  }
  if (ncells != G.max_ig) { set_error(std::string(who) + "file holds another grid's cell count"); return SWM_ERR_DIMS; }
  std::vector<Header> hd((size_t)nvar);
  for (int v = 0; v < nvar; ++v) {
    if (!read_at(F.f, (long long)(2 + v) * lrecl, r.data(), 64)) { set_error(std::string(who) + "short file"); return SWM_ERR_ARG; }
    std::string nm(reinterpret_cast<const char*>(r.data()), kNameLen);
    nm.erase(nm.find_last_not_of(' ') + 1);
    hd[v].name = nm;
    hd[v].ndims = get<int32_t>(r, kNameLen);
    for (int d = 0; d < kMaxDims; ++d) hd[v].dims[d] = get<int32_t>(r, kNameLen + 4 + 4 * d);
    hd[v].start = get<int32_t>(r, kNameLen + 4 + 4 * kMaxDims);
  }
  const int per_rec = lrecl / real_bytes;
  // unpack_data_record_{1d,2d}: the variable by name, its dimension count and
  // (stricter than the reference, which reads var_dims values into an array
  // of the model's size) its extents
  auto find = [&](const char* name, int ndims, const int* dims, int* start) -> swm_status {
    for (const Header& x : hd) {
      if (x.name != name) continue;
      if (x.ndims != ndims) { set_error(std::string(who) + "variable " + name + " has another rank"); return SWM_ERR_DIMS; }
      int len = 1;
      for (int d = 0; d < ndims; ++d) {
        if (x.dims[d] != dims[d]) { set_error(std::string(who) + "variable " + name + " has other extents"); return SWM_ERR_DIMS; }
        len *= dims[d];
      }
      if (x.start < 1 || x.start - 1 + len > per_rec) { set_error(std::string(who) + "variable " + name + " outside the record"); return SWM_ERR_ARG; }
      *start = x.start - 1;
      return SWM_OK;
    }
    set_error(std::string(who) + "variable " + name + " does not exist");
    return SWM_ERR_ARG;
  };
  const int d_h[1] = {1}, d_t[1] = {nt}, d_u[2] = {2, 2};
  int s_h = 0, s_t = 0, s_u = 0;
  swm_status st;
  if ((st = find("H", 1, d_h, &s_h)) || (st = find("TRACER", 1, d_t, &s_t)) || (st = find("U", 2, d_u, &s_u))) return st;
  if (nsecond) {
    *nsecond = 0;
    if (restart) {
      if (!read_at(F.f, (long long)(nhead - 1) * lrecl, r.data(), 8)) { set_error(std::string(who) + "short file"); return SWM_ERR_ARG; }
      *nsecond = get<int64_t>(r, 0);
    }
  }
  // all data records at once, then the read loop of swm_restart.F:35-45
  std::vector<unsigned char> data((size_t)G.max_ig * lrecl);
  if (!read_at(F.f, (long long)nhead * lrecl, data.data(), data.size())) { set_error(std::string(who) + "short file"); return SWM_ERR_ARG; }
  auto val = [&](int npt, int k) -> double {
    const unsigned char* p = data.data() + (size_t)(npt - 1) * lrecl + (size_t)k * real_bytes;
    if (real_bytes == 8) {
      double x;
      std::memcpy(&x, p, 8);
      return x;
    }
    float x;
    std::memcpy(&x, p, 4);
    return (double)x;
  };
  const size_t pl = (size_t)G.iim * G.jjm, ns = pl * 10;
  for (size_t n = 0; n < ns; ++n) {
    const int npt = G.index[n];
    if (npt == 0) continue;   // read_record skips npt = 0 (io_utilities.F:706)
    if (npt < 1 || npt > G.max_ig) { set_error(std::string(who) + "grid cell index out of range"); return SWM_ERR_STATE; }
    const size_t kp = n / pl, ij = n % pl;
    h[n] = val(npt, s_h);
    for (int q = 0; q < nt; ++q) tracer[((size_t)q + (size_t)nt * kp) * pl + ij] = val(npt, s_t + q);
    for (int c = 0; c < 4; ++c) u[4 * n + c] = val(npt, s_u + c);
  }
  return SWM_OK;
}

// write_swm_restart's file name, swm_restart.F:122-145
swm_status restart_name(long long nsecond, char* buf, int len) {
  if (!buf || len < 38 || nsecond < 0 || nsecond / 86400 > 999999) {
    set_error("swm_restart_name: bad argument");
    return SWM_ERR_ARG;
  }
  const long long day = nsecond / 86400, rem = nsecond - 86400 * day;
  const long long hr = rem / 3600, mn = (rem - 3600 * hr) / 60, sc = rem - 3600 * hr - 60 * mn;
  std::snprintf(buf, (size_t)len, "restarts/swm_restart.d%06lld.%02lld:%02lld:%02lld", day, hr, mn, sc);
  return SWM_OK;
}

}  // namespace swm

extern "C" {

swm_status swm_restart_name(long long nsecond, char* buf, int len) { return swm::restart_name(nsecond, buf, len); }

swm_status swm_restart_file_write(const swm_grid* grid, const char* path, int ntracer, long long nsecond,
                                  const double* h, const double* u, const double* tracer) {
  if (!grid || !path || !h || !u || ntracer < 0 || (ntracer > 0 && !tracer)) {
    swm::set_error("swm_restart_file_write: bad argument");
    return SWM_ERR_ARG;
  }
  return swm::restart_write(swm::cell_index(swm::grid_of(grid)), path, ntracer, nsecond, h, u, tracer);
}

swm_status swm_restart_file_read(const swm_grid* grid, const char* path, int ntracer, long long* nsecond,
                                 double* h, double* u, double* tracer) {
  if (!grid || !path || !h || !u || ntracer < 0 || (ntracer > 0 && !tracer)) {
    swm::set_error("swm_restart_file_read: bad argument");
    return SWM_ERR_ARG;
  }
  return swm::restart_read(swm::cell_index(swm::grid_of(grid)), path, ntracer, nsecond, h, u, tracer);
}

}  // extern "C"
