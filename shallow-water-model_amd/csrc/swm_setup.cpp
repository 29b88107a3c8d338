// swm_setup.cpp -- host-side FP64 set-up of the icosahedral grid.
//
// Follows the reference algorithms step for step (same floating-point
// operation order) so that grid points, halo tables and metrics agree with the
// reference bit for bit or to round-off; each routine cites what it follows.
// Compiled without floating-point contraction.
#include "swm_setup.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <stdexcept>

namespace swm {

namespace {

inline int ipow2(int n) { return 1 << n; }

// ---------------------------------------------------------------------------
// grid points: recursive bisection of the icosahedron
// (utilities/grid_generation/generate_grid.F:69-130 and set_p1_p2 :327-463)
// ---------------------------------------------------------------------------
void bisect(const double* a, const double* b, double* out) {
  double t[3];
  for (int k = 0; k < 3; ++k) t[k] = 0.5 * (a[k] + b[k]);
  double s = 0.0;
  for (int k = 0; k < 3; ++k) s = s + t[k] * t[k];
  double r = std::sqrt(s);
  for (int k = 0; k < 3; ++k) out[k] = t[k] / r;
}

}  // namespace

std::vector<double> generate_points(int g) {
  if (g < 1 || g > 10) throw std::runtime_error("generate_points: level out of range");
  const int jm = 2 + ipow2(g);
  const int im = 5 * (2 + 2 * ipow2(g));
  const int im5 = im / 5;
  std::vector<double> p1(3 * (size_t)jm * jm, 0.0), p2(3 * (size_t)im * jm, 0.0);
  auto P1 = [&](int i, int j) { return &p1[3 * ((size_t)(i - 1) + (size_t)jm * (j - 1))]; };
  auto P2 = [&](int i, int j) { return &p2[3 * ((size_t)(i - 1) + (size_t)im * (j - 1))]; };
  const double alpha = 1.25663706143591729538;      // 2 pi / 5
  const double plat = 0.463647609;                  // latitude of the first pentagon ring
  const double beta = 0.55357435889704525150853;    // tilt of the square block

  double* q = P1(1, 1);
  q[0] = std::cos(-0.5 * alpha) * std::cos(plat);
  q[1] = std::sin(-0.5 * alpha) * std::cos(plat);
  q[2] = std::sin(plat);
  q = P1(jm - 1, 1);
  q[0] = std::cos(plat); q[1] = 0.0; q[2] = std::sin(-plat);
  q = P1(1, jm - 1);
  q[0] = 0.0; q[1] = 0.0; q[2] = 1.0;
  q = P1(jm - 1, jm - 1);
  q[0] = std::cos(0.5 * alpha) * std::cos(plat);
  q[1] = std::sin(0.5 * alpha) * std::cos(plat);
  q[2] = std::sin(plat);

  for (int level = 1; level <= g; ++level) {
    const int o1 = (jm - 2) / ipow2(level), o2 = (jm - 2) / ipow2(level - 1);
    for (int j = o2 + 1; j <= jm - 1; j += o2)
      for (int i = 1; i <= jm - 1; i += o2) bisect(P1(i, j), P1(i, j - o2), P1(i, j - o1));
    for (int j = o2 + 1; j <= jm - 1; j += o2)
      for (int i = o2 + 1; i <= jm - 1; i += o2) bisect(P1(i, j), P1(i - o2, j - o2), P1(i - o1, j - o1));
    for (int j = 1; j <= jm - 1; j += o2)
      for (int i = o2 + 1; i <= jm - 1; i += o2) bisect(P1(i, j), P1(i - o2, j), P1(i - o1, j));
  }
  // move the square block down to the equator
  const double cmb = std::cos(-beta), smb = std::sin(-beta), cb = std::cos(beta), sb = std::sin(beta);
  for (int j = 1; j <= jm; ++j)
    for (int i = 1; i <= jm; ++i) {
      double* v = P1(i, j);
      double x = cmb * v[0] - smb * v[2], y = v[1], z = smb * v[0] + cmb * v[2];
      v[0] = x; v[1] = y; v[2] = z;
    }
  // upper half of the first panel, rotated towards the pole
  for (int j = 1; j <= jm - 1; ++j)
    for (int i = 1; i <= jm - 1; ++i) {
      const double* v = P1(i, j);
      double* o = P2(i, j);
      o[0] = cb * v[0] - sb * v[2]; o[1] = v[1]; o[2] = sb * v[0] + cb * v[2];
    }
  // lower half of the first panel
  const double ca = std::cos(0.5 * alpha), sa = std::sin(0.5 * alpha);
  for (int j = 1; j <= jm - 1; ++j)
    for (int i = 1; i <= jm - 1; ++i) {
      const double* v = P1(i, j);
      double t0 = cmb * v[0] - smb * v[2], t1 = v[1], t2 = smb * v[0] + cmb * v[2];
      double* o = P2((jm - 2) + i, j);
      o[0] = ca * t0 - sa * t1; o[1] = sa * t0 + ca * t1; o[2] = t2;
    }
  // remaining four panels by rotation about the polar axis
  for (int panel = 1; panel <= 4; ++panel) {
    const double ang = alpha * (double)panel;
    const double cp = std::cos(ang), sp = std::sin(ang);
    for (int j = 1; j <= jm - 1; ++j)
      for (int i = 1; i <= im5 - 1; ++i) {
        const double* v = P2(i, j);
        double* o = P2(im5 * panel + i, j);
        o[0] = cp * v[0] - sp * v[1]; o[1] = sp * v[0] + cp * v[1]; o[2] = v[2];
      }
  }
  q = P2(1, jm - 1); q[0] = 0.0; q[1] = 0.0; q[2] = 1.0;
  q = P2(im - 1, 1); q[0] = 0.0; q[1] = 0.0; q[2] = -1.0;

  // global list order: NP, SP, then rows j, panels, columns i
  // (gather_scatter.F:317-357 reduce_to_vector_global)
  const int max_ig = 2 + 10 * ipow2(g) * ipow2(g);
  std::vector<double> out(3 * (size_t)max_ig);
  size_t n = 0;
  auto push = [&](const double* v) { out[n++] = v[0]; out[n++] = v[1]; out[n++] = v[2]; };
  push(P2(1, jm - 1));
  push(P2(im - 1, 1));
  for (int j = 2; j <= jm - 1; ++j)
    for (int panel = 0; panel <= 4; ++panel)
      for (int i = 2; i <= im5 - 1; ++i) push(P2(im5 * panel + i, j));
  return out;
}

// ---------------------------------------------------------------------------
// topology and halo tables (utilities/layer_2/wrap_data.F:34-1424,2328-2513)
//
// The reference builds message lists between the 10 panels from a global
// "face" graph (cell names, neighbour lists, sorted vertex triples and edge
// pairs) and applies them as copies followed by fix-ups.  We rebuild the same
// lists and apply them to an array of slot identities; the final identities
// are the gather table.
// ---------------------------------------------------------------------------
namespace {

struct Face {
  int name = 0, ntot = 0, subdomain = 0;
  int nghbr[6] = {0, 0, 0, 0, 0, 0};
  int vertex[2][3] = {{0, 0, 0}, {0, 0, 0}};
  int edge[3][2] = {{0, 0}, {0, 0}, {0, 0}};
};

template <int N>
void sort_small(int* v) {  // ascending bubble sort (wrap_data.F:397-425)
  for (int it = 0; it < N - 1; ++it)
    for (int n = 0; n < N - 1; ++n)
      if (v[n + 1] < v[n]) std::swap(v[n], v[n + 1]);
}

struct Topology {
  int g, im, jm, im5, iim, jjm;
  std::vector<Face> glob;                 // (im, jm)
  std::vector<Face> face1, face2;         // (iim, jjm, 10)
  Face& G(int i, int j) { return glob[(size_t)(i - 1) + (size_t)im * (j - 1)]; }
  Face& F1(int i, int j, int k) { return face1[(size_t)(i - 1) + (size_t)iim * ((j - 1) + (size_t)jjm * (k - 1))]; }
  Face& F2(int i, int j, int k) { return face2[(size_t)(i - 1) + (size_t)iim * ((j - 1) + (size_t)jjm * (k - 1))]; }

  explicit Topology(int g_) : g(g_) {
    jm = 2 + ipow2(g);
    im = 5 * (2 + 2 * ipow2(g));
    im5 = im / 5;
    iim = 2 + ipow2(g);
    jjm = iim;
    glob.assign((size_t)im * jm, Face());
    build_global();
    build_panels();
  }

  // closed-form fill of the global halo (wrap_data.F:305-357)
  void wrap_cell() {
    const int p = im - 1, q = jm - 1, m5 = 2 * jm - 2;
    for (int panel = 0; panel <= 4; ++panel) {
      const int o1 = m5 * (panel % 5), o2 = m5 * ((panel + 1) % 5);
      G(o1 + 1, q) = G(1, q);
      G(o1 + m5 - 1, 1) = G(p, 1);
      for (int ij = 2; ij <= q; ++ij) G(o1 + ij, jm) = G(o2 + 2, q + 2 - ij);
      for (int ij = q + 1; ij <= m5 - 1; ++ij) G(o1 + ij, jm) = G(o2 - q + 1 + ij, 2);
      for (int ij = 2; ij <= q; ++ij) G(o1 + m5, ij) = G(o2 + m5 + 1 - ij, 2);
      for (int ij = 1; ij <= q - 1; ++ij) G(o2 + 1, ij) = G(o1 + q + 1 - ij, q);
      for (int ij = 2; ij <= q; ++ij) G(o2 + ij, 1) = G(o1 + q - 1 + ij, q);
      for (int ij = q + 1; ij <= m5 - 2; ++ij) G(o2 + ij, 1) = G(o1 + m5 - 1, m5 - ij);
    }
  }

  // wrap_data.F:165-297 set_connections and :428-475 set_subdomain_number
  void build_global() {
    G(1, jm - 1).name = 1;
    G(im - 1, 1).name = 2;
    int count = 3;
    for (int j = 2; j <= jm - 1; ++j)
      for (int panel = 0; panel <= 4; ++panel)
        for (int i = 2; i <= im5 - 1; ++i) G(im5 * panel + i, j).name = count++;
    wrap_cell();
    static const int di5[5] = {-1, -1, 0, 1, 0}, dj5[5] = {0, -1, -1, 0, 1};
    static const int di6[6] = {-1, -1, 0, 1, 1, 0}, dj6[6] = {0, -1, -1, 0, 1, 1};
    Face& np = G(1, jm - 1);
    np.ntot = 5;
    for (int panel = 0; panel <= 4; ++panel) np.nghbr[panel] = G(im5 * panel + 2, jm - 1).name;
    Face& sp = G(im - 1, 1);
    sp.ntot = 5;
    for (int panel = 1; panel <= 5; ++panel) sp.nghbr[panel - 1] = G(im5 * panel - 1, 2).name;
    for (int j = 2; j <= jm - 1; ++j)
      for (int panel = 0; panel <= 4; ++panel)
        for (int i = 2; i <= im5 - 1; ++i) {
          const int ii = im5 * panel + i;
          Face& f = G(ii, j);
          if ((i == jm - 1 || i == im5 - 1) && j == jm - 1) {
            f.ntot = 5;
            for (int n = 0; n < 5; ++n) f.nghbr[n] = G(ii + di5[n], j + dj5[n]).name;
          } else {
            f.ntot = 6;
            for (int n = 0; n < 6; ++n) f.nghbr[n] = G(ii + di6[n], j + dj6[n]).name;
          }
        }
    wrap_cell();
    for (int j = 2; j <= jm - 1; ++j)
      for (int panel = 0; panel <= 4; ++panel)
        for (int i = 2; i <= im5 - 1; ++i) {
          Face& f = G(im5 * panel + i, j);
          for (int n = 0; n < 2; ++n) {
            f.vertex[n][0] = f.name; f.vertex[n][1] = f.nghbr[n]; f.vertex[n][2] = f.nghbr[n + 1];
            sort_small<3>(f.vertex[n]);
          }
          for (int n = 0; n < 3; ++n) {
            f.edge[n][0] = f.name; f.edge[n][1] = f.nghbr[n];
            sort_small<2>(f.edge[n]);
          }
        }
    wrap_cell();
    // subdomains: one northern (isd=0) and one southern (isd=1) rhombus per panel
    G(1, jm - 1).subdomain = 1;
    G(im - 1, 1).subdomain = 10;
    int nsd = 0;
    for (int panel = 0; panel <= 4; ++panel)
      for (int isd = 0; isd <= 1; ++isd) {
        ++nsd;
        const int ii = im5 * panel + (iim - 2) * isd;
        for (int j = 2; j <= jjm - 1; ++j)
          for (int i = ii + 2; i <= ii + iim - 1; ++i) G(i, j).subdomain = nsd;
      }
    wrap_cell();
  }

  // wrap_data.F:577-638
  void correct_face(int ref, int twist, Face& f) {
    if (f.ntot == 5) {
      if (f.name != 1 && f.name != 2) {
        for (auto& v : f.vertex) for (int& x : v) x = -999;
        for (auto& e : f.edge) for (int& x : e) x = -999;
      }
      return;
    }
    const int N = f.ntot;
    int nref = -1;
    for (int n = 0; n < N && nref < 0; ++n)
      if (f.nghbr[n] == ref) nref = n + 1;
    if (nref < 0) throw std::runtime_error("halo tables: reference face not found");
    for (int n = 1; n <= 2; ++n) {
      f.vertex[n - 1][0] = f.name;
      f.vertex[n - 1][1] = f.nghbr[(nref + twist + n) % N];
      f.vertex[n - 1][2] = f.nghbr[(nref + twist + n + 1) % N];
      sort_small<3>(f.vertex[n - 1]);
    }
    for (int n = 1; n <= 3; ++n) {
      f.edge[n - 1][0] = f.name;
      f.edge[n - 1][1] = f.nghbr[(nref + twist + n) % N];
      sort_small<2>(f.edge[n - 1]);
    }
  }

  // wrap_data.F:483-569 (one block per rhombus: isdm=2, jsdm=1)
  void build_panels() {
    face1.assign((size_t)iim * jjm * 10, Face());
    int nsd = 0;
    for (int panel = 0; panel <= 4; ++panel)
      for (int isd = 0; isd <= 1; ++isd) {
        ++nsd;
        for (int jj = 1; jj <= jjm; ++jj)
          for (int ii = 1; ii <= iim; ++ii) F1(ii, jj, nsd) = G(im5 * panel + (iim - 2) * isd + ii, jj);
        F1(iim, jjm, nsd).name = -999;  // (iim-1,jjm-1) is a pentagon on every rhombus
      }
    face2 = face1;
    for (int k = 1; k <= 10; ++k) {
      for (int jj = 1; jj <= jjm - 1; ++jj) correct_face(F2(2, jj, k).name, 1, F2(1, jj, k));
      for (int ii = 2; ii <= iim - 1; ++ii) correct_face(F2(ii, 2, k).name, 5, F2(ii, 1, k));
      for (int jj = 2; jj <= jjm - 1; ++jj) correct_face(F2(iim - 1, jj, k).name, 4, F2(iim, jj, k));
      for (int ii = 2; ii <= iim - 1; ++ii) correct_face(F2(ii, jjm - 1, k).name, 2, F2(ii, jjm, k));
      if (F2(iim - 1, jjm - 1, k).ntot == 6) correct_face(F2(iim - 1, jjm - 1, k).name, 3, F2(iim, jjm, k));
    }
  }
};

struct Message {
  int sender, seq;
  int nsd1, i1, j1, di1, dj1;   // sender block, start, stride
  int nsd2, i2, j2, di2, dj2;   // receiver block, start, stride
  int len;
  int up_v[2][3], up_e[3][3];   // unpack offsets (i, j, component) or -999
};

bool eq3(const int* a, const int* b) { return a[0] == b[0] && a[1] == b[1] && a[2] == b[2]; }
bool eq2(const int* a, const int* b) { return a[0] == b[0] && a[1] == b[1]; }

void set_unpack(Topology& T, int nsd2, const Face& src, int off_i, int off_j, Message& m) {
  for (int n1 = 0; n1 < 2; ++n1) {
    bool found = false;
    for (int jj = 1; jj <= T.jjm; ++jj)
      for (int ii = 1; ii <= T.iim; ++ii)
        for (int n2 = 0; n2 < 2; ++n2)
          if (eq3(T.F2(ii, jj, nsd2).vertex[n2], src.vertex[n1])) {
            found = true;
            m.up_v[n1][0] = ii - off_i; m.up_v[n1][1] = jj - off_j; m.up_v[n1][2] = n2 + 1;
          }
    if (!found) m.up_v[n1][0] = m.up_v[n1][1] = m.up_v[n1][2] = -999;
  }
  for (int n1 = 0; n1 < 3; ++n1) {
    bool found = false;
    for (int jj = 1; jj <= T.jjm; ++jj)
      for (int ii = 1; ii <= T.iim; ++ii)
        for (int n2 = 0; n2 < 3; ++n2)
          if (eq2(T.F2(ii, jj, nsd2).edge[n2], src.edge[n1])) {
            found = true;
            m.up_e[n1][0] = ii - off_i; m.up_e[n1][1] = jj - off_j; m.up_e[n1][2] = n2 + 1;
          }
    if (!found) m.up_e[n1][0] = m.up_e[n1][1] = m.up_e[n1][2] = -999;
  }
}

void single(std::vector<Message>& L, int& seq, int s, int is, int js, int d, int id, int jd,
            std::array<int, 6> upv, std::array<int, 9> upe) {
  Message m{};
  m.sender = s; m.seq = seq++;
  m.nsd1 = s; m.i1 = is; m.j1 = js; m.di1 = 0; m.dj1 = 0;
  m.nsd2 = d; m.i2 = id; m.j2 = jd; m.di2 = 0; m.dj2 = 0;
  m.len = 1;
  for (int a = 0; a < 2; ++a) for (int b = 0; b < 3; ++b) m.up_v[a][b] = upv[3 * a + b];
  for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) m.up_e[a][b] = upe[3 * a + b];
  L.push_back(m);
}

std::vector<Message> build_messages(Topology& T) {
  std::vector<Message> L;
  int seq = 0;
  const int iim = T.iim, jjm = T.jjm;
  // rows and columns (wrap_data.F:646-809)
  const int i_nghbr[4] = {1, 2, iim, 2}, j_nghbr[4] = {2, 1, 2, jjm};
  const int i_start[4] = {2, 2, iim - 1, 2}, j_start[4] = {2, 2, 2, jjm - 1};
  const int di[4] = {0, 1, 0, 1}, dj[4] = {1, 0, 1, 0};
  for (int nsd = 1; nsd <= 10; ++nsd)
    for (int nm = 0; nm < 4; ++nm) {
      Message m{};
      m.sender = nsd; m.seq = seq++;
      const int nb = T.F1(i_nghbr[nm], j_nghbr[nm], nsd).subdomain;
      m.nsd1 = nsd; m.i1 = i_start[nm]; m.j1 = j_start[nm]; m.di1 = di[nm]; m.dj1 = dj[nm];
      m.len = iim - 2;
      const Face& f1 = T.F1(m.i1, m.j1, nsd);
      const Face& f2 = T.F1(m.i1 + m.di1, m.j1 + m.dj1, nsd);
      int a_i = -1, a_j = -1, b_i = -1, b_j = -1;
      for (int jj = 1; jj <= jjm; ++jj)
        for (int ii = 1; ii <= iim; ++ii) {
          const int nm_ = T.F1(ii, jj, nb).name;
          if (nm_ == f1.name) { a_i = ii; a_j = jj; }
          if (nm_ == f2.name) { b_i = ii; b_j = jj; }
        }
      if (a_i < 0 || b_i < 0) throw std::runtime_error("halo tables: message start not found");
      m.nsd2 = nb; m.i2 = a_i; m.j2 = a_j; m.di2 = b_i - a_i; m.dj2 = b_j - a_j;
      if (f2.ntot == 6) set_unpack(T, nb, f2, b_i, b_j, m);
      else set_unpack(T, nb, f1, a_i, a_j, m);
      L.push_back(m);
    }
  // diagonal messages (wrap_data.F:838-956): with one block per rhombus every
  // block is of type 0 and every (iim-1,jjm-1) is a pentagon, so none is sent.
  const std::array<int, 9> no_e = {-999, -999, -999, -999, -999, -999, -999, -999, -999};
  // northern hemisphere vertex fixes (wrap_data.F:962-990)
  for (int panel = 1; panel <= 4; ++panel) {
    const int s = (2 * (panel + 2) + 1) % 10, d = (2 * panel + 1) % 10;
    single(L, seq, s, 2, jjm - 1, d, 1, jjm, {1, 0, 1, -999, -999, -999}, no_e);
  }
  // southern hemisphere vertex fixes (wrap_data.F:1029-1057)
  for (int panel = 0; panel <= 3; ++panel) {
    const int s = (2 * (panel + 2) + 1) % 10 + 1, d = (2 * panel + 1) % 10 + 1;
    single(L, seq, s, iim - 1, 2, d, iim, 1, {-999, -999, -999, 0, 1, 2}, no_e);
  }
  // messages towards the pole blocks (wrap_data.F:1228-1330)
  single(L, seq, 5, 2, jjm - 1, 1, 1, jjm, {1, 0, 1, -999, -999, -999},
         {0, 0, 3, 1, 0, 1, -999, -999, -999});
  single(L, seq, 7, 2, jjm - 1, 1, iim, 1, {1 - iim, jjm - 2, 1, -999, -999, -999},
         {1 - iim, jjm - 2, 2, 1 - iim, jjm - 1, 2, -999, -999, -999});
  single(L, seq, 4, iim - 1, 2, 10, iim, 1, {-999, -999, -999, 0, 1, 2},
         {-999, -999, -999, 0, 1, 3, 0, 0, 1});
  single(L, seq, 6, iim - 1, 2, 10, 1, jjm, {-999, -999, -999, iim - 2, 1 - jjm, 2},
         {-999, -999, -999, iim - 1, 1 - jjm, 2, iim - 2, 1 - jjm, 2});
  // pole values to the other blocks (wrap_data.F:1335-1419)
  const std::array<int, 6> no_v = {-999, -999, -999, -999, -999, -999};
  for (int panel = 1; panel <= 4; ++panel) single(L, seq, 1, 1, jjm - 1, 2 * panel + 1, 1, jjm - 1, no_v, no_e);
  for (int panel = 0; panel <= 3; ++panel) single(L, seq, 10, iim - 1, 1, 2 * panel + 2, iim - 1, 1, no_v, no_e);
  // applied sender by sender, each sender's list in creation order
  std::stable_sort(L.begin(), L.end(), [](const Message& a, const Message& b) {
    return a.sender != b.sender ? a.sender < b.sender : a.seq < b.seq;
  });
  return L;
}

}  // namespace

void build_halo_tables(int g, HaloTable& face, HaloTable& vrtx, HaloTable& edge) {
  Topology T(g);
  const std::vector<Message> L = build_messages(T);
  const int iim = T.iim, jjm = T.jjm;
  auto slot = [&](int i, int j, int k) { return (i - 1) + iim * ((j - 1) + jjm * (k - 1)); };
  const int nslot = iim * jjm * 10;

  auto run = [&](int ncomp, HaloTable& t) {
    std::vector<int32_t> id((size_t)ncomp * nslot);
    for (size_t k = 0; k < id.size(); ++k) id[k] = (int32_t)k;
    for (const Message& m : L)
      for (int n = 0; n < m.len; ++n) {
        const int i1 = m.i1 + m.di1 * n, j1 = m.j1 + m.dj1 * n;
        const int i2 = m.i2 + m.di2 * n, j2 = m.j2 + m.dj2 * n;
        if (ncomp == 1) {
          id[slot(i2, j2, m.nsd2)] = id[slot(i1, j1, m.nsd1)];
        } else {
          for (int c = 0; c < ncomp; ++c) {
            const int* up = (ncomp == 2) ? m.up_v[c] : m.up_e[c];
            if (up[2] == -999) continue;
            id[(size_t)ncomp * slot(i2 + up[0], j2 + up[1], m.nsd2) + (up[2] - 1)] =
                id[(size_t)ncomp * slot(i1, j1, m.nsd1) + c];
          }
        }
      }
    // fix-ups after the copies (wrap_data.F:2421-2513)
    for (int k = 1; k <= 10; ++k) {
      if (ncomp == 1) {
        if (k != 1 && k != 10) {
          id[slot(iim, 1, k)] = id[slot(2, 2, k)];
          id[slot(1, jjm, k)] = id[slot(2, 2, k)];
        }
        id[slot(iim, jjm, k)] = id[slot(iim - 1, jjm, k)];
      } else if (ncomp == 2) {
        if (k == 1) id[2 * (size_t)slot(1, jjm, k) + 1] = -1;
        if (k == 10) id[2 * (size_t)slot(iim, 1, k) + 0] = -1;
        if (k % 2 == 1) id[2 * (size_t)slot(iim, jjm, k) + 1] = id[2 * (size_t)slot(iim, jjm, k) + 0];
        else id[2 * (size_t)slot(iim, jjm, k) + 0] = id[2 * (size_t)slot(iim, jjm, k) + 1];
      } else {
        id[3 * (size_t)slot(1, 1, k) + 1] = -1;
        id[3 * (size_t)slot(iim - 1, 1, k) + 2] = -1;
        id[3 * (size_t)slot(1, jjm - 1, k) + 0] = -1;
        id[3 * (size_t)slot(iim, jjm, k) + 1] = -1;
      }
    }
    t.ncomp = ncomp;
    t.dst.clear(); t.src.clear();
    for (size_t k = 0; k < id.size(); ++k)
      if (id[k] != (int32_t)k) { t.dst.push_back((int32_t)k); t.src.push_back(id[k]); }
    // the gather may run in place only if no source is also a destination
    std::vector<char> isdst(id.size(), 0);
    for (int32_t d : t.dst) isdst[d] = 1;
    for (int32_t s : t.src)
      if (s >= 0 && isdst[s]) throw std::runtime_error("halo tables: chained copy");
  };
  run(1, face);
  run(2, vrtx);
  run(3, edge);
}

// closed-form face halo of the multigrid (multigrid_solver.F:1105-1158)
HaloTable build_simple_face_table(int g) {
  const int iim = 2 + ipow2(g), jjm = iim, p = iim - 1, q = jjm - 1;
  auto slot = [&](int i, int j, int k) { return (i - 1) + iim * ((j - 1) + jjm * (k - 1)); };
  std::vector<int32_t> id((size_t)iim * jjm * 10);
  for (size_t k = 0; k < id.size(); ++k) id[k] = (int32_t)k;
  for (int nsd = 3; nsd <= 9; nsd += 2) id[slot(1, q, nsd)] = id[slot(1, q, 1)];
  for (int nsd = 2; nsd <= 8; nsd += 2) id[slot(p, 1, nsd)] = id[slot(p, 1, 10)];
  // the reference uses array sections: every right-hand side is read before
  // the section is stored; sources are interior cells, so elementwise copies agree
  for (int nsd = 1; nsd <= 9; nsd += 2) {
    const int a = (nsd + 7) % 10 + 1, b = (nsd + 8) % 10 + 1, c = (nsd + 0) % 10 + 1, d = (nsd + 1) % 10 + 1;
    for (int j = 1; j <= q - 1; ++j) id[slot(1, j, nsd)] = id[slot(p - (j - 1), q, a)];
    for (int i = 2; i <= p; ++i) id[slot(i, 1, nsd)] = id[slot(i, q, b)];
    for (int j = 2; j <= q; ++j) id[slot(p + 1, j, nsd)] = id[slot(2, j, c)];
    for (int i = 2; i <= p; ++i) id[slot(i, q + 1, nsd)] = id[slot(2, q - (i - 2), d)];
  }
  for (int nsd = 2; nsd <= 10; nsd += 2) {
    const int a = (nsd + 8) % 10 + 1, b = (nsd + 7) % 10 + 1, c = (nsd + 1) % 10 + 1, d = (nsd + 0) % 10 + 1;
    for (int j = 2; j <= q; ++j) id[slot(1, j, nsd)] = id[slot(p, j, a)];
    for (int i = 1; i <= p - 1; ++i) id[slot(i, 1, nsd)] = id[slot(p, q - (i - 1), b)];
    for (int j = 2; j <= q; ++j) id[slot(p + 1, j, nsd)] = id[slot(p - (j - 2), 2, c)];
    for (int i = 2; i <= p; ++i) id[slot(i, q + 1, nsd)] = id[slot(i, 2, d)];
  }
  id[slot(1, q + 1, 1)] = id[slot(2, q, 5)];
  id[slot(p + 1, 1, 1)] = id[slot(2, q, 7)];
  id[slot(p + 1, 1, 10)] = id[slot(p, 2, 4)];
  id[slot(1, q + 1, 10)] = id[slot(p, 2, 6)];
  HaloTable t;
  t.ncomp = 1;
  for (size_t k = 0; k < id.size(); ++k)
    if (id[k] != (int32_t)k) { t.dst.push_back((int32_t)k); t.src.push_back(id[k]); }
  return t;
}

void apply_halo(const HaloTable& t, double* x, int nlayer, int64_t layer_stride) {
  for (int l = 0; l < nlayer; ++l) {
    double* xl = x + (int64_t)l * layer_stride;
    for (size_t k = 0; k < t.dst.size(); ++k) xl[t.dst[k]] = t.src[k] < 0 ? 0.0 : xl[t.src[k]];
  }
}

void apply_halo_strided(const HaloTable& t, double* x, int stride, int offset) {
  for (size_t k = 0; k < t.dst.size(); ++k)
    x[(int64_t)stride * t.dst[k] + offset] = t.src[k] < 0 ? 0.0 : x[(int64_t)stride * t.src[k] + offset];
}

}  // namespace swm
