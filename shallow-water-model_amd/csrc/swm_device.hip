// swm_device.hip -- device model context, time step driver and the C ABI of
// the device path.  The kernels live in swm_kernels.hip (same translation
// unit, included below).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/swm_mi355x.h"
#include "swm_internal.h"
#include "swm_kernels.h"
#include "swm_kernels.hip"
#include "swm_tlad.hip"
#include "swm_remap.hip"
#include "swm_diss.hip"

namespace {

using swm::dev::FieldSet;
using swm::dev::Geo;
using swm::dev::Lev;

constexpr int BLK = 256;

#define HIP_TRY(expr)                                                                       \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess) {                                                                 \
      swm::set_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " #expr);    \
      return SWM_ERR_HIP;                                                                   \
    }                                                                                       \
  } while (0)

inline unsigned nblk(size_t n) { return (unsigned)((n + BLK - 1) / BLK); }
inline unsigned nblk_b(size_t n, int b) { return (unsigned)((n + b - 1) / b); }

struct DevTable {
  int n = 0;
  int* dst = nullptr;
  int* src = nullptr;
  // host copy of the entries and the span of a field (components x slots),
  // for the halo shadow check
  std::vector<int> hdst, hsrc;
  size_t span = 0;
  // the table inverted for producer-side halo writes (swm::dev::Push)
  int4* hp = nullptr;
  int* pe = nullptr;
  swm::dev::Push push() const { return swm::dev::Push{hp, pe}; }
};

// one 256-byte aligned device arena for all buffers of a context
struct Arena {
  char* base = nullptr;
  size_t used = 0, cap = 0;
  template <class T>
  T* take(size_t n) {
    const size_t bytes = ((n * sizeof(T) + 255) / 256) * 256;
    T* p = reinterpret_cast<T*>(base + used);
    used += bytes;
    return p;
  }
};

}  // namespace

namespace {

// one multigrid level of a view: weights, reciprocal centre weight, halo
// map and the level's alpha / beta / smoother scratch (two fields each)
template <class R>
struct LevelT {
  int iim = 0, jjm = 0, ns = 0;
  R *lw = nullptr, *lwt = nullptr, *alpha = nullptr, *beta = nullptr, *temp = nullptr;
  int* src = nullptr;
  swm::dev::LevT<R> L{};
};

// the device model in one storage type: metrics, multigrid hierarchy, state
// and scratch of the time step.  R = double: the FP64 path; R = float: the
// mixed path (thickness-like fields stored less h0, vorticity less f).
template <class R>
struct View {
  swm::dev::GeoT<R> geo{};
  std::vector<LevelT<R>> lev;
  R *h[2] = {}, *eta[2] = {}, *div[2] = {}, *u[2] = {};
  std::vector<R*> tr[2], told;
  R *eta_f[3] = {}, *div_f[3] = {}, *msh_f[3] = {};
  R* h_f[3] = {};                 // thickness tendencies (centred / upwind advection)
  std::vector<R*> tr_f[3], tflux;  // mass-times-tracer tendencies and tracer edge fluxes
  R *relative = nullptr, *ke = nullptr, *tlow = nullptr, *mass_flux = nullptr;   // tlow: low-order tracer (FCT)
  R *hc = nullptr, *flx1 = nullptr, *anti = nullptr, *mss_low = nullptr, *beta = nullptr, *f1 = nullptr,
    *f2 = nullptr, *ml = nullptr, *bdn = nullptr, *bup = nullptr, *vf = nullptr, *dcurl = nullptr, *rhs = nullptr;
  R* state_begin = nullptr;   // state and scratch are contiguous from here
  size_t state_words = 0;
};

}  // namespace

// incremental remapping (SWM_ADV_REMAP, FP64): the remapping quantities as
// device planes, the step's fields and transports, the tracer pointer
// tables and the pinned flag of a negative thickness
struct RemapDev {
  void* mem = nullptr;
  double *cellp = nullptr, *edgep = nullptr, *cornp = nullptr;
  double *mx = nullptr, *my = nullptr, *mp = nullptr, *tx = nullptr, *ty = nullptr, *tp = nullptr;
  double *dp = nullptr, *mflux = nullptr, *mtflux = nullptr;
  double** ptab = nullptr;   // tracers of levels 1, 2, then tendencies of slots 1..3 (nt each)
  int* flag_host = nullptr;
  int* flag_dev = nullptr;
  int limit_min = 1, limit_max = 1;
};

struct swm_ctx {
  int g = 0, iim = 0, jjm = 0, ns = 0, nt = 0, max_ig = 0;
  double dt = 0, grav = 0;
  int precision = SWM_FP64, device = 0;
  int advection = SWM_ADV_FCT;
  hipStream_t stream = nullptr;
  hipStream_t stream2 = nullptr;   // the tracer transport beside the dynamics (overlap)
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  hipEvent_t ev_norm = nullptr;    // the convergence test's norms are in host memory
  int overlap = 3;             // tracer transport on stream2: 0 off, 1 from the mass update, n >= 2 from level g - n + 1 of the first V-cycle
  // overlap: work issued by the first V-cycle of a step when it reaches
  // level hook_level on the way down (the tracer transport, started where
  // the cycle's launches become latency-bound)
  std::function<void()> hook;
  int hook_level = 0;
  // mixed step, one member, absolute criterion: the velocity kernel of the
  // step, queued at a test of the solve before the host waits for the
  // norms (post_solve, given the iterate's buffer); it runs only if the
  // device's copy of the test (k_conv_go into go) passes.  post_done: the
  // last test's launch stands (the solve ended there), so the step does not
  // launch the kernel again
  std::function<void(const double*)> post_solve;
  bool post_done = false;
  Arena arena;             // shared: metrics, halo tables, multigrid weights and maps
  size_t arena_bytes = 0;
  // ensemble members (same grid and settings).  Member m's state and scratch
  // is one block at marena.base + m * mb; the pointers below address member
  // 0's block (or, while one member is selected for host-side calls, that
  // member's), and launches over the state cover nz members.
  Arena marena;
  int nmem = 1, nz = 1, sel = 0;
  size_t mb = 0;
  std::vector<char> member_init;   // member m has been initialised
  bool h0_set = false;             // mixed path: thickness offset fixed by a first member (ensembles)
  // fixed schedule (> 0): mixed solves run exactly this many correction
  // cycles with no host-side test, so that a step can be captured in a
  // graph; the residuals are tracked on the device instead
  int fixed_cycles = 0;
  int fixed_start = 0;             // cycles of the fixed schedule's solves before two solutions are held
  // workgroup sizes of the step's stencil kernels and halo fills, and of the
  // multigrid smoother / transfer kernels (64, 128 or 256; an autotuner's
  // choice for ensembles)
  int blk_step = 256, blk_mg = 64;
  long long epoch = 0;             // changes whenever captured launches would differ (initialisation, settings)
  double* track = nullptr;
  double* fct_rec = nullptr;       // FP64 FCT step: the limiter coefficients go here (linearised model's base steps)
  double* tr_rec = nullptr;        // FP64 FCT step: the tracer limiter coefficients (3 planes per tracer, 6 planes apart)
  double* rmean = nullptr;         // mixed path: sums of area * rhs of the two fields (mean removal)
  double* go = nullptr;            // mixed path: the device's copy of the solve's convergence test
  double* part_rhs = nullptr;      // mixed path: their per-block partial sums, from the tendencies kernel         // per member: max mean(r^2) of the two fields, count of solves above the criterion
  // FP64 metrics (both paths: set-up, FP64 residual, sums)
  double *area, *area_inv, *cai, *sub_area, *dc, *nrm, *tng, *vw, *sw, *lw, *f, *hs, *mask, *area_mask, *pxyz;
  double total_area = 0.0, radius = 0.0;
  DevTable face, vrtx, edge;
  View<double> d;       // FP64 model (the FP64 path; initial state of the mixed path)
  View<float> m;        // mixed path
  float* mx_r = nullptr;   // mixed elliptic: FP32 defect (2 planes)
  float* mx_e = nullptr;   // mixed elliptic: FP32 correction (2 planes)
  double h0 = 0.0;         // mixed path: thickness offset
  double* xout = nullptr;  // streamfunction and velocity potential (FP64, both paths)
  double* xalt = nullptr;  // mixed path: second FP64 iterate of the defect correction
  double* xprev = nullptr;  // mixed path: the solution before xout (extrapolated warm start)
  double* xprev2 = nullptr; // mixed path: the solution before xprev (quadratic extrapolation)
  int nsol = 0;             // mixed path: solutions held since the initialisation (xout, xprev, xprev2), at most 3
  double* part_slots = nullptr;   // per-block partial sums, one block per 256 slots, up to four columns
  double* gauge = nullptr;        // mixed path: area-weighted change of the solution's mean, two fields
  double* r2 = nullptr;
  double *part = nullptr, *sums = nullptr;
  double* part_budget = nullptr;   // partial sums of the budgets, (8 + BUDGET_MAXTR) x npart
  double* host_sums = nullptr;   // pinned, 64 values per member
  double* host_sums_dev = nullptr;   // its device address (kernels write there directly)
  int npart = 0;
  // time stepping
  int step_count = 0, n3 = 1, n4 = 2, n1f = 1, n2f = 2, n3f = 3;
  long long nsecond = 0;   // model time in whole seconds (time_manager_module.F nsecond_dyn)
  // global cell numbering and real-cell mask of the slots (restart files)
  std::vector<int32_t> cell_index;
  std::vector<double> cell_mask;
  double w1 = 0, w2 = 0, w3 = 0;
  int last_vcycles = 0;
  // cycles of the last FP64 two-field solve: `first` cycles of both fields
  // (2, or the fixed count), then `extra` cycles of the fields in `bad_mask`
  // (bit 0 psi, bit 1 chi); fixed: the fixed schedule (no halo fill between)
  struct SolveSched { int first = 0, extra = 0, bad_mask = 0, fixed = 0; } last_sched;
  int lin_exact = 1;           // TL/AD: the derivative of the forward's own cycles (1) or of converged solves (0)
  int halo_shadow = 0;         // every halo fill checked against the host's (swm_set_halo_shadow)
  int fuse_push = 1;           // FCT step: producer kernels store the halo copies (SWM_FUSE_HALO_PUSH)
  int fuse_anti = 1;           // FCT step: the final mass kernel corrects the antidiffusive fluxes (SWM_FUSE_MASS_ANTI)
  bool anti_ok = false;        // the fill tables allow it (anti_fusion_check)
  // the next step's corner thickness, queued while the host waits on the
  // solve's convergence test (swm_step sets next_follows when another step
  // of the same call follows; corners_ready: queued, the next step skips it)
  bool next_follows = false, corners_ready = false;
  // the last step left out the kinetic energy of its new velocity; the next
  // step's mass-flux kernel forms it (producer-side halo writes only)
  bool ke_deferred = false;
  swm_status layout_status = SWM_OK;   // halo table set-up
  long long shadow_checks = 0, shadow_bad = 0;
  bool initialized = false;
  RemapDev rm;
  double* diss = nullptr;      // dissipation operator geometry (swm_init_diss_op), 36 planes
  // elliptic solve settings (reference: mean(r^2) <= 1e-20, 12 cycles) and
  // the mean squared residuals / right-hand sides of the last solve
  double solver_tol = 1.0e-20;
  int solver_relative = 0;
  int solver_max_cycles = 12;
  int solver_stall = 0;        // also stop where a cycle no longer halves the residual (FP64 floor)
  int mixed_extrapolate = 2;   // mixed path: start from the quadratic (2) or linear (1) extrapolation, or x_n (0)
  int mixed_min_cycles = 1;    // mixed path: cycles before the first check
  int mixed_fine_sweeps = 2;   // mixed path: smoothing sweeps before / after the coarse correction on the finest level
  int mixed_half = 1;          // mixed path: half-precision weights in the smoother and restriction above the coarse levels
  int fuse_mg = 1;             // multigrid: an injection also performs the coarse level's first sweep (SWM_FUSE_INJECT_SWEEP)
  int mixed_coarse_direct = 1; // mixed path: the coarsest levels as one dense solve of level 3 (1) or 4 (2, where level 5 exists), or the V-cycle (0)
  float* cd_m = nullptr;       // its FP32 matrix (cd_n x cd_n) and the real cells of level 3
  int* cd_cell = nullptr;
  int cd_n = 0;
  float* cd_b = nullptr;       // the dense solve's right-hand side in its cell order (member block, 2 planes)
  float* cd4_m = nullptr;      // level 4 (levels 5 and above exist): its matrix and real cells
  int* cd4_cell = nullptr;
  int cd4_n = 0;
  double last_resid[2] = {0.0, 0.0};
  double last_rhs_ms[2] = {0.0, 0.0};
  std::vector<double> rhs_ms;   // per member and field (relative criterion)
  long long solves = 0, vcycles_total = 0, checks_total = 0;   // since the last initialisation
  // timing
  hipEvent_t ev_a = nullptr, ev_b = nullptr;
  double last_step_ms = 0.0;
  std::vector<hipEvent_t> kev;   // pairs around the profiled kernel
  size_t kev_used = 0;
  int profile = 0;                 // > 0: every profile-th finest-level smoother launch is timed
  long long prof_seq = 0;
  double kernel_bytes = 0.0;
  int kernel_pair_launches = 0;   // launches covered by the timed event pairs
  // tangent-linear / adjoint window (allocated on first use)
  struct TlAd* tl = nullptr;
};

namespace {

swm_status upload(double* d, const std::vector<double>& h, hipStream_t s) {
  HIP_TRY(hipMemcpyAsync(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, s));
  return SWM_OK;
}

// reference layout (ncomp, slots) -> planes (ncomp, slots) with plane index
// remapped by perm(component)
std::vector<double> to_planes(const std::vector<double>& a, int ncomp, int ns, const std::vector<int>& perm,
                              int nplanes) {
  std::vector<double> out((size_t)nplanes * ns, 0.0);
  for (int s = 0; s < ns; ++s)
    for (int c = 0; c < ncomp; ++c) {
      const int p = perm[c];
      if (p >= 0) out[(size_t)p * ns + s] = a[(size_t)ncomp * s + c];
    }
  return out;
}

swm_status make_table(swm_ctx* C, const swm::HaloTable& t, int ns, DevTable& d) {
  // table entries are (component, slot) with the component fastest; the
  // device index is component*ns + slot
  std::vector<int> dst(t.dst.size()), src(t.src.size());
  for (size_t k = 0; k < t.dst.size(); ++k) {
    dst[k] = (t.dst[k] % t.ncomp) * ns + t.dst[k] / t.ncomp;
    src[k] = t.src[k] < 0 ? -1 : (t.src[k] % t.ncomp) * ns + t.src[k] / t.ncomp;
  }
  d.n = (int)dst.size();
  d.hdst = dst;
  d.hsrc = src;
  d.span = (size_t)t.ncomp * ns;
  d.dst = C->arena.take<int>(dst.size() + 1);
  d.src = C->arena.take<int>(src.size() + 1);
  // the push words and entries (layout in swm_kernels.h, Push): entries
  // grouped by source slot, in table order within a slot
  std::vector<int4> hp(ns, int4{0, 0, 0, 0});
  std::vector<int> cnt(ns, 0);
  int nover = 0;
  for (size_t k = 0; k < dst.size(); ++k) {
    const int g = dst[k] % ns, c = dst[k] / ns;
    if (c > 2) { swm::set_error("halo table: more than three components"); return SWM_ERR_ARG; }
    hp[g].x |= (src[k] >= 0 ? 1 : 2) << (4 + 2 * c);
    if (src[k] >= 0) {
      const int q = src[k] % ns;
      if (dst[k] >= (1 << 28)) { swm::set_error("halo table: grid too large for the push entries"); return SWM_ERR_ARG; }
      const int ent = (int)((unsigned)(src[k] / ns) << 28) | dst[k];
      if (cnt[q] == 0) hp[q].y = ent;
      else if (cnt[q] == 1) hp[q].z = ent;
      else if (cnt[q] == 2) hp[q].w = ent;
      else nover += 1;
      cnt[q] += 1;
      if (cnt[q] > 15) { swm::set_error("halo table: more than 15 copies of one slot"); return SWM_ERR_ARG; }
    }
  }
  if (nover >= (1 << 22)) { swm::set_error("halo table: too many entries for the push words"); return SWM_ERR_ARG; }
  std::vector<int> pe((size_t)nover + 1, 0), seen(ns, 0);
  std::vector<int> start(ns, 0);
  for (int q = 0, o = 0; q < ns; ++q) {
    start[q] = o;
    if (cnt[q] > 3) o += cnt[q] - 3;
    hp[q].x |= cnt[q] | ((cnt[q] > 3 ? start[q] : 0) << 10);
  }
  for (size_t k = 0; k < dst.size(); ++k) {
    if (src[k] < 0) continue;
    const int q = src[k] % ns;
    const int r = seen[q]++;
    if (r >= 3) pe[(size_t)start[q] + r - 3] = (int)((unsigned)(src[k] / ns) << 28) | dst[k];
  }
  d.hp = C->arena.take<int4>(ns);
  d.pe = C->arena.take<int>(pe.size());
  if (d.n && C->arena.base) {
    HIP_TRY(hipMemcpy(d.dst, dst.data(), dst.size() * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.src, src.data(), src.size() * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.hp, hp.data(), hp.size() * sizeof(int4), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.pe, pe.data(), pe.size() * sizeof(int), hipMemcpyHostToDevice));
  }
  return SWM_OK;
}

// producer-side halo writes in the step (SWM_FUSE_HALO_PUSH): FCT step only,
// and not while every fill is checked against the host's
inline bool halo_push(const swm_ctx* C) {
  return C->fuse_push && !C->halo_shadow && C->advection == SWM_ADV_FCT;
}

// launch grid over the state of the nz members in play (z = member)
inline dim3 zgrid(const swm_ctx* C, unsigned x, unsigned y = 1) { return dim3(x, y, (unsigned)C->nz); }

// halo fills of member state fields
template <class R>
void gather_launch(swm_ctx* C, const DevTable& t, const std::vector<R*>& fields) {
  for (size_t a = 0; a < fields.size(); a += 8) {
    swm::dev::FieldSetT<R> fs{};
    fs.mb = C->mb;
    int nf = 0;
    for (size_t b = a; b < fields.size() && b < a + 8; ++b) fs.p[nf++] = fields[b];
    if (t.n)
      hipLaunchKernelGGL(swm::dev::k_gather<R>, zgrid(C, nblk_b(t.n, C->blk_step), nf), dim3(C->blk_step), 0, C->stream,
                         t.dst, t.src, t.n, fs);
  }
}
// the halo shadow check (swm_set_halo_shadow): each field of every member in
// play is copied to the host before and after the device fill, the host
// applies the same table to its copy, and the two are compared bit for bit
template <class R>
void gather_shadow(swm_ctx* C, const DevTable& t, const std::vector<R*>& fields) {
  const size_t span = t.span;
  std::vector<std::vector<R>> pre(fields.size() * (size_t)C->nz, std::vector<R>(span));
  auto at = [&](R* f, int z) { return (R*)((char*)f + (size_t)z * C->mb); };
  bool ok = hipStreamSynchronize(C->stream) == hipSuccess;
  for (size_t f = 0; ok && f < fields.size(); ++f)
    for (int z = 0; ok && z < C->nz; ++z)
      ok = hipMemcpy(pre[f * C->nz + z].data(), at(fields[f], z), span * sizeof(R), hipMemcpyDeviceToHost) == hipSuccess;
  gather_launch<R>(C, t, fields);
  ok = ok && hipStreamSynchronize(C->stream) == hipSuccess;
  std::vector<R> post(span);
  for (size_t f = 0; ok && f < fields.size(); ++f)
    for (int z = 0; ok && z < C->nz; ++z) {
      std::vector<R>& x = pre[f * C->nz + z];
      ok = hipMemcpy(post.data(), at(fields[f], z), span * sizeof(R), hipMemcpyDeviceToHost) == hipSuccess;
      for (int k = 0; k < t.n; ++k) x[(size_t)t.hdst[k]] = t.hsrc[k] < 0 ? R(0) : x[(size_t)t.hsrc[k]];
      for (size_t q = 0; ok && q < span; ++q)
        if (std::memcmp(&x[q], &post[q], sizeof(R)) != 0) C->shadow_bad += 1;
      C->shadow_checks += 1;
    }
  if (!ok) C->shadow_bad += 1;   // a copy failed: report it through the same count
}
template <class R>
void gather_list(swm_ctx* C, const DevTable& t, const std::vector<R*>& fields) {
  if (C->halo_shadow && t.n) gather_shadow<R>(C, t, fields);
  else gather_launch<R>(C, t, fields);
}
template <class R>
void gather(swm_ctx* C, const DevTable& t, std::initializer_list<R*> fields) {
  gather_list<R>(C, t, std::vector<R*>(fields));
}

// x = v on n values of member state, for the members in play
template <class R>
void fill(swm_ctx* C, R* x, size_t n, R v) {
  hipLaunchKernelGGL(swm::dev::k_fill<R>, zgrid(C, nblk(n)), dim3(BLK), 0, C->stream, x, n, v, C->mb);
}
// x = v on n values of one buffer outside the member blocks
template <class R>
void fill_plain(swm_ctx* C, R* x, size_t n, R v) {
  hipLaunchKernelGGL(swm::dev::k_fill<R>, dim3(nblk(n)), dim3(BLK), 0, C->stream, x, n, v, (size_t)0);
}
void copy_state(swm_ctx* C, double* y, const double* x, size_t n) {
  hipLaunchKernelGGL(swm::dev::k_copy, zgrid(C, nblk(n)), dim3(BLK), 0, C->stream, y, x, n, C->mb);
}
// compensated sums of nf columns of block partials, per member in play
void sum_final(swm_ctx* C, const double* part, int nb, double* out, int nf) {
  hipLaunchKernelGGL(swm::dev::k_sum_final, zgrid(C, (unsigned)nf), dim3(swm::dev::SUM_THREADS), 0, C->stream, part,
                     nb, out, C->mb, nf, (const double*)nullptr, 0, (double*)nullptr);
}
// two sets of columns in one launch (nf1 of part, nf2 of part2)
void sum_final2(swm_ctx* C, const double* part, int nb, double* out, int nf1, const double* part2, int nb2,
                double* out2, int nf2) {
  hipLaunchKernelGGL(swm::dev::k_sum_final, zgrid(C, (unsigned)(nf1 + nf2)), dim3(swm::dev::SUM_THREADS), 0, C->stream,
                     part, nb, out, C->mb, nf1, part2, nb2, out2);
}

// ------------------------------------------------------------- multigrid --
template <class R, int NV>
void launch_relax(swm_ctx* C, const LevelT<R>& L, bool zero, const R* beta, const R* a, R* t) {
  const dim3 grid = zgrid(C, nblk_b(L.ns, C->blk_mg)), blk(C->blk_mg);
  if constexpr (sizeof(R) == 4) {
    if (!zero && L.L.hw && C->mixed_half) {
      hipLaunchKernelGGL((swm::dev::k_relax_h<NV>), grid, blk, 0, C->stream, L.L, beta, a, t);
      return;
    }
  }
  if (zero) hipLaunchKernelGGL((swm::dev::k_relax<R, NV, true>), grid, blk, 0, C->stream, L.L, beta, a, t);
  else hipLaunchKernelGGL((swm::dev::k_relax<R, NV, false>), grid, blk, 0, C->stream, L.L, beta, a, t);
}

// algorithmic bytes of one sweep at a level per real cell of the model
// grid: the weights (seven words, or seven halves for the half-precision
// sweep) and per field beta, alpha (not read from zero) and the result
template <class R>
double sweep_bytes(const swm_ctx* C, const LevelT<R>& L, bool zero, int nvar) {
  const bool half = sizeof(R) == 4 && L.L.hw && C->mixed_half && !zero;
  return (double)C->max_ig * ((half ? 7.0 * 2.0 : 7.0 * sizeof(R)) + (zero ? 2.0 : 3.0) * nvar * sizeof(R));
}

// iters Jacobi sweeps at level l (an even count leaves the result in alpha);
// zero_start: alpha is known to be zero on entry
template <class R>
void relax(swm_ctx* C, View<R>& V, int l, int iters, R* alpha, const R* beta, int nvar, bool zero_start) {
  LevelT<R>& L = V.lev[l - 1];
  R* a = alpha;
  R* t = L.temp;
  // swm_profile: a pair of finest-level sweeps (neither from zero) is timed
  // between two events, which halves the events' share in the per-launch time
  const bool prof = C->profile && l == C->g && iters == 2 && !zero_start && C->kev_used + 2 <= C->kev.size() &&
                    C->prof_seq++ % C->profile == 0;
  if (prof) {
    (void)hipEventRecord(C->kev[C->kev_used++], C->stream);
    // algorithmic bytes of a finest-level sweep per real cell: the Laplacian
    // weights and per field beta, alpha and the result (neighbour values hit
    // in cache), twice
    C->kernel_bytes += 2.0 * sweep_bytes(C, L, false, nvar);
    C->kernel_pair_launches += 2;
  }
  for (int it = 0; it < iters; ++it) {
    const bool zero = zero_start && it == 0;
    if (nvar == 2) launch_relax<R, 2>(C, L, zero, beta, a, t);
    else launch_relax<R, 1>(C, L, zero, beta, a, t);
    std::swap(a, t);
  }
  if (prof) (void)hipEventRecord(C->kev[C->kev_used++], C->stream);
}

// one sweep at level l from a into out (zero: a is zero and is not read, so
// out may be a itself)
template <class R>
void relax_once(swm_ctx* C, View<R>& V, int l, const R* a, R* out, const R* beta, int nvar, bool zero) {
  LevelT<R>& L = V.lev[l - 1];
  if (nvar == 2) launch_relax<R, 2>(C, L, zero, beta, a, out);
  else launch_relax<R, 1>(C, L, zero, beta, a, out);
}

// multigrid_solver.F:476-591.  The coarse levels' alpha is zero at the start
// of a cycle (the reference clears it); their first sweep is told so instead
// of clearing memory.  Coarse beta is written by the injection at every slot
// the smoother reads.  top_zero: the finest alpha is zero on entry as well
// (a correction cycle of the mixed path).  fine_sweeps = 1 (with top_zero
// only): one sweep before and one after the coarse correction on the finest
// level instead of two; the first, from zero, writes alpha in place, the
// second leaves the result in the level's scratch.  Returns the finest
// level's result (alpha, or that scratch).
template <class R>
R* vcycle(swm_ctx* C, View<R>& V, int nvar, R* alpha, const R* beta, bool top_zero = false, int fine_sweeps = 2,
          bool pre_done = false, size_t pre_off = 0) {
  const int G = C->g;
  const int top = G < swm::dev::CO_MAXLEV ? G : swm::dev::CO_MAXLEV;   // levels 1..top in one workgroup
  // mixed path: the level of the dense coarse solve (levels below it run no
  // sweeps), else top
  int dtop = top;
  bool direct = false;
  if constexpr (sizeof(R) == 4) {
    if (C->mixed_coarse_direct == 2 && C->cd4_n && G > top + 1) { dtop = top + 1; direct = true; }
    else if (C->mixed_coarse_direct && C->cd_n && G > top) direct = true;
  }
  auto A = [&](int l) { return l == G ? alpha : V.lev[l - 1].alpha; };
  auto B = [&](int l) { return l == G ? const_cast<R*>(beta) : V.lev[l - 1].beta; };
  // finest level with another sweep count (mixed correction cycles): the
  // iterate moves between alpha and the level's scratch
  const bool own = fine_sweeps != 2 && top_zero && G > top;
  R* fa = alpha;
  R* fo = V.lev[G - 1].temp;
  auto fine = [&](int n, bool zero) {
    for (int i = 0; i < n; ++i) {
      if (zero && i == 0) {
        relax_once<R>(C, V, G, fa, fa, B(G), nvar, true);   // from zero: in place
      } else {
        relax_once<R>(C, V, G, fa, fo, B(G), nvar, false);
        std::swap(fa, fo);
      }
    }
  };
  // the injection into a level that runs its own sweeps (above the coarse
  // kernel's) also performs that level's first sweep from zero, into the
  // level's scratch (swm_set_fusion)
  const bool fuse = C->fuse_mg;
  auto inject = [&](int l, const R* xa, bool half_ok) {
    LevelT<R>& Fl = V.lev[l - 1];
    LevelT<R>& Cl = V.lev[l - 2];
    R* z0 = fuse && l - 1 > dtop ? Cl.temp : nullptr;
    // into the level of a dense coarse solve: its right-hand side in the
    // solve's cell order as well
    R* xc = nullptr;
    if constexpr (sizeof(R) == 4)
      if (direct && l - 1 == dtop) xc = C->cd_b;
    const dim3 grid = zgrid(C, nblk_b(Cl.ns, C->blk_mg)), blk(C->blk_mg);
    if constexpr (sizeof(R) == 4) {
      if (half_ok && Fl.L.hw && C->mixed_half && top_zero) {
        if (nvar == 2)
          hipLaunchKernelGGL((swm::dev::k_inject_h<2>), grid, blk, 0, C->stream, Fl.L, Cl.L, xa, B(l), B(l - 1), z0, xc);
        else
          hipLaunchKernelGGL((swm::dev::k_inject_h<1>), grid, blk, 0, C->stream, Fl.L, Cl.L, xa, B(l), B(l - 1), z0, xc);
        return;
      }
    }
    if (nvar == 2)
      hipLaunchKernelGGL((swm::dev::k_inject<R, 2>), grid, blk, 0, C->stream, Fl.L, Cl.L, xa, B(l), B(l - 1), z0, xc);
    else
      hipLaunchKernelGGL((swm::dev::k_inject<R, 1>), grid, blk, 0, C->stream, Fl.L, Cl.L, xa, B(l), B(l - 1), z0, xc);
  };
  for (int l = G; l > dtop; --l) {
    if (C->hook && l == C->hook_level) {
      std::function<void()> h = std::move(C->hook);
      C->hook = nullptr;
      h();
    }
    if (l == G && own) {
      fine(fine_sweeps, true);
      inject(l, fa, false);
      continue;
    }
    if (l == G && top_zero && pre_done)
      // the first sweep from zero is in the level's scratch already (the
      // defect kernel wrote it): the second one
      relax_once<R>(C, V, G, V.lev[G - 1].temp + pre_off, alpha, B(G), nvar, false);
    else if (l < G && fuse)
      // the first sweep from zero is in the level's scratch (the injection
      // wrote it): the second one
      relax_once<R>(C, V, l, V.lev[l - 1].temp, A(l), B(l), nvar, false);
    else
      relax<R>(C, V, l, 2, A(l), B(l), nvar, l < G || top_zero);
    inject(l, A(l), true);
  }
  if constexpr (sizeof(R) == 4) {
    if (direct && dtop == top + 1) {
      // level 4: eight rows per workgroup of 512
      const dim3 grid = zgrid(C, (unsigned)((C->cd4_n + 7) / 8)), blk(512);
      const int n4 = V.lev[dtop - 1].ns;
      if (nvar == 2)
        hipLaunchKernelGGL((swm::dev::k_coarse_direct<2, swm::dev::CD_MAXN4>), grid, blk, 0, C->stream, C->cd4_m,
                           C->cd4_cell, C->cd4_n, n4, C->cd_b, A(dtop), C->mb);
      else
        hipLaunchKernelGGL((swm::dev::k_coarse_direct<1, swm::dev::CD_MAXN4>), grid, blk, 0, C->stream, C->cd4_m,
                           C->cd4_cell, C->cd4_n, n4, C->cd_b, A(dtop), C->mb);
    } else if (direct) {
      const dim3 grid = zgrid(C, (unsigned)((C->cd_n + 3) / 4)), blk(256);
      const int n3 = V.lev[top - 1].ns;
      if (nvar == 2)
        hipLaunchKernelGGL((swm::dev::k_coarse_direct<2, swm::dev::CD_MAXN>), grid, blk, 0, C->stream, C->cd_m, C->cd_cell,
                           C->cd_n, n3, C->cd_b, A(top), C->mb);
      else
        hipLaunchKernelGGL((swm::dev::k_coarse_direct<1, swm::dev::CD_MAXN>), grid, blk, 0, C->stream, C->cd_m, C->cd_cell,
                           C->cd_n, n3, C->cd_b, A(top), C->mb);
    }
  }
  swm::dev::CoarseArgs<R> P{};
  for (int l = 1; l <= top; ++l) P.lev[l - 1] = V.lev[l - 1].L;
  P.nlev = top;
  P.nvar = nvar;
  P.beta = B(top);
  P.alpha = A(top);
  P.zero = (top < G || top_zero) ? 1 : 0;
  P.mb = C->mb;
  if (!direct) hipLaunchKernelGGL(swm::dev::k_coarse_cycle<R>, zgrid(C, 1), dim3(swm::dev::CO_THREADS), 0, C->stream, P);
  for (int l = dtop; l <= G - 1; ++l) {
    const bool f = l + 1 == G && own;
    R* up = f ? fa : A(l + 1);
    if (nvar == 2)
      hipLaunchKernelGGL((swm::dev::k_interp<R, 2>), zgrid(C, nblk_b(V.lev[l].ns, C->blk_mg)), dim3(C->blk_mg), 0, C->stream,
                         V.lev[l - 1].L, V.lev[l].L, A(l), up);
    else
      hipLaunchKernelGGL((swm::dev::k_interp<R, 1>), zgrid(C, nblk_b(V.lev[l].ns, C->blk_mg)), dim3(C->blk_mg), 0, C->stream,
                         V.lev[l - 1].L, V.lev[l].L, A(l), up);
    if (f) {
      fine(fine_sweeps, false);
      return fa;
    }
    relax<R>(C, V, l + 1, 2, A(l + 1), B(l + 1), nvar, false);
  }
  return alpha;
}

// step.F:132-141 of the next step, queued early: on the second stream with
// the tracer transport, or else before the host waits on a convergence test
// of this step's solve, so that the GPU has work while the host reads the
// norms.  The corner thickness of the next step depends only on this step's
// new thickness and its halo (FCT: both complete before the solve), and
// nothing in the rest of this step reads or writes it.  The next step then
// skips it.
template <class R>
void corners_next(swm_ctx* C, View<R>& V) {
  const int n4 = C->n4 - 1;
  const bool push = halo_push(C);
  swm::dev::GeoT<R> Gp = V.geo;
  if (push) Gp.pv = C->vrtx.push();
  hipLaunchKernelGGL(swm::dev::k_avg_corners<R>, zgrid(C, nblk_b(C->ns, C->blk_step)), dim3(C->blk_step), 0, C->stream,
                     Gp, V.geo.sw, V.h[n4], V.hc);
  if (!push) gather<R>(C, C->vrtx, {V.hc});
}
void corners_for_next(swm_ctx* C) {
  if (!C->next_follows || C->corners_ready || C->advection != SWM_ADV_FCT) return;
  if (C->precision == SWM_MIXED) corners_next<float>(C, C->m);
  else corners_next<double>(C, C->d);
  C->corners_ready = true;
}

// sums of x0*w and x1*w into C->sums[0..1] (block partials, then a
// compensated fixed-order sum)
template <class R>
void sum2(swm_ctx* C, const R* x0, const R* x1, const R* w, int n) {
  hipLaunchKernelGGL(swm::dev::k_sum_partial<R>, zgrid(C, C->npart, 2), dim3(BLK), 0, C->stream, x0, x1, w, n, C->part,
                     C->mb);
  sum_final(C, C->part, C->npart, C->sums, 2);
}

// sums[0..n) of each member in play to out[m * n + b]
swm_status read_sums(swm_ctx* C, int n, double* out) {
  if (C->nz == 1) {
    HIP_TRY(hipMemcpyAsync(C->host_sums, C->sums, n * sizeof(double), hipMemcpyDeviceToHost, C->stream));
  } else {
    HIP_TRY(hipMemcpy2DAsync(C->host_sums, n * sizeof(double), C->sums, C->mb, n * sizeof(double), C->nz,
                             hipMemcpyDeviceToHost, C->stream));
  }
  HIP_TRY(hipStreamSynchronize(C->stream));
  for (int b = 0; b < n * C->nz; ++b) out[b] = C->host_sums[b];
  return SWM_OK;
}

// convergence test of elliptic_solver.F:203-239 on the mean squared
// residual of field v of member m: mean(r^2) <= tol (the reference's
// absolute form), or, when set relative, sqrt(mean(r^2) / mean(rhs^2)) <= tol
bool converged(const swm_ctx* C, int m, int v, double ms) {
  if (C->solver_relative) return ms <= C->solver_tol * C->solver_tol * C->rhs_ms[2 * m + v];
  return ms <= C->solver_tol;
}

// the mean squared residuals of fields vars of every member in play (sums
// of the per-block partials in part, nb blocks): ms[m * 2 + b] for vars[b];
// last_resid takes the largest over the members
// before_wait (one member): work queued behind the sums, before the host
// waits for them
swm_status member_norms(swm_ctx* C, const double* part, int nb, const std::vector<int>& vars, double* ms,
                        const std::function<void()>& before_wait = {}) {
  const int nv = (int)vars.size();
  std::vector<double> raw((size_t)nv * C->nz);
  if (C->nz == 1 && C->host_sums_dev) {
    // one member: the final sums go straight to pinned host memory (no copy
    // launch before the host reads them); the host waits for them, not for
    // the next step's corners queued behind them
    sum_final(C, part, nb, C->host_sums_dev, nv);
    HIP_TRY(hipEventRecord(C->ev_norm, C->stream));
    if (before_wait) before_wait();
    corners_for_next(C);
    HIP_TRY(hipEventSynchronize(C->ev_norm));
    for (int b = 0; b < nv; ++b) raw[b] = C->host_sums[b];
  } else {
    sum_final(C, part, nb, C->sums, nv);
    swm_status st = read_sums(C, nv, raw.data());
    if (st) return st;
  }
  for (int b = 0; b < nv; ++b) C->last_resid[vars[b]] = 0.0;
  for (int m = 0; m < C->nz; ++m)
    for (int b = 0; b < nv; ++b)
      if (!std::isfinite(raw[(size_t)m * nv + b])) {
        swm::set_error("non-finite residual norm in the elliptic solve (member " + std::to_string(m) +
                       ", field " + std::to_string(vars[b]) + "): the state holds NaN or Inf");
        return SWM_ERR_NAN;
      }
  for (int m = 0; m < C->nz; ++m)
    for (int b = 0; b < nv; ++b) {
      const double v = raw[(size_t)m * nv + b] / (double)C->max_ig;
      ms[2 * m + b] = v;
      C->last_resid[vars[b]] = std::max(C->last_resid[vars[b]], v);
    }
  return SWM_OK;
}

// mean squares of the right-hand sides over real cells (relative criterion)
template <class R>
swm_status rhs_mean_squares(swm_ctx* C, const R* rhs, int nvar) {
  if (!C->solver_relative) return SWM_OK;
  const int ns = C->ns;
  hipLaunchKernelGGL(swm::dev::k_sumsq_mixed<R>, zgrid(C, C->npart, nvar), dim3(BLK), 0, C->stream, rhs,
                     rhs + (size_t)(nvar - 1) * ns, C->mask, ns, C->part, C->mb);
  sum_final(C, C->part, C->npart, C->sums, nvar);
  std::vector<double> ms((size_t)nvar * C->nz);
  swm_status st = read_sums(C, nvar, ms.data());
  if (st) return st;
  C->rhs_ms.assign(2 * (size_t)C->nz, 0.0);
  for (int m = 0; m < C->nz; ++m)
    for (int v = 0; v < nvar; ++v) C->rhs_ms[2 * m + v] = ms[(size_t)m * nvar + v] / (double)C->max_ig;
  for (int v = 0; v < nvar; ++v) C->last_rhs_ms[v] = C->rhs_ms[v];
  return SWM_OK;
}

// elliptic_solver.F:40-195 on the FP64 fields rhs / xout.  With several
// members, a field takes further cycles (for every member) while any member
// misses the criterion.
swm_status elliptic(swm_ctx* C, double* rhs, double* xout, int nvar) {
  const int ns = C->ns;
  View<double>& V = C->d;
  for (int v = 0; v < nvar; ++v) gather<double>(C, C->face, {xout + (size_t)v * ns, rhs + (size_t)v * ns});
  sum2<double>(C, rhs, nvar > 1 ? rhs + ns : rhs, C->area_mask, ns);
  hipLaunchKernelGGL(swm::dev::k_sub_mean<double>, zgrid(C, nblk(ns)), dim3(BLK), 0, C->stream, ns, C->sums,
                     C->total_area, rhs, nvar > 1 ? rhs + ns : C->r2, C->mb);
  swm_status st0 = rhs_mean_squares<double>(C, rhs, nvar);
  if (st0) return st0;
  const unsigned nb = nblk(ns);
  if (C->fixed_cycles > 0 && nvar == 2) {
    // fixed schedule: exactly fixed_cycles V-cycles, no host-side test; the
    // final mean squared residuals go to the per-member record
    for (int it = 0; it < C->fixed_cycles; ++it) vcycle<double>(C, V, 2, xout, rhs);
    for (int v = 0; v < 2; ++v) gather<double>(C, C->face, {xout + (size_t)v * ns});
    hipLaunchKernelGGL(swm::dev::k_resid2<2>, zgrid(C, nb), dim3(BLK), 0, C->stream, V.geo, xout, rhs, C->part_slots);
    sum_final(C, C->part_slots, (int)nb, C->sums, 2);
    hipLaunchKernelGGL(swm::dev::k_track_resid, zgrid(C, 1), dim3(64), 0, C->stream, C->sums, 1.0 / (double)C->max_ig,
                       C->solver_tol, C->track, C->mb);
    C->last_vcycles = C->fixed_cycles;
    C->last_sched = {C->fixed_cycles, 0, 3, 1};
    C->solves += 1; C->vcycles_total += C->fixed_cycles; C->checks_total += 1;
    return SWM_OK;
  }
  for (int it = 0; it < 2; ++it) vcycle<double>(C, V, nvar, xout, rhs);
  for (int v = 0; v < nvar; ++v) gather<double>(C, C->face, {xout + (size_t)v * ns});
  auto norms = [&](const std::vector<int>& vars, double* ms) -> swm_status {
    const int v0 = vars[0];
    if (vars.size() == 2)
      hipLaunchKernelGGL(swm::dev::k_resid2<2>, zgrid(C, nb), dim3(BLK), 0, C->stream, V.geo, xout, rhs, C->part_slots);
    else
      hipLaunchKernelGGL(swm::dev::k_resid2<1>, zgrid(C, nb), dim3(BLK), 0, C->stream, V.geo, xout + (size_t)v0 * ns,
                         rhs + (size_t)v0 * ns, C->part_slots);
    return member_norms(C, C->part_slots, (int)nb, vars, ms);
  };
  std::vector<int> all;
  for (int v = 0; v < nvar; ++v) all.push_back(v);
  const int M = C->nz;
  std::vector<double> ms(2 * (size_t)M), prev(2 * (size_t)M);
  swm_status st = norms(all, ms.data());
  if (st) return st;
  std::vector<int> bad;
  for (int v = 0; v < nvar; ++v) {
    bool ok = true;
    for (int m = 0; m < M; ++m) {
      prev[2 * m + v] = ms[2 * m + v];
      if (!converged(C, m, v, ms[2 * m + v])) ok = false;
    }
    if (!ok) bad.push_back(v);
  }
  int iter = 0;
  bool any = !bad.empty();
  while (any && iter < C->solver_max_cycles - 2) {
    ++iter;
    if (bad.size() == 2) vcycle<double>(C, V, 2, xout, rhs);
    else vcycle<double>(C, V, 1, xout + (size_t)bad[0] * ns, rhs + (size_t)bad[0] * ns);
    for (int v : bad) gather<double>(C, C->face, {xout + (size_t)v * ns});
    st = norms(bad, ms.data());
    if (st) return st;
    any = false;
    for (int m = 0; m < M; ++m)
      for (size_t b = 0; b < bad.size(); ++b) {
        const int v = bad[b];
        const double now = ms[2 * m + b];
        // with solver_stall, a cycle that no longer halves the residual ends
        // the solve: the FP64 floor of the residual is reached
        const bool stalled = C->solver_stall && now > 0.25 * prev[2 * m + v];
        prev[2 * m + v] = now;
        if (!converged(C, m, v, now) && !stalled) any = true;
      }
  }
  C->last_vcycles = 2 + iter;
  {
    int mask = 0;
    for (int v : bad) mask |= 1 << v;
    C->last_sched = {2, iter, mask, 0};
  }
  C->solves += 1; C->vcycles_total += 2 + iter; C->checks_total += 1 + iter;
  if (any) {
    swm::set_error("elliptic solve did not converge (residual above the criterion after " +
                   std::to_string(C->solver_max_cycles) + " V-cycles; the reference stops here, "
                   "elliptic_solver.F:176-178)");
    return SWM_ERR_NOT_CONVERGED;
  }
  return SWM_OK;
}

// Mixed path: the same solve as FP64 defect correction.  The FP32 right-hand
// side (mean removed with FP64 sums) is the problem.  The solve starts from
// the quadratic extrapolation 3 x_n - 3 x_(n-1) + x_(n-2) of the last three
// solutions (the linear 2 x_n - x_(n-1) with two, the last one alone after
// an initialisation), third-order accurate in dt where the reference's warm
// start is first-order, and forms its FP64 defect r = rhs - L(x).  A cycle
// runs one FP32 V-cycle on L(e) = r from e = 0, fills the halos of e, and
// one pass (k_defect) adds e to the iterate and forms the new FP64 defect
// with the sums of r^2.  One cycle, then the criterion before each further
// cycle, at most solver_max_cycles cycles (the start and the cycles before
// the first check are settings: swm_set_mixed_start).  The defect is formed
// afresh in FP64 every cycle, so the FP32 cycles only ever see a small,
// accurately known residual.  The iterates move between FP64 buffers that
// are renamed at the end: xout holds the solution, xprev and xprev2 the two
// before.
swm_status elliptic_mixed(swm_ctx* C, float* rhs, double* xout, int nb_rhs, const double** bsum_out) {
  const int ns = C->ns;
  View<float>& V = C->m;
  const swm::dev::Geo& G64 = C->d.geo;
  float* r = C->mx_r;
  float* e = C->mx_e;
  const unsigned nb = nblk(ns);
  // k_defect forms two slots per thread (ns is even: iim is): its launches
  // have nbd blocks, and nbs is the block count of the partial sums of d^2
  // that the last defect pass left
  const unsigned nbd = nblk(ns / 2);
  unsigned nbs = nbd;
  const size_t mb = C->mb;
  // The iterate's halos are consistent (every defect pass stores the whole
  // iterate, halos included) and the solve reads the right-hand side at real
  // cells only, so neither is filled here.  The area-weighted means
  // (elliptic_solver.F:80-92): FP64 sums of the tendencies kernel's block
  // partials; the consumers take the mean off on the fly.  The relative
  // criterion needs the mean-free right-hand side in memory for its norms.
  const bool fixed = C->fixed_cycles > 0;
  // the start: 3 x_n - 3 x_(n-1) + x_(n-2) with three solutions held, the
  // linear 2 x_n - x_(n-1) with two, x_n with one (or as set)
  const int order = std::min(C->mixed_extrapolate, C->nsol - 1);
  // with the quadratic start (and the absolute criterion) the means' sums
  // are formed in one launch with the start's gauge sums
  const bool late_mean = order >= 2 && !(C->solver_relative && !fixed);
  if (!late_mean) sum_final(C, C->part_rhs, nb_rhs, C->rmean, 2);
  const double* bsum = C->rmean;
  if (C->solver_relative && !fixed) {
    hipLaunchKernelGGL(swm::dev::k_sub_mean<float>, zgrid(C, nblk(ns)), dim3(BLK), 0, C->stream, ns, C->rmean,
                       C->total_area, rhs, rhs + ns, mb);
    bsum = nullptr;
    swm_status st0 = rhs_mean_squares<float>(C, rhs, 2);
    if (st0) return st0;
  }
  *bsum_out = bsum;
  const double total = C->total_area;
  // With the reference's two finest sweeps, each defect pass also forms the
  // next correction cycle's first sweep (from zero) in the finest level's
  // scratch, field v in plane v, and the cycle starts at its second sweep.
  const float* fine_lwt = (C->mixed_fine_sweeps == 2 && C->g > swm::dev::CO_MAXLEV) ? V.lev[C->g - 1].lwt : nullptr;
  auto first_e = [&](int, size_t o) -> float* { return fine_lwt ? V.lev[C->g - 1].temp + o : nullptr; };
  // the extrapolated start goes to xalt; then the iterates alternate between
  // xalt and xprev2, whose content is no longer needed once the start is
  // formed (xprev is kept: it is the next solve's xprev2)
  double* blk[3] = {xout, C->xalt, C->xprev2};
  double* cur[2] = {blk[1], blk[1] + ns};
  double* oth[2] = {blk[2], blk[2] + ns};
  // the mean change of the extrapolated start still to be taken off, per field
  bool shift[2] = {false, false};
  const double inv_area = 1.0 / C->total_area;
  // last: no further cycle follows (the fixed schedule's last one), or none
  // is expected (the criterion is tested next): the defect's FP32 copy and
  // the next cycle's first sweep are not stored; redefect forms them if a
  // cycle does follow
  auto launch_defect = [&](int nv, int v0, bool last) {
    const size_t o = (size_t)v0 * ns;
    const double* gg = shift[v0] ? C->gauge + v0 : nullptr;
    float* e0 = last ? nullptr : first_e(v0, o);
    float* rr = last ? nullptr : r + o;
    if (nv == 2)
      hipLaunchKernelGGL((swm::dev::k_defect<2, true>), zgrid(C, nbd), dim3(BLK), 0, C->stream, G64, cur[0], e, oth[0],
                         rhs, rr, C->part_slots, gg, inv_area, bsum, total, fine_lwt, e0);
    else
      hipLaunchKernelGGL((swm::dev::k_defect<1, true>), zgrid(C, nbd), dim3(BLK), 0, C->stream, G64, cur[v0], e + o,
                         oth[v0], rhs + o, rr, C->part_slots, gg, inv_area, bsum ? bsum + v0 : nullptr, total,
                         fine_lwt, e0);
    nbs = nbd;
    for (int v = v0; v < v0 + nv; ++v) {
      std::swap(cur[v], oth[v]);
      shift[v] = false;
    }
  };
  auto redefect = [&](int nv, int v0) {
    const size_t o = (size_t)v0 * ns;
    if (nv == 2)
      hipLaunchKernelGGL((swm::dev::k_defect<2, false>), zgrid(C, nbd), dim3(BLK), 0, C->stream, G64, cur[0], e, cur[0],
                         rhs, r, C->part_slots, (const double*)nullptr, inv_area, bsum, total, fine_lwt, first_e(0, 0));
    else
      hipLaunchKernelGGL((swm::dev::k_defect<1, false>), zgrid(C, nbd), dim3(BLK), 0, C->stream, G64, cur[v0], e + o,
                         cur[v0], rhs + o, r + o, C->part_slots, (const double*)nullptr, inv_area,
                         bsum ? bsum + v0 : nullptr, total, fine_lwt, first_e(v0, o));
    nbs = nbd;
  };
  // one FP32 correction cycle on fields v0.. (nv of them: both, or one);
  // the correction lands in e or, with another finest sweep count, in that
  // level's scratch (same plane layout)
  auto cycle = [&](int nv, int v0, bool last) {
    // the two fields' planes are contiguous: cur[0] + ns == cur[1] holds while they move together
    float* res = vcycle<float>(C, V, nv, e + (size_t)v0 * ns, r + (size_t)v0 * ns, true, C->mixed_fine_sweeps,
                               fine_lwt != nullptr, (size_t)v0 * ns);
    e = res - (size_t)v0 * ns;
    if (nv == 2) gather<float>(C, C->face, {e, e + ns});
    else gather<float>(C, C->face, {e + (size_t)v0 * ns});
    launch_defect(nv, v0, last);
    e = C->mx_e;
  };
  const bool extrap = order >= 1;
  if (order >= 2) {
    // the quadratic start pointwise, then its defect with one read per
    // stencil point (measured: 25 us faster at g=9 than forming the start at
    // the seven points of every stencil, three reads each)
    hipLaunchKernelGGL(swm::dev::k_extrap_start, zgrid(C, nb), dim3(BLK), 0, C->stream, ns, C->area_mask, blk[0],
                       C->xprev, blk[2], blk[1], C->part_slots, mb);
    if (late_mean)
      sum_final2(C, C->part_rhs, nb_rhs, C->rmean, 2, C->part_slots + 2 * (size_t)nb, (int)nb, C->gauge, 2);
    // (its d^2 sums, nbd blocks, go to columns 0 and 1 below the start's
    // gauge sums in columns 2 and 3 of nb blocks each)
    hipLaunchKernelGGL((swm::dev::k_defect<2, false>), zgrid(C, nbd), dim3(BLK), 0, C->stream, G64, blk[1], e, blk[1],
                       rhs, r, C->part_slots, (const double*)nullptr, inv_area, bsum, total, fine_lwt, first_e(0, 0));
    nbs = nbd;
  } else {
    hipLaunchKernelGGL(swm::dev::k_defect_extrap<2>, zgrid(C, nb), dim3(BLK), 0, C->stream, G64, C->area_mask, blk[0],
                       extrap ? C->xprev : blk[0], blk[1], rhs, r, C->part_slots, bsum, total, fine_lwt, first_e(0, 0));
    nbs = nb;
  }
  if (extrap) {
    if (!late_mean) sum_final(C, C->part_slots + 2 * (size_t)nb, (int)nb, C->gauge, 2);
    shift[0] = shift[1] = true;
  }
  int iter;
  bool any = false;
  const double* post_at = nullptr;   // the iterate the last test's queued velocity kernel read
  if (fixed) {
    // the fixed schedule: no host-side test; the last defect's mean squares
    // go to the member's residual record on the device
    // (the start-up cycles until the warm start has its full order)
    iter = C->nsol > C->mixed_extrapolate ? C->fixed_cycles : std::max(C->fixed_cycles, C->fixed_start);
    for (int it = 0; it < iter; ++it) cycle(2, 0, it == iter - 1);
    sum_final(C, C->part_slots, (int)nbs, C->sums, 2);
    hipLaunchKernelGGL(swm::dev::k_track_resid, zgrid(C, 1), dim3(64), 0, C->stream, C->sums, 1.0 / (double)C->max_ig,
                       C->solver_tol, C->track, mb);
  } else {
    // the defect before the first test is stored lazily (a converged solve
    // needs neither its FP32 copy nor the next first sweep)
    for (int it = 0; it < C->mixed_min_cycles; ++it) cycle(2, 0, it == C->mixed_min_cycles - 1);
    bool lazy = C->mixed_min_cycles > 0;
    const int M = C->nz;
    std::vector<double> ms(2 * (size_t)M);
    // the step's velocity kernel queued behind a test of both fields whose
    // iterate is in its final buffer (field 1 next to field 0, no mean
    // change left to take off): it runs if the device's copy of the test
    // passes, so the GPU need not wait for the host's decision
    auto post = [&]() {
      if (!C->post_solve || M != 1 || C->solver_relative || cur[1] != cur[0] + ns || shift[0] || shift[1]) return;
      hipLaunchKernelGGL(swm::dev::k_conv_go, dim3(1), dim3(64), 0, C->stream, C->host_sums_dev,
                         (double)C->max_ig, C->solver_tol, C->go);
      C->post_solve(cur[0]);
      post_at = cur[0];
    };
    swm_status st = member_norms(C, C->part_slots, (int)nbs, {0, 1}, ms.data(), post);
    if (st) return st;
    std::vector<int> bad;
    for (int v = 0; v < 2; ++v) {
      bool ok = true;
      for (int m = 0; m < M; ++m)
        if (!converged(C, m, v, ms[2 * m + v])) ok = false;
      if (!ok) bad.push_back(v);
    }
    iter = C->mixed_min_cycles;
    any = !bad.empty();
    while (any && iter < C->solver_max_cycles) {
      ++iter;
      const int nv = (int)bad.size(), v0 = bad[0];
      if (nv == 2 && cur[1] != cur[0] + ns) {
        // the fields ended on different buffers: bring field 1 next to field 0
        copy_state(C, cur[0] + ns, cur[1], (size_t)ns);
        oth[1] = cur[1];
        cur[1] = cur[0] + ns;
      }
      if (lazy) redefect(nv, v0);
      cycle(nv, v0, true);
      lazy = true;
      post_at = nullptr;
      st = member_norms(C, C->part_slots, (int)nbs, bad, ms.data(), nv == 2 ? post : std::function<void()>{});
      if (st) return st;
      any = false;
      for (int m = 0; m < M; ++m)
        for (int b = 0; b < nv; ++b)
          if (!converged(C, m, bad[b], ms[2 * m + b])) any = true;
    }
  }
  // a field that converged without a correction keeps the extrapolated
  // mean change: taken off here
  for (int v = 0; v < 2; ++v)
    if (shift[v])
      hipLaunchKernelGGL(swm::dev::k_add_scaled, zgrid(C, nb), dim3(BLK), 0, C->stream, (size_t)ns, C->gauge + v,
                         -inv_area, cur[v], mb);
  // both fields into one buffer (the other field's plane there is free),
  // which becomes xout; the previous solutions become xprev and xprev2
  if (cur[1] != cur[0] + ns) copy_state(C, cur[0] + ns, cur[1], (size_t)ns);
  double* fin = cur[0];
  C->post_done = post_at != nullptr && !any && post_at == fin;
  C->xprev2 = C->xprev;
  C->xprev = blk[0];
  C->xout = fin;
  C->xalt = fin == blk[1] ? blk[2] : blk[1];
  C->nsol = std::min(C->nsol + 1, 3);
  C->last_vcycles = iter;
  C->solves += 1; C->vcycles_total += iter; C->checks_total += fixed ? 0 : 1 + iter;
  if (any) {
    swm::set_error("elliptic solve did not converge (residual above the criterion after " +
                   std::to_string(C->solver_max_cycles) + " correction cycles)");
    return SWM_ERR_NOT_CONVERGED;
  }
  return SWM_OK;
}

void pointers(swm_ctx* C) {   // model_control.F:28-48
  const int sc = C->step_count;
  C->n4 = (sc + 1) % 2 + 1; C->n3 = sc % 2 + 1;
  C->n3f = (sc + 2) % 3 + 1; C->n2f = (sc + 1) % 3 + 1; C->n1f = sc % 3 + 1;
}
void weights(swm_ctx* C) {    // model_control.F:58-89
  const int sc = C->step_count;
  if (sc <= 1) { C->w3 = 1.0; C->w2 = 0.0; C->w1 = 0.0; }
  if (sc == 2) { C->w3 = 1.5; C->w2 = -0.5; C->w1 = 0.0; }
  if (sc >= 3) { C->w3 = 1.916666666666666666666; C->w2 = -1.333333333333333333333; C->w1 = 0.416666666666666666666; }
}

template <class R>
// after_step: the step has just filled the halos of the new velocity and
// (FCT) of the new thickness, and left the old velocity level as it was; a
// halo fill copies values, so filling those again changes nothing and is
// left out
void wrap_prognostics(swm_ctx* C, View<R>& V, bool after_step = false) {   // step.F:438-493
  const int n4 = C->n4 - 1;
  const int ns = C->ns;
  // with producer-side halo writes the step's kernels have stored every
  // halo copy this fill would make, and the tracer transport filled its own
  if (after_step && halo_push(C)) return;
  // inv_h (step.F:448-452) is not formed: nothing on the path reads it
  std::vector<R*> f = {V.eta[n4], V.div[n4], V.relative, V.ke};
  if (!after_step || C->advection != SWM_ADV_FCT) f.push_back(V.h[n4]);
  for (int q = 0; q < C->nt; ++q) f.push_back(V.tr[n4][q]);
  gather_list<R>(C, C->face, f);
  if (!after_step) gather<R>(C, C->vrtx, {V.u[0], V.u[0] + 2 * (size_t)ns, V.u[1], V.u[1] + 2 * (size_t)ns});
}

// get_remap_flux (step.F:687-771): remap_mass_and_tracers with the halo
// fills of the reference between its parts, then the ABM3 updates
void remap_step(swm_ctx* C, View<double>& V) {
  RemapDev& D = C->rm;
  const int ns = C->ns, nt = C->nt;
  const unsigned nb = nblk_b(ns, C->blk_step);
  hipStream_t S = C->stream;
  const int n3 = C->n3 - 1, n4 = C->n4 - 1;
  swm::dev::RemapArgs A{C->iim, C->jjm, ns, nt, D.cellp, D.edgep, D.cornp, C->radius, D.limit_min, D.limit_max,
                        D.flag_dev};
  const size_t np = (size_t)ns;
  double* const* tr3 = nt ? D.ptab + (size_t)n3 * nt : nullptr;
  double* const* tr4 = nt ? D.ptab + (size_t)n4 * nt : nullptr;
  hipLaunchKernelGGL(swm::dev::k_rm_construct, zgrid(C, nb), dim3(C->blk_step), 0, S, A, V.h[n3],
                     (const double* const*)tr3, D.mx, D.my, D.mp, D.tx, D.ty, D.tp);
  std::vector<double*> ff = {D.mx, D.my, D.mp};
  for (int q = 0; q < nt; ++q) {
    ff.push_back(D.tx + q * np);
    ff.push_back(D.ty + q * np);
    ff.push_back(D.tp + q * np);
  }
  gather_list<double>(C, C->face, ff);
  hipLaunchKernelGGL(swm::dev::k_rm_departure, zgrid(C, nb), dim3(C->blk_step), 0, S, A, C->dt, V.u[n3],
                     V.u[n3] + 2 * np, D.dp);
  gather<double>(C, C->vrtx, {D.dp, D.dp + 2 * np, D.dp + 4 * np});
  hipLaunchKernelGGL(swm::dev::k_rm_transport, zgrid(C, nb), dim3(C->blk_step), 0, S, A, D.dp, D.mx, D.my, D.mp, D.tx,
                     D.ty, D.tp, D.mflux, D.mtflux);
  std::vector<double*> ef = {D.mflux};
  for (int q = 0; q < nt; ++q) ef.push_back(D.mtflux + 3 * q * np);
  gather_list<double>(C, C->edge, ef);
  const int a = C->n3f - 1, b = C->n2f - 1, c = C->n1f - 1;
  double* const* tfa = nt ? D.ptab + (size_t)(2 + a) * nt : nullptr;
  double* const* tfb = nt ? D.ptab + (size_t)(2 + b) * nt : nullptr;
  double* const* tfc = nt ? D.ptab + (size_t)(2 + c) * nt : nullptr;
  // step.F:705 declares 1/dt real*4 for the tendencies; the edge flux of
  // remap.F:1578 divides in FP64
  const double dti_step = (double)(float)(1.0 / C->dt);
  hipLaunchKernelGGL(swm::dev::k_rm_update, zgrid(C, nb), dim3(C->blk_step), 0, S, A, C->dt, 1.0 / C->dt, dti_step,
                     C->w3, C->w2, C->w1, V.h[n3], (const double* const*)tr3, D.mflux, D.mtflux, V.h_f[a], V.h_f[b],
                     V.h_f[c], tfa, (const double* const*)tfb, (const double* const*)tfc, V.h[n4], tr4, V.mass_flux);
}

// step.F:114-212 step_dt (FCT scheme, get_fct_flux step.F:611-678)
template <class R>
swm_status step_dt(swm_ctx* C, View<R>& V) {
  const int ns = C->ns;
  const unsigned nb = nblk_b(ns, C->blk_step);
  hipStream_t S = C->stream;
  const swm::dev::GeoT<R>& G = V.geo;
  const int n3 = C->n3 - 1, n4 = C->n4 - 1;
  const R dt = (R)C->dt, w3 = (R)C->w3, w2 = (R)C->w2, w1 = (R)C->w1, grav = (R)C->grav;
  R* m3 = V.msh_f[C->n3f - 1];
  R* m2 = V.msh_f[C->n2f - 1];
  R* m1 = V.msh_f[C->n1f - 1];
  const bool fct = C->advection == SWM_ADV_FCT;
  const bool overlap = C->overlap && C->nt > 0 && fct;
  // producer-side halo writes (FCT step): the kernels below that form a
  // field whose halo the step fills store the halo copies themselves with
  // Gp, and those fills are left out
  const bool push = halo_push(C);
  swm::dev::GeoT<R> Gp = G;
  if (push) { Gp.pf = C->face.push(); Gp.pv = C->vrtx.push(); Gp.pe = C->edge.push(); }

  if (C->corners_ready) {
    C->corners_ready = false;   // queued by the previous step (corners_next)
  } else {
    hipLaunchKernelGGL(swm::dev::k_avg_corners<R>, zgrid(C, nb), dim3(C->blk_step), 0, S, Gp, G.sw, V.h[n3], V.hc);
    if (!push) gather<R>(C, C->vrtx, {V.hc});
  }
  if (C->advection == SWM_ADV_REMAP) {
    // get_remap_flux (step.F:687-771), FP64 contexts only
    if constexpr (sizeof(R) == 8) remap_step(C, V);
  } else if (C->advection != SWM_ADV_FCT) {
    // get_centered_flux / get_upwind_flux (step.F:495-602): the centred or
    // upstream mass flux (the other one lands in scratch), tracer fluxes,
    // then step_surface_height and step_tracer (step.F:221-327)
    const bool cen = C->advection == SWM_ADV_CENTERED;
    hipLaunchKernelGGL(swm::dev::k_mass_fluxes<R>, zgrid(C, nb), dim3(C->blk_step), 0, S, G, V.u[n3], V.h[n3], V.hc,
                       cen ? V.flx1 : V.mass_flux, cen ? V.mass_flux : V.flx1, (const R*)nullptr, (R*)nullptr, (R*)nullptr);
    for (int q = 0; q < C->nt; ++q) {
      if (cen)
        hipLaunchKernelGGL((swm::dev::k_tr_flux<R, true>), zgrid(C, nb), dim3(C->blk_step), 0, S, G, V.mass_flux, V.tr[n3][q],
                           V.tflux[q]);
      else
        hipLaunchKernelGGL((swm::dev::k_tr_flux<R, false>), zgrid(C, nb), dim3(C->blk_step), 0, S, G, V.mass_flux, V.tr[n3][q],
                           V.tflux[q]);
    }
    std::vector<R*> ef = {V.mass_flux};
    for (int q = 0; q < C->nt; ++q) ef.push_back(V.tflux[q]);
    gather_list<R>(C, C->edge, ef);
    const int a = C->n3f - 1, b = C->n2f - 1, c = C->n1f - 1;
    hipLaunchKernelGGL(swm::dev::k_height_update<R>, zgrid(C, nb), dim3(C->blk_step), 0, S, G, dt, w3, w2, w1, V.mass_flux,
                       V.h[n3], V.h_f[a], V.h_f[b], V.h_f[c], V.h[n4]);
    for (int q = 0; q < C->nt; ++q)
      hipLaunchKernelGGL(swm::dev::k_tracer_update<R>, zgrid(C, nb), dim3(C->blk_step), 0, S, G, dt, w3, w2, w1, V.tflux[q],
                         V.tr[n3][q], V.h[n3], V.h[n4], V.tr_f[a][q], V.tr_f[b][q], V.tr_f[c][q], V.tr[n4][q]);
  } else {
  // mass FCT (mass_fct.F:23-539)
  // with the divergence-curl flux of the old level (k_vort_curl's second
  // half, fused: the same u, corner thickness, h and eta), and the old
  // level's kinetic energy where the previous step of this call left it out
  hipLaunchKernelGGL(swm::dev::k_mass_fluxes<R>, zgrid(C, nb), dim3(C->blk_step), 0, S, Gp, V.u[n3], V.h[n3], V.hc, V.flx1, m3,
                     (const R*)V.eta[n3], V.dcurl, C->ke_deferred ? V.ke : (R*)nullptr);
  C->ke_deferred = false;
  if (!push) gather<R>(C, C->edge, {V.flx1, m3});
  hipLaunchKernelGGL(swm::dev::k_mass_low_beta<R>, zgrid(C, nb), dim3(C->blk_step), 0, S, Gp, dt, w3, w2, w1, V.flx1, m3, m2, m1,
                     V.h[n3], V.anti, V.mss_low, V.beta);
  if (!push) gather<R>(C, C->face, {V.beta});
  if constexpr (sizeof(R) == 8) {
    if (C->fct_rec)
      hipLaunchKernelGGL(swm::dev::k_fct_coef, zgrid(C, nb), dim3(C->blk_step), 0, S, G, V.beta, V.anti, C->fct_rec);
  }
  // the correction of the antidiffusive fluxes, or (SWM_FUSE_MASS_ANTI)
  // the final mass kernel corrects the values it reads itself
  const bool anti_fused = C->fuse_anti && C->anti_ok;
  if (!anti_fused) {
    hipLaunchKernelGGL(swm::dev::k_mass_anti<R>, zgrid(C, nb), dim3(C->blk_step), 0, S, Gp, V.beta, V.anti);
    if (!push) gather<R>(C, C->edge, {V.anti});
  }
  // with the first tracer's edge fluxes (k_tr_fluxes fused)
  swm::dev::TrFluxArgs<R> tf0{};
  if (C->nt > 0) tf0 = {V.tr[n3][0], V.tr[n4][0], V.told[0], V.f1, V.f2, w3, w2, w1};
  // with the vorticity edge flux of the new mass flux (k_vort_curl's first
  // half, fused)
  const swm::dev::VortFluxArgs<R> vfa{V.h[n3], V.eta[n3], V.vf};
  hipLaunchKernelGGL(swm::dev::k_mass_final<R>, zgrid(C, nb), dim3(C->blk_step), 0, S, Gp, dt, V.flx1, V.anti, V.mss_low,
                     V.mass_flux, V.h[n4], tf0, vfa, anti_fused ? (const R*)V.beta : (const R*)nullptr);
  // the tracer limiter fills the halo of the new mass in place (tracer_fct.F:344)
  if (!push) gather<R>(C, C->face, {V.h[n4]});
  }
  // Tracer FCT (tracer_fct.F:22-670).  The tracers depend on the new mass
  // and the mass flux only, and nothing after them but the final halo fill
  // reads them: with overlap they run on the second stream beside the
  // vorticity, divergence and elliptic solve, started here (mode 1) or by
  // the first V-cycle when it reaches a level of latency-bound launches
  // (mode >= 2: level g - mode + 1, through C->hook).
  swm_status tr_st = SWM_OK;
  auto tracers = [&]() {
    const hipStream_t main_stream = C->stream;
    if (overlap) {
      hipError_t he = hipEventRecord(C->ev_fork, main_stream);
      if (he == hipSuccess) he = hipStreamWaitEvent(C->stream2, C->ev_fork, 0);
      if (he != hipSuccess) {
        swm::set_error(std::string("tracer overlap: ") + hipGetErrorString(he));
        tr_st = SWM_ERR_HIP;
        return;
      }
      C->stream = C->stream2;
    }
    const hipStream_t St = C->stream;
    for (int q = 0; q < (fct ? C->nt : 0); ++q) {
      R* t3 = V.tr[n3][q];
      R* t4 = V.tr[n4][q];
      if (q > 0)
        hipLaunchKernelGGL(swm::dev::k_tr_fluxes<R>, zgrid(C, nb), dim3(C->blk_step), 0, St, G, w3, w2, w1, t3, t4,
                           V.told[q], V.mass_flux, V.f1, V.f2);
      if (q > 0 || !push) gather<R>(C, C->edge, {V.f1, V.f2});   // the first tracer's: k_mass_final's copies
      hipLaunchKernelGGL(swm::dev::k_tr_low<R>, zgrid(C, nb), dim3(C->blk_step), 0, St, G, dt, V.h[n3], t3, V.f1,
                         V.h[n4], V.ml, V.tlow);
      gather<R>(C, C->face, {V.ml, V.tlow});
      hipLaunchKernelGGL(swm::dev::k_tr_limiter<R>, zgrid(C, nb), dim3(C->blk_step), 0, St, G, dt, t3, V.ml, V.tlow,
                         V.mss_low, V.f1, V.f2, V.anti, V.bdn, V.bup);
      gather<R>(C, C->face, {V.bdn, V.bup});
      if constexpr (sizeof(R) == 8) {
        if (C->tr_rec)
          hipLaunchKernelGGL(swm::dev::k_tr_coef, zgrid(C, nb), dim3(C->blk_step), 0, St, G, V.bdn, V.bup, V.anti,
                             C->tr_rec + (size_t)(6 * q + 3) * ns);
      }
      // the correction of the antidiffusive fluxes, or (SWM_FUSE_MASS_ANTI)
      // the final tracer kernel corrects the values it reads itself
      const bool tr_fused = C->fuse_anti && C->anti_ok;
      if (!tr_fused) {
        hipLaunchKernelGGL(swm::dev::k_tr_anti<R>, zgrid(C, nb), dim3(C->blk_step), 0, St, G, V.bdn, V.bup, V.anti);
        gather<R>(C, C->edge, {V.anti});
      }
      hipLaunchKernelGGL(swm::dev::k_tr_final<R>, zgrid(C, nb), dim3(C->blk_step), 0, St, G, dt, V.ml, V.anti, V.h[n4],
                         t4, V.told[q], tr_fused ? (const R*)V.bdn : (const R*)nullptr,
                         tr_fused ? (const R*)V.bup : (const R*)nullptr);
    }
    // with producer-side halo writes the end-of-step fill (wrap_prognostics)
    // has nothing else left: the tracers' halos are filled here instead
    if (push) {
      std::vector<R*> tf;
      for (int q = 0; q < C->nt; ++q) tf.push_back(V.tr[n4][q]);
      gather_list<R>(C, C->face, tf);
    }
    // on the second stream the next step's corner thickness joins the
    // tracers (this step's last reader of it, k_mass_fluxes, is before the
    // fork; the next step waits for the join)
    if (overlap) corners_for_next(C);
    if (overlap) {
      const hipError_t he = hipEventRecord(C->ev_join, C->stream2);
      if (he != hipSuccess) {
        swm::set_error(std::string("tracer overlap: ") + hipGetErrorString(he));
        tr_st = SWM_ERR_HIP;
      }
      C->stream = main_stream;
    }
  };
  const int hook_level = C->g - C->overlap + 1;
  if (overlap && C->overlap >= 2 && hook_level > swm::dev::CO_MAXLEV) {
    C->hook = tracers;
    C->hook_level = hook_level;
  } else {
    tracers();
  }
  // vorticity and divergence (step.F:143-163, :336-429)
  // (FCT: both fluxes formed with the mass fluxes above)
  if (!fct) {
    hipLaunchKernelGGL(swm::dev::k_vort_curl<R>, zgrid(C, nb), dim3(C->blk_step), 0, S, Gp, V.mass_flux, V.h[n3],
                       V.eta[n3], V.hc, V.u[n3], V.vf, V.dcurl);
  }
  if (!push) gather<R>(C, C->edge, {V.vf, V.dcurl});
  // mixed path: the kernel also forms the block sums of the right-hand
  // sides' area-weighted means, which the solve takes off on the fly
  constexpr bool mixed = sizeof(R) == 4;
  hipLaunchKernelGGL((swm::dev::k_tendencies<R, mixed>), zgrid(C, nb), dim3(C->blk_step), 0, S, Gp, dt, grav, w3, w2, w1,
                     V.vf, V.dcurl, V.ke, V.h[n3], V.eta[n3], V.div[n3], V.eta_f[C->n3f - 1], V.eta_f[C->n2f - 1],
                     V.eta_f[C->n1f - 1], V.div_f[C->n3f - 1], V.div_f[C->n2f - 1], V.div_f[C->n1f - 1], V.eta[n4],
                     V.div[n4], V.relative, V.rhs, V.rhs + ns, mixed ? C->area_mask : nullptr,
                     mixed ? C->part_rhs : nullptr);
  // streamfunction and velocity potential (step.F:170-184)
  swm_status st;
  const double* bsum = nullptr;
  C->post_done = false;
  if constexpr (sizeof(R) == 8) {
    st = elliptic(C, V.rhs, C->xout, 2);
  } else {
    // the velocity kernel below, for the solve to queue before its tests
    C->post_solve = [&](const double* x) {
      hipLaunchKernelGGL(swm::dev::k_velocity<R>, zgrid(C, nb), dim3(C->blk_step), 0, S, Gp, x, x + ns, V.u[n4], V.rhs,
                         bsum, C->total_area, V.eta[n4], (const double*)C->go);
    };
    st = elliptic_mixed(C, V.rhs, C->xout, (int)nb, &bsum);
    C->post_solve = nullptr;
  }
  if (C->hook) {   // a solve that did not reach the hook's level
    std::function<void()> h = std::move(C->hook);
    C->hook = nullptr;
    h();
  }
  if (st) return st;
  if (tr_st) return tr_st;
  if constexpr (mixed) {
    // the new vorticity (the right-hand side less its mean) with the velocity
    // (unless the solve's last test queued it and it stands)
    if (!C->post_done)
      hipLaunchKernelGGL(swm::dev::k_velocity<R>, zgrid(C, nb), dim3(C->blk_step), 0, S, Gp, C->xout, C->xout + ns,
                         V.u[n4], V.rhs, bsum, C->total_area, V.eta[n4], (const double*)nullptr);
  } else {
    hipLaunchKernelGGL(swm::dev::k_post_elliptic<R>, zgrid(C, nb), dim3(C->blk_step), 0, S, Gp, V.rhs, V.eta[n4]);
    hipLaunchKernelGGL(swm::dev::k_velocity<R>, zgrid(C, nb), dim3(C->blk_step), 0, S, Gp, C->xout, C->xout + ns,
                       V.u[n4], (const R*)nullptr, (const double*)nullptr, 0.0, (R*)nullptr, (const double*)nullptr);
  }
  if (!push) gather<R>(C, C->vrtx, {V.u[n4], V.u[n4] + 2 * (size_t)ns});
  // the kinetic energy's halo is filled with the other fields' by
  // wrap_prognostics, which follows the step directly (with producer-side
  // halo writes: by this kernel).  With producer-side halo writes and a
  // next step in this call, that step's mass-flux kernel forms it instead.
  if (push && C->next_follows) C->ke_deferred = true;
  else hipLaunchKernelGGL(swm::dev::k_center_ke<R>, zgrid(C, nb), dim3(C->blk_step), 0, S, Gp, V.u[n4], V.ke);
  if (overlap) HIP_TRY(hipStreamWaitEvent(S, C->ev_join, 0));   // the tracers before the halo fill
  return SWM_OK;
}

// reference layout <-> planes for corner vectors (space, corner, slot)
void u_to_planes(const double* ref, double* planes, int ns) {
  for (int s = 0; s < ns; ++s)
    for (int ic = 0; ic < 2; ++ic)
      for (int sp = 0; sp < 2; ++sp) planes[(size_t)(ic + 2 * sp) * ns + s] = ref[sp + 2 * (ic + 2 * (size_t)s)];
}
void u_from_planes(const double* planes, double* ref, int ns) {
  for (int s = 0; s < ns; ++s)
    for (int ic = 0; ic < 2; ++ic)
      for (int sp = 0; sp < 2; ++sp) ref[sp + 2 * (ic + 2 * (size_t)s)] = planes[(size_t)(ic + 2 * sp) * ns + s];
}


// splitmix64 counter-based generator (same sequence as swm_mi355x/ics.py)
struct SplitMix {
  uint64_t state;
  explicit SplitMix(uint64_t s) : state(s) {}
  uint64_t next() {
    state += 0x9E3779B97F4A7C15ull;
    uint64_t z = state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  double normal() {
    const double u1 = std::max(uniform(), 1e-300), u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
  }
};

// zero the whole state (the reference's module arrays start at zero), the
// solver's warm start included, and the clock and step counter
// (time_manager_module.F:60-72); keep_hs leaves the surface height as it is
}  // namespace
namespace swm {
void shift_members(swm_ctx* C, ptrdiff_t d);
}  // namespace swm
namespace {

// every member's latest time level into both time levels (h, eta, div, u,
// tracers; FP64 and FP32 views), so that after the step counter restarts,
// the next step of each member is a forward Euler step from its latest
// state.  Works on all members whatever member a scope has selected.
template <class R>
swm_status latest_into_both_levels_view(swm_ctx* C, View<R>& V) {
  const int lv = C->n4 - 1, ot = 1 - lv;
  const size_t ns = C->ns;
  auto cp = [&](R* dst, const R* src, size_t n) -> swm_status {
    if (!dst || !src) return SWM_OK;
    HIP_TRY(hipMemcpy2DAsync(dst, C->mb, src, C->mb, n * sizeof(R), (size_t)C->nmem, hipMemcpyDeviceToDevice,
                             C->stream));
    return SWM_OK;
  };
  swm_status st;
  if ((st = cp(V.h[ot], V.h[lv], ns)) || (st = cp(V.eta[ot], V.eta[lv], ns)) || (st = cp(V.div[ot], V.div[lv], ns)) ||
      (st = cp(V.u[ot], V.u[lv], 4 * ns)))
    return st;
  for (int q = 0; q < C->nt; ++q)
    if ((st = cp(V.tr[ot][q], V.tr[lv][q], ns))) return st;
  return SWM_OK;
}
swm_status latest_into_both_levels(swm_ctx* C) {
  if (C->step_count == 0) return SWM_OK;
  const ptrdiff_t d = C->nz == 1 ? (ptrdiff_t)C->sel * (ptrdiff_t)C->mb : 0;   // undo a member scope
  const int nz = C->nz;
  swm::shift_members(C, -d);
  C->nz = C->nmem;
  swm_status st = latest_into_both_levels_view<double>(C, C->d);
  if (!st && C->precision == SWM_MIXED) st = latest_into_both_levels_view<float>(C, C->m);
  swm::shift_members(C, d);
  C->nz = nz;
  return st;
}

void reset_state(swm_ctx* C, bool keep_hs = false) {
  // the other members continue from their latest state (the step counter
  // restarts for all of them)
  if (C->nmem > 1) (void)latest_into_both_levels(C);
  C->solves = C->vcycles_total = C->checks_total = 0;
  fill<double>(C, C->d.state_begin, C->d.state_words, 0.0);
  if (C->precision == SWM_MIXED) fill<float>(C, C->m.state_begin, C->m.state_words, 0.0f);
  if (!keep_hs) fill_plain<double>(C, C->hs, C->ns, 0.0);
  fill<double>(C, C->xout, 2 * (size_t)C->ns, 0.0);
  fill<double>(C, C->xalt, 2 * (size_t)C->ns, 0.0);
  fill<double>(C, C->xprev, 2 * (size_t)C->ns, 0.0);
  fill<double>(C, C->xprev2, 2 * (size_t)C->ns, 0.0);
  C->nsol = 0;
  C->step_count = 0;
  C->nsecond = 0;
  C->initialized = false;
  C->member_init[C->sel] = 0;
  C->epoch += 1;
  fill<double>(C, C->track, 4, 0.0);
  pointers(C);
}

// FP64 field -> FP32 storage of the mixed path (y = x - off - c), and back
void to32(swm_ctx* C, const double* x, float* y, size_t n, const double* off = nullptr, double c = 0.0) {
  hipLaunchKernelGGL(swm::dev::k_to_f32, dim3(nblk(n)), dim3(BLK), 0, C->stream, n, x, off, c, y);
}
void to64(swm_ctx* C, const float* x, double* y, size_t n, const double* off = nullptr, double c = 0.0) {
  hipLaunchKernelGGL(swm::dev::k_to_f64, dim3(nblk(n)), dim3(BLK), 0, C->stream, n, x, off, c, y);
}

// h, hs, u (planes) and tracers of time level 1 are on the device: derive the
// rest as the reference initial conditions do (initialize.F:418-468 and
// operators_hl.F:153-266), then fill the halos as wrap_prognostics.  The
// mixed path starts from this FP64 state: h less its area-weighted mean h0,
// eta less f, the rest as is.
swm_status finish_init(swm_ctx* C) {
  const int ns = C->ns;
  hipStream_t S = C->stream;
  const unsigned nb = nblk(ns);
  View<double>& D = C->d;
  copy_state(C, D.h[1], D.h[0], (size_t)ns);
  copy_state(C, D.u[1], D.u[0], 4 * (size_t)ns);
  for (int t = 0; t < 2; ++t) {
    gather<double>(C, C->face, {D.h[t]});
    gather<double>(C, C->vrtx, {D.u[t], D.u[t] + 2 * (size_t)ns});
  }
  gather<double>(C, C->face, {C->hs});
  gather<double>(C, C->vrtx, {D.u[0], D.u[0] + 2 * (size_t)ns});
  hipLaunchKernelGGL(swm::dev::k_curl_div, dim3(nb), dim3(BLK), 0, S, D.geo, D.u[0], D.relative, D.div[0]);
  hipLaunchKernelGGL(swm::dev::k_center_ke<double>, dim3(nb), dim3(BLK), 0, S, D.geo, D.u[0], D.ke);
  gather<double>(C, C->face, {D.relative, D.div[0], D.ke});
  copy_state(C, D.div[1], D.div[0], (size_t)ns);
  hipLaunchKernelGGL(swm::dev::k_post_elliptic<double>, dim3(nb), dim3(BLK), 0, S, D.geo, D.relative, D.eta[0]);
  copy_state(C, D.eta[1], D.eta[0], (size_t)ns);
  for (int q = 0; q < C->nt; ++q)
    copy_state(C, D.tr[1][q], D.tr[0][q], (size_t)ns);
  wrap_prognostics<double>(C, D);
  if (C->precision == SWM_MIXED) {
    View<float>& M = C->m;
    // the thickness offset: the area-weighted mean thickness; the members
    // of an ensemble share the one of the member initialised first
    if (C->nmem == 1 || !C->h0_set) {
      sum2<double>(C, D.h[0], D.h[0], C->area_mask, ns);
      double sh[2];
      swm_status st = read_sums(C, 2, sh);
      if (st) return st;
      C->h0 = sh[0] / C->total_area;
      C->h0_set = true;
    }
    M.geo.h0 = (float)C->h0;
    to32(C, C->hs, const_cast<float*>(M.geo.hs), ns);
    for (int t = 0; t < 2; ++t) {
      to32(C, D.h[t], M.h[t], ns, nullptr, C->h0);
      to32(C, D.eta[t], M.eta[t], ns, C->f);
      to32(C, D.div[t], M.div[t], ns);
      to32(C, D.u[t], M.u[t], 4 * (size_t)ns);
      for (int q = 0; q < C->nt; ++q) to32(C, D.tr[t][q], M.tr[t][q], ns);
    }
    to32(C, D.relative, M.relative, ns);
    to32(C, D.ke, M.ke, ns);
  }
  HIP_TRY(hipStreamSynchronize(S));
  HIP_TRY(hipGetLastError());
  C->member_init[C->sel] = 1;
  C->initialized = std::all_of(C->member_init.begin(), C->member_init.end(), [](char c) { return c != 0; });
  return SWM_OK;
}


// ---------------------------------------- host-array operator entry points --
// (ncomp, iim, jjm, 10) reference layout <-> ncomp device planes
void ref_to_planes(const double* ref, int ncomp, int ns, double* planes) {
  for (int s = 0; s < ns; ++s)
    for (int c = 0; c < ncomp; ++c) planes[(size_t)c * ns + s] = ref[(size_t)ncomp * s + c];
}
void planes_to_ref(const double* planes, int ncomp, int ns, double* ref) {
  for (int s = 0; s < ns; ++s)
    for (int c = 0; c < ncomp; ++c) ref[(size_t)ncomp * s + c] = planes[(size_t)c * ns + s];
}

template <class... T>
bool all_set(const T*... p) {
  return ((p != nullptr) && ...);
}

// device buffers of one operator call (allocated per call, freed on return;
// the model state and the step's scratch are not touched).  ns is the
// number of slots of the call's arrays (nx * ny * panels).
struct OpIO {
  swm_ctx* C;
  size_t ns;
  std::vector<double*> bufs;
  std::vector<double> host;
  ~OpIO() {
    for (double* p : bufs) (void)hipFree(p);
  }
  double* dev(int nplanes) {
    double* p = nullptr;
    if (hipMalloc(&p, (size_t)nplanes * ns * sizeof(double)) != hipSuccess) return nullptr;
    bufs.push_back(p);
    return p;
  }
  // reference layout (lead, slots), lead leading entries per slot -> planes;
  // plane p takes leading entry perm[p] (identity when perm is empty)
  double* put(const double* ref, int lead, int nplanes, const std::vector<int>& perm = {}) {
    double* d = dev(nplanes);
    if (!d) return nullptr;
    host.assign((size_t)nplanes * ns, 0.0);
    for (size_t s = 0; s < ns; ++s)
      for (int p = 0; p < nplanes; ++p) host[(size_t)p * ns + s] = ref[(size_t)lead * s + (perm.empty() ? p : perm[p])];
    if (hipMemcpy(d, host.data(), host.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    return d;
  }
  double* zeros(int nplanes) {
    double* d = dev(nplanes);
    if (d) fill_plain<double>(C, d, (size_t)nplanes * ns, 0.0);
    return d;
  }
  swm_status get(double* ref, const double* d, int lead, const std::vector<int>& perm = {}) {
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(C->stream));
    host.assign((size_t)lead * ns, 0.0);
    HIP_TRY(hipMemcpy(host.data(), d, host.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t s = 0; s < ns; ++s)
      for (int p = 0; p < lead; ++p) ref[(size_t)lead * s + (perm.empty() ? p : perm[p])] = host[(size_t)p * ns + s];
    return SWM_OK;
  }
};

// plane order ic + 2*space of a corner vector (space, corner): leading index
const std::vector<int> kCornerVec = {0, 2, 1, 3};
// planes kk + 3*(a + 2*ic), a = 0, 1, of normal / tangent (ndata=3, r3=3, corner)
const std::vector<int> kFrame = {0, 1, 2, 3, 4, 5, 9, 10, 11, 12, 13, 14};

// (iim, jjm, nvar, 10) <-> nvar planes
void vars_to_planes(const double* ref, int nvar, int pl, double* planes) {
  for (int k = 0; k < 10; ++k)
    for (int v = 0; v < nvar; ++v)
      std::memcpy(planes + (size_t)v * pl * 10 + (size_t)k * pl, ref + ((size_t)v + (size_t)nvar * k) * pl, pl * sizeof(double));
}
void planes_to_vars(const double* planes, int nvar, int pl, double* ref) {
  for (int k = 0; k < 10; ++k)
    for (int v = 0; v < nvar; ++v)
      std::memcpy(ref + ((size_t)v + (size_t)nvar * k) * pl, planes + (size_t)v * pl * 10 + (size_t)k * pl, pl * sizeof(double));
}
}  // namespace

namespace swm {

// The dense coarse solve of the mixed path (k_coarse_direct): the real cells
// of level l (interior slots panel by panel, then the two poles) and the
// level's Laplacian (the smoother's operator: l_weights(7) at the cell,
// l_weights(1..6) at its neighbours through the level's halo map; the
// poles' forms leave one out, as k_relax), bordered by ones so that it is
// regular: B = [[A, 1], [1^T, 0]], row-major m x m with m = n + 1.
void coarse_bordered(const std::vector<double>& lw, int l, std::vector<double>& B, std::vector<int>& cell) {
  const int iim = 2 + (1 << l), jjm = iim, ns = iim * jjm * 10;
  auto SL = [&](int i, int j, int k) { return (i - 1) + iim * ((j - 1) + jjm * (k - 1)); };
  std::vector<int> src(ns);
  for (int n = 0; n < ns; ++n) src[n] = n;
  const swm::HaloTable t = swm::build_simple_face_table(l);
  for (size_t q = 0; q < t.dst.size(); ++q) src[t.dst[q]] = t.src[q];
  cell.clear();
  for (int k = 1; k <= 10; ++k)
    for (int j = 2; j <= jjm - 1; ++j)
      for (int i = 2; i <= iim - 1; ++i) cell.push_back(SL(i, j, k));
  cell.push_back(SL(1, jjm - 1, 1));
  cell.push_back(SL(iim - 1, 1, 10));
  const int n = (int)cell.size(), m = n + 1;
  std::vector<int> idx(ns, -1);
  for (int r = 0; r < n; ++r) idx[cell[r]] = r;
  B.assign((size_t)m * m, 0.0);
  for (int r = 0; r < n; ++r) {
    const int s = cell[r];
    const int i = s % iim + 1, j = (s / iim) % jjm + 1, k = s / (iim * jjm) + 1;
    const bool np = i == 1 && j == jjm - 1, sp = i == iim - 1 && j == 1;
    int nb[6] = {SL(i - 1, j, k), SL(i - 1, j - 1, k), SL(i, j - 1, k), SL(i + 1, j, k), SL(i + 1, j + 1, k),
                 SL(i, j + 1, k)};
    if (np) { nb[0] = -1; nb[1] = SL(iim, 1, k); }
    if (sp) { nb[1] = SL(1, jjm, k); nb[2] = -1; }
    B[(size_t)r * m + r] += lw[7 * (size_t)s + 6];
    for (int q = 0; q < 6; ++q) {
      if (nb[q] < 0) continue;
      const int c = idx[src[nb[q]]];
      if (c >= 0) B[(size_t)r * m + c] += lw[7 * (size_t)s + q];
    }
    B[(size_t)r * m + n] = 1.0;
    B[(size_t)n * m + r] = 1.0;
  }
}

// level 3: the FP32 n x n block of B's inverse by Gauss-Jordan elimination
// with partial pivoting in FP64 on the host (0.1 s)
void coarse_direct_matrix(const std::vector<double>& lw, int l, std::vector<float>& M, std::vector<int>& cell) {
  std::vector<double> B;
  coarse_bordered(lw, l, B, cell);
  const int n = (int)cell.size(), m = n + 1;
  std::vector<double> X((size_t)m * m, 0.0);
  for (int r = 0; r < m; ++r) X[(size_t)r * m + r] = 1.0;
  for (int c = 0; c < m; ++c) {
    int p = c;
    for (int r = c + 1; r < m; ++r)
      if (std::fabs(B[(size_t)r * m + c]) > std::fabs(B[(size_t)p * m + c])) p = r;
    if (p != c)
      for (int q = 0; q < m; ++q) {
        std::swap(B[(size_t)p * m + q], B[(size_t)c * m + q]);
        std::swap(X[(size_t)p * m + q], X[(size_t)c * m + q]);
      }
    const double d = 1.0 / B[(size_t)c * m + c];
    for (int q = 0; q < m; ++q) { B[(size_t)c * m + q] *= d; X[(size_t)c * m + q] *= d; }
    for (int r = 0; r < m; ++r) {
      if (r == c) continue;
      const double f = B[(size_t)r * m + c];
      if (f == 0.0) continue;
      for (int q = 0; q < m; ++q) { B[(size_t)r * m + q] -= f * B[(size_t)c * m + q]; X[(size_t)r * m + q] -= f * X[(size_t)c * m + q]; }
    }
  }
  M.assign((size_t)n * n, 0.0f);
  for (int r = 0; r < n; ++r)
    for (int c = 0; c < n; ++c) M[(size_t)r * n + c] = (float)X[(size_t)r * m + c];
}

// level 4 (2563 x 2563): the same inverse by in-place Gauss-Jordan on the
// device (k_gj_pivot / k_gj_update, FP64, about 0.1 s), the row swaps undone
// as column swaps on the host; its FP32 n x n block into Md
swm_status coarse_direct_matrix_device(swm_ctx* C, const std::vector<double>& lw, int l, float* Md, int* celld,
                                       int nexp) {
  std::vector<double> B;
  std::vector<int> cell;
  coarse_bordered(lw, l, B, cell);
  const int n = (int)cell.size(), m = n + 1;
  if (n != nexp) { swm::set_error("coarse solve: level size"); return SWM_ERR_DIMS; }
  double *A = nullptr, *rowb = nullptr, *colb = nullptr;
  int* perm = nullptr;
  swm_status st = SWM_OK;
  auto chk = [&](hipError_t e) {
    if (e != hipSuccess && st == SWM_OK) {
      swm::set_error(std::string("coarse solve set-up: ") + hipGetErrorString(e));
      st = SWM_ERR_HIP;
    }
    return st == SWM_OK;
  };
  if (chk(hipMalloc(&A, B.size() * sizeof(double))) && chk(hipMalloc(&rowb, m * sizeof(double))) &&
      chk(hipMalloc(&colb, m * sizeof(double))) && chk(hipMalloc(&perm, m * sizeof(int))) &&
      chk(hipMemcpyAsync(A, B.data(), B.size() * sizeof(double), hipMemcpyHostToDevice, C->stream))) {
    const unsigned nbu = (unsigned)(((size_t)m * m + 255) / 256);
    for (int c = 0; c < m; ++c) {
      hipLaunchKernelGGL(swm::dev::k_gj_pivot, dim3(1), dim3(swm::dev::GJ_THREADS), 0, C->stream, A, m, c, rowb, colb, perm);
      hipLaunchKernelGGL(swm::dev::k_gj_update, dim3(nbu), dim3(256), 0, C->stream, A, m, c, (const double*)rowb,
                         (const double*)colb);
    }
    std::vector<int> pv(m);
    if (chk(hipGetLastError()) &&
        chk(hipMemcpyAsync(B.data(), A, B.size() * sizeof(double), hipMemcpyDeviceToHost, C->stream)) &&
        chk(hipMemcpyAsync(pv.data(), perm, m * sizeof(int), hipMemcpyDeviceToHost, C->stream)) &&
        chk(hipStreamSynchronize(C->stream))) {
      for (int c = m - 1; c >= 0; --c)
        if (pv[c] != c)
          for (int r = 0; r < m; ++r) std::swap(B[(size_t)r * m + c], B[(size_t)r * m + pv[c]]);
      std::vector<float> M((size_t)n * n);
      bool finite = true;
      for (int r = 0; r < n; ++r)
        for (int c = 0; c < n; ++c) {
          const double x = B[(size_t)r * m + c];
          finite = finite && std::isfinite(x);
          M[(size_t)r * n + c] = (float)x;
        }
      if (!finite && st == SWM_OK) { swm::set_error("coarse solve set-up: singular level-4 operator"); st = SWM_ERR_NAN; }
      if (st == SWM_OK && chk(hipMemcpy(Md, M.data(), M.size() * sizeof(float), hipMemcpyHostToDevice)))
        chk(hipMemcpy(celld, cell.data(), cell.size() * sizeof(int), hipMemcpyHostToDevice));
    }
  }
  if (A) (void)hipFree(A);
  if (rowb) (void)hipFree(rowb);
  if (colb) (void)hipFree(colb);
  if (perm) (void)hipFree(perm);
  return st;
}

// buffers of a view: multigrid level weights and maps in the shared arena
// (the maps shared with the FP64 view when given); the levels' fields, the
// state and scratch in the member block
template <class R>
void layout_view(swm_ctx* C, View<R>& V, const View<double>* maps) {
  Arena& A = C->arena;
  Arena& MA = C->marena;
  const size_t ns = C->ns;
  V.lev.resize(C->g);
  for (int l = 1; l <= C->g; ++l) {
    LevelT<R>& L = V.lev[l - 1];
    L.iim = L.jjm = 2 + (1 << l);
    L.ns = L.iim * L.jjm * 10;
    const size_t nl = L.ns;
    L.lw = A.take<R>(7 * nl); L.lwt = A.take<R>(nl);
    L.alpha = MA.take<R>(2 * nl); L.beta = MA.take<R>(2 * nl); L.temp = MA.take<R>(2 * nl);
    L.src = maps ? maps->lev[l - 1].src : A.take<int>(nl);
    L.L = swm::dev::LevT<R>{L.iim, L.jjm, L.ns, L.lw, L.lwt, L.src, 0};
  }
  auto take = [&](size_t n) { return MA.take<R>(n); };
  const size_t before = MA.used;
  for (int t = 0; t < 2; ++t) {
    V.h[t] = take(ns); V.eta[t] = take(ns); V.div[t] = take(ns); V.u[t] = take(4 * ns);
    V.tr[t].clear();
    for (int q = 0; q < C->nt; ++q) V.tr[t].push_back(take(ns));
  }
  V.told.clear();
  for (int q = 0; q < C->nt; ++q) V.told.push_back(take(ns));
  for (int t = 0; t < 3; ++t) { V.eta_f[t] = take(ns); V.div_f[t] = take(ns); V.msh_f[t] = take(3 * ns); }
  for (int t = 0; t < 3; ++t) {
    V.h_f[t] = take(ns);
    V.tr_f[t].clear();
    for (int q = 0; q < C->nt; ++q) V.tr_f[t].push_back(take(ns));
  }
  V.tflux.clear();
  for (int q = 0; q < C->nt; ++q) V.tflux.push_back(take(3 * ns));
  V.relative = take(ns); V.ke = take(ns); V.tlow = take(ns); V.mass_flux = take(3 * ns);
  V.hc = take(2 * ns); V.flx1 = take(3 * ns); V.anti = take(3 * ns); V.mss_low = take(ns);
  V.beta = take(ns); V.f1 = take(3 * ns); V.f2 = take(3 * ns); V.ml = take(ns); V.bdn = take(ns); V.bup = take(ns);
  V.vf = take(3 * ns); V.dcurl = take(3 * ns); V.rhs = take(2 * ns);
  V.state_begin = V.h[0];
  V.state_words = (MA.used - before) / sizeof(R);
}

// FP64 metric planes (reference layout -> planes), tables, levels and views
void layout_all(swm_ctx* C, const swm::Grid& Gh) {
  Arena& A = C->arena;
  const size_t ns = C->ns;
  auto take = [&](size_t n) { return A.take<double>(n); };
  C->area = take(ns); C->area_inv = take(ns); C->cai = take(2 * ns); C->sub_area = take(6 * ns); C->dc = take(6 * ns);
  C->nrm = take(12 * ns); C->tng = take(12 * ns); C->vw = take(12 * ns); C->sw = take(6 * ns); C->lw = take(7 * ns);
  C->f = take(ns); C->hs = take(ns); C->mask = take(ns); C->area_mask = take(ns); C->pxyz = take(3 * ns);
  for (auto tb : {std::make_pair(&Gh.face, &C->face), std::make_pair(&Gh.vrtx, &C->vrtx), std::make_pair(&Gh.edge, &C->edge)})
    if (C->layout_status == SWM_OK) C->layout_status = make_table(C, *tb.first, C->ns, *tb.second);
  // the member block: solver iterates, partial sums, then the views' state
  auto mt = [&](size_t n) { return C->marena.take<double>(n); };
  C->xout = mt(2 * ns); C->r2 = mt(ns); C->xalt = mt(2 * ns); C->xprev = mt(2 * ns); C->xprev2 = mt(2 * ns); C->gauge = mt(2);
  C->part_slots = mt(4 * (size_t)nblk(ns));
  C->part = mt(2 * (size_t)C->npart); C->sums = mt(64);
  C->part_budget = mt((8 + swm::dev::BUDGET_MAXTR) * (size_t)C->npart);
  C->track = mt(4);
  C->rmean = mt(4);
  C->go = mt(1);
  C->part_rhs = mt(2 * (size_t)nblk_b(ns, 64));
  layout_view<double>(C, C->d, nullptr);
  C->d.geo = swm::dev::Geo{C->iim, C->jjm, C->ns, C->area, C->area_inv, C->cai, C->sub_area, C->dc, C->nrm, C->tng,
                           C->vw, C->sw, C->lw, C->f, C->hs, C->mask, 0.0, nullptr};
  if (C->precision == SWM_MIXED) {
    // the FP64 defects' symmetric Laplacian coefficients and masked area_inv
    C->d.geo.sym = take(3 * ns);
    C->d.geo.aim = take(ns);
    // FP32 copies of the metric planes, same order as the FP64 ones
    auto tf = [&](size_t n) { return A.take<float>(n); };
    swm::dev::GeoT<float>& g = C->m.geo;
    float *area = tf(ns), *ai = tf(ns), *cai = tf(2 * ns), *sa = tf(6 * ns), *dc = tf(6 * ns), *nrm = tf(12 * ns),
          *tng = tf(12 * ns), *vw = tf(12 * ns), *sw = tf(6 * ns), *lw = tf(7 * ns), *f = tf(ns), *hs = tf(ns),
          *mask = tf(ns);
    g = swm::dev::GeoT<float>{C->iim, C->jjm, C->ns, area, ai, cai, sa, dc, nrm, tng, vw, sw, lw, f, hs, mask, 0.0f, f};
    g.ndc = tf(12 * ns);
    g.tdc = tf(12 * ns);
    layout_view<float>(C, C->m, &C->d);
    for (int l = swm::dev::CO_MAXLEV + 1; l <= C->g; ++l) {
      LevelT<float>& L = C->m.lev[l - 1];
      L.L.hw = A.take<_Float16>(6 * (size_t)L.ns);
      L.L.hcw = A.take<_Float16>((size_t)L.ns);
    }
    if (C->g > swm::dev::CO_MAXLEV) {
      const int e = 1 << swm::dev::CO_MAXLEV;
      C->cd_n = 10 * e * e + 2;
      C->cd_m = A.take<float>((size_t)C->cd_n * C->cd_n);
      C->cd_cell = A.take<int>((size_t)C->cd_n);
    }
    if (C->g > swm::dev::CO_MAXLEV + 1) {
      const int e = 2 << swm::dev::CO_MAXLEV;
      C->cd4_n = 10 * e * e + 2;
      C->cd4_m = A.take<float>((size_t)C->cd4_n * C->cd4_n);
      C->cd4_cell = A.take<int>((size_t)C->cd4_n);
    }
    C->mx_r = C->marena.take<float>(2 * ns);
    C->mx_e = C->marena.take<float>(2 * ns);
    if (C->cd_n) C->cd_b = C->marena.take<float>(2 * (size_t)std::max(C->cd_n, C->cd4_n));
  }
}

// every pointer into the member block moves by d bytes (selects another
// member's block for host-side calls, and back)
template <class T>
void shift_ptr(T*& p, ptrdiff_t d) {
  if (p) p = reinterpret_cast<T*>(reinterpret_cast<char*>(p) + d);
}
template <class R>
void shift_view(View<R>& V, ptrdiff_t d) {
  for (LevelT<R>& L : V.lev) { shift_ptr(L.alpha, d); shift_ptr(L.beta, d); shift_ptr(L.temp, d); }
  for (int t = 0; t < 2; ++t) {
    shift_ptr(V.h[t], d); shift_ptr(V.eta[t], d); shift_ptr(V.div[t], d); shift_ptr(V.u[t], d);
    for (R*& p : V.tr[t]) shift_ptr(p, d);
  }
  for (R*& p : V.told) shift_ptr(p, d);
  for (int t = 0; t < 3; ++t) {
    shift_ptr(V.eta_f[t], d); shift_ptr(V.div_f[t], d); shift_ptr(V.msh_f[t], d); shift_ptr(V.h_f[t], d);
    for (R*& p : V.tr_f[t]) shift_ptr(p, d);
  }
  for (R*& p : V.tflux) shift_ptr(p, d);
  for (R** p : {&V.relative, &V.ke, &V.tlow, &V.mass_flux, &V.hc, &V.flx1, &V.anti, &V.mss_low, &V.beta, &V.f1,
                &V.f2, &V.ml, &V.bdn, &V.bup, &V.vf, &V.dcurl, &V.rhs, &V.state_begin})
    shift_ptr(*p, d);
}
void shift_members(swm_ctx* C, ptrdiff_t d) {
  if (d == 0) return;
  shift_view(C->d, d);
  if (C->precision == SWM_MIXED) shift_view(C->m, d);
  for (double** p : {&C->xout, &C->r2, &C->xalt, &C->xprev, &C->xprev2, &C->gauge, &C->part_slots, &C->part, &C->sums,
                     &C->part_budget, &C->track, &C->rmean, &C->part_rhs, &C->go})
    shift_ptr(*p, d);
  shift_ptr(C->mx_r, d);
  shift_ptr(C->mx_e, d);
}

// host-side calls address one member: its block's pointers and launches over
// that member alone, for the scope's lifetime
struct MemberScope {
  swm_ctx* C;
  ptrdiff_t d;
  explicit MemberScope(swm_ctx* c) : C(c), d((ptrdiff_t)c->sel * (ptrdiff_t)c->mb) {
    shift_members(C, d);
    C->nz = 1;
  }
  ~MemberScope() {
    shift_members(C, -d);
    C->nz = C->nmem;
  }
};

swm_status up64(double* d, const std::vector<double>& h) {
  HIP_TRY(hipMemcpy(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
  return SWM_OK;
}

template <class R>
void launch_budgets(swm_ctx* C, View<R>& V, int nq) {
  const int n4 = C->n4 - 1;
  swm::dev::BudgetIn<R> B{};
  B.h = V.h[n4]; B.ke = V.ke; B.eta = V.eta[n4]; B.div = V.div[n4];
  for (int q = 0; q < C->nt && q < swm::dev::BUDGET_MAXTR; ++q) B.tr[q] = V.tr[n4][q];
  B.area = C->area; B.mask = C->mask; B.hs = C->hs;
  B.f = sizeof(R) == 4 ? C->f : nullptr;
  B.h0 = sizeof(R) == 4 ? C->h0 : 0.0;
  B.grav = C->grav;
  B.ns = C->ns;
  hipLaunchKernelGGL(swm::dev::k_budgets<R>, dim3(C->npart, nq), dim3(BLK), 0, C->stream, B, C->part_budget);
  sum_final(C, C->part_budget, C->npart, C->sums, nq);
}

// init_remap's quantities (built on the host) as device planes, the step's
// remapping fields and the pinned flag (SWM_ADV_REMAP)
swm_status remap_setup(swm_ctx* C, const swm::Grid& Gh) {
  const swm::RemapGeometry& R = swm::remap_geometry(Gh);
  RemapDev& D = C->rm;
  const size_t ns = (size_t)C->ns, nt = (size_t)C->nt;
  const size_t ne = 3 * ns, nc = 2 * ns;
  std::vector<double> cellp((size_t)swm::dev::RM_CELL * ns), edgep((size_t)swm::dev::RM_EDGE * ne),
      cornp((size_t)swm::dev::RM_CORN * nc);
  for (size_t n = 0; n < ns; ++n) {
    for (int c = 0; c < 2; ++c)
      for (int nv = 0; nv < 7; ++nv) cellp[(size_t)(c * 7 + nv) * ns + n] = R.gwgt_x3[c + 3 * (nv + 7 * n)];
    cellp[14 * ns + n] = R.garea_inv[n];
    cellp[15 * ns + n] = R.xav[n];
    cellp[16 * ns + n] = R.yav[n];
    cellp[17 * ns + n] = R.xxav[n];
    cellp[18 * ns + n] = R.xyav[n];
    cellp[19 * ns + n] = R.yyav[n];
    for (int nv = 0; nv < 6; ++nv)
      for (int c = 0; c < 2; ++c) cellp[(size_t)(20 + 2 * nv + c) * ns + n] = R.c_x3[c + 2 * (nv + 6 * n)];
    cellp[32 * ns + n] = R.area_inv_x3[n];
    static const int nd0[3] = {0, 0, 2}, nc0[3] = {0, 1, 1};   // the edge's midpoint w(nd0, :, nc0)
    for (int e = 0; e < 3; ++e) {
      const size_t h = e + 3 * n, d = (size_t)e * ns + n;
      auto put = [&](int plane, double v) { edgep[(size_t)plane * ne + d] = v; };
      for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 3; ++c) put(r * 3 + c, R.trans_x1_to_x4[r + 3 * (c + 3 * h)]);
      for (int nd = 0; nd < 4; ++nd)
        for (int r = 0; r < 2; ++r)
          for (int c = 0; c < 2; ++c) put(6 + 4 * nd + r + 2 * c, R.trans_x4_to_x3[r + 2 * (c + 2 * (nd + 4 * h))]);
      put(22, R.cl_x4[2 * h]);
      put(23, R.cr_x4[2 * h]);
      put(24, R.tl_x4[2 * h]); put(25, R.tl_x4[2 * h + 1]);
      put(26, R.tr_x4[2 * h]); put(27, R.tr_x4[2 * h + 1]);
      put(28, R.bl_x4[2 * h]); put(29, R.bl_x4[2 * h + 1]);
      put(30, R.br_x4[2 * h]); put(31, R.br_x4[2 * h + 1]);
      for (int nd = 0; nd < 4; ++nd)
        for (int c = 0; c < 2; ++c) put(32 + 2 * nd + c, R.p_x4[c + 2 * (nd + 4 * h)]);
      for (int r = 0; r < 3; ++r) put(40 + r, Gh.w[9 * (2 * n + nc0[e]) + nd0[e] + 3 * r]);
    }
    for (int ic = 0; ic < 2; ++ic) {
      const size_t h = 2 * n + ic, d = (size_t)ic * ns + n;
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 2; ++c) cornp[(size_t)(r + 3 * c) * nc + d] = Gh.trans_to_x1[9 * h + r + 3 * c];
      for (int r = 0; r < 3; ++r) cornp[(size_t)(6 + r) * nc + d] = Gh.c[3 * h + r];
    }
  }
  auto al = [](size_t bytes) { return (bytes + 255) / 256 * 256; };
  const size_t b_cell = al(cellp.size() * 8), b_edge = al(edgep.size() * 8), b_corn = al(cornp.size() * 8);
  const size_t b_cf = al(ns * 8), b_tf = al(std::max<size_t>(nt, 1) * ns * 8), b_dp = al(3 * nc * 8),
               b_mf = al(ne * 8), b_mtf = al(std::max<size_t>(nt, 1) * ne * 8), b_pt = al(std::max<size_t>(5 * nt, 1) * 8);
  const size_t total = b_cell + b_edge + b_corn + 3 * b_cf + 3 * b_tf + b_dp + b_mf + b_mtf + b_pt;
  HIP_TRY(hipMalloc(&D.mem, total));
  char* p = static_cast<char*>(D.mem);
  auto take = [&](size_t b) { char* q = p; p += b; return q; };
  D.cellp = (double*)take(b_cell); D.edgep = (double*)take(b_edge); D.cornp = (double*)take(b_corn);
  D.mx = (double*)take(b_cf); D.my = (double*)take(b_cf); D.mp = (double*)take(b_cf);
  D.tx = (double*)take(b_tf); D.ty = (double*)take(b_tf); D.tp = (double*)take(b_tf);
  D.dp = (double*)take(b_dp); D.mflux = (double*)take(b_mf); D.mtflux = (double*)take(b_mtf);
  D.ptab = (double**)take(b_pt);
  HIP_TRY(hipMemcpy(D.cellp, cellp.data(), cellp.size() * 8, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(D.edgep, edgep.data(), edgep.size() * 8, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(D.cornp, cornp.data(), cornp.size() * 8, hipMemcpyHostToDevice));
  HIP_TRY(hipMemset(D.mx, 0, 3 * b_cf + 3 * b_tf + b_dp + b_mf + b_mtf));
  std::vector<double*> tab;
  for (int t = 0; t < 2; ++t)
    for (size_t q = 0; q < nt; ++q) tab.push_back(C->d.tr[t][q]);
  for (int t = 0; t < 3; ++t)
    for (size_t q = 0; q < nt; ++q) tab.push_back(C->d.tr_f[t][q]);
  if (nt) HIP_TRY(hipMemcpy(D.ptab, tab.data(), tab.size() * sizeof(double*), hipMemcpyHostToDevice));
  HIP_TRY(hipHostMalloc(&D.flag_host, 64, hipHostMallocDefault));
  HIP_TRY(hipHostGetDevicePointer((void**)&D.flag_dev, D.flag_host, 0));
  *D.flag_host = 0;
  return SWM_OK;
}

// Whether the final mass kernel can take over set_corrected_mss_anti
// (SWM_FUSE_MASS_ANTI), from the host copies of the fill tables: every
// forward edge of a real cell that lies in the halo (edge 0 at (i+1, j),
// 1 at (i+1, j+1), 2 at (i, j+1)) is a fill destination that copies an edge
// owned by the slot the face fill copies there (or is set to zero), so that
// its corrected value follows from the halo copies of beta; and each real
// pole's two own edges copy edges whose upstream neighbour is the pole.
// The pole edges' source slots go to psrc.
bool anti_fusion_check(const swm_ctx* C, int psrc[4]) {
  const int iim = C->iim, jjm = C->jjm, ns = C->ns;
  auto SLh = [&](int i, int j, int k) { return (i - 1) + iim * ((j - 1) + jjm * (k - 1)); };
  std::vector<int> fsrc(ns, -1);
  for (size_t q = 0; q < C->face.hdst.size(); ++q) fsrc[C->face.hdst[q] % ns] = C->face.hsrc[q];
  std::vector<int> esrc((size_t)3 * ns, -2);   // -2: not a destination
  for (size_t q = 0; q < C->edge.hdst.size(); ++q) esrc[C->edge.hdst[q]] = C->edge.hsrc[q];
  auto inside = [&](int i, int j) { return i >= 2 && i <= iim - 1 && j >= 2 && j <= jjm - 1; };
  auto fwd_ok = [&](int i, int j, int k) {
    const int nbi[3] = {i + 1, i + 1, i}, nbj[3] = {j, j + 1, j + 1};
    for (int e = 0; e < 3; ++e) {
      if (inside(nbi[e], nbj[e])) continue;
      const int n = SLh(nbi[e], nbj[e], k);
      const int src = esrc[(size_t)e * ns + n];
      if (src == -2) return false;
      if (src >= 0 && src % ns != fsrc[n]) return false;
    }
    return true;
  };
  for (int k = 1; k <= 10; ++k)
    for (int j = 2; j <= jjm - 1; ++j)
      for (int i = 2; i <= iim - 1; ++i)
        if ((i == iim - 1 || j == jjm - 1) && !fwd_ok(i, j, k)) return false;
  const int poles[2][3] = {{1, jjm - 1, 1}, {iim - 1, 1, 10}};
  const int own[2][2] = {{1, 2}, {0, 1}};
  for (int p = 0; p < 2; ++p) {
    const int i = poles[p][0], j = poles[p][1], k = poles[p][2];
    const int sp = SLh(i, j, k);
    if (!fwd_ok(i, j, k)) return false;
    for (int q = 0; q < 2; ++q) {
      const int src = esrc[(size_t)own[p][q] * ns + sp];
      if (src < 0) return false;
      const int s2 = src % ns, c2 = src / ns;
      const int i2 = s2 % iim + 1, j2 = (s2 / iim) % jjm + 1, k2 = s2 / (iim * jjm) + 1;
      if (!inside(i2, j2)) return false;
      const int up = c2 == 0 ? SLh(i2 - 1, j2, k2) : c2 == 1 ? SLh(i2 - 1, j2 - 1, k2) : SLh(i2, j2 - 1, k2);
      if (up != sp && fsrc[up] != sp) return false;
      psrc[2 * p + q] = s2;
    }
  }
  return true;
}

swm_status create_device(swm_ctx* C, const swm::Grid& Gh) {
  HIP_TRY(hipSetDevice(C->device));
  HIP_TRY(hipStreamCreateWithFlags(&C->stream, hipStreamNonBlocking));
  HIP_TRY(hipStreamCreateWithFlags(&C->stream2, hipStreamNonBlocking));
  HIP_TRY(hipEventCreateWithFlags(&C->ev_fork, hipEventDisableTiming));
  HIP_TRY(hipEventCreateWithFlags(&C->ev_join, hipEventDisableTiming));
  HIP_TRY(hipEventCreateWithFlags(&C->ev_norm, hipEventDisableTiming));
  HIP_TRY(hipEventCreate(&C->ev_a));
  HIP_TRY(hipEventCreate(&C->ev_b));
  // sizing pass, then the shared arena and the member blocks
  C->arena = Arena{};
  C->marena = Arena{};
  layout_all(C, Gh);
  C->arena_bytes = C->arena.used + 4096;
  C->mb = ((C->marena.used + 4095) / 4096) * 4096;
  const size_t member_bytes = C->mb;
  C->arena = Arena{};
  C->marena = Arena{};
  HIP_TRY(hipMalloc(&C->arena.base, C->arena_bytes));
  C->arena.cap = C->arena_bytes;
  HIP_TRY(hipMalloc(&C->marena.base, member_bytes * (size_t)C->nmem));
  C->marena.cap = member_bytes;
  C->face = C->vrtx = C->edge = DevTable{};
  layout_all(C, Gh);
  if (C->layout_status != SWM_OK) return C->layout_status;
  if (C->arena.used > C->arena.cap || C->marena.used > C->marena.cap) {
    swm::set_error("swm_create: arena overflow");
    return SWM_ERR_HIP;
  }
  C->d.geo.mb = C->mb;
  C->m.geo.mb = C->mb;
  {
    int psrc[4] = {0, 0, 0, 0};
    C->anti_ok = anti_fusion_check(C, psrc);
    for (int q = 0; q < 4; ++q) { C->d.geo.psrc[q] = psrc[q]; C->m.geo.psrc[q] = psrc[q]; }
  }
  for (LevelT<double>& L : C->d.lev) L.L.mb = C->mb;
  for (LevelT<float>& L : C->m.lev) L.L.mb = C->mb;
  C->nz = C->nmem;
  C->member_init.assign(C->nmem, 0);
  HIP_TRY(hipHostMalloc(&C->host_sums, 64 * (size_t)C->nmem * sizeof(double), hipHostMallocDefault));
  HIP_TRY(hipHostGetDevicePointer((void**)&C->host_sums_dev, C->host_sums, 0));
  const int ns = C->ns;
  auto id = [](int n) { std::vector<int> v(n); for (int k = 0; k < n; ++k) v[k] = k; return v; };
  swm_status st;
  if ((st = up64(C->pxyz, to_planes(Gh.p, 3, ns, id(3), 3)))) return st;
  if ((st = up64(C->area, Gh.area))) return st;
  if ((st = up64(C->area_inv, Gh.area_inv))) return st;
  if ((st = up64(C->cai, to_planes(Gh.corner_area_inv, 2, ns, id(2), 2)))) return st;
  if ((st = up64(C->sub_area, to_planes(Gh.sub_area, 6, ns, id(6), 6)))) return st;
  if ((st = up64(C->dc, to_planes(Gh.dc, 6, ns, id(6), 6)))) return st;
  if ((st = up64(C->sw, to_planes(Gh.s_weights, 6, ns, id(6), 6)))) return st;
  if ((st = up64(C->vw, to_planes(Gh.v_weights, 12, ns, id(12), 12)))) return st;
  if ((st = up64(C->lw, to_planes(Gh.l_weights, 7, ns, id(7), 7)))) return st;
  {
    // normal/tangent (kk, a, ic) with a = 0..2: keep a = 0, 1
    std::vector<int> perm(18, -1);
    for (int ic = 0; ic < 2; ++ic)
      for (int a = 0; a < 2; ++a)
        for (int kk = 0; kk < 3; ++kk) perm[kk + 3 * (a + 3 * ic)] = kk + 3 * (a + 2 * ic);
    if ((st = up64(C->nrm, to_planes(Gh.normal, 18, ns, perm, 12)))) return st;
    if ((st = up64(C->tng, to_planes(Gh.tangent, 18, ns, perm, 12)))) return st;
  }
  if ((st = up64(C->f, Gh.f))) return st;
  if ((st = up64(C->hs, std::vector<double>(ns, 0.0)))) return st;
  if ((st = up64(C->mask, Gh.center_mask))) return st;
  {
    std::vector<double> am(ns);
    double tot = 0.0;
    for (int n = 0; n < ns; ++n) { am[n] = Gh.area[n] * Gh.center_mask[n]; tot = tot + am[n] * 1.0; }
    C->total_area = tot;
    if ((st = up64(C->area_mask, am))) return st;
  }
  // multigrid levels
  for (int l = 1; l <= C->g; ++l) {
    LevelT<double>& L = C->d.lev[l - 1];
    const size_t nl = L.ns;
    const std::vector<double>& lw = Gh.levels[l - 1].lap_w;
    if ((st = up64(L.lw, to_planes(lw, 7, (int)nl, id(7), 7)))) return st;
    std::vector<double> lwt(nl);
    for (size_t n = 0; n < nl; ++n) lwt[n] = lw[6 + 7 * n] != 0.0 ? 1.0 / (lw[6 + 7 * n] - 0.0) : 0.0;
    if ((st = up64(L.lwt, lwt))) return st;
    swm::HaloTable t = swm::build_simple_face_table(l);
    std::vector<int> src(nl);
    for (size_t n = 0; n < nl; ++n) src[n] = (int)n;
    for (size_t k = 0; k < t.dst.size(); ++k) src[t.dst[k]] = t.src[k];
    HIP_TRY(hipMemcpy(L.src, src.data(), nl * sizeof(int), hipMemcpyHostToDevice));
  }
  const unsigned nb = nblk(ns);
  hipStream_t S = C->stream;
  if (C->precision == SWM_MIXED) {
    // FP32 metrics, level weights and reciprocal centre weights
    const swm::dev::GeoT<float>& g = C->m.geo;
    auto cv = [&](const double* x, const float* y, size_t n) {
      hipLaunchKernelGGL(swm::dev::k_to_f32, dim3(nblk(n)), dim3(BLK), 0, S, n, x, (const double*)nullptr, 0.0,
                         const_cast<float*>(y));
    };
    cv(C->area, g.area, ns); cv(C->area_inv, g.area_inv, ns); cv(C->cai, g.cai, 2 * (size_t)ns);
    cv(C->sub_area, g.sub_area, 6 * (size_t)ns); cv(C->dc, g.dc, 6 * (size_t)ns); cv(C->nrm, g.nrm, 12 * (size_t)ns);
    cv(C->tng, g.tng, 12 * (size_t)ns); cv(C->vw, g.vw, 12 * (size_t)ns); cv(C->sw, g.sw, 6 * (size_t)ns);
    cv(C->lw, g.lw, 7 * (size_t)ns); cv(C->f, g.f, ns); cv(C->mask, g.mask, ns);
    hipLaunchKernelGGL(swm::dev::k_premul_dc, dim3(nblk(ns)), dim3(BLK), 0, S, ns, C->nrm, C->dc, const_cast<float*>(g.ndc));
    hipLaunchKernelGGL(swm::dev::k_premul_dc, dim3(nblk(ns)), dim3(BLK), 0, S, ns, C->tng, C->dc, const_cast<float*>(g.tdc));
    for (int l = 1; l <= C->g; ++l) {
      const LevelT<double>& L = C->d.lev[l - 1];
      cv(L.lw, C->m.lev[l - 1].lw, 7 * (size_t)L.ns);
      cv(L.lwt, C->m.lev[l - 1].lwt, (size_t)L.ns);
    }
    {
      // symmetric edge coefficients of the FP64 defects: area * l_weights(e+1)
      // at interior slots, halos by the edge table (components re-indexed
      // across panel seams), and area_inv at real cells
      const std::vector<double>& lw = Gh.l_weights;   // the model grid's, as G.lw
      std::vector<double> sym(3 * (size_t)ns, 0.0), aim(ns, 0.0);
      for (int k = 0; k < 10; ++k)
        for (int j = 2; j <= C->jjm - 1; ++j)
          for (int i = 2; i <= C->iim - 1; ++i) {
            const size_t sl = (size_t)(i - 1) + (size_t)C->iim * ((j - 1) + (size_t)C->jjm * k);
            for (int e = 0; e < 3; ++e) sym[(size_t)e * ns + sl] = Gh.area[sl] * lw[7 * sl + e];
          }
      const swm::HaloTable& t = Gh.edge;
      for (size_t q = 0; q < t.dst.size(); ++q) {
        const size_t d = (size_t)(t.dst[q] % t.ncomp) * ns + t.dst[q] / t.ncomp;
        sym[d] = t.src[q] < 0 ? 0.0 : sym[(size_t)(t.src[q] % t.ncomp) * ns + t.src[q] / t.ncomp];
      }
      for (int n = 0; n < ns; ++n) aim[n] = Gh.center_mask[n] == 1.0 ? Gh.area_inv[n] : 0.0;
      HIP_TRY(hipMemcpy(const_cast<double*>(C->d.geo.sym), sym.data(), sym.size() * sizeof(double), hipMemcpyHostToDevice));
      HIP_TRY(hipMemcpy(const_cast<double*>(C->d.geo.aim), aim.data(), aim.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    for (int l = swm::dev::CO_MAXLEV + 1; l <= C->g; ++l) {
      // half-precision weights of the levels above the coarse three:
      // l_weights(1..6) scaled by 1 / their largest magnitude, and
      // 0.8 / (S l_weights(7))
      LevelT<float>& L = C->m.lev[l - 1];
      const std::vector<double>& lw = Gh.levels[l - 1].lap_w;
      const size_t nl = L.ns;
      double mx = 0.0;
      for (size_t n = 0; n < nl; ++n)
        for (int q = 0; q < 6; ++q) mx = std::max(mx, std::fabs(lw[7 * n + q]));
      const double S = mx > 0.0 ? 1.0 / mx : 1.0;
      std::vector<_Float16> hw(6 * nl), hc(nl);
      for (size_t n = 0; n < nl; ++n) {
        for (int q = 0; q < 6; ++q) hw[(size_t)q * nl + n] = (_Float16)(S * lw[7 * n + q]);
        const double w7 = lw[7 * n + 6];
        hc[n] = (_Float16)(w7 != 0.0 ? 0.8 / (S * w7) : 0.0);
      }
      HIP_TRY(hipMemcpy(const_cast<_Float16*>(L.L.hw), hw.data(), hw.size() * sizeof(_Float16), hipMemcpyHostToDevice));
      HIP_TRY(hipMemcpy(const_cast<_Float16*>(L.L.hcw), hc.data(), hc.size() * sizeof(_Float16), hipMemcpyHostToDevice));
      L.L.hs = (float)S;
      L.L.his = (float)(1.0 / S);
    }
    if (C->cd_n) {
      std::vector<float> M;
      std::vector<int> cell;
      coarse_direct_matrix(Gh.levels[swm::dev::CO_MAXLEV - 1].lap_w, swm::dev::CO_MAXLEV, M, cell);
      if ((int)cell.size() != C->cd_n || C->cd_n > swm::dev::CD_MAXN) {
        swm::set_error("coarse solve: level size");
        return SWM_ERR_DIMS;
      }
      HIP_TRY(hipMemcpy(C->cd_m, M.data(), M.size() * sizeof(float), hipMemcpyHostToDevice));
      HIP_TRY(hipMemcpy(C->cd_cell, cell.data(), cell.size() * sizeof(int), hipMemcpyHostToDevice));
    }
    if (C->cd4_n) {
      if (C->cd4_n > swm::dev::CD_MAXN4) { swm::set_error("coarse solve: level size"); return SWM_ERR_DIMS; }
      swm_status cst = coarse_direct_matrix_device(C, Gh.levels[swm::dev::CO_MAXLEV].lap_w, swm::dev::CO_MAXLEV + 1,
                                                   C->cd4_m, C->cd4_cell, C->cd4_n);
      if (cst) return cst;
    }
    (void)nb;
  }
  // zero every member block (the reference initialises its module arrays to 0)
  fill_plain<double>(C, reinterpret_cast<double*>(C->marena.base), member_bytes / 8 * (size_t)C->nmem, 0.0);
  HIP_TRY(hipStreamSynchronize(S));
  if (C->advection == SWM_ADV_REMAP && (st = remap_setup(C, Gh))) return st;
  HIP_TRY(hipGetLastError());
  return SWM_OK;
}

}  // namespace swm

namespace swm {
void tlad_free(swm_ctx* C);
}

extern "C" {

swm_status swm_create(const swm_grid* grid, const swm_config* cfg, swm_ctx** out) {
  return swm_create_members(grid, cfg, 1, out);
}

namespace {
std::mutex g_selftest_mutex;
int g_selftest_done[64] = {};   // 1: passed on that device
}  // namespace

swm_status swm_self_test(int device) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
    swm::set_error("swm_self_test: no such HIP device");
    return SWM_ERR_NO_DEVICE;
  }
  std::lock_guard<std::mutex> lock(g_selftest_mutex);
  if (device < 64 && g_selftest_done[device]) return SWM_OK;
  int prev = 0;
  HIP_TRY(hipGetDevice(&prev));
  HIP_TRY(hipSetDevice(device));
  int* d = nullptr;
  HIP_TRY(hipMalloc(&d, 2 * sizeof(int)));
  bool ok = true;
  for (int b : {64, 128, 256, 512, 1024}) {
    int h[2] = {0, 0};
    hipLaunchKernelGGL(swm::dev::k_block_size_probe, dim3(2), dim3(b), 0, 0, d);
    if (hipMemcpy(h, d, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess || h[0] != 1 || h[1] != 1) ok = false;
  }
  (void)hipFree(d);
  (void)hipSetDevice(prev);
  if (!ok) {
    swm::set_error("swm_self_test: the kernels' workgroup size, read from the implicit kernel arguments, differs "
                   "from blockDim (the code object is not version 5)");
    return SWM_ERR_HIP;
  }
  if (device < 64) g_selftest_done[device] = 1;
  return SWM_OK;
}

swm_status swm_create_members(const swm_grid* grid, const swm_config* cfg, int nmem, swm_ctx** out) {
  if (!grid || !cfg || !out || nmem < 1 || nmem > 4096 || cfg->ntracer < 0 || cfg->ntracer > 64 || !(cfg->dt > 0.0) ||
      (cfg->precision != SWM_FP64 && cfg->precision != SWM_MIXED) || cfg->advection < SWM_ADV_FCT ||
      cfg->advection > SWM_ADV_REMAP) {
    swm::set_error("swm_create: bad argument");
    return SWM_ERR_ARG;
  }
  if (cfg->advection == SWM_ADV_REMAP && (cfg->precision != SWM_FP64 || nmem != 1)) {
    swm::set_error("swm_create: incremental remapping runs in single-member FP64 contexts");
    return SWM_ERR_ARG;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= cfg->device) {
    swm::set_error("swm_create: no HIP device");
    return SWM_ERR_NO_DEVICE;
  }
  swm_status st0 = swm_self_test(cfg->device);
  if (st0) return st0;
  const swm::Grid& Gh = swm::grid_of(grid);
  swm_ctx* C = new swm_ctx;
  C->g = Gh.g; C->iim = Gh.iim; C->jjm = Gh.jjm; C->ns = Gh.iim * Gh.jjm * 10; C->nt = cfg->ntracer;
  C->max_ig = Gh.max_ig;
  C->dt = cfg->dt; C->grav = cfg->grav; C->precision = cfg->precision; C->device = cfg->device;
  C->advection = cfg->advection;
  C->npart = 1024;
  C->radius = Gh.radius;
  C->cell_index = Gh.center_index;
  C->cell_mask = Gh.center_mask;
  C->nmem = nmem;
  swm_status st = swm::create_device(C, Gh);
  if (st) {
    swm_destroy(C);
    return st;
  }
  *out = C;
  return SWM_OK;
}

void swm_destroy(swm_ctx* C) {
  if (!C) return;
  // teardown: errors here have no caller to report to
  (void)hipSetDevice(C->device);
  if (C->stream) (void)hipStreamSynchronize(C->stream);
  for (hipEvent_t e : C->kev) (void)hipEventDestroy(e);
  if (C->stream2) (void)hipStreamSynchronize(C->stream2);
  if (C->ev_fork) (void)hipEventDestroy(C->ev_fork);
  if (C->ev_join) (void)hipEventDestroy(C->ev_join);
  if (C->ev_norm) (void)hipEventDestroy(C->ev_norm);
  if (C->stream2) (void)hipStreamDestroy(C->stream2);
  if (C->ev_a) (void)hipEventDestroy(C->ev_a);
  if (C->ev_b) (void)hipEventDestroy(C->ev_b);
  if (C->arena.base) (void)hipFree(C->arena.base);
  if (C->marena.base) (void)hipFree(C->marena.base);
  if (C->host_sums) (void)hipHostFree(C->host_sums);
  if (C->rm.mem) (void)hipFree(C->rm.mem);
  if (C->diss) (void)hipFree(C->diss);
  if (C->rm.flag_host) (void)hipHostFree(C->rm.flag_host);
  swm::tlad_free(C);
  if (C->stream) (void)hipStreamDestroy(C->stream);
  delete C;
}

namespace swm {
// state from host arrays in the reference layout; hs == nullptr with keep_hs
// leaves the surface height as it is, without keep_hs sets it to zero
swm_status init_from_host(swm_ctx* C, const double* h, const double* hs, const double* u, const double* tracer,
                          bool keep_hs) {
  HIP_TRY(hipSetDevice(C->device));
  MemberScope scope(C);
  const int ns = C->ns;
  hipStream_t S = C->stream;
  std::vector<double> up(4 * (size_t)ns), tq((size_t)C->nt * ns);
  u_to_planes(u, up.data(), ns);
  for (int q = 0; q < C->nt; ++q)
    for (int k = 0; k < 10; ++k)
      std::memcpy(&tq[(size_t)q * ns + (size_t)k * C->iim * C->jjm],
                  tracer + ((size_t)q + (size_t)C->nt * k) * C->iim * C->jjm, (size_t)C->iim * C->jjm * sizeof(double));
  reset_state(C, keep_hs);
  HIP_TRY(hipMemcpyAsync(C->d.h[0], h, ns * sizeof(double), hipMemcpyHostToDevice, S));
  HIP_TRY(hipMemcpyAsync(C->d.u[0], up.data(), 4 * (size_t)ns * sizeof(double), hipMemcpyHostToDevice, S));
  if (hs) HIP_TRY(hipMemcpyAsync(C->hs, hs, ns * sizeof(double), hipMemcpyHostToDevice, S));
  for (int q = 0; q < C->nt; ++q)
    HIP_TRY(hipMemcpyAsync(C->d.tr[0][q], &tq[(size_t)q * ns], ns * sizeof(double), hipMemcpyHostToDevice, S));
  HIP_TRY(hipStreamSynchronize(S));
  return finish_init(C);
}
}  // namespace swm

swm_status swm_init_state(swm_ctx* C, const double* h, const double* hs, const double* u, const double* tracer) {
  if (!C || !h || !u || (C->nt > 0 && !tracer)) { swm::set_error("swm_init_state: bad argument"); return SWM_ERR_ARG; }
  return swm::init_from_host(C, h, hs, u, tracer, false);
}

// write_swm_restart (swm/source/swm_restart.F:113-175): the current time
// level's h, tracers and corner velocity of every real cell and the clock
swm_status swm_write_restart(swm_ctx* C, const char* path) {
  if (!C) { swm::set_error("swm_write_restart: bad argument"); return SWM_ERR_ARG; }
  if (!C->member_init[C->sel]) { swm::set_error("swm_write_restart: state not initialised"); return SWM_ERR_STATE; }
  char name[64];
  if (!path) {
    swm_status st = swm::restart_name(C->nsecond, name, (int)sizeof name);
    if (st) return st;
    path = name;
  }
  const size_t ns = C->ns;
  std::vector<double> h(ns), u(4 * ns), tr(std::max<size_t>(1, (size_t)C->nt * ns));
  swm_status st;
  if ((st = swm_get_field(C, "h", h.data())) || (st = swm_get_field(C, "u", u.data())) ||
      (C->nt > 0 && (st = swm_get_field(C, "tracer", tr.data()))))
    return st;
  const swm::CellIndex ci{C->iim, C->jjm, C->max_ig, C->cell_index.data(), C->cell_mask.data()};
  return swm::restart_write(ci, path, C->nt, C->nsecond, h.data(), u.data(), tr.data());
}

// read_swm_restart (swm/source/swm_restart.F:25-81): h, tracers and corner
// velocity of every slot from its global cell's record, then the state
// derived as at initialisation; the surface height stays as it is, the clock
// takes the file's time stamp and the step counter restarts at zero
swm_status swm_read_restart(swm_ctx* C, const char* path) {
  if (!C) { swm::set_error("swm_read_restart: bad argument"); return SWM_ERR_ARG; }
  if (!path) path = "restarts/swm_restart";
  const size_t ns = C->ns;
  std::vector<double> h(ns, 0.0), u(4 * ns, 0.0), tr(std::max<size_t>(1, (size_t)C->nt * ns), 0.0);
  const swm::CellIndex ci{C->iim, C->jjm, C->max_ig, C->cell_index.data(), C->cell_mask.data()};
  long long ts = 0;
  swm_status st = swm::restart_read(ci, path, C->nt, &ts, h.data(), u.data(), tr.data());
  if (st) return st;
  if ((st = swm::init_from_host(C, h.data(), nullptr, u.data(), tr.data(), true))) return st;
  C->nsecond = ts;
  return SWM_OK;
}

long long swm_time_seconds(const swm_ctx* C) { return C ? C->nsecond : -1; }

// Seeded synthetic state generated on the device (same construction as the
// host version in swm_mi355x/ics.py): harmonic sums of degree 1..lmax for
// vorticity, divergence (10 % RMS) and thickness perturbation, corner velocity
// from the model's own gradient scaled to max |u| = umax, Gaussian tracer bumps.
namespace swm {
swm_status init_synthetic(swm_ctx* C, int seed_index, int lmax, double umax, double h0, double h_rms, int nbumps,
                          double bump_radius, int pert_index, double pert_rel) {
  HIP_TRY(hipSetDevice(C->device));
  MemberScope scope(C);
  const int ns = C->ns;
  hipStream_t S = C->stream;
  const unsigned nb = nblk(ns);
  const uint64_t s0 = 20261203ull + 100ull * (uint64_t)seed_index;
  // coefficients in the order of the host version: stream s0+1 (vorticity),
  // s0+2 (divergence), s0+3 (thickness); l = 1..lmax, m = -l..l
  const int nlm = (lmax + 1) * (lmax + 1) - 1;
  std::vector<double> coef(3 * (size_t)nlm);
  for (int set = 0; set < 3; ++set) {
    SplitMix r(s0 + 1 + set);
    for (int l = 1; l <= lmax; ++l)
      for (int m = -l; m <= l; ++m) coef[3 * (size_t)(l * l - 1 + l + m) + set] = r.normal();
  }
  reset_state(C);
  // scratch: psi -> xout, zeta -> f1, chi -> xout+ns, delta -> f2, hp -> ml
  double* dcoef = C->d.flx1;   // 3*ns >= 3*nlm for every supported level
  if ((size_t)3 * nlm > (size_t)3 * ns) { swm::set_error("swm_init_synthetic: lmax too large for the grid"); return SWM_ERR_ARG; }
  HIP_TRY(hipMemcpyAsync(dcoef, coef.data(), coef.size() * sizeof(double), hipMemcpyHostToDevice, S));
  double *psi = C->xout, *chi = C->xout + ns, *zeta = C->d.f1, *delta = C->d.f2, *hp = C->d.ml;
  hipLaunchKernelGGL(swm::dev::k_harmonics, dim3(nb), dim3(BLK), 0, S, ns, C->pxyz, lmax, dcoef, psi, zeta, chi,
                     delta, hp);
  const double a2 = C->radius * C->radius;
  hipLaunchKernelGGL(swm::dev::k_scale, dim3(nb), dim3(BLK), 0, S, psi, (size_t)ns, a2);
  hipLaunchKernelGGL(swm::dev::k_scale, dim3(nb), dim3(BLK), 0, S, chi, (size_t)ns, a2);
  auto sumsq = [&](const double* x, double* out) -> swm_status {
    hipLaunchKernelGGL(swm::dev::k_sumsq_partial, dim3(C->npart), dim3(BLK), 0, S, x, C->mask, ns, C->part);
    sum_final(C, C->part, C->npart, C->sums, 1);
    HIP_TRY(hipMemcpyAsync(C->host_sums, C->sums, sizeof(double), hipMemcpyDeviceToHost, S));
    HIP_TRY(hipStreamSynchronize(S));
    *out = C->host_sums[0];
    return SWM_OK;
  };
  double sz, sd, sh;
  swm_status st;
  if ((st = sumsq(zeta, &sz)) || (st = sumsq(delta, &sd)) || (st = sumsq(hp, &sh))) return st;
  const double nreal = (double)C->max_ig;
  if (pert_index >= 0 && pert_rel != 0.0) {
    // a member perturbation: vorticity harmonics of the same degrees from
    // stream s0 + 1000 + pert_index, scaled to pert_rel times the RMS of the
    // vorticity, added through the streamfunction
    std::vector<double> pc(3 * (size_t)nlm, 0.0);
    SplitMix rp(s0 + 1000 + (uint64_t)pert_index);
    for (int l = 1; l <= lmax; ++l)
      for (int m = -l; m <= l; ++m) pc[3 * (size_t)(l * l - 1 + l + m)] = rp.normal();
    HIP_TRY(hipMemcpyAsync(dcoef, pc.data(), pc.size() * sizeof(double), hipMemcpyHostToDevice, S));
    double *ppsi = C->d.vf, *pzeta = C->d.vf + ns, *pchi = C->d.vf + 2 * (size_t)ns, *pdel = C->d.dcurl,
           *php = C->d.dcurl + ns;
    hipLaunchKernelGGL(swm::dev::k_harmonics, dim3(nb), dim3(BLK), 0, S, ns, C->pxyz, lmax, dcoef, ppsi, pzeta, pchi,
                       pdel, php);
    double sp;
    if ((st = sumsq(pzeta, &sp))) return st;
    const double c = pert_rel * std::sqrt(sz / sp);
    hipLaunchKernelGGL(swm::dev::k_axpy, dim3(nb), dim3(BLK), 0, S, (size_t)ns, c * a2, ppsi, psi);
  }
  const double rz = std::sqrt(sz / nreal), rd = std::sqrt(sd / nreal), rh = std::sqrt(sh / nreal);
  hipLaunchKernelGGL(swm::dev::k_synth_finish, dim3(nb), dim3(BLK), 0, S, C->d.geo, h0, h_rms / rh, hp, C->d.h[0],
                     0.1 * rz / rd, chi);
  hipLaunchKernelGGL(swm::dev::k_velocity<double>, dim3(nb), dim3(BLK), 0, S, C->d.geo, psi, chi, C->d.u[0], (const double*)nullptr, (const double*)nullptr, 0.0, (double*)nullptr, (const double*)nullptr);
  hipLaunchKernelGGL(swm::dev::k_speedmax_partial, dim3(C->npart), dim3(BLK), 0, S, C->d.geo, C->d.u[0], C->part);
  std::vector<double> pm(C->npart);
  HIP_TRY(hipMemcpyAsync(pm.data(), C->part, C->npart * sizeof(double), hipMemcpyDeviceToHost, S));
  HIP_TRY(hipStreamSynchronize(S));
  double vmax = 0.0;
  for (double v : pm) vmax = std::max(vmax, v);
  hipLaunchKernelGGL(swm::dev::k_scale, dim3(nblk(4 * (size_t)ns)), dim3(BLK), 0, S, C->d.u[0], 4 * (size_t)ns, umax / vmax);
  // tracer bumps: stream s0+4, one set of centres per tracer
  SplitMix rt(s0 + 4);
  for (int q = 0; q < C->nt; ++q) {
    std::vector<double> cen(3 * (size_t)std::max(1, nbumps));
    for (int b = 0; b < nbumps; ++b) {
      const double zc = 2.0 * rt.uniform() - 1.0, lc = 2.0 * M_PI * rt.uniform();
      const double rc = std::sqrt(std::max(0.0, 1.0 - zc * zc));
      cen[3 * b] = rc * std::cos(lc); cen[3 * b + 1] = rc * std::sin(lc); cen[3 * b + 2] = zc;
    }
    HIP_TRY(hipMemcpyAsync(C->sums, cen.data(), cen.size() * sizeof(double), hipMemcpyHostToDevice, S));
    hipLaunchKernelGGL(swm::dev::k_bumps, dim3(nb), dim3(BLK), 0, S, ns, C->pxyz, nbumps, C->sums, C->radius,
                       bump_radius, C->d.tr[0][q]);
    HIP_TRY(hipStreamSynchronize(S));
  }
  fill<double>(C, C->xout, 2 * (size_t)ns, 0.0);   // the solver's warm start begins at zero
  HIP_TRY(hipStreamSynchronize(S));
  return finish_init(C);
}
}  // namespace swm

swm_status swm_init_synthetic(swm_ctx* C, int seed_index, int lmax, double umax, double h0, double h_rms,
                              int nbumps, double bump_radius) {
  if (!C || lmax < 1 || lmax > 400 || nbumps < 0 || nbumps > 64) { swm::set_error("swm_init_synthetic: bad argument"); return SWM_ERR_ARG; }
  return swm::init_synthetic(C, seed_index, lmax, umax, h0, h_rms, nbumps, bump_radius, -1, 0.0);
}

swm_status swm_init_synthetic_member(swm_ctx* C, int seed_index, int lmax, double umax, double h0, double h_rms,
                                     int nbumps, double bump_radius, int member_seed, double vort_rel) {
  if (!C || lmax < 1 || lmax > 400 || nbumps < 0 || nbumps > 64 || member_seed < 0 || !(vort_rel >= 0.0) ||
      vort_rel > 1.0) {
    swm::set_error("swm_init_synthetic_member: bad argument");
    return SWM_ERR_ARG;
  }
  return swm::init_synthetic(C, seed_index, lmax, umax, h0, h_rms, nbumps, bump_radius, member_seed, vort_rel);
}

namespace swm {
// one time step of every member, queued on the context's stream (no host
// synchronisation unless the solver's convergence test needs one)
swm_status step_once(swm_ctx* C) {
  C->step_count += 1;
  C->nsecond = (long long)((double)C->nsecond + C->dt);   // time_manager_module.F:96, truncated
  pointers(C);
  weights(C);
  swm_status st;
  if (C->precision == SWM_MIXED) {
    st = step_dt<float>(C, C->m);
    if (st) return st;
    wrap_prognostics<float>(C, C->m, true);
  } else {
    st = step_dt<double>(C, C->d);
    if (st) return st;
    wrap_prognostics<double>(C, C->d, true);
  }
  return SWM_OK;
}
}  // namespace swm

swm_status swm_step(swm_ctx* C, int nsteps) {
  if (!C || nsteps < 0) { swm::set_error("swm_step: bad argument"); return SWM_ERR_ARG; }
  if (!C->initialized) { swm::set_error("swm_step: state not initialised (every member)"); return SWM_ERR_STATE; }
  HIP_TRY(hipSetDevice(C->device));
  C->kev_used = 0;
  C->kernel_bytes = 0.0;
  C->kernel_pair_launches = 0;
  HIP_TRY(hipEventRecord(C->ev_a, C->stream));
  const long long bad0 = C->shadow_bad;
  auto remap_stop = [&]() {
    swm::set_error("remapping: a thickness below -1e-10 at step " + std::to_string(C->step_count) +
                   " (remap_prep stops the reference there)");
    *C->rm.flag_host = 0;
    return SWM_ERR_STATE;
  };
  // the corners queued ahead belong to this call's next step only
  struct Ahead {
    swm_ctx* C;
    ~Ahead() { C->next_follows = false; C->corners_ready = false; C->ke_deferred = false; }
  } ahead{C};
  C->corners_ready = false;
  C->ke_deferred = false;
  for (int n = 0; n < nsteps; ++n) {
    C->next_follows = n + 1 < nsteps;
    swm_status st = swm::step_once(C);
    if (st) return st;
    if (C->rm.flag_host && *(volatile int*)C->rm.flag_host) return remap_stop();
    if (C->shadow_bad != bad0) {
      swm::set_error("halo shadow check: " + std::to_string(C->shadow_bad - bad0) + " values of the device's halo "
                     "fills differ from the host's fill with the same tables (step " + std::to_string(C->step_count) +
                     ")");
      return SWM_ERR_STATE;
    }
  }
  HIP_TRY(hipEventRecord(C->ev_b, C->stream));
  HIP_TRY(hipEventSynchronize(C->ev_b));
  float ms = 0.0f;
  HIP_TRY(hipEventElapsedTime(&ms, C->ev_a, C->ev_b));
  C->last_step_ms = ms;
  HIP_TRY(hipGetLastError());
  if (C->rm.flag_host && *(volatile int*)C->rm.flag_host) return remap_stop();
  return SWM_OK;
}

swm_status swm_set_remap_limiter(swm_ctx* C, int limit_min, int limit_max) {
  if (!C || C->advection != SWM_ADV_REMAP) {
    swm::set_error("swm_set_remap_limiter: needs a context with incremental remapping");
    return SWM_ERR_ARG;
  }
  C->rm.limit_min = limit_min ? 1 : 0;
  C->rm.limit_max = limit_max ? 1 : 0;
  return SWM_OK;
}

swm_status swm_get_field(swm_ctx* C, const char* name, double* out) {
  if (!C || !name || !out) { swm::set_error("swm_get_field: bad argument"); return SWM_ERR_ARG; }
  HIP_TRY(hipSetDevice(C->device));
  swm::MemberScope scope(C);
  const int ns = C->ns;
  const int n4 = C->n4 - 1;
  const std::string s(name);
  const bool mixed = C->precision == SWM_MIXED;
  // mixed path: FP32 storage back to FP64 values in a scratch plane first
  double* tmp = C->d.f1;   // 3 planes of FP64 scratch, unused by the mixed step
  auto plane = [&](const double* d64, const float* d32, const double* off = nullptr, double c = 0.0) -> const double* {
    if (!mixed) return d64;
    to64(C, d32, tmp, ns, off, c);
    return tmp;
  };
  const double* src = nullptr;
  if (s == "h") src = plane(C->d.h[n4], C->m.h[n4], nullptr, C->h0);
  else if (s == "eta") src = plane(C->d.eta[n4], C->m.eta[n4], C->f);
  else if (s == "div") src = plane(C->d.div[n4], C->m.div[n4]);
  else if (s == "relative") src = plane(C->d.relative, C->m.relative);
  else if (s == "ke") src = plane(C->d.ke, C->m.ke);
  else if (s == "psi") src = C->xout;
  else if (s == "chi") src = C->xout + ns;
  else if (s == "hs") src = C->hs;
  if (src) {
    HIP_TRY(hipStreamSynchronize(C->stream));
    HIP_TRY(hipMemcpy(out, src, ns * sizeof(double), hipMemcpyDeviceToHost));
    return SWM_OK;
  }
  if (s == "u") {
    std::vector<double> p(4 * (size_t)ns);
    for (int c = 0; c < 4; ++c) {
      const double* q = plane(C->d.u[n4] + (size_t)c * ns, mixed ? C->m.u[n4] + (size_t)c * ns : nullptr);
      HIP_TRY(hipStreamSynchronize(C->stream));
      HIP_TRY(hipMemcpy(p.data() + (size_t)c * ns, q, ns * sizeof(double), hipMemcpyDeviceToHost));
    }
    u_from_planes(p.data(), out, ns);
    return SWM_OK;
  }
  if (s == "tracer") {
    std::vector<double> p((size_t)ns);
    const size_t pl = (size_t)C->iim * C->jjm;
    for (int q = 0; q < C->nt; ++q) {
      const double* t = plane(C->d.tr[n4][q], mixed ? C->m.tr[n4][q] : nullptr);
      HIP_TRY(hipStreamSynchronize(C->stream));
      HIP_TRY(hipMemcpy(p.data(), t, ns * sizeof(double), hipMemcpyDeviceToHost));
      for (int k = 0; k < 10; ++k) std::memcpy(out + ((size_t)q + (size_t)C->nt * k) * pl, &p[k * pl], pl * sizeof(double));
    }
    return SWM_OK;
  }
  swm::set_error("swm_get_field: unknown field " + s);
  return SWM_ERR_ARG;
}

swm_status swm_set_mixed_smoother(swm_ctx* C, int fine_sweeps, int half_weights) {
  if (!C || fine_sweeps < 1 || fine_sweeps > 4) {
    swm::set_error("swm_set_mixed_smoother: 1 to 4 sweeps");
    return SWM_ERR_ARG;
  }
  C->mixed_fine_sweeps = fine_sweeps;
  C->mixed_half = half_weights ? 1 : 0;
  C->epoch += 1;
  return SWM_OK;
}

swm_status swm_set_overlap(swm_ctx* C, int mode) {
  if (!C || mode < 0) { swm::set_error("swm_set_overlap: bad argument"); return SWM_ERR_ARG; }
  C->overlap = mode;
  C->epoch += 1;
  return SWM_OK;
}

swm_status swm_set_mixed_coarse(swm_ctx* C, int direct) {
  if (!C || C->precision != SWM_MIXED) { swm::set_error("swm_set_mixed_coarse: needs a mixed-precision context"); return SWM_ERR_ARG; }
  if (direct < 0 || direct > 2) { swm::set_error("swm_set_mixed_coarse: direct is 0, 1 or 2"); return SWM_ERR_ARG; }
  C->mixed_coarse_direct = direct;
  C->epoch += 1;
  return SWM_OK;
}

swm_status swm_set_fusion(swm_ctx* C, int flags) {
  if (!C || (flags & ~SWM_FUSE_ALL)) { swm::set_error("swm_set_fusion: bad argument"); return SWM_ERR_ARG; }
  C->fuse_mg = (flags & SWM_FUSE_INJECT_SWEEP) ? 1 : 0;
  C->fuse_push = (flags & SWM_FUSE_HALO_PUSH) ? 1 : 0;
  C->fuse_anti = (flags & SWM_FUSE_MASS_ANTI) ? 1 : 0;
  C->epoch += 1;
  return SWM_OK;
}

int swm_members(const swm_ctx* C) { return C ? C->nmem : 0; }

swm_status swm_ctx_layout(const swm_ctx* C, int* iim, int* jjm, int* ntracer) {
  if (!C) { swm::set_error("swm_ctx_layout: bad argument"); return SWM_ERR_ARG; }
  if (iim) *iim = C->iim;
  if (jjm) *jjm = C->jjm;
  if (ntracer) *ntracer = C->nt;
  return SWM_OK;
}

swm_status swm_select_member(swm_ctx* C, int member) {
  if (!C || member < 0 || member >= C->nmem) { swm::set_error("swm_select_member: no such member"); return SWM_ERR_ARG; }
  C->sel = member;
  return SWM_OK;
}

swm_status swm_set_fixed_cycles(swm_ctx* C, int ncycles, int nstart) {
  if (!C || ncycles < 0 || ncycles > 12 || nstart < 0 || nstart > 24) {
    swm::set_error("swm_set_fixed_cycles: bad argument");
    return SWM_ERR_ARG;
  }
  if (ncycles > 0 && C->solver_relative) {
    swm::set_error("swm_set_fixed_cycles: the fixed schedule takes the absolute criterion");
    return SWM_ERR_ARG;
  }
  C->fixed_cycles = ncycles;
  C->fixed_start = nstart;
  C->epoch += 1;
  return SWM_OK;
}

swm_status swm_solver_track(swm_ctx* C, double* out) {
  if (!C || !out) { swm::set_error("swm_solver_track: bad argument"); return SWM_ERR_ARG; }
  HIP_TRY(hipSetDevice(C->device));
  HIP_TRY(hipStreamSynchronize(C->stream));
  swm::MemberScope scope(C);
  HIP_TRY(hipMemcpy(out, C->track, 3 * sizeof(double), hipMemcpyDeviceToHost));
  return SWM_OK;
}

swm_status swm_synchronize(swm_ctx* C) {
  if (!C) return SWM_ERR_ARG;
  HIP_TRY(hipStreamSynchronize(C->stream));
  return SWM_OK;
}

int swm_step_count(const swm_ctx* C) { return C ? C->step_count : -1; }
int swm_last_vcycles(const swm_ctx* C) { return C ? C->last_vcycles : -1; }

swm_status swm_set_solver(swm_ctx* C, double tol, int relative, int max_cycles) {
  if (!C || !(tol >= 0.0) || (tol == 0.0 && !relative) || max_cycles < 2 || max_cycles > 1000) {
    swm::set_error("swm_set_solver: need tol > 0 (or 0 with relative) and 2 <= max_cycles <= 1000");
    return SWM_ERR_ARG;
  }
  if (relative && C->fixed_cycles > 0) {
    swm::set_error("swm_set_solver: the fixed schedule (swm_set_fixed_cycles) takes the absolute criterion");
    return SWM_ERR_ARG;
  }
  C->solver_tol = tol;
  C->solver_relative = relative ? 1 : 0;
  C->solver_max_cycles = max_cycles;
  C->solver_stall = tol == 0.0 ? 1 : 0;
  C->epoch += 1;
  return SWM_OK;
}

swm_status swm_set_mixed_start(swm_ctx* C, int extrapolate, int min_cycles) {
  if (!C || min_cycles < 0 || min_cycles > C->solver_max_cycles || extrapolate < 0 || extrapolate > 2) {
    swm::set_error("swm_set_mixed_start: need 0 <= extrapolate <= 2 and 0 <= min_cycles <= max_cycles");
    return SWM_ERR_ARG;
  }
  C->mixed_extrapolate = extrapolate;
  C->mixed_min_cycles = min_cycles;
  C->epoch += 1;
  return SWM_OK;
}

swm_status swm_last_residual(const swm_ctx* C, double* resid_ms, double* rhs_ms) {
  if (!C) return SWM_ERR_ARG;
  for (int v = 0; v < 2; ++v) {
    if (resid_ms) resid_ms[v] = C->last_resid[v];
    if (rhs_ms) rhs_ms[v] = C->last_rhs_ms[v];
  }
  return SWM_OK;
}


swm_status swm_budgets(swm_ctx* C, double* out, int nout) {
  if (!C || !out) return SWM_ERR_ARG;
  const int nt = std::min(C->nt, (int)swm::dev::BUDGET_MAXTR);
  const int nq = 8 + nt;
  if (nout < nq) {
    swm::set_error("swm_budgets: out needs 8 + ntracer entries (at most 8 + 16)");
    return SWM_ERR_ARG;
  }
  if (!C->member_init[C->sel]) { swm::set_error("swm_budgets: state not initialised"); return SWM_ERR_STATE; }
  HIP_TRY(hipSetDevice(C->device));
  swm::MemberScope scope(C);
  if (C->precision == SWM_MIXED) swm::launch_budgets<float>(C, C->m, nq);
  else swm::launch_budgets<double>(C, C->d, nq);
  std::vector<double> sums(nq);
  HIP_TRY(hipMemcpyAsync(sums.data(), C->sums, nq * sizeof(double), hipMemcpyDeviceToHost, C->stream));
  HIP_TRY(hipStreamSynchronize(C->stream));
  HIP_TRY(hipGetLastError());
  const double inv = 1.0 / C->total_area;
  for (int q = 0; q < nq; ++q) out[q] = sums[q] * inv;
  out[4] = out[1] + out[2] + out[3];
  return SWM_OK;
}

swm_status swm_solver_stats(const swm_ctx* C, long long* solves, long long* vcycles, long long* checks) {
  if (!C) return SWM_ERR_ARG;
  if (solves) *solves = C->solves;
  if (vcycles) *vcycles = C->vcycles_total;
  if (checks) *checks = C->checks_total;
  return SWM_OK;
}

swm_status swm_profile(swm_ctx* C, int enable) {
  if (!C) return SWM_ERR_ARG;
  HIP_TRY(hipSetDevice(C->device));
  C->profile = enable > 0 ? enable : 0;
  C->prof_seq = 0;
  if (C->profile && C->kev.empty()) {
    C->kev.resize(4096);
    for (hipEvent_t& e : C->kev) HIP_TRY(hipEventCreate(&e));
  }
  return SWM_OK;
}

swm_status swm_timing(const swm_ctx* C, double* step_ms, double* kernel_ms_mean, double* kernel_bytes,
                      int* kernel_launches) {
  if (!C) return SWM_ERR_ARG;
  if (step_ms) *step_ms = C->last_step_ms;
  double tot = 0.0;
  int n = 0;
  for (size_t k = 0; k + 1 < C->kev_used; k += 2) {
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, C->kev[k], C->kev[k + 1]));
    tot += ms;
    ++n;
  }
  const int nl = n ? C->kernel_pair_launches : 0;
  if (kernel_ms_mean) *kernel_ms_mean = nl ? tot / nl : 0.0;
  if (kernel_bytes) *kernel_bytes = nl ? C->kernel_bytes / nl : 0.0;
  if (kernel_launches) *kernel_launches = nl;
  return SWM_OK;
}

swm_status swm_total_mass(swm_ctx* C, double* mass) {
  if (!C || !mass) return SWM_ERR_ARG;
  HIP_TRY(hipSetDevice(C->device));
  swm::MemberScope scope(C);
  const int n4 = C->n4 - 1;
  double m[2];
  if (C->precision == SWM_MIXED) {
    // sum of (h' + h0) area = sum h' area + h0 * total area
    hipLaunchKernelGGL((swm::dev::k_sum_partial<float, double>), zgrid(C, C->npart, 2), dim3(BLK), 0, C->stream,
                       C->m.h[n4], C->m.h[n4], C->area_mask, C->ns, C->part, C->mb);
    sum_final(C, C->part, C->npart, C->sums, 2);
    swm_status st = read_sums(C, 2, m);
    if (st) return st;
    *mass = m[0] + C->h0 * C->total_area;
    return SWM_OK;
  }
  sum2<double>(C, C->d.h[n4], C->d.h[n4], C->area_mask, C->ns);
  swm_status st = read_sums(C, 2, m);
  if (st) return st;
  *mass = m[0];
  return SWM_OK;
}

// Microbench of single operators on the context's resident metrics and
// state: reps back-to-back launches of the kernel the operator entry point
// runs, timed with HIP events on the context's stream.  The output goes to a
// scratch plane of the step; the state is not changed.
swm_status swm_operator_bench(swm_ctx* C, int op, int reps, double* us_per_launch, double* bytes_per_launch) {
  if (!C || op < 0 || op > 6 || reps < 1 || !us_per_launch || !bytes_per_launch) {
    swm::set_error("swm_operator_bench: bad argument (op 0..6, reps >= 1)");
    return SWM_ERR_ARG;
  }
  if (C->precision != SWM_FP64 || !C->initialized) {
    swm::set_error("swm_operator_bench: needs an initialised FP64 context");
    return SWM_ERR_STATE;
  }
  if ((op == 4 || op == 5) && (C->advection != SWM_ADV_REMAP || C->step_count < 1)) {
    swm::set_error("swm_operator_bench: the remapping kernels need a remapping context after a step");
    return SWM_ERR_STATE;
  }
  if (op == 6 && !C->diss) {
    swm::set_error("swm_operator_bench: the dissipation kernel needs swm_init_diss_op");
    return SWM_ERR_STATE;
  }
  HIP_TRY(hipSetDevice(C->device));
  swm::MemberScope scope(C);
  View<double>& V = C->d;
  const swm::dev::Geo& G = V.geo;
  const int lev = C->n4 - 1;
  const unsigned nb = nblk(C->ns);
  double* out = V.flx1;
  if (op == 3)   // the halo fill works in place on a copy of the thickness
    HIP_TRY(hipMemcpyAsync(out, V.h[lev], C->ns * sizeof(double), hipMemcpyDeviceToDevice, C->stream));
  RemapDev& D = C->rm;
  const swm::dev::RemapArgs RA{C->iim, C->jjm, C->ns, C->nt, D.cellp, D.edgep, D.cornp, C->radius, D.limit_min,
                               D.limit_max, D.flag_dev};
  const unsigned nbs = nblk_b(C->ns, C->blk_step);
  double* const* tr3 = C->nt ? D.ptab + (size_t)(C->n3 - 1) * C->nt : nullptr;
  auto launch = [&]() {
    switch (op) {
      case 4:
        hipLaunchKernelGGL(swm::dev::k_rm_transport, zgrid(C, nbs), dim3(C->blk_step), 0, C->stream, RA, D.dp, D.mx,
                           D.my, D.mp, D.tx, D.ty, D.tp, D.mflux, D.mtflux);
        break;
      case 5:
        hipLaunchKernelGGL(swm::dev::k_rm_construct, zgrid(C, nbs), dim3(C->blk_step), 0, C->stream, RA,
                           V.h[C->n3 - 1], (const double* const*)tr3, D.mx, D.my, D.mp, D.tx, D.ty, D.tp);
        break;
      case 6:
        hipLaunchKernelGGL(swm::dev::k_diss, dim3(nbs), dim3(C->blk_step), 0, C->stream, C->iim, C->jjm, C->ns, C->diss,
                           C->cai, 1.0, V.u[lev], V.flx1);
        break;
      case 0: hipLaunchKernelGGL(swm::dev::k_op_div_edge, dim3(nb), dim3(BLK), 0, C->stream, G, V.mass_flux, out); break;
      case 1: hipLaunchKernelGGL(swm::dev::k_curl_div, dim3(nb), dim3(BLK), 0, C->stream, G, V.u[lev], out,
                                 (double*)nullptr); break;
      case 2: hipLaunchKernelGGL(swm::dev::k_op_lap, dim3(nb), dim3(BLK), 0, C->stream, G, V.h[lev], out); break;
      default: gather<double>(C, C->face, {out}); break;
    }
  };
  launch();   // warm
  HIP_TRY(hipEventRecord(C->ev_a, C->stream));
  for (int r = 0; r < reps; ++r) launch();
  HIP_TRY(hipEventRecord(C->ev_b, C->stream));
  HIP_TRY(hipEventSynchronize(C->ev_b));
  float ms = 0.0f;
  HIP_TRY(hipEventElapsedTime(&ms, C->ev_a, C->ev_b));
  *us_per_launch = 1e3 * (double)ms / reps;
  // algorithmic FP64 bytes per real cell: edge-flux divergence 3 own edges
  // + 3 neighbour edges, area_inv, result (40 B); curl of a corner vector
  // 12 v_weights, 4 velocity components, area_inv, result (144 B);
  // Laplacian 7 weights, centre value, result (72 B); halo fill 16 B per
  // ghost slot (value read and written).  Remapping transports: 43 edge
  // planes of 3 edges (1032 B), the departure points of 2 corners (48 B),
  // the mass fields mp, mx, my (24 B) and per tracer tp, tx, ty (24 B) read,
  // the mass and per tracer the mass-tracer transports of 3 edges written
  // (24 B each): 1128 + 48 nt B.  Remapping fields: thickness and tracers
  // (8 + 8 nt B), 32 cell planes (256 B) read, mx, my, mp and per tracer
  // tx, ty, tp written (24 + 24 nt B): 288 + 32 nt B.  Dissipation: 4
  // corner-vector planes and 36 geometry planes and 2 corner-area planes
  // read, 4 planes written (368 B).
  static const double per_cell[3] = {40.0, 144.0, 72.0};
  const double nt = (double)C->nt;
  if (op < 3) *bytes_per_launch = per_cell[op] * (double)C->max_ig;
  else if (op == 3) *bytes_per_launch = 16.0 * (double)C->face.n;
  else if (op == 4) *bytes_per_launch = (1128.0 + 48.0 * nt) * (double)C->max_ig;
  else if (op == 5) *bytes_per_launch = (288.0 + 32.0 * nt) * (double)C->max_ig;
  else *bytes_per_launch = 368.0 * (double)C->max_ig;
  return SWM_OK;
}

// global_reductions.F:77-98 global_real_sum: sum(X * MASK) over n values,
// here as FP64 block partials (1024 blocks, strided, then a tree) and a
// compensated fixed-order sum of the partials, so the result is the same on
// every call and within about one rounding of the exact sum.
constexpr int64_t kSumMaxN = (int64_t)1 << 30;   // the partial-sum kernel indexes with int
constexpr int kSumParts = 1024;

void launch_global_sum(hipStream_t S, const double* x, const double* mask, int n, double* part, double* out) {
  hipLaunchKernelGGL((swm::dev::k_sum_partial<double, double>), dim3(kSumParts, 1), dim3(BLK), 0, S, x, x, mask, n,
                     part, (size_t)0);
  hipLaunchKernelGGL(swm::dev::k_sum_final, dim3(1), dim3(swm::dev::SUM_THREADS), 0, S, (const double*)part, kSumParts,
                     out, (size_t)0, 1, (const double*)nullptr, 0, (double*)nullptr);
}

swm_status swm_global_real_sum_device(swm_ctx* C, const double* x, const double* mask, int64_t n, double* out) {
  if (!C || (n > 0 && !x) || !out || n < 0 || n > kSumMaxN) {
    swm::set_error("swm_global_real_sum_device: need x (n > 0), out and 0 <= n <= 2^30");
    return SWM_ERR_ARG;
  }
  HIP_TRY(hipSetDevice(C->device));
  // the context's partial-sum buffer (its first member block) holds the partials
  launch_global_sum(C->stream, x, mask, (int)n, C->part, out);
  HIP_TRY(hipGetLastError());
  return SWM_OK;
}

swm_status swm_global_real_sum(swm_ctx* C, const double* x, const double* mask, int64_t n, double* out) {
  if (!C || (n > 0 && !x) || !out || n < 0 || n > kSumMaxN) {
    swm::set_error("swm_global_real_sum: need x (n > 0), out and 0 <= n <= 2^30");
    return SWM_ERR_ARG;
  }
  HIP_TRY(hipSetDevice(C->device));
  double *dx = nullptr, *dm = nullptr, *dout = nullptr;
  const size_t bytes = (size_t)std::max<int64_t>(n, 1) * sizeof(double);
  swm_status st = SWM_OK;
  if (hipMalloc(&dx, bytes) != hipSuccess || (mask && hipMalloc(&dm, bytes) != hipSuccess) ||
      hipMalloc(&dout, sizeof(double)) != hipSuccess) {
    swm::set_error("swm_global_real_sum: device allocation failed");
    st = SWM_ERR_HIP;
  }
  if (!st && n > 0 && (hipMemcpy(dx, x, (size_t)n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
                       (mask && hipMemcpy(dm, mask, (size_t)n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess))) {
    swm::set_error("swm_global_real_sum: copy to the device failed");
    st = SWM_ERR_HIP;
  }
  if (!st) st = swm_global_real_sum_device(C, dx, dm, n, dout);
  if (!st && (hipStreamSynchronize(C->stream) != hipSuccess ||
              hipMemcpy(out, dout, sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)) {
    swm::set_error("swm_global_real_sum: the sum failed on the device");
    st = SWM_ERR_HIP;
  }
  (void)hipFree(dx);
  (void)hipFree(dm);
  (void)hipFree(dout);
  return st;
}

// Allocation-free global sum, for callers that cannot allocate device memory
// (code under graph capture, serving loops).  The block partials go to a
// file-scope device array of kSumParts doubles (8 KB per GPU, held until the
// process ends) instead of a context buffer; every partial is written by its
// block on each call, so nothing needs clearing.  Calls on one GPU share the
// array: calls that overlap in time on one GPU race.
static __device__ __align__(256) double scratch[kSumParts];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kMaxDevices] = {};   // the array's address on each device
}  // namespace

swm_status swm_global_real_sum_noalloc(const double* x, const double* mask, int64_t n, double* out, int device,
                                       void* stream) {
  if ((n > 0 && !x) || !out || n < 0) {
    swm::set_error("swm_global_real_sum_noalloc: need x (n > 0), out and n >= 0");
    return SWM_ERR_ARG;
  }
  if (n > kSumMaxN) {
    swm::set_error("swm_global_real_sum_noalloc: n = " + std::to_string(n) +
                   " exceeds the maximum of the allocation-free sum (2^30 values)");
    return SWM_ERR_ARG;
  }
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev || device >= kMaxDevices) {
    swm::set_error("swm_global_real_sum_noalloc: no such device");
    return SWM_ERR_ARG;
  }
  int prev = 0;
  HIP_TRY(hipGetDevice(&prev));
  HIP_TRY(hipSetDevice(device));
  double* part = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
      void* p = nullptr;
      const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p) {
        (void)hipSetDevice(prev);
        swm::set_error(std::string("swm_global_real_sum_noalloc: hipGetSymbolAddress failed: ") + hipGetErrorString(e));
        return SWM_ERR_HIP;
      }
      g_scratch_ptr[device] = static_cast<double*>(p);
    }
    part = g_scratch_ptr[device];
  }
  launch_global_sum(static_cast<hipStream_t>(stream), x, mask, (int)n, part, out);
  const hipError_t e = hipGetLastError();
  (void)hipSetDevice(prev);
  if (e != hipSuccess) {
    swm::set_error(std::string("swm_global_real_sum_noalloc: launch failed: ") + hipGetErrorString(e));
    return SWM_ERR_HIP;
  }
  return SWM_OK;
}

#define OP_BEGIN(C)                                                              \
  if (!(C)) { swm::set_error("null context"); return SWM_ERR_ARG; }             \
  HIP_TRY(hipSetDevice((C)->device));                                           \
  HIP_TRY(hipStreamSynchronize((C)->stream));                                   \
  swm::MemberScope member_scope_(C);
#define OP_ARGS(...)                                                             \
  if (!all_set(__VA_ARGS__)) { swm::set_error("null array argument"); return SWM_ERR_ARG; }
#define OP_BUF(x)                                                                \
  if (!(x)) { swm::set_error("operator buffers: device allocation or copy failed"); return SWM_ERR_HIP; }
// extents of a single-operator call: one-cell padding, at least one interior
// cell per panel, and the fixed counts of this grid (3 data, 2 corners,
// 2 space components, 3 edges, 7 Laplacian weights, 3 frame components)
#define OP_DIMS(npad, nx, ny, nsdm, ok)                                           \
  if ((npad) != 1 || (nx) < 3 || (ny) < 3 || (nsdm) < 1 || !(ok)) {              \
    swm::set_error("operator extents: need npad = 1, nx, ny >= 3, nsdm >= 1 and the grid's fixed counts"); \
    return SWM_ERR_DIMS;                                                        \
  }                                                                             \
  OpIO io{C, (size_t)(nx) * (ny) * (nsdm), {}, {}};                             \
  swm::dev::Geo G{};                                                            \
  G.iim = (nx); G.jjm = (ny); G.ns = (int)io.ns;                                \
  const unsigned nb_ = nblk(io.ns);

swm_status swm_op_curl_of_a_vector(swm_ctx* C, int npad, int nx, int ny, int nsdm, int ndata, int nspace, int ncorners,
                                   const double* area_inv, const double* v_weights, const double* xin, double* xout) {
  OP_BEGIN(C);
  OP_ARGS(area_inv, v_weights, xin, xout);
  OP_DIMS(npad, nx, ny, nsdm, ndata == 3 && nspace == 2 && ncorners == 2);
  G.area_inv = io.put(area_inv, 1, 1);
  G.vw = io.put(v_weights, 12, 12);
  double *x = io.put(xin, 4, 4, kCornerVec), *y = io.dev(1);
  OP_BUF(G.area_inv && G.vw && x && y);
  hipLaunchKernelGGL(swm::dev::k_curl_div, dim3(nb_), dim3(BLK), 0, C->stream, G, x, y, (double*)nullptr);
  return io.get(xout, y, 1);
}

swm_status swm_op_div_of_a_vector(swm_ctx* C, int npad, int nx, int ny, int nsdm, int ndata, int nspace, int ncorners,
                                  const double* area_inv, const double* v_weights, const double* xin, double* xout) {
  OP_BEGIN(C);
  OP_ARGS(area_inv, v_weights, xin, xout);
  OP_DIMS(npad, nx, ny, nsdm, ndata == 3 && nspace == 2 && ncorners == 2);
  G.area_inv = io.put(area_inv, 1, 1);
  G.vw = io.put(v_weights, 12, 12);
  double *x = io.put(xin, 4, 4, kCornerVec), *y = io.dev(1);
  OP_BUF(G.area_inv && G.vw && x && y);
  hipLaunchKernelGGL(swm::dev::k_curl_div, dim3(nb_), dim3(BLK), 0, C->stream, G, x, (double*)nullptr, y);
  return io.get(xout, y, 1);
}

swm_status swm_op_div_of_an_edge_flux(swm_ctx* C, int npad, int nx, int ny, int nsdm, int nedges,
                                      const double* area_inv, const double* xin, double* xout) {
  OP_BEGIN(C);
  OP_ARGS(area_inv, xin, xout);
  OP_DIMS(npad, nx, ny, nsdm, nedges == 3);
  G.area_inv = io.put(area_inv, 1, 1);
  double *x = io.put(xin, 3, 3), *y = io.dev(1);
  OP_BUF(G.area_inv && x && y);
  hipLaunchKernelGGL(swm::dev::k_op_div_edge, dim3(nb_), dim3(BLK), 0, C->stream, G, x, y);
  return io.get(xout, y, 1);
}

swm_status swm_op_laplacian_of_a_scalar(swm_ctx* C, int npad, int nx, int ny, int nsdm, int nweights,
                                        const double* l_weights, const double* xin, double* xout) {
  OP_BEGIN(C);
  OP_ARGS(l_weights, xin, xout);
  OP_DIMS(npad, nx, ny, nsdm, nweights == 7);
  G.lw = io.put(l_weights, 7, 7);
  double *x = io.put(xin, 1, 1), *y = io.dev(1);
  OP_BUF(G.lw && x && y);
  hipLaunchKernelGGL(swm::dev::k_op_lap, dim3(nb_), dim3(BLK), 0, C->stream, G, x, y);
  return io.get(xout, y, 1);
}

swm_status swm_op_k_cross_v(swm_ctx* C, int nx, int ny, int nsdm, int nspace, int ncorners, const double* xin,
                            double* xout) {
  OP_BEGIN(C);
  OP_ARGS(xin, xout);
  OP_DIMS(1, nx, ny, nsdm, nspace == 2 && ncorners == 2);
  (void)G;
  double *x = io.put(xin, 4, 4, kCornerVec), *y = io.dev(4);
  OP_BUF(x && y);
  hipLaunchKernelGGL(swm::dev::k_op_cross, dim3(nb_), dim3(BLK), 0, C->stream, (int)io.ns, x, y);
  return io.get(xout, y, 4, kCornerVec);
}

swm_status swm_op_gradient(swm_ctx* C, int npad, int nx, int ny, int nsdm, int ndata, int nspace, int ncorners,
                           const double* corner_area_inv, const double* v_weights, const double* xin, double* xout) {
  OP_BEGIN(C);
  OP_ARGS(corner_area_inv, v_weights, xin, xout);
  OP_DIMS(npad, nx, ny, nsdm, ndata == 3 && nspace == 2 && ncorners == 2);
  G.cai = io.put(corner_area_inv, 2, 2);
  G.vw = io.put(v_weights, 12, 12);
  double *x = io.put(xin, 1, 1), *y = io.dev(4);
  OP_BUF(G.cai && G.vw && x && y);
  hipLaunchKernelGGL(swm::dev::k_op_gradient, dim3(nb_), dim3(BLK), 0, C->stream, G, x, y);
  return io.get(xout, y, 4, kCornerVec);
}

swm_status swm_op_average_to_cell_corners(swm_ctx* C, int npad, int nx, int ny, int nsdm, int ndata, int ncorners,
                                          const double* s_weights, const double* xin, double* xout, int flag,
                                          int min_value) {
  OP_BEGIN(C);
  OP_ARGS(s_weights, xin, xout);
  OP_DIMS(npad, nx, ny, nsdm, ndata == 3 && ncorners == 2 && min_value >= -1 && min_value <= 1);
  if (min_value == 0) {
    // present and .false.: the reference leaves the interior undefined
    swm::set_error("average_to_cell_corners: min_value = .false. leaves the result undefined");
    return SWM_ERR_ARG;
  }
  G.sw = io.put(s_weights, 6, 6);
  double *x = io.put(xin, 1, 1), *y = io.dev(2);
  OP_BUF(G.sw && x && y);
  const int mode = min_value == 1 ? 2 : (flag ? 0 : 1);
  hipLaunchKernelGGL(swm::dev::k_op_avg_corners, dim3(nb_), dim3(BLK), 0, C->stream, G, mode, x, y);
  return io.get(xout, y, 2);
}

swm_status swm_op_calculate_center_kinetic_energy(swm_ctx* C, int npad, int nx, int ny, int nsdm, int ndata,
                                                  int nspace, int ncorners, const double* sub_area,
                                                  const double* area_inv, const double* ke_c, double* ke) {
  OP_BEGIN(C);
  OP_ARGS(sub_area, area_inv, ke_c, ke);
  OP_DIMS(npad, nx, ny, nsdm, ndata == 3 && nspace == 2 && ncorners == 2);
  G.sub_area = io.put(sub_area, 6, 6);
  G.area_inv = io.put(area_inv, 1, 1);
  double *x = io.put(ke_c, 2, 2), *y = io.dev(1);
  OP_BUF(G.sub_area && G.area_inv && x && y);
  hipLaunchKernelGGL(swm::dev::k_op_center_ke, dim3(nb_), dim3(BLK), 0, C->stream, G, x, y);
  return io.get(ke, y, 1);
}

swm_status swm_op_calculate_velocity(swm_ctx* C, int npad, int nx, int ny, int nsdm, int ndata, int nspace,
                                     int ncorners, const double* corner_area_inv, const double* v_weights,
                                     const double* psi, const double* chi, double* xout) {
  OP_BEGIN(C);
  OP_ARGS(corner_area_inv, v_weights, psi, chi, xout);
  OP_DIMS(npad, nx, ny, nsdm, ndata == 3 && nspace == 2 && ncorners == 2);
  G.cai = io.put(corner_area_inv, 2, 2);
  G.vw = io.put(v_weights, 12, 12);
  double *p = io.put(psi, 1, 1), *c = io.put(chi, 1, 1), *y = io.dev(4);
  OP_BUF(G.cai && G.vw && p && c && y);
  hipLaunchKernelGGL(swm::dev::k_velocity<double>, dim3(nb_), dim3(BLK), 0, C->stream, G, p, c, y, (const double*)nullptr, (const double*)nullptr, 0.0, (double*)nullptr, (const double*)nullptr);
  return io.get(xout, y, 4, kCornerVec);
}

swm_status swm_op_centered_mass_flux(swm_ctx* C, int npad, int nx, int ny, int nsdm, int ndata, int ncorners,
                                     int nedges, int nspace, int r3, const double* dc, const double* normal,
                                     const double* mass_c, const double* u, const double* mass, double* flux) {
  OP_BEGIN(C);
  OP_ARGS(dc, normal, mass_c, u, mass, flux);
  OP_DIMS(npad, nx, ny, nsdm, ndata == 3 && ncorners == 2 && nedges == 3 && nspace == 2 && r3 == 3);
  G.dc = io.put(dc, 6, 6);
  G.nrm = io.put(normal, 18, 12, kFrame);
  double *mc = io.put(mass_c, 2, 2), *uu = io.put(u, 4, 4, kCornerVec), *m = io.put(mass, 1, 1), *up = io.dev(3),
         *y = io.dev(3);
  OP_BUF(G.dc && G.nrm && mc && uu && m && up && y);
  hipLaunchKernelGGL(swm::dev::k_mass_fluxes, dim3(nb_), dim3(BLK), 0, C->stream, G, uu, m, mc, up, y, (const double*)nullptr,
                     (double*)nullptr, (double*)nullptr);
  return io.get(flux, y, 3);
}

swm_status swm_op_upstream_mass_flux(swm_ctx* C, int npad, int nx, int ny, int nsdm, int ndata, int ncorners,
                                     int nedges, int nspace, int r3, const double* dc, const double* normal,
                                     const double* u, const double* mass, double* flux) {
  OP_BEGIN(C);
  OP_ARGS(dc, normal, u, mass, flux);
  OP_DIMS(npad, nx, ny, nsdm, ndata == 3 && ncorners == 2 && nedges == 3 && nspace == 2 && r3 == 3);
  G.dc = io.put(dc, 6, 6);
  G.nrm = io.put(normal, 18, 12, kFrame);
  double *uu = io.put(u, 4, 4, kCornerVec), *m = io.put(mass, 1, 1), *mc = io.zeros(2), *y = io.dev(3),
         *cen = io.dev(3);
  OP_BUF(G.dc && G.nrm && uu && m && mc && y && cen);
  hipLaunchKernelGGL(swm::dev::k_mass_fluxes, dim3(nb_), dim3(BLK), 0, C->stream, G, uu, m, mc, y, cen, (const double*)nullptr,
                     (double*)nullptr, (double*)nullptr);
  return io.get(flux, y, 3);
}

swm_status swm_op_centered_vorticity_flux(swm_ctx* C, int npad, int nx, int ny, int nsdm, int nedges,
                                          const double* mass_flux, const double* mass, const double* xin,
                                          double* xout) {
  OP_BEGIN(C);
  OP_ARGS(mass_flux, mass, xin, xout);
  OP_DIMS(npad, nx, ny, nsdm, nedges == 3);
  double *mf = io.put(mass_flux, 3, 3), *m = io.put(mass, 1, 1), *e = io.put(xin, 1, 1), *y = io.dev(3);
  OP_BUF(mf && m && e && y);
  hipLaunchKernelGGL(swm::dev::k_op_vort_flux, dim3(nb_), dim3(BLK), 0, C->stream, G, mf, m, e, y);
  return io.get(xout, y, 3);
}

swm_status swm_op_centered_divergence_curl(swm_ctx* C, int npad, int nx, int ny, int nsdm, int ndata, int ncorners,
                                           int nedges, int nspace, int r3, const double* dc, const double* tangent,
                                           const double* mass_c, const double* eta, const double* u,
                                           const double* mass, double* flux) {
  OP_BEGIN(C);
  OP_ARGS(dc, tangent, mass_c, eta, u, mass, flux);
  OP_DIMS(npad, nx, ny, nsdm, ndata == 3 && ncorners == 2 && nedges == 3 && nspace == 2 && r3 == 3);
  G.dc = io.put(dc, 6, 6);
  G.tng = io.put(tangent, 18, 12, kFrame);
  double *mc = io.put(mass_c, 2, 2), *e = io.put(eta, 1, 1), *uu = io.put(u, 4, 4, kCornerVec),
         *m = io.put(mass, 1, 1), *y = io.dev(3);
  OP_BUF(G.dc && G.tng && mc && e && uu && m && y);
  hipLaunchKernelGGL(swm::dev::k_op_div_curl, dim3(nb_), dim3(BLK), 0, C->stream, G, mc, e, uu, m, y);
  return io.get(flux, y, 3);
}

// The two flux-corrected transport routines under their own argument lists,
// on the whole sphere of the context's grid (iim, jjm and nsdm = 10 of the
// grid, one layer), with the step's kernels.  Levels of a multi-level host
// array are moved to and from device planes slot by slot.
}  // extern "C"
namespace {
// level t of a host array (lead, iim*jjm, nlev, 10) as (lead, ns) in slot order
std::vector<double> take_level(const double* a, int lead, int pl, int nlev, int t) {
  std::vector<double> out((size_t)lead * pl * 10);
  for (int k = 0; k < 10; ++k)
    for (int q = 0; q < pl; ++q)
      for (int e = 0; e < lead; ++e)
        out[(size_t)e + lead * ((size_t)q + (size_t)pl * k)] = a[(size_t)e + lead * ((size_t)q + (size_t)pl * (t + (size_t)nlev * k))];
  return out;
}
void put_level(double* a, const std::vector<double>& v, int lead, int pl, int nlev, int t) {
  for (int k = 0; k < 10; ++k)
    for (int q = 0; q < pl; ++q)
      for (int e = 0; e < lead; ++e)
        a[(size_t)e + lead * ((size_t)q + (size_t)pl * (t + (size_t)nlev * k))] = v[(size_t)e + lead * ((size_t)q + (size_t)pl * k)];
}
}  // namespace
extern "C" {

swm_status swm_op_set_mass_fct(swm_ctx* C, int iim, int jjm, int nlm, int nsdm, int ntprog, int nttend, int ncorners,
                               int ndata, int nedges, int nspace, int npad, int r3, int n3, int n4,
                               const int* tend_index, const double* tend_weights, double dt, const double* area_inv,
                               const double* area, const double* dc, const double* normal, double* mss,
                               const double* mss_c, const double* u, double* mss_low, double* mss_flux_horz,
                               double* msh_f, const double* mss_flux_forced, const double* mask) {
  OP_BEGIN(C);
  OP_ARGS(tend_index, tend_weights, area_inv, area, dc, normal, mss, mss_c, u, mss_low, mss_flux_horz, msh_f);
  if (mss_flux_forced || mask) {
    swm::set_error("swm_op_set_mass_fct: the optional mss_flux_forced and mask are not supported "
                   "(step.F passes neither)");
    return SWM_ERR_ARG;
  }
  bool ti_ok = true;
  for (int t = 0; t < 3; ++t) ti_ok = ti_ok && tend_index[t] >= 1 && tend_index[t] <= 3;
  ti_ok = ti_ok && tend_index[0] != tend_index[1] && tend_index[0] != tend_index[2] && tend_index[1] != tend_index[2];
  if (iim != C->iim || jjm != C->jjm || nsdm != 10 || nlm != 1 || ntprog != 2 || nttend != 3 || ncorners != 2 ||
      ndata != 3 || nedges != 3 || nspace != 2 || npad != 1 || r3 != 3 || n3 < 1 || n3 > 2 || n4 < 1 || n4 > 2 ||
      n3 == n4 || !ti_ok) {
    swm::set_error("swm_op_set_mass_fct: the whole grid (iim, jjm, 10 panels, one layer), the model's fixed counts, "
                   "two time levels and a permutation of the three tendency slots");
    return SWM_ERR_DIMS;
  }
  const int pl = iim * jjm;
  OpIO io{C, (size_t)C->ns, {}, {}};
  swm::dev::Geo G = C->d.geo;   // the grid's extents and halo maps; metrics from the arguments
  G.area_inv = io.put(area_inv, 1, 1);
  G.area = io.put(area, 1, 1);
  G.dc = io.put(dc, 6, 6);
  G.nrm = io.put(normal, 18, 12, kFrame);
  const std::vector<double> m3h = take_level(mss, 1, pl, 2, n3 - 1);
  double* h3 = io.put(m3h.data(), 1, 1);
  double *hc = io.put(mss_c, 2, 2), *uu = io.put(u, 4, 4, kCornerVec);
  double* hist[3];
  for (int t = 0; t < 3; ++t) {
    const std::vector<double> v = take_level(msh_f, 3, pl, 3, t);
    hist[t] = io.put(v.data(), 3, 3);
  }
  double *flx1 = io.zeros(3), *anti = io.zeros(3), *low = io.zeros(1), *beta = io.zeros(1), *mflux = io.zeros(3),
         *h4 = io.zeros(1);
  OP_BUF(G.area_inv && G.area && G.dc && G.nrm && h3 && hc && uu && hist[0] && hist[1] && hist[2] && flx1 && anti &&
         low && beta && mflux && h4);
  double *ma = hist[tend_index[0] - 1], *mb = hist[tend_index[1] - 1], *mc = hist[tend_index[2] - 1];
  const double w3 = tend_weights[0], w2 = tend_weights[1], w1 = tend_weights[2];
  const unsigned nb = nblk(C->ns);
  hipStream_t S = C->stream;
  // mass_fct.F:98-133: upstream and centred fluxes, low order, beta down,
  // limited antidiffusive flux, flux-corrected mass and total flux
  hipLaunchKernelGGL(swm::dev::k_mass_fluxes<double>, dim3(nb), dim3(BLK), 0, S, G, uu, h3, hc, flx1, ma,
                     (const double*)nullptr, (double*)nullptr, (double*)nullptr);
  gather<double>(C, C->edge, {flx1, ma});
  hipLaunchKernelGGL(swm::dev::k_mass_low_beta<double>, dim3(nb), dim3(BLK), 0, S, G, dt, w3, w2, w1, flx1, ma, mb, mc,
                     h3, anti, low, beta);
  gather<double>(C, C->face, {beta});
  hipLaunchKernelGGL(swm::dev::k_mass_anti<double>, dim3(nb), dim3(BLK), 0, S, G, beta, anti);
  gather<double>(C, C->edge, {anti});
  hipLaunchKernelGGL(swm::dev::k_mass_final<double>, dim3(nb), dim3(BLK), 0, S, G, dt, flx1, anti, low, mflux, h4,
                     swm::dev::TrFluxArgs<double>{}, swm::dev::VortFluxArgs<double>{},
                     (const double*)nullptr);
  swm_status st;
  std::vector<double> v((size_t)C->ns);
  if ((st = io.get(v.data(), h4, 1))) return st;
  put_level(mss, v, 1, pl, 2, n4 - 1);
  if ((st = io.get(mss_low, low, 1))) return st;
  if ((st = io.get(mss_flux_horz, mflux, 3))) return st;
  v.assign(3 * (size_t)C->ns, 0.0);
  if ((st = io.get(v.data(), ma, 3))) return st;
  put_level(msh_f, v, 3, pl, 3, tend_index[0] - 1);
  return SWM_OK;
}

swm_status swm_op_set_tracer_fct(swm_ctx* C, int iim, int jjm, int nlm, int nsdm, int nedges, double dt,
                                 const double* area_inv, const double* area, const double* trc, const double* trc_high,
                                 const double* mss, const double* mss_low, const double* mss_horz,
                                 const double* mss_flux_horz, double* trc_fct, double* trc_flux_horz,
                                 double* msstrc_low_order, const double* mask) {
  OP_BEGIN(C);
  OP_ARGS(area_inv, area, trc, trc_high, mss, mss_low, mss_horz, mss_flux_horz, trc_fct);
  if (mask) {
    swm::set_error("swm_op_set_tracer_fct: the optional mask is not supported (step.F does not pass it)");
    return SWM_ERR_ARG;
  }
  if (iim != C->iim || jjm != C->jjm || nsdm != 10 || nlm != 1 || nedges != 3) {
    swm::set_error("swm_op_set_tracer_fct: the whole grid (iim, jjm, 10 panels, one layer) and 3 edges");
    return SWM_ERR_DIMS;
  }
  OpIO io{C, (size_t)C->ns, {}, {}};
  swm::dev::Geo G = C->d.geo;
  G.area_inv = io.put(area_inv, 1, 1);
  G.area = io.put(area, 1, 1);
  double *t3 = io.put(trc, 1, 1), *th = io.put(trc_high, 1, 1), *h3 = io.put(mss, 1, 1), *low = io.put(mss_low, 1, 1),
         *h4 = io.put(mss_horz, 1, 1), *mflux = io.put(mss_flux_horz, 3, 3);
  double *f1 = io.zeros(3), *f2 = io.zeros(3), *anti = io.zeros(3), *ml = io.zeros(1), *tlow = io.zeros(1),
         *bdn = io.zeros(1), *bup = io.zeros(1), *t4 = io.zeros(1), *told = io.zeros(1);
  OP_BUF(G.area_inv && G.area && t3 && th && h3 && low && h4 && mflux && f1 && f2 && anti && ml && tlow && bdn && bup &&
         t4 && told);
  const unsigned nb = nblk(C->ns);
  hipStream_t S = C->stream;
  // tracer_fct.F:98-159: upstream and centred tracer fluxes (the centred one
  // of trc_high: weights 0, 1, 0 select it exactly), low order, the limiter
  // after the halo fills of trc and mss_horz, beta up and down, the limited
  // antidiffusive flux, the flux-corrected tracer
  hipLaunchKernelGGL(swm::dev::k_tr_fluxes<double>, dim3(nb), dim3(BLK), 0, S, G, 0.0, 1.0, 0.0, t3, th, t3, mflux, f1,
                     f2);
  gather<double>(C, C->edge, {f1, f2});
  gather<double>(C, C->face, {t3, h4});
  hipLaunchKernelGGL(swm::dev::k_tr_low<double>, dim3(nb), dim3(BLK), 0, S, G, dt, h3, t3, f1, h4, ml, tlow);
  gather<double>(C, C->face, {ml, tlow});
  hipLaunchKernelGGL(swm::dev::k_tr_limiter<double>, dim3(nb), dim3(BLK), 0, S, G, dt, t3, ml, tlow, low, f1, f2, anti,
                     bdn, bup);
  gather<double>(C, C->face, {bdn, bup});
  hipLaunchKernelGGL(swm::dev::k_tr_anti<double>, dim3(nb), dim3(BLK), 0, S, G, bdn, bup, anti);
  gather<double>(C, C->edge, {anti});
  hipLaunchKernelGGL(swm::dev::k_tr_final<double>, dim3(nb), dim3(BLK), 0, S, G, dt, ml, anti, h4, t4, told,
                     (const double*)nullptr, (const double*)nullptr);
  swm_status st;
  if ((st = io.get(trc_fct, t4, 1))) return st;
  if (msstrc_low_order && (st = io.get(msstrc_low_order, ml, 1))) return st;
  if (trc_flux_horz) {
    hipLaunchKernelGGL(swm::dev::k_axpy, dim3(nblk(3 * (size_t)C->ns)), dim3(BLK), 0, S, 3 * (size_t)C->ns, 1.0, anti, f1);
    if ((st = io.get(trc_flux_horz, f1, 3))) return st;
  }
  return SWM_OK;
}

swm_status swm_op_remap_mass_and_tracers(swm_ctx* C, int iim, int jjm, int nmass, int nsdm, int ntrace, double* mass,
                                         double* trace, double dt, const double* uc, double* edgeflux) {
  OP_BEGIN(C);
  OP_ARGS(mass, uc);
  if (ntrace < 0 || (ntrace > 0 && !trace) || !(dt > 0.0)) {
    swm::set_error("swm_op_remap_mass_and_tracers: bad argument");
    return SWM_ERR_ARG;
  }
  if (C->advection != SWM_ADV_REMAP || !C->rm.mem) {
    swm::set_error("swm_op_remap_mass_and_tracers: needs a context made with SWM_ADV_REMAP");
    return SWM_ERR_ARG;
  }
  if (iim != C->iim || jjm != C->jjm || nsdm != 10 || nmass != 1) {
    swm::set_error("swm_op_remap_mass_and_tracers: the whole grid (iim, jjm, 10 panels) and one mass field");
    return SWM_ERR_DIMS;
  }
  const size_t ns = (size_t)C->ns;
  const int nq = ntrace > 0 ? ntrace : 1;
  OpIO io{C, ns, {}, {}};
  double* h = io.put(mass, 1, 1);
  double* u = io.put(uc, 4, 4, kCornerVec);
  double *tr = io.dev(nq), *ptab = io.dev(1), *mx = io.zeros(1), *my = io.zeros(1), *mp = io.zeros(1),
         *tx = io.zeros(nq), *ty = io.zeros(nq), *tp = io.zeros(nq), *dp = io.zeros(6), *mflux = io.zeros(3),
         *mtflux = io.zeros(3 * nq), *hout = io.zeros(1), *tout = io.zeros(nq), *ef = io.zeros(3);
  OP_BUF(h && u && tr && ptab && mx && my && mp && tx && ty && tp && dp && mflux && mtflux && hout && tout && ef);
  // tracers (iim, jjm, ntrace, 1, 10) -> planes; the plane pointer table
  std::vector<double> planes((size_t)ntrace * ns);
  const size_t pl = (size_t)iim * jjm;
  for (int k = 0; k < 10; ++k)
    for (int q = 0; q < ntrace; ++q)
      std::memcpy(&planes[(size_t)q * ns + k * pl], &trace[((size_t)q + (size_t)ntrace * k) * pl], pl * sizeof(double));
  std::vector<double*> tab(nq);
  for (int q = 0; q < nq; ++q) tab[q] = tr + (size_t)q * ns;
  if (ntrace) HIP_TRY(hipMemcpy(tr, planes.data(), planes.size() * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(ptab, tab.data(), tab.size() * sizeof(double*), hipMemcpyHostToDevice));
  RemapDev& D = C->rm;
  *D.flag_host = 0;
  swm::dev::RemapArgs A{C->iim, C->jjm, (int)ns, ntrace, D.cellp, D.edgep, D.cornp, C->radius, D.limit_min,
                        D.limit_max, D.flag_dev};
  const unsigned nb = nblk_b(ns, C->blk_step);
  hipStream_t S = C->stream;
  const double* const* t3 = reinterpret_cast<const double* const*>(ptab);
  // remap.F:1340-1580 as get_remap_flux calls it: prep and fields, halo
  // fills, departure points (after the velocity's own halo fill), triangles
  // and transports, their halo fills, the updated fields and the edge flux
  gather<double>(C, C->vrtx, {u, u + 2 * ns});
  hipLaunchKernelGGL(swm::dev::k_rm_construct, zgrid(C, nb), dim3(C->blk_step), 0, S, A, h, t3, mx, my, mp, tx, ty, tp);
  std::vector<double*> ff = {mx, my, mp};
  for (int q = 0; q < ntrace; ++q) {
    ff.push_back(tx + q * ns);
    ff.push_back(ty + q * ns);
    ff.push_back(tp + q * ns);
  }
  gather_list<double>(C, C->face, ff);
  hipLaunchKernelGGL(swm::dev::k_rm_departure, zgrid(C, nb), dim3(C->blk_step), 0, S, A, dt, u, u + 2 * ns, dp);
  gather<double>(C, C->vrtx, {dp, dp + 2 * ns, dp + 4 * ns});
  hipLaunchKernelGGL(swm::dev::k_rm_transport, zgrid(C, nb), dim3(C->blk_step), 0, S, A, dp, mx, my, mp, tx, ty, tp,
                     mflux, mtflux);
  std::vector<double*> efl = {mflux};
  for (int q = 0; q < ntrace; ++q) efl.push_back(mtflux + 3 * q * ns);
  gather_list<double>(C, C->edge, efl);
  hipLaunchKernelGGL(swm::dev::k_rm_fields, zgrid(C, nb), dim3(C->blk_step), 0, S, A, 1.0 / dt, h, t3, mflux, mtflux,
                     hout, tout, ef);
  HIP_TRY(hipStreamSynchronize(S));
  if (*(volatile int*)D.flag_host) {
    *D.flag_host = 0;
    swm::set_error("swm_op_remap_mass_and_tracers: a mass below -1e-10 (remap_prep stops the reference there)");
    return SWM_ERR_STATE;
  }
  swm_status st;
  if ((st = io.get(mass, hout, 1))) return st;
  if (edgeflux && (st = io.get(edgeflux, ef, 3))) return st;
  if (ntrace) {
    HIP_TRY(hipMemcpy(planes.data(), tout, planes.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int k = 0; k < 10; ++k)
      for (int q = 0; q < ntrace; ++q)
        std::memcpy(&trace[((size_t)q + (size_t)ntrace * k) * pl], &planes[(size_t)q * ns + k * pl], pl * sizeof(double));
  }
  return SWM_OK;
}

swm_status swm_init_diss_op(swm_ctx* C, const swm_grid* grid) {
  OP_BEGIN(C);
  if (!grid) { swm::set_error("swm_init_diss_op: bad argument"); return SWM_ERR_ARG; }
  const swm::Grid& Gh = swm::grid_of(grid);
  if (Gh.iim != C->iim || Gh.jjm != C->jjm) {
    swm::set_error("swm_init_diss_op: the grid is not the context's");
    return SWM_ERR_DIMS;
  }
  swm::DissGeometry D;
  swm::build_diss_geometry(Gh, D);
  const size_t ns = (size_t)C->ns;
  std::vector<double> planes((size_t)swm::dev::DS_PLANES * ns);
  for (size_t n = 0; n < ns; ++n) {
    for (int q = 0; q < 24; ++q) planes[(size_t)q * ns + n] = D.t[q + 24 * n];
    for (int ic = 0; ic < 2; ++ic)
      for (int id = 0; id < 3; ++id) {
        const size_t m = 2 * (id + 3 * (ic + 2 * n));
        planes[(size_t)(24 + id + 3 * ic) * ns + n] = D.magt[m] + D.magt[m + 1];
        planes[(size_t)(30 + id + 3 * ic) * ns + n] = D.magn[id + 3 * (ic + 2 * n)];
      }
  }
  if (!C->diss) HIP_TRY(hipMalloc(&C->diss, planes.size() * sizeof(double)));
  HIP_TRY(hipMemcpy(C->diss, planes.data(), planes.size() * sizeof(double), hipMemcpyHostToDevice));
  return SWM_OK;
}

swm_status swm_op_dissipation(swm_ctx* C, int norder, double visc, const double* xin, double* xout) {
  OP_BEGIN(C);
  OP_ARGS(xin, xout);
  if (norder < 1) { swm::set_error("swm_op_dissipation: norder >= 1"); return SWM_ERR_ARG; }
  if (!C->diss) { swm::set_error("swm_op_dissipation: swm_init_diss_op first"); return SWM_ERR_STATE; }
  const size_t ns = (size_t)C->ns;
  OpIO io{C, ns, {}, {}};
  double* a = io.put(xin, 4, 4, kCornerVec);
  double* b = io.zeros(4);
  OP_BUF(a && b);
  const unsigned nb = nblk_b(ns, C->blk_step);
  hipStream_t S = C->stream;
  // norder applications, each after the corner halo fill (the vertex wrap of
  // :279); the last one times -visc (norder even) or +visc (:436-441)
  for (int it = 0; it < norder; ++it) {
    gather<double>(C, C->vrtx, {a, a + 2 * ns});
    const double f = it + 1 < norder ? 1.0 : (norder % 2 == 0 ? -visc : visc);
    hipLaunchKernelGGL(swm::dev::k_diss, dim3(nb), dim3(C->blk_step), 0, S, C->iim, C->jjm, (int)ns, C->diss, C->cai, f,
                       a, b);
    std::swap(a, b);
  }
  return io.get(xout, a, 4, kCornerVec);
}

swm_status swm_wrap(swm_ctx* C, int kind, int nspace, int nlayer, double* x) {
  OP_BEGIN(C);
  OpIO io{C, (size_t)C->ns, {}, {}};
  if (!x || kind < 0 || kind > 2 || nlayer < 1 || nspace < 1 || nspace > (kind == 0 ? 1 : (kind == 1 ? 4 : 2))) {
    swm::set_error("swm_wrap: kind must be 0 (face), 1 (vrtx, nspace 1..4) or 2 (edge, nspace 1..2)");
    return SWM_ERR_ARG;
  }
  const int ncomp = kind == 0 ? 1 : (kind == 1 ? 2 : 3);
  const DevTable& t = kind == 0 ? C->face : (kind == 1 ? C->vrtx : C->edge);
  const int ns = C->ns, pl = C->iim * C->jjm;
  double* d = io.dev(ncomp);
  OP_BUF(d);
  // reference layout (nspace, ncomp, iim, jjm, nlayer, 10); the gather table
  // moves every space component of an entry together
  std::vector<double> planes((size_t)ncomp * ns);
  auto at = [&](int sp, int c, int q, int l, int k) -> double& {
    return x[sp + (size_t)nspace * (c + (size_t)ncomp * (q + (size_t)pl * (l + (size_t)nlayer * k)))];
  };
  for (int l = 0; l < nlayer; ++l)
    for (int sp = 0; sp < nspace; ++sp) {
      for (int k = 0; k < 10; ++k)
        for (int q = 0; q < pl; ++q)
          for (int c = 0; c < ncomp; ++c) planes[(size_t)c * ns + (size_t)k * pl + q] = at(sp, c, q, l, k);
      HIP_TRY(hipMemcpy(d, planes.data(), planes.size() * sizeof(double), hipMemcpyHostToDevice));
      gather<double>(C, t, {d});
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipStreamSynchronize(C->stream));
      HIP_TRY(hipMemcpy(planes.data(), d, planes.size() * sizeof(double), hipMemcpyDeviceToHost));
      for (int k = 0; k < 10; ++k)
        for (int q = 0; q < pl; ++q)
          for (int c = 0; c < ncomp; ++c) at(sp, c, q, l, k) = planes[(size_t)c * ns + (size_t)k * pl + q];
    }
  return SWM_OK;
}

swm_status swm_elliptic(swm_ctx* C, int nvar, double* rhs, double* xout) {
  OP_BEGIN(C);
  OpIO io{C, (size_t)C->ns, {}, {}};
  if (!rhs || !xout || nvar < 1 || nvar > 2) {
    swm::set_error("swm_elliptic: nvar must be 1 or 2");
    return SWM_ERR_ARG;
  }
  const int pl = C->iim * C->jjm;
  std::vector<double> b((size_t)nvar * C->ns);
  double *r = io.dev(nvar), *x = io.dev(nvar);
  OP_BUF(r && x);
  vars_to_planes(rhs, nvar, pl, b.data());
  HIP_TRY(hipMemcpy(r, b.data(), b.size() * sizeof(double), hipMemcpyHostToDevice));
  vars_to_planes(xout, nvar, pl, b.data());
  HIP_TRY(hipMemcpy(x, b.data(), b.size() * sizeof(double), hipMemcpyHostToDevice));
  const swm_status st = elliptic(C, r, x, nvar);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(C->stream));
  HIP_TRY(hipMemcpy(b.data(), r, b.size() * sizeof(double), hipMemcpyDeviceToHost));
  planes_to_vars(b.data(), nvar, pl, rhs);
  HIP_TRY(hipMemcpy(b.data(), x, b.size() * sizeof(double), hipMemcpyDeviceToHost));
  planes_to_vars(b.data(), nvar, pl, xout);
  return st;
}

swm_status swm_multigrid(swm_ctx* C, int nvar, const double* beta, double* alpha) {
  OP_BEGIN(C);
  OpIO io{C, (size_t)C->ns, {}, {}};
  if (!beta || !alpha || nvar < 1 || nvar > 2) {
    swm::set_error("swm_multigrid: nvar must be 1 or 2");
    return SWM_ERR_ARG;
  }
  const int pl = C->iim * C->jjm;
  std::vector<double> b((size_t)nvar * C->ns);
  double *be = io.dev(nvar), *al = io.dev(nvar);
  OP_BUF(be && al);
  vars_to_planes(beta, nvar, pl, b.data());
  HIP_TRY(hipMemcpy(be, b.data(), b.size() * sizeof(double), hipMemcpyHostToDevice));
  vars_to_planes(alpha, nvar, pl, b.data());
  HIP_TRY(hipMemcpy(al, b.data(), b.size() * sizeof(double), hipMemcpyHostToDevice));
  vcycle<double>(C, C->d, nvar, al, be);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(C->stream));
  HIP_TRY(hipMemcpy(b.data(), al, b.size() * sizeof(double), hipMemcpyDeviceToHost));
  planes_to_vars(b.data(), nvar, pl, alpha);
  return SWM_OK;
}

}  // extern "C"

// ===================================================================
// Tangent-linear and adjoint model of a window of FP64 steps with any of
// the three advection schemes, and optionally the tracers (the kernels are
// in swm_tlad.hip).
//
// swm_adj_forward runs the window from the initial state, keeping a
// snapshot of the model state every `every` steps (and at the end).  The
// TL and AD sweeps then recompute one segment of base states at a time
// from its snapshot (h, eta and u at the start of each step, and what the
// scheme and the tracers need: the coefficients of the linearised step) and
// run the linear steps over it, forward for TL, backward for AD.  Each
// step's elliptic solve is linearised as the base step's own V-cycles
// (exact_lin), or as a converged inverse solved to a relative residual of
// 1e-13, whose transpose uses the symmetry of area times the Laplacian,
// L^T = A L A^-1.
// ===================================================================
struct TlAd {
  // TL or AD state: two time levels of h, eta, div, u (4 planes), three
  // ABM3 history slots of the h, eta and div tendencies, and (FCT) three of
  // the centred edge mass fluxes
  double *h[2], *eta[2], *div[2], *u[2], *hf[3], *ef[3], *df[3], *mf[3];
  // scratch planes
  double *hc, *F, *vf, *dc, *ke, *X, *rhs, *xs, *psia, *bhc, *bF, *zeros, *am_inv;
  double *fl, *an, *ml, *bsc;         // FCT: upstream flux, antidiffusive flux, low-order mass, limiter scratch
  int seg_planes = 6;                 // per base state: h, eta, u (4); FCT: + limiter coefficients (3), mass flux (3)
  double* mem = nullptr;
  // trajectory
  int every = 0, nsteps = 0;
  size_t snap_words = 0;
  std::vector<double*> ckpt;          // snapshots at window steps 0, every, 2 every, ..., and the end
  std::vector<long long> ckpt_meta;   // step_count, nsecond per snapshot
  double* seg = nullptr;              // (every + 1) base states (seg_planes each)
  bool have_window = false;
  swm::dev::Geo lin;                  // metrics with hs = 0 and f = 0
  // FCT: edges whose frozen limiter coefficient is below 1, counted over the
  // base steps recomputed by the last TL or AD sweep
  unsigned long long* lim_count = nullptr;
  long long lim_steps = 0;
  // exact linearisation of the solves: the schedule of each base step of the
  // segment, the adjoint multigrid buffers (per level: iterate, right-hand
  // side, scratch; two fields) and the coarse cycle's transpose
  std::vector<swm_ctx::SolveSched> sched;
  double* mg_mem = nullptr;
  double *mg_a[24] = {}, *mg_b[24] = {}, *mg_t[24] = {};
  double* coarse_mt = nullptr;
  int coarse_n = 0;
  // tracers (the *_tracer entry points): per tracer two levels, the AB3
  // extrapolation's oldest level (FCT) and three tendency history slots
  // (centred, upstream); scratch for the fluxes; the frozen tracer limiter's
  // count as for the mass
  bool tracers = false;
  double* trmem = nullptr;
  std::vector<double*> tr[2], tro, trf[3];
  double *tf1 = nullptr, *tf2 = nullptr, *tan = nullptr, *tml = nullptr;
  unsigned long long* tlim_count = nullptr;
  long long tlim_steps = 0;
};

namespace swm {

void tlad_free(swm_ctx* C) {
  if (!C->tl) return;
  TlAd* T = C->tl;
  for (double* p : T->ckpt) (void)hipFree(p);
  if (T->seg) (void)hipFree(T->seg);
  if (T->mem) (void)hipFree(T->mem);
  if (T->lim_count) (void)hipFree(T->lim_count);
  if (T->mg_mem) (void)hipFree(T->mg_mem);
  if (T->coarse_mt) (void)hipFree(T->coarse_mt);
  if (T->trmem) (void)hipFree(T->trmem);
  if (T->tlim_count) (void)hipFree(T->tlim_count);
  delete T;
  C->tl = nullptr;
}

namespace {

constexpr double kLinTol = 1.0e-13;   // relative residual of the linear solves
constexpr int kLinMaxCycles = 80;

swm_status tlad_alloc(swm_ctx* C) {
  if (C->tl) return SWM_OK;
  const size_t ns = C->ns;
  TlAd* T = new TlAd;
  C->tl = T;
  const size_t words = ns * (2 * (1 + 1 + 1 + 4) + 9 + 9 + 2 + 3 + 3 + 3 + 1 + 1 + 2 + 2 + 2 + 2 + 3 + 1 + 1 + 3 + 3 + 1 + 1);
  HIP_TRY(hipMalloc(&T->mem, words * sizeof(double)));
  HIP_TRY(hipMalloc(&T->lim_count, sizeof(unsigned long long)));
  HIP_TRY(hipMemsetAsync(T->lim_count, 0, sizeof(unsigned long long), C->stream));
  double* p = T->mem;
  auto take = [&](size_t n) { double* q = p; p += n; return q; };
  for (int t = 0; t < 2; ++t) { T->h[t] = take(ns); T->eta[t] = take(ns); T->div[t] = take(ns); T->u[t] = take(4 * ns); }
  for (int t = 0; t < 3; ++t) { T->hf[t] = take(ns); T->ef[t] = take(ns); T->df[t] = take(ns); }
  for (int t = 0; t < 3; ++t) T->mf[t] = take(3 * ns);
  T->hc = take(2 * ns); T->F = take(3 * ns); T->vf = take(3 * ns); T->dc = take(3 * ns); T->ke = take(ns);
  T->X = take(ns); T->rhs = take(2 * ns); T->xs = take(2 * ns); T->psia = take(2 * ns); T->bhc = take(2 * ns);
  T->bF = take(3 * ns); T->zeros = take(ns); T->am_inv = take(ns);
  T->fl = take(3 * ns); T->an = take(3 * ns); T->ml = take(ns); T->bsc = take(ns);
  fill<double>(C, T->mem, words, 0.0);
  T->lin = C->d.geo;
  T->lin.hs = T->zeros;
  T->lin.f = T->zeros;
  // 1 / area at real cells (the transpose of the solve scales by A^-1)
  hipLaunchKernelGGL(swm::dev::k_mul_mask, dim3(nblk(ns)), dim3(BLK), 0, C->stream, (int)ns, C->mask, C->area_inv,
                     C->mask, T->am_inv);
  T->snap_words = C->d.state_words + 2 * ns;
  if (C->nt > 0) {
    const size_t tw = ns * (6 * (size_t)C->nt + 10);
    HIP_TRY(hipMalloc(&T->trmem, tw * sizeof(double)));
    HIP_TRY(hipMalloc(&T->tlim_count, sizeof(unsigned long long)));
    HIP_TRY(hipMemsetAsync(T->tlim_count, 0, sizeof(unsigned long long), C->stream));
    double* t = T->trmem;
    for (int q = 0; q < C->nt; ++q) {
      for (int l = 0; l < 2; ++l) { T->tr[l].push_back(t); t += ns; }
      T->tro.push_back(t); t += ns;
      for (int l = 0; l < 3; ++l) { T->trf[l].push_back(t); t += ns; }
    }
    T->tf1 = t; t += 3 * ns;
    T->tf2 = t; t += 3 * ns;
    T->tan = t; t += 3 * ns;
    T->tml = t;
    fill<double>(C, T->trmem, tw, 0.0);
  }
  HIP_TRY(hipStreamSynchronize(C->stream));
  return SWM_OK;
}

void clear_state(swm_ctx* C) {
  TlAd* T = C->tl;
  const size_t ns = C->ns;
  fill<double>(C, T->h[0], (size_t)ns * 2 * (1 + 1 + 1 + 4) + 18 * ns, 0.0);   // the state planes are contiguous
  if (T->trmem) fill<double>(C, T->trmem, ns * (6 * (size_t)C->nt + 10), 0.0);
}

swm_status snapshot(swm_ctx* C, double* dst) {
  HIP_TRY(hipMemcpyAsync(dst, C->d.state_begin, C->d.state_words * sizeof(double), hipMemcpyDeviceToDevice, C->stream));
  HIP_TRY(hipMemcpyAsync(dst + C->d.state_words, C->xout, 2 * (size_t)C->ns * sizeof(double), hipMemcpyDeviceToDevice,
                         C->stream));
  return SWM_OK;
}
swm_status restore(swm_ctx* C, const double* src, long long sc, long long nsec) {
  HIP_TRY(hipMemcpyAsync(C->d.state_begin, src, C->d.state_words * sizeof(double), hipMemcpyDeviceToDevice, C->stream));
  HIP_TRY(hipMemcpyAsync(C->xout, src + C->d.state_words, 2 * (size_t)C->ns * sizeof(double), hipMemcpyDeviceToDevice,
                         C->stream));
  C->step_count = (int)sc;
  C->nsecond = nsec;
  pointers(C);
  return SWM_OK;
}

// one forward step of the model (the nonlinear base)
swm_status base_step(swm_ctx* C) {
  C->step_count += 1;
  C->nsecond = (long long)((double)C->nsecond + C->dt);
  pointers(C);
  weights(C);
  swm_status st = step_dt<double>(C, C->d);
  if (st) return st;
  wrap_prognostics<double>(C, C->d);
  return SWM_OK;
}

// base state slot q of the segment buffer: h, eta, u (4 planes); FCT: the
// step's frozen limiter coefficients (3 planes) and total mass flux (3)
double* seg_h(swm_ctx* C, int q) { return C->tl->seg + (size_t)q * C->tl->seg_planes * C->ns; }
double* seg_e(swm_ctx* C, int q) { return seg_h(C, q) + C->ns; }
double* seg_u(swm_ctx* C, int q) { return seg_h(C, q) + 2 * (size_t)C->ns; }
double* seg_c(swm_ctx* C, int q) { return seg_h(C, q) + 6 * (size_t)C->ns; }
double* seg_F(swm_ctx* C, int q) { return seg_h(C, q) + 9 * (size_t)C->ns; }
// tracer tq of base state q (tracer runs): the level the step reads (t3),
// the older level (t4) and the AB3 extrapolation's oldest (told), then
// (FCT) the step's three frozen limiter coefficient planes
double* seg_tr(swm_ctx* C, int q, int tq) {
  return seg_h(C, q) + (size_t)((C->advection == SWM_ADV_FCT ? 12 : 6) + 6 * tq) * C->ns;
}

swm_status record_base(swm_ctx* C, int q) {
  const int n3 = C->n4 - 1;   // the level the next step reads (after pointers of the last step: n4)
  const size_t ns = C->ns;
  HIP_TRY(hipMemcpyAsync(seg_h(C, q), C->d.h[n3], ns * sizeof(double), hipMemcpyDeviceToDevice, C->stream));
  HIP_TRY(hipMemcpyAsync(seg_e(C, q), C->d.eta[n3], ns * sizeof(double), hipMemcpyDeviceToDevice, C->stream));
  HIP_TRY(hipMemcpyAsync(seg_u(C, q), C->d.u[n3], 4 * ns * sizeof(double), hipMemcpyDeviceToDevice, C->stream));
  if (C->tl->tracers) {
    const int n4 = C->n3 - 1;   // the other level: the next step's older one
    for (int tq = 0; tq < C->nt; ++tq) {
      double* b = seg_tr(C, q, tq);
      HIP_TRY(hipMemcpyAsync(b, C->d.tr[n3][tq], ns * sizeof(double), hipMemcpyDeviceToDevice, C->stream));
      HIP_TRY(hipMemcpyAsync(b + ns, C->d.tr[n4][tq], ns * sizeof(double), hipMemcpyDeviceToDevice, C->stream));
      HIP_TRY(hipMemcpyAsync(b + 2 * ns, C->d.told[tq], ns * sizeof(double), hipMemcpyDeviceToDevice, C->stream));
    }
  }
  return SWM_OK;
}

// base states of segment sidx: from its snapshot, len steps; slot q holds
// the state before window step sidx*every + q + 1 (q = len: after the last)
swm_status recompute_segment(swm_ctx* C, int sidx, int len) {
  TlAd* T = C->tl;
  swm_status st;
  if ((st = restore(C, T->ckpt[sidx], T->ckpt_meta[2 * sidx], T->ckpt_meta[2 * sidx + 1]))) return st;
  if ((st = record_base(C, 0))) return st;
  const bool fct = C->advection == SWM_ADV_FCT;
  for (int q = 0; q < len; ++q) {
    // FCT: the step records its limiter coefficients (step_dt), then its
    // total mass flux is kept
    if (fct) C->fct_rec = seg_c(C, q);
    if (fct && T->tracers) C->tr_rec = seg_tr(C, q, 0);
    st = base_step(C);
    C->fct_rec = nullptr;
    C->tr_rec = nullptr;
    if (st) return st;
    if (fct && T->tracers) {
      for (int tq = 0; tq < C->nt; ++tq) {
        const size_t n3 = 3 * (size_t)C->ns;
        hipLaunchKernelGGL(swm::dev::k_count_below_one, dim3((unsigned)((n3 + BLK - 1) / BLK)), dim3(BLK), 0,
                           C->stream, n3, (const double*)(seg_tr(C, q, tq) + 3 * (size_t)C->ns), T->tlim_count);
      }
      T->tlim_steps += 1;
    }
    T->sched[q] = C->last_sched;
    if (fct) {
      HIP_TRY(hipMemcpyAsync(seg_F(C, q), C->d.mass_flux, 3 * (size_t)C->ns * sizeof(double), hipMemcpyDeviceToDevice,
                             C->stream));
      const size_t n3 = 3 * (size_t)C->ns;
      hipLaunchKernelGGL(swm::dev::k_count_below_one, dim3((unsigned)((n3 + BLK - 1) / BLK)), dim3(BLK), 0, C->stream,
                         n3, (const double*)seg_c(C, q), T->lim_count);
      T->lim_steps += 1;
    }
    if ((st = record_base(C, q + 1))) return st;
  }
  return SWM_OK;
}

// linear elliptic solve (FP64 solver, relative residual kLinTol) of two
// fields: rhs (gathered and mean-removed in place), x warm start and result
swm_status lin_solve(swm_ctx* C, double* rhs, double* x) {
  const double tol = C->solver_tol;
  const int rel = C->solver_relative, mc = C->solver_max_cycles, stall = C->solver_stall;
  const int lv = C->last_vcycles, fc = C->fixed_cycles;
  C->solver_tol = kLinTol; C->solver_relative = 1; C->solver_max_cycles = kLinMaxCycles; C->solver_stall = 1;
  C->fixed_cycles = 0;
  swm_status st = elliptic(C, rhs, x, 2);
  C->solver_tol = tol; C->solver_relative = rel; C->solver_max_cycles = mc; C->solver_stall = stall;
  C->last_vcycles = lv;
  C->fixed_cycles = fc;
  return st;
}

// the step's coefficients from base state q: corner thickness, and the
// base mass flux the vorticity flux takes (centred or upstream, recomputed;
// FCT: the recorded total flux); returns the mass flux
const double* base_coeffs(swm_ctx* C, int q) {
  TlAd* T = C->tl;
  const unsigned nb = nblk(C->ns);
  const swm::dev::Geo& G = C->d.geo;
  hipLaunchKernelGGL(swm::dev::k_avg_corners<double>, dim3(nb), dim3(BLK), 0, C->stream, G, G.sw, seg_h(C, q), T->bhc);
  gather<double>(C, C->vrtx, {T->bhc});
  if (C->advection == SWM_ADV_FCT) return seg_F(C, q);
  // k_mass_fluxes: the upstream flux first, the centred one second
  const bool cen = C->advection == SWM_ADV_CENTERED;
  hipLaunchKernelGGL(swm::dev::k_mass_fluxes<double>, dim3(nb), dim3(BLK), 0, C->stream, G, seg_u(C, q), seg_h(C, q),
                     T->bhc, cen ? T->vf : T->bF, cen ? T->bF : T->vf, (const double*)nullptr, (double*)nullptr,
                     (double*)nullptr);
  gather<double>(C, C->edge, {T->bF});
  return T->bF;
}

// halo transposes of a set of fields
void gather_adj(swm_ctx* C, const DevTable& t, std::initializer_list<double*> fields) {
  std::vector<double*> f(fields);
  for (size_t a = 0; a < f.size(); a += 8) {
    swm::dev::FieldSet fs{};
    int nf = 0;
    for (size_t b = a; b < f.size() && b < a + 8; ++b) fs.p[nf++] = f[b];
    if (!t.n) continue;
    hipLaunchKernelGGL(swm::dev::k_gather_adj, dim3(nblk(t.n), nf), dim3(BLK), 0, C->stream, t.dst, t.src, t.n, fs);
    hipLaunchKernelGGL(swm::dev::k_gather_adj_zero, dim3(nblk(t.n), nf), dim3(BLK), 0, C->stream, t.dst, t.n, fs);
  }
}

// y -= am * sum(y) / A over all slots: the transpose of the mean removal
void mean_removal_adj(swm_ctx* C, double* y) {
  const int ns = C->ns;
  hipLaunchKernelGGL(swm::dev::k_sum_partial<double>, dim3(C->npart, 1), dim3(BLK), 0, C->stream, y, y,
                     (const double*)nullptr, ns, C->part, C->mb);
  sum_final(C, C->part, C->npart, C->sums, 1);
  hipLaunchKernelGGL(swm::dev::k_sub_am, dim3(nblk(ns)), dim3(BLK), 0, C->stream, ns, C->area_mask, C->sums,
                     1.0 / C->total_area, y);
}

// ---------------------------------------- exact linearisation of the solve
// The FP64 step's solve (elliptic) is: halo fills of the warm start and the
// right-hand side, mean removal of the right-hand side, `first` V-cycles of
// both fields, then `extra` cycles of the fields that missed the criterion,
// with halo fills of the iterate between.  Every kernel of a V-cycle is
// linear with the level's weights as coefficients, so for the cycle counts
// of the base step the solve is a linear map of (warm start, right-hand
// side).  Its tangent-linear is the same sequence on the perturbations (the
// warm start's perturbation carried from the previous step), its adjoint
// the transposed kernels in reverse order.  This is the derivative of the
// model as it runs, whatever its criterion; the solves' dependence on the
// warm start is part of it.  Needs a level above the coarse cycle's.
bool exact_lin(const swm_ctx* C) { return C->lin_exact && C->g > swm::dev::CO_MAXLEV; }

// adjoint level buffers and the coarse cycle's transpose, once per context:
// the transpose is probed from the device's own coarse kernel, one unit
// right-hand side at a time (column j of M is the cycle's answer to e_j,
// which is row j of M^T)
swm_status exact_prepare(swm_ctx* C) {
  TlAd* T = C->tl;
  if (T->mg_mem) return SWM_OK;
  View<double>& V = C->d;
  const int G = C->g, top = swm::dev::CO_MAXLEV;
  size_t words = 0;
  for (int l = top; l <= G; ++l) words += 3 * 2 * (size_t)V.lev[l - 1].ns;
  HIP_TRY(hipMalloc(&T->mg_mem, words * sizeof(double)));
  fill_plain<double>(C, T->mg_mem, words, 0.0);
  double* p = T->mg_mem;
  for (int l = top; l <= G; ++l) {
    const size_t n2 = 2 * (size_t)V.lev[l - 1].ns;
    T->mg_a[l] = p; p += n2;
    T->mg_b[l] = p; p += n2;
    T->mg_t[l] = p; p += n2;
  }
  const int n = V.lev[top - 1].ns;
  T->coarse_n = n;
  HIP_TRY(hipMalloc(&T->coarse_mt, (size_t)n * n * sizeof(double)));
  swm::dev::CoarseArgs<double> P{};
  for (int l = 1; l <= top; ++l) P.lev[l - 1] = V.lev[l - 1].L;
  P.nlev = top;
  P.nvar = 1;
  P.beta = V.lev[top - 1].beta;
  P.alpha = V.lev[top - 1].alpha;
  P.zero = 1;
  P.mb = C->mb;
  for (int j = 0; j < n; ++j) {
    hipLaunchKernelGGL(swm::dev::k_unit, dim3(nblk(n)), dim3(BLK), 0, C->stream, V.lev[top - 1].beta, n, j);
    hipLaunchKernelGGL(swm::dev::k_coarse_cycle<double>, dim3(1), dim3(swm::dev::CO_THREADS), 0, C->stream, P);
    HIP_TRY(hipMemcpyAsync(T->coarse_mt + (size_t)j * n, V.lev[top - 1].alpha, n * sizeof(double),
                           hipMemcpyDeviceToDevice, C->stream));
  }
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(C->stream));
  return SWM_OK;
}

// TL of the step's solve with the base step's schedule: rhs (gathered and
// mean-removed in place), x the warm start's perturbation and the result
void solve_sched_tl(swm_ctx* C, double* rhs, double* x, const swm_ctx::SolveSched& sc) {
  const int ns = C->ns;
  View<double>& V = C->d;
  for (int v = 0; v < 2; ++v) gather<double>(C, C->face, {x + (size_t)v * ns, rhs + (size_t)v * ns});
  sum2<double>(C, rhs, rhs + ns, C->area_mask, ns);
  hipLaunchKernelGGL(swm::dev::k_sub_mean<double>, zgrid(C, nblk(ns)), dim3(BLK), 0, C->stream, ns, C->sums,
                     C->total_area, rhs, rhs + ns, C->mb);
  for (int it = 0; it < sc.first; ++it) vcycle<double>(C, V, 2, x, rhs);
  for (int v = 0; v < 2; ++v) gather<double>(C, C->face, {x + (size_t)v * ns});
  if (sc.fixed) return;
  for (int it = 0; it < sc.extra; ++it) {
    if (sc.bad_mask == 3) vcycle<double>(C, V, 2, x, rhs);
    else {
      const int b = sc.bad_mask == 1 ? 0 : 1;
      vcycle<double>(C, V, 1, x + (size_t)b * ns, rhs + (size_t)b * ns);
    }
    for (int v = 0; v < 2; ++v)
      if (sc.bad_mask >> v & 1) gather<double>(C, C->face, {x + (size_t)v * ns});
  }
}

// transpose of vcycle<double> (the FP64 cycle: top_zero off, two finest
// sweeps) on nvar fields: xb the adjoint of the finest iterate (out: of its
// value on entry), bb += the adjoint of the finest right-hand side
template <int NV>
void vcycle_ad_nv(swm_ctx* C, double* xb, double* bb) {
  TlAd* T = C->tl;
  View<double>& V = C->d;
  const int G = C->g, top = swm::dev::CO_MAXLEV;
  auto Lv = [&](int l) -> const Lev& { return V.lev[l - 1].L; };
  auto Ab = [&](int l) { return l == G ? xb : T->mg_a[l]; };
  auto Bb = [&](int l) { return l == G ? bb : T->mg_b[l]; };
  auto Tb = [&](int l) { return T->mg_t[l]; };
  auto grid_of = [&](int l) { return dim3(nblk_b(V.lev[l - 1].ns, C->blk_mg)); };
  const dim3 blk(C->blk_mg);
  hipStream_t S = C->stream;
  // a sweep out = R(a, b) at level l, transposed: ob consumed, ab and bb added
  auto relax_ad = [&](int l, double* ob, double* ab, double* b, bool zero) {
    if (zero) hipLaunchKernelGGL((swm::dev::k_relax_ad<NV, true>), grid_of(l), blk, 0, S, Lv(l), ob, ab, b);
    else hipLaunchKernelGGL((swm::dev::k_relax_ad<NV, false>), grid_of(l), blk, 0, S, Lv(l), ob, ab, b);
  };
  const bool fuse = C->fuse_mg;
  // the way up, reversed: level l+1's two sweeps (alpha -> temp -> alpha),
  // then the interpolation from level l
  for (int l = G - 1; l >= top; --l) {
    relax_ad(l + 1, Ab(l + 1), Tb(l + 1), Bb(l + 1), false);
    relax_ad(l + 1, Tb(l + 1), Ab(l + 1), Bb(l + 1), false);
    hipLaunchKernelGGL((swm::dev::k_interp_ad<NV>), grid_of(l + 1), blk, 0, S, Lv(l), Lv(l + 1), Ab(l), Ab(l + 1));
  }
  // the coarse cycle from zero: bb(top) += M^T ab(top), whose iterate's
  // adjoint is then spent
  {
    const int n = T->coarse_n;
    hipLaunchKernelGGL(swm::dev::k_coarse_ad, dim3((unsigned)((n + 3) / 4), NV), dim3(256), 0, S, T->coarse_mt, n,
                       Ab(top), Bb(top));
    fill_plain<double>(C, Ab(top), (size_t)NV * V.lev[top - 1].ns, 0.0);
  }
  // the way down, reversed: the injection into level l-1 (with its fused
  // first sweep of level l-1 where that level runs its own sweeps), then
  // level l's sweeps
  for (int l = top + 1; l <= G; ++l) {
    double* cz = fuse && l - 1 > top ? Tb(l - 1) : nullptr;
    hipLaunchKernelGGL((swm::dev::k_inject_ad<NV>), grid_of(l - 1), blk, 0, S, Lv(l), Lv(l - 1), Ab(l), Bb(l),
                       Bb(l - 1), cz);
    if (l == G) {
      relax_ad(l, Ab(l), Tb(l), Bb(l), false);
      relax_ad(l, Tb(l), Ab(l), Bb(l), false);
    } else if (fuse) {
      // the second sweep from the injection's first one (in temp)
      relax_ad(l, Ab(l), Tb(l), Bb(l), false);
    } else {
      // two sweeps, the first from zero: alpha -> temp -> alpha
      relax_ad(l, Ab(l), Tb(l), Bb(l), false);
      relax_ad(l, Tb(l), Ab(l), Bb(l), true);
    }
  }
}
void vcycle_ad(swm_ctx* C, int nvar, double* xb, double* bb) {
  if (nvar == 2) vcycle_ad_nv<2>(C, xb, bb);
  else vcycle_ad_nv<1>(C, xb, bb);
}

// AD of solve_sched_tl up to (not including) the mean removal and the first
// halo fills: xb the adjoint of the solution after the solve's last halo fill
// (out: of the iterate before the cycles); bb = the adjoint of the
// mean-removed right-hand side
void solve_sched_ad(swm_ctx* C, double* xb, double* bb, const swm_ctx::SolveSched& sc) {
  const int ns = C->ns;
  fill<double>(C, bb, 2 * (size_t)ns, 0.0);
  if (!sc.fixed) {
    for (int it = sc.extra - 1; it >= 0; --it) {
      for (int v = 0; v < 2; ++v)
        if (sc.bad_mask >> v & 1) gather_adj(C, C->face, {xb + (size_t)v * ns});
      if (sc.bad_mask == 3) vcycle_ad(C, 2, xb, bb);
      else {
        const int b = sc.bad_mask == 1 ? 0 : 1;
        vcycle_ad(C, 1, xb + (size_t)b * ns, bb + (size_t)b * ns);
      }
    }
  }
  gather_adj(C, C->face, {xb, xb + ns});
  for (int it = 0; it < sc.first; ++it) vcycle_ad(C, 2, xb, bb);
}

// ------------------------------------------------------- tracer transport
// TL of the step's tracer transport (after the new thickness): FCT with the
// base step's frozen tracer limiter (tracer_fct.F), or the centred /
// upstream tracer fluxes and step_tracer (step.F:278-327, 495-602); bF is
// the base mass flux the tracers take (FCT: the recorded total flux)
void tl_tracers(swm_ctx* C, int q, const double* bF) {
  TlAd* T = C->tl;
  const size_t ns = C->ns;
  const unsigned nb = nblk(ns);
  hipStream_t S = C->stream;
  const swm::dev::Geo& G = C->d.geo;
  const int n3 = C->n3 - 1, n4 = C->n4 - 1;
  const int a = C->n3f - 1, b = C->n2f - 1, c = C->n1f - 1;
  const double dt = C->dt, w3 = C->w3, w2 = C->w2, w1 = C->w1;
  for (int tq = 0; tq < C->nt; ++tq) {
    const double* bt = seg_tr(C, q, tq);
    const double* bt4n = seg_tr(C, q + 1, tq);
    if (C->advection == SWM_ADV_FCT) {
      hipLaunchKernelGGL(swm::dev::k_tl_tr_fluxes, dim3(nb), dim3(BLK), 0, S, G, w3, w2, w1, bF, bt, bt + ns,
                         bt + 2 * ns, T->F, T->tr[n3][tq], T->tr[n4][tq], T->tro[tq], T->tf1, T->tf2);
      gather<double>(C, C->edge, {T->tf1, T->tf2});
      hipLaunchKernelGGL(swm::dev::k_tl_tr_low, dim3(nb), dim3(BLK), 0, S, G, dt, seg_h(C, q), bt, T->h[n3],
                         T->tr[n3][tq], T->tf1, T->tml);
      hipLaunchKernelGGL(swm::dev::k_tl_tr_anti, dim3(nblk(3 * ns)), dim3(BLK), 0, S, 3 * ns,
                         (const double*)(bt + 3 * ns), T->tf1, T->tf2, T->tan);
      gather<double>(C, C->edge, {T->tan});
      hipLaunchKernelGGL(swm::dev::k_tl_tr_final, dim3(nb), dim3(BLK), 0, S, G, dt, seg_h(C, q + 1), bt4n, T->tml,
                         T->tan, T->h[n4], T->tr[n4][tq], T->tro[tq]);
    } else {
      if (C->advection == SWM_ADV_CENTERED)
        hipLaunchKernelGGL(swm::dev::k_tl_tr_flux<true>, dim3(nb), dim3(BLK), 0, S, G, bF, bt, T->F, T->tr[n3][tq],
                           T->tf1);
      else
        hipLaunchKernelGGL(swm::dev::k_tl_tr_flux<false>, dim3(nb), dim3(BLK), 0, S, G, bF, bt, T->F, T->tr[n3][tq],
                           T->tf1);
      gather<double>(C, C->edge, {T->tf1});
      hipLaunchKernelGGL(swm::dev::k_tl_tracer_update, dim3(nb), dim3(BLK), 0, S, G, dt, w3, w2, w1, T->tf1, bt,
                         seg_h(C, q), seg_h(C, q + 1), bt4n, T->tr[n3][tq], T->h[n3], T->h[n4], T->trf[a][tq],
                         T->trf[b][tq], T->trf[c][tq], T->tr[n4][tq]);
    }
  }
}

// transpose of tl_tracers: adds into the mass flux's, the thicknesses' and
// the tracer levels' adjoints
void ad_tracers(swm_ctx* C, int q, const double* bF) {
  TlAd* T = C->tl;
  const size_t ns = C->ns;
  const unsigned nb = nblk(ns);
  hipStream_t S = C->stream;
  const swm::dev::Geo& G = C->d.geo;
  const int n3 = C->n3 - 1, n4 = C->n4 - 1;
  const int a = C->n3f - 1, b = C->n2f - 1, c = C->n1f - 1;
  const double dt = C->dt, w3 = C->w3, w2 = C->w2, w1 = C->w1;
  for (int tq = 0; tq < C->nt; ++tq) {
    const double* bt = seg_tr(C, q, tq);
    const double* bt4n = seg_tr(C, q + 1, tq);
    if (C->advection == SWM_ADV_FCT) {
      fill<double>(C, T->tml, ns, 0.0);
      fill<double>(C, T->tan, 3 * ns, 0.0);
      fill<double>(C, T->tf1, 3 * ns, 0.0);
      fill<double>(C, T->tf2, 3 * ns, 0.0);
      hipLaunchKernelGGL(swm::dev::k_ad_tr_final, dim3(nb), dim3(BLK), 0, S, G, dt, seg_h(C, q + 1), bt4n,
                         T->tr[n4][tq], T->tro[tq], T->tml, T->tan, T->h[n4]);
      gather_adj(C, C->edge, {T->tan});
      hipLaunchKernelGGL(swm::dev::k_ad_tr_anti, dim3(nblk(3 * ns)), dim3(BLK), 0, S, 3 * ns,
                         (const double*)(bt + 3 * ns), T->tan, T->tf1, T->tf2);
      hipLaunchKernelGGL(swm::dev::k_ad_tr_low, dim3(nb), dim3(BLK), 0, S, G, dt, seg_h(C, q), bt, T->tml, T->h[n3],
                         T->tr[n3][tq], T->tf1);
      gather_adj(C, C->edge, {T->tf1, T->tf2});
      hipLaunchKernelGGL(swm::dev::k_ad_tr_fluxes, dim3(nb), dim3(BLK), 0, S, G, w3, w2, w1, bF, bt, bt + ns,
                         bt + 2 * ns, T->tf1, T->tf2, T->F, T->tr[n3][tq], T->tr[n4][tq], T->tro[tq]);
    } else {
      fill<double>(C, T->tf1, 3 * ns, 0.0);
      hipLaunchKernelGGL(swm::dev::k_ad_tracer_update, dim3(nb), dim3(BLK), 0, S, G, dt, w3, w2, w1, bt, seg_h(C, q),
                         seg_h(C, q + 1), bt4n, T->tr[n4][tq], T->trf[a][tq], T->trf[b][tq], T->trf[c][tq],
                         T->tr[n3][tq], T->h[n3], T->h[n4], T->tf1);
      gather_adj(C, C->edge, {T->tf1});
      if (C->advection == SWM_ADV_CENTERED)
        hipLaunchKernelGGL(swm::dev::k_ad_tr_flux<true>, dim3(nb), dim3(BLK), 0, S, G, bF, bt, T->tf1, T->F,
                           T->tr[n3][tq]);
      else
        hipLaunchKernelGGL(swm::dev::k_ad_tr_flux<false>, dim3(nb), dim3(BLK), 0, S, G, bF, bt, T->tf1, T->F,
                           T->tr[n3][tq]);
    }
  }
}

// ---------------------------------------------------------------- TL step
// window step with base state q -> q+1; pointers / weights of the model step
swm_status tl_step(swm_ctx* C, int q) {
  TlAd* T = C->tl;
  const size_t ns = C->ns;
  const unsigned nb = nblk(ns);
  hipStream_t S = C->stream;
  const swm::dev::Geo& G = C->d.geo;
  const int n3 = C->n3 - 1, n4 = C->n4 - 1;
  const int a = C->n3f - 1, b = C->n2f - 1, c = C->n1f - 1;
  const double dt = C->dt, w3 = C->w3, w2 = C->w2, w1 = C->w1;
  const double* bF = base_coeffs(C, q);
  double *dh3 = T->h[n3], *de3 = T->eta[n3], *du3 = T->u[n3];
  // kinetic energy of u3 (computed at the end of the previous step)
  hipLaunchKernelGGL(swm::dev::k_tl_center_ke, dim3(nb), dim3(BLK), 0, S, G, seg_u(C, q), du3, T->ke);
  gather<double>(C, C->face, {T->ke});
  hipLaunchKernelGGL(swm::dev::k_avg_corners<double>, dim3(nb), dim3(BLK), 0, S, G, G.sw, dh3, T->hc);
  gather<double>(C, C->vrtx, {T->hc});
  if (C->advection == SWM_ADV_FCT) {
    // mass FCT with the base step's limiter coefficients: upstream and
    // centred fluxes, antidiffusive flux and low-order mass, the limited
    // flux, then the total flux and the new thickness
    hipLaunchKernelGGL(swm::dev::k_tl_upwind_flux, dim3(nb), dim3(BLK), 0, S, G, seg_u(C, q), seg_h(C, q), du3, dh3, T->fl);
    hipLaunchKernelGGL(swm::dev::k_tl_mass_flux, dim3(nb), dim3(BLK), 0, S, G, seg_u(C, q), T->bhc, du3, T->hc, T->mf[a]);
    gather<double>(C, C->edge, {T->fl, T->mf[a]});
    hipLaunchKernelGGL(swm::dev::k_mass_low_beta<double>, dim3(nb), dim3(BLK), 0, S, G, dt, w3, w2, w1, T->fl, T->mf[a],
                       T->mf[b], T->mf[c], dh3, T->an, T->ml, T->bsc);
    hipLaunchKernelGGL(swm::dev::k_fct_scale, dim3(nblk(3 * ns)), dim3(BLK), 0, S, 3 * ns, seg_c(C, q), T->an);
    gather<double>(C, C->edge, {T->an});
    hipLaunchKernelGGL(swm::dev::k_mass_final<double>, dim3(nb), dim3(BLK), 0, S, G, dt, T->fl, T->an, T->ml, T->F,
                       T->h[n4], swm::dev::TrFluxArgs<double>{}, swm::dev::VortFluxArgs<double>{},
                     (const double*)nullptr);
    gather<double>(C, C->face, {T->h[n4]});
    if (T->tracers) tl_tracers(C, q, bF);
  } else {
    if (C->advection == SWM_ADV_CENTERED)
      hipLaunchKernelGGL(swm::dev::k_tl_mass_flux, dim3(nb), dim3(BLK), 0, S, G, seg_u(C, q), T->bhc, du3, T->hc, T->F);
    else
      hipLaunchKernelGGL(swm::dev::k_tl_upwind_flux, dim3(nb), dim3(BLK), 0, S, G, seg_u(C, q), seg_h(C, q), du3, dh3,
                         T->F);
    gather<double>(C, C->edge, {T->F});
    hipLaunchKernelGGL(swm::dev::k_tl_height, dim3(nb), dim3(BLK), 0, S, G, dt, w3, w2, w1, T->F, dh3, T->hf[a],
                       T->hf[b], T->hf[c], seg_h(C, q + 1), T->h[n4]);
    if (T->tracers) tl_tracers(C, q, bF);
  }
  hipLaunchKernelGGL(swm::dev::k_tl_vort_curl, dim3(nb), dim3(BLK), 0, S, G, bF, seg_h(C, q), seg_e(C, q), T->bhc,
                     seg_u(C, q), T->F, dh3, de3, T->hc, du3, T->vf, T->dc);
  gather<double>(C, C->edge, {T->vf, T->dc});
  hipLaunchKernelGGL(swm::dev::k_tendencies<double>, dim3(nb), dim3(BLK), 0, S, T->lin, dt, C->grav, w3, w2, w1, T->vf,
                     T->dc, T->ke, dh3, de3, T->div[n3], T->ef[a], T->ef[b], T->ef[c], T->df[a], T->df[b], T->df[c],
                     T->eta[n4], T->div[n4], T->X, T->rhs, T->rhs + ns, (const double*)nullptr, (double*)nullptr);
  if (exact_lin(C)) {
    solve_sched_tl(C, T->rhs, T->xs, T->sched[q]);   // xs: the warm start's perturbation, carried
  } else {
    swm_status st = lin_solve(C, T->rhs, T->xs);
    if (st) return st;
  }
  hipLaunchKernelGGL(swm::dev::k_post_elliptic<double>, dim3(nb), dim3(BLK), 0, S, T->lin, T->rhs, T->eta[n4]);
  hipLaunchKernelGGL(swm::dev::k_velocity<double>, dim3(nb), dim3(BLK), 0, S, G, T->xs, T->xs + ns, T->u[n4], (const double*)nullptr, (const double*)nullptr, 0.0, (double*)nullptr, (const double*)nullptr);
  gather<double>(C, C->vrtx, {T->u[n4], T->u[n4] + 2 * ns});
  // wrap_prognostics
  gather<double>(C, C->face, {T->h[n4], T->eta[n4], T->div[n4]});
  gather<double>(C, C->vrtx, {T->u[0], T->u[0] + 2 * ns, T->u[1], T->u[1] + 2 * ns});
  if (T->tracers) {
    std::vector<double*> f;
    for (int tq = 0; tq < C->nt; ++tq) f.push_back(T->tr[n4][tq]);
    gather_list<double>(C, C->face, f);
  }
  return SWM_OK;
}

// ---------------------------------------------------------------- AD step
// transpose of tl_step, same pointers / weights
swm_status ad_step(swm_ctx* C, int q) {
  TlAd* T = C->tl;
  const size_t ns = C->ns;
  const unsigned nb = nblk(ns);
  hipStream_t S = C->stream;
  const swm::dev::Geo& G = C->d.geo;
  const int n3 = C->n3 - 1, n4 = C->n4 - 1;
  const int a = C->n3f - 1, b = C->n2f - 1, c = C->n1f - 1;
  const double dt = C->dt, w3 = C->w3, w2 = C->w2, w1 = C->w1;
  const double* bF = base_coeffs(C, q);
  // wrap_prognostics
  if (T->tracers)
    for (int tq = 0; tq < C->nt; ++tq) gather_adj(C, C->face, {T->tr[n4][tq]});
  gather_adj(C, C->vrtx, {T->u[0], T->u[0] + 2 * ns, T->u[1], T->u[1] + 2 * ns});
  gather_adj(C, C->face, {T->h[n4], T->eta[n4], T->div[n4]});
  // u4 = W velocity(psi, chi)
  gather_adj(C, C->vrtx, {T->u[n4], T->u[n4] + 2 * ns});
  fill<double>(C, T->psia, 2 * ns, 0.0);
  hipLaunchKernelGGL(swm::dev::k_ad_velocity, dim3(nb), dim3(BLK), 0, S, G, T->u[n4], T->psia, T->psia + ns);
  if (exact_lin(C)) {
    // the solution's adjoint: the velocity's and (xs, carried) the next
    // step's warm start's; through the solve's cycles, transposed
    hipLaunchKernelGGL(swm::dev::k_axpy, dim3(nblk(2 * ns)), dim3(BLK), 0, S, 2 * ns, 1.0, T->psia, T->xs);
    solve_sched_ad(C, T->xs, T->rhs, T->sched[q]);
    hipLaunchKernelGGL(swm::dev::k_axpy, dim3(nb), dim3(BLK), 0, S, ns, 1.0, T->eta[n4], T->rhs);
    fill<double>(C, T->eta[n4], ns, 0.0);
    mean_removal_adj(C, T->rhs);
    mean_removal_adj(C, T->rhs + ns);
    gather_adj(C, C->face, {T->rhs, T->rhs + ns});
    gather_adj(C, C->face, {T->xs, T->xs + ns});   // xs: the warm start's adjoint, to the previous step
  } else {
  // the solve's final halo fill, then psi_a / A, solved: rhs_a = A z
  gather_adj(C, C->face, {T->psia, T->psia + ns});
  hipLaunchKernelGGL(swm::dev::k_mul_mask, dim3(nb), dim3(BLK), 0, S, (int)ns, T->psia, T->am_inv, C->mask, T->psia);
  hipLaunchKernelGGL(swm::dev::k_mul_mask, dim3(nb), dim3(BLK), 0, S, (int)ns, T->psia + ns, T->am_inv, C->mask,
                     T->psia + ns);
  swm_status st = lin_solve(C, T->psia, T->xs);
  if (st) return st;
  // rhs_a = A z (real cells) + the adjoint of eta4 = rhs1 (+ f)
  hipLaunchKernelGGL(swm::dev::k_mul_mask, dim3(nb), dim3(BLK), 0, S, (int)ns, T->xs, C->area, C->mask, T->rhs);
  hipLaunchKernelGGL(swm::dev::k_mul_mask, dim3(nb), dim3(BLK), 0, S, (int)ns, T->xs + ns, C->area, C->mask,
                     T->rhs + ns);
  hipLaunchKernelGGL(swm::dev::k_axpy, dim3(nb), dim3(BLK), 0, S, ns, 1.0, T->eta[n4], T->rhs);
  fill<double>(C, T->eta[n4], ns, 0.0);
  // the solve's mean removal and its first halo fill of rhs
  mean_removal_adj(C, T->rhs);
  mean_removal_adj(C, T->rhs + ns);
  gather_adj(C, C->face, {T->rhs, T->rhs + ns});
  }
  // tendencies: rhs2 = d4 joins the new level's adjoint
  hipLaunchKernelGGL(swm::dev::k_axpy, dim3(nb), dim3(BLK), 0, S, ns, 1.0, T->rhs + ns, T->div[n4]);
  fill<double>(C, T->vf, 3 * ns, 0.0);
  fill<double>(C, T->dc, 3 * ns, 0.0);
  fill<double>(C, T->X, ns, 0.0);
  hipLaunchKernelGGL(swm::dev::k_ad_tendencies, dim3(nb), dim3(BLK), 0, S, G, dt, w3, w2, w1, T->rhs, T->div[n4],
                     T->ef[a], T->ef[b], T->ef[c], T->df[a], T->df[b], T->df[c], T->eta[n3], T->div[n3], T->vf, T->dc,
                     T->X);
  fill<double>(C, T->ke, ns, 0.0);
  hipLaunchKernelGGL(swm::dev::k_ad_fold_x, dim3(nb), dim3(BLK), 0, S, (int)ns, C->grav, T->X, T->ke, T->h[n3]);
  // vorticity flux and divergence curl
  gather_adj(C, C->edge, {T->vf, T->dc});
  fill<double>(C, T->F, 3 * ns, 0.0);
  fill<double>(C, T->hc, 2 * ns, 0.0);
  hipLaunchKernelGGL(swm::dev::k_ad_vort_curl, dim3(nb), dim3(BLK), 0, S, G, bF, seg_h(C, q), seg_e(C, q), T->bhc,
                     seg_u(C, q), T->vf, T->dc, T->F, T->h[n3], T->eta[n3], T->hc, T->u[n3]);
  // the tracers read the new thickness and the mass flux
  if (T->tracers) ad_tracers(C, q, bF);
  if (C->advection == SWM_ADV_FCT) {
    // transpose of the mass FCT with the frozen limiter (tl_step)
    gather_adj(C, C->face, {T->h[n4]});
    fill<double>(C, T->fl, 3 * ns, 0.0);
    fill<double>(C, T->an, 3 * ns, 0.0);
    fill<double>(C, T->ml, ns, 0.0);
    hipLaunchKernelGGL(swm::dev::k_ad_fct_final, dim3(nb), dim3(BLK), 0, S, G, dt, T->F, T->h[n4], T->fl, T->an, T->ml);
    gather_adj(C, C->edge, {T->an});
    hipLaunchKernelGGL(swm::dev::k_fct_scale, dim3(nblk(3 * ns)), dim3(BLK), 0, S, 3 * ns, seg_c(C, q), T->an);
    hipLaunchKernelGGL(swm::dev::k_ad_fct_low, dim3(nb), dim3(BLK), 0, S, G, dt, w3, w2, w1, T->an, T->ml, T->mf[a],
                       T->mf[b], T->mf[c], T->fl, T->h[n3]);
    gather_adj(C, C->edge, {T->fl, T->mf[a]});
    hipLaunchKernelGGL(swm::dev::k_ad_upwind_flux, dim3(nb), dim3(BLK), 0, S, G, seg_u(C, q), seg_h(C, q), T->fl,
                       T->u[n3], T->h[n3]);
    hipLaunchKernelGGL(swm::dev::k_ad_mass_flux, dim3(nb), dim3(BLK), 0, S, G, seg_u(C, q), T->bhc, T->mf[a], T->u[n3],
                       T->hc);
    // the slot's value before this step is read by no later step
    fill<double>(C, T->mf[a], 3 * ns, 0.0);
  } else {
    // height update
    hipLaunchKernelGGL(swm::dev::k_ad_height, dim3(nb), dim3(BLK), 0, S, G, dt, w3, w2, w1, T->h[n4], T->hf[a],
                       T->hf[b], T->hf[c], seg_h(C, q + 1), T->h[n3], T->F);
    // mass flux
    gather_adj(C, C->edge, {T->F});
    if (C->advection == SWM_ADV_CENTERED)
      hipLaunchKernelGGL(swm::dev::k_ad_mass_flux, dim3(nb), dim3(BLK), 0, S, G, seg_u(C, q), T->bhc, T->F, T->u[n3],
                         T->hc);
    else
      hipLaunchKernelGGL(swm::dev::k_ad_upwind_flux, dim3(nb), dim3(BLK), 0, S, G, seg_u(C, q), seg_h(C, q), T->F,
                         T->u[n3], T->h[n3]);
  }
  // corner thickness
  gather_adj(C, C->vrtx, {T->hc});
  hipLaunchKernelGGL(swm::dev::k_ad_avg_corners, dim3(nb), dim3(BLK), 0, S, G, T->hc, T->h[n3]);
  // kinetic energy of u3
  gather_adj(C, C->face, {T->ke});
  hipLaunchKernelGGL(swm::dev::k_ad_center_ke, dim3(nb), dim3(BLK), 0, S, G, seg_u(C, q), T->ke, T->u[n3]);
  return SWM_OK;
}

// pointers / weights of window step n (1-based) of a window started at
// step_count 0
void set_step(swm_ctx* C, int n) {
  C->step_count = n;
  pointers(C);
  weights(C);
}

// ------------------------------------------------------ initial state map
// control (zeta, delta, h) -> state: (psi, chi) = L^-1 P (zeta, delta),
// u = W velocity(psi, chi), the model's initialisation from h and u
// (finish_init): both levels, eta = f + W curl(u), div = W div(u)
// dtr: the initial tracers' perturbation (device planes, tracer runs; both
// levels, as finish_init copies them)
swm_status init_tl(swm_ctx* C, const double* dz, const double* dd, const double* dh, const double* dtr) {
  TlAd* T = C->tl;
  const size_t ns = C->ns;
  const unsigned nb = nblk(ns);
  hipStream_t S = C->stream;
  const swm::dev::Geo& G = C->d.geo;
  HIP_TRY(hipMemcpyAsync(T->rhs, dz, ns * sizeof(double), hipMemcpyDeviceToDevice, S));
  HIP_TRY(hipMemcpyAsync(T->rhs + ns, dd, ns * sizeof(double), hipMemcpyDeviceToDevice, S));
  fill<double>(C, T->xs, 2 * ns, 0.0);
  swm_status st = lin_solve(C, T->rhs, T->xs);
  if (st) return st;
  clear_state(C);
  hipLaunchKernelGGL(swm::dev::k_velocity<double>, dim3(nb), dim3(BLK), 0, S, G, T->xs, T->xs + ns, T->u[0], (const double*)nullptr, (const double*)nullptr, 0.0, (double*)nullptr, (const double*)nullptr);
  HIP_TRY(hipMemcpyAsync(T->h[0], dh, ns * sizeof(double), hipMemcpyDeviceToDevice, S));
  // finish_init: both levels, wrapped; curl / div of level 0 (wrapped again)
  HIP_TRY(hipMemcpyAsync(T->h[1], T->h[0], ns * sizeof(double), hipMemcpyDeviceToDevice, S));
  HIP_TRY(hipMemcpyAsync(T->u[1], T->u[0], 4 * ns * sizeof(double), hipMemcpyDeviceToDevice, S));
  for (int t = 0; t < 2; ++t) {
    gather<double>(C, C->face, {T->h[t]});
    gather<double>(C, C->vrtx, {T->u[t], T->u[t] + 2 * ns});
  }
  gather<double>(C, C->vrtx, {T->u[0], T->u[0] + 2 * ns});
  hipLaunchKernelGGL(swm::dev::k_curl_div, dim3(nb), dim3(BLK), 0, S, G, T->u[0], T->eta[0], T->div[0]);
  gather<double>(C, C->face, {T->eta[0], T->div[0]});
  HIP_TRY(hipMemcpyAsync(T->div[1], T->div[0], ns * sizeof(double), hipMemcpyDeviceToDevice, S));
  HIP_TRY(hipMemcpyAsync(T->eta[1], T->eta[0], ns * sizeof(double), hipMemcpyDeviceToDevice, S));
  // wrap_prognostics of the initial state (level 1 is the new level after
  // initialisation: step_count 0)
  gather<double>(C, C->face, {T->h[1], T->eta[1], T->div[1]});
  gather<double>(C, C->vrtx, {T->u[0], T->u[0] + 2 * ns, T->u[1], T->u[1] + 2 * ns});
  if (T->tracers && dtr)
    for (int tq = 0; tq < C->nt; ++tq) {
      for (int l = 0; l < 2; ++l)
        HIP_TRY(hipMemcpyAsync(T->tr[l][tq], dtr + (size_t)tq * ns, ns * sizeof(double), hipMemcpyDeviceToDevice, S));
      gather<double>(C, C->face, {T->tr[1][tq]});
    }
  return SWM_OK;
}

// transpose of init_tl into gz, gd, gh and (tracer runs) gtr (device planes)
swm_status init_ad(swm_ctx* C, double* gz, double* gd, double* gh, double* gtr) {
  TlAd* T = C->tl;
  const size_t ns = C->ns;
  const unsigned nb = nblk(ns);
  hipStream_t S = C->stream;
  const swm::dev::Geo& G = C->d.geo;
  if (T->tracers && gtr)
    for (int tq = 0; tq < C->nt; ++tq) {
      gather_adj(C, C->face, {T->tr[1][tq]});
      double* g = gtr + (size_t)tq * ns;
      HIP_TRY(hipMemcpyAsync(g, T->tr[0][tq], ns * sizeof(double), hipMemcpyDeviceToDevice, S));
      hipLaunchKernelGGL(swm::dev::k_axpy, dim3(nb), dim3(BLK), 0, S, ns, 1.0, T->tr[1][tq], g);
    }
  gather_adj(C, C->vrtx, {T->u[0], T->u[0] + 2 * ns, T->u[1], T->u[1] + 2 * ns});
  gather_adj(C, C->face, {T->h[1], T->eta[1], T->div[1]});
  hipLaunchKernelGGL(swm::dev::k_axpy, dim3(nb), dim3(BLK), 0, S, ns, 1.0, T->eta[1], T->eta[0]);
  hipLaunchKernelGGL(swm::dev::k_axpy, dim3(nb), dim3(BLK), 0, S, ns, 1.0, T->div[1], T->div[0]);
  gather_adj(C, C->face, {T->eta[0], T->div[0]});
  hipLaunchKernelGGL(swm::dev::k_ad_curl_div, dim3(nb), dim3(BLK), 0, S, G, T->eta[0], T->div[0], T->u[0]);
  gather_adj(C, C->vrtx, {T->u[0], T->u[0] + 2 * ns});
  for (int t = 1; t >= 0; --t) {
    gather_adj(C, C->vrtx, {T->u[t], T->u[t] + 2 * ns});
    gather_adj(C, C->face, {T->h[t]});
  }
  hipLaunchKernelGGL(swm::dev::k_axpy, dim3(nblk(4 * ns)), dim3(BLK), 0, S, 4 * ns, 1.0, T->u[1], T->u[0]);
  HIP_TRY(hipMemcpyAsync(gh, T->h[0], ns * sizeof(double), hipMemcpyDeviceToDevice, S));
  hipLaunchKernelGGL(swm::dev::k_axpy, dim3(nb), dim3(BLK), 0, S, ns, 1.0, T->h[1], gh);
  // u = velocity(psi, chi)
  fill<double>(C, T->psia, 2 * ns, 0.0);
  hipLaunchKernelGGL(swm::dev::k_ad_velocity, dim3(nb), dim3(BLK), 0, S, G, T->u[0], T->psia, T->psia + ns);
  // the solution is also the first step's warm start (exact linearisation)
  if (exact_lin(C))
    hipLaunchKernelGGL(swm::dev::k_axpy, dim3(nblk(2 * ns)), dim3(BLK), 0, S, 2 * ns, 1.0, T->xs, T->psia);
  gather_adj(C, C->face, {T->psia, T->psia + ns});
  hipLaunchKernelGGL(swm::dev::k_mul_mask, dim3(nb), dim3(BLK), 0, S, (int)ns, T->psia, T->am_inv, C->mask, T->psia);
  hipLaunchKernelGGL(swm::dev::k_mul_mask, dim3(nb), dim3(BLK), 0, S, (int)ns, T->psia + ns, T->am_inv, C->mask,
                     T->psia + ns);
  fill<double>(C, T->xs, 2 * ns, 0.0);
  swm_status st = lin_solve(C, T->psia, T->xs);
  if (st) return st;
  hipLaunchKernelGGL(swm::dev::k_mul_mask, dim3(nb), dim3(BLK), 0, S, (int)ns, T->xs, C->area, C->mask, gz);
  hipLaunchKernelGGL(swm::dev::k_mul_mask, dim3(nb), dim3(BLK), 0, S, (int)ns, T->xs + ns, C->area, C->mask, gd);
  mean_removal_adj(C, gz);
  mean_removal_adj(C, gd);
  gather_adj(C, C->face, {gz, gd});
  clear_state(C);
  return SWM_OK;
}

// nonlinear initial state from the control (the base of init_tl); tr the
// initial tracers (device planes, or null: zero)
swm_status init_from_vorticity(swm_ctx* C, const double* z, const double* d, const double* h, const double* tr) {
  TlAd* T = C->tl;
  const size_t ns = C->ns;
  hipStream_t S = C->stream;
  reset_state(C, true);
  if (tr)
    for (int tq = 0; tq < C->nt; ++tq)
      HIP_TRY(hipMemcpyAsync(C->d.tr[0][tq], tr + (size_t)tq * ns, ns * sizeof(double), hipMemcpyDeviceToDevice, S));
  HIP_TRY(hipMemcpyAsync(T->rhs, z, ns * sizeof(double), hipMemcpyDeviceToDevice, S));
  HIP_TRY(hipMemcpyAsync(T->rhs + ns, d, ns * sizeof(double), hipMemcpyDeviceToDevice, S));
  fill<double>(C, T->xs, 2 * ns, 0.0);
  swm_status st = lin_solve(C, T->rhs, T->xs);
  if (st) return st;
  hipLaunchKernelGGL(swm::dev::k_velocity<double>, dim3(nblk(ns)), dim3(BLK), 0, S, C->d.geo, T->xs, T->xs + ns,
                     C->d.u[0], (const double*)nullptr, (const double*)nullptr, 0.0, (double*)nullptr, (const double*)nullptr);
  HIP_TRY(hipMemcpyAsync(C->d.h[0], h, ns * sizeof(double), hipMemcpyDeviceToDevice, S));
  // the model's first solve starts from this streamfunction / velocity potential
  HIP_TRY(hipMemcpyAsync(C->xout, T->xs, 2 * ns * sizeof(double), hipMemcpyDeviceToDevice, S));
  return finish_init(C);
}

// host (reference layout) <-> device planes of the TL / AD state
// The copies in are synchronous host copies, which do not wait for work on
// the context's (non-blocking) stream: the stream is drained first, so that
// clear_state's fills, queued just before, cannot land on the inputs.
swm_status state_in(swm_ctx* C, int lev, const double* h, const double* e, const double* d, const double* u) {
  TlAd* T = C->tl;
  const size_t ns = C->ns;
  HIP_TRY(hipStreamSynchronize(C->stream));
  std::vector<double> up(4 * ns);
  u_to_planes(u, up.data(), (int)ns);
  HIP_TRY(hipMemcpy(T->h[lev], h, ns * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(T->eta[lev], e, ns * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(T->div[lev], d, ns * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(T->u[lev], up.data(), 4 * ns * sizeof(double), hipMemcpyHostToDevice));
  return SWM_OK;
}
swm_status state_out(swm_ctx* C, int lev, double* h, double* e, double* d, double* u) {
  TlAd* T = C->tl;
  const size_t ns = C->ns;
  HIP_TRY(hipStreamSynchronize(C->stream));
  std::vector<double> up(4 * ns);
  HIP_TRY(hipMemcpy(h, T->h[lev], ns * sizeof(double), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(e, T->eta[lev], ns * sizeof(double), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(d, T->div[lev], ns * sizeof(double), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(up.data(), T->u[lev], 4 * ns * sizeof(double), hipMemcpyDeviceToHost));
  u_from_planes(up.data(), u, (int)ns);
  return SWM_OK;
}

// tracers (iim, jjm, ntracer, 10) <-> planes (one per tracer)
void tracer_to_planes(const swm_ctx* C, const double* ref, double* planes) {
  const size_t pan = (size_t)C->iim * C->jjm, ns = C->ns;
  for (int q = 0; q < C->nt; ++q)
    for (int k = 0; k < 10; ++k)
      std::memcpy(planes + q * ns + k * pan, ref + ((size_t)q + (size_t)C->nt * k) * pan, pan * sizeof(double));
}
void tracer_from_planes(const swm_ctx* C, const double* planes, double* ref) {
  const size_t pan = (size_t)C->iim * C->jjm, ns = C->ns;
  for (int q = 0; q < C->nt; ++q)
    for (int k = 0; k < 10; ++k)
      std::memcpy(ref + ((size_t)q + (size_t)C->nt * k) * pan, planes + q * ns + k * pan, pan * sizeof(double));
}
swm_status tracer_in(swm_ctx* C, int lev, const double* tr) {   // (drains the stream as state_in)
  TlAd* T = C->tl;
  const size_t ns = C->ns;
  HIP_TRY(hipStreamSynchronize(C->stream));
  std::vector<double> tp((size_t)C->nt * ns);
  tracer_to_planes(C, tr, tp.data());
  for (int tq = 0; tq < C->nt; ++tq)
    HIP_TRY(hipMemcpy(T->tr[lev][tq], tp.data() + tq * ns, ns * sizeof(double), hipMemcpyHostToDevice));
  return SWM_OK;
}
swm_status tracer_out(swm_ctx* C, int lev, double* tr) {
  TlAd* T = C->tl;
  const size_t ns = C->ns;
  HIP_TRY(hipStreamSynchronize(C->stream));
  std::vector<double> tp((size_t)C->nt * ns);
  for (int tq = 0; tq < C->nt; ++tq)
    HIP_TRY(hipMemcpy(tp.data() + tq * ns, T->tr[lev][tq], ns * sizeof(double), hipMemcpyDeviceToHost));
  tracer_from_planes(C, tp.data(), tr);
  return SWM_OK;
}

swm_status window_check(swm_ctx* C, const char* who) {
  if (C->nmem != 1) {
    swm::set_error(std::string(who) + ": the linearised model runs on single-member contexts");
    return SWM_ERR_ARG;
  }
  if (C->precision != SWM_FP64) {
    swm::set_error(std::string(who) + ": the linearised model is the FP64 step");
    return SWM_ERR_ARG;
  }
  if (C->advection == SWM_ADV_REMAP) {
    swm::set_error(std::string(who) + ": the linearised model has no incremental remapping");
    return SWM_ERR_ARG;
  }
  return SWM_OK;
}

// level read by window step 1 (step_count 1) and written by the last step
int first_level() { return 1 % 2 + 1 - 1; }
int last_level(int nsteps) { return (nsteps + 1) % 2 + 1 - 1; }

}  // namespace
}  // namespace swm

extern "C" {

swm_status swm_init_vorticity_tracer(swm_ctx* C, const double* zeta, const double* delta, const double* h,
                                     const double* tracer) {
  if (!C || !zeta || !delta || !h || C->nmem != 1) {
    swm::set_error("swm_init_vorticity: bad argument");
    return SWM_ERR_ARG;
  }
  HIP_TRY(hipSetDevice(C->device));
  swm_status st;
  if ((st = swm::tlad_alloc(C))) return st;
  const size_t ns = C->ns;
  const size_t nt = tracer ? (size_t)C->nt : 0;
  double* buf = nullptr;
  HIP_TRY(hipMalloc(&buf, (3 + nt) * ns * sizeof(double)));
  HIP_TRY(hipMemcpy(buf, zeta, ns * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(buf + ns, delta, ns * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(buf + 2 * ns, h, ns * sizeof(double), hipMemcpyHostToDevice));
  if (nt) {
    std::vector<double> tp(nt * ns);
    swm::tracer_to_planes(C, tracer, tp.data());
    HIP_TRY(hipMemcpy(buf + 3 * ns, tp.data(), nt * ns * sizeof(double), hipMemcpyHostToDevice));
  }
  st = swm::init_from_vorticity(C, buf, buf + ns, buf + 2 * ns, nt ? buf + 3 * ns : nullptr);
  (void)hipStreamSynchronize(C->stream);
  (void)hipFree(buf);
  return st;
}

swm_status swm_init_vorticity(swm_ctx* C, const double* zeta, const double* delta, const double* h) {
  return swm_init_vorticity_tracer(C, zeta, delta, h, nullptr);
}

swm_status swm_adj_forward(swm_ctx* C, int nsteps, int every) {
  if (!C || nsteps < 1 || every < 1) { swm::set_error("swm_adj_forward: bad argument"); return SWM_ERR_ARG; }
  swm_status st;
  if ((st = swm::window_check(C, "swm_adj_forward"))) return st;
  if (!C->initialized || C->step_count != 0) {
    swm::set_error("swm_adj_forward: the window starts from a freshly initialised state (step count 0)");
    return SWM_ERR_STATE;
  }
  HIP_TRY(hipSetDevice(C->device));
  if ((st = swm::tlad_alloc(C))) return st;
  TlAd* T = C->tl;
  for (double* p : T->ckpt) (void)hipFree(p);
  T->ckpt.clear();
  T->ckpt_meta.clear();
  if (T->seg) { (void)hipFree(T->seg); T->seg = nullptr; }
  T->have_window = false;
  const int nck = (nsteps + every - 1) / every + 1;
  for (int k = 0; k < nck; ++k) {
    double* p = nullptr;
    HIP_TRY(hipMalloc(&p, T->snap_words * sizeof(double)));
    T->ckpt.push_back(p);
  }
  T->seg_planes = (C->advection == SWM_ADV_FCT ? 12 : 6) + 6 * C->nt;   // tracer runs: 6 planes per tracer
  HIP_TRY(hipMalloc(&T->seg, (size_t)(every + 1) * T->seg_planes * C->ns * sizeof(double)));
  T->every = every;
  T->nsteps = nsteps;
  T->sched.assign((size_t)every, swm_ctx::SolveSched{});
  HIP_TRY(hipEventRecord(C->ev_a, C->stream));
  for (int n = 0; n < nsteps; ++n) {
    if (n % every == 0) {
      if ((st = swm::snapshot(C, T->ckpt[n / every]))) return st;
      T->ckpt_meta.push_back(C->step_count);
      T->ckpt_meta.push_back(C->nsecond);
    }
    if ((st = swm::base_step(C))) return st;
  }
  // the final state, restored after the sweeps
  if ((st = swm::snapshot(C, T->ckpt[nck - 1]))) return st;
  T->ckpt_meta.push_back(C->step_count);
  T->ckpt_meta.push_back(C->nsecond);
  HIP_TRY(hipEventRecord(C->ev_b, C->stream));
  HIP_TRY(hipEventSynchronize(C->ev_b));
  float ms = 0.0f;
  HIP_TRY(hipEventElapsedTime(&ms, C->ev_a, C->ev_b));
  C->last_step_ms = ms;
  T->have_window = true;
  return SWM_OK;
}

// TL of the window: mode 0, inputs the state perturbation (h, eta, div, u)
// read by the first step; mode 1, inputs the control perturbation (zeta,
// delta, h; in3 unused) through the initial state map.  Outputs the state
// perturbation after the last step.  With in_tr / out_tr (tracer runs) the
// tracers' perturbations go in and out too.
static swm_status tl_run_impl(swm_ctx* C, int mode, const double* in0, const double* in1, const double* in2,
                              const double* in3, const double* in_tr, double* out_h, double* out_eta,
                              double* out_div, double* out_u, double* out_tr) {
  if (!C->tl || !C->tl->have_window) { swm::set_error("swm_tl_run: no window (swm_adj_forward first)"); return SWM_ERR_STATE; }
  HIP_TRY(hipSetDevice(C->device));
  TlAd* T = C->tl;
  const size_t ns = C->ns;
  swm_status st;
  T->tracers = in_tr != nullptr;
  swm::clear_state(C);
  HIP_TRY(hipMemsetAsync(T->lim_count, 0, sizeof(unsigned long long), C->stream));
  T->lim_steps = 0;
  if (T->tlim_count) HIP_TRY(hipMemsetAsync(T->tlim_count, 0, sizeof(unsigned long long), C->stream));
  T->tlim_steps = 0;
  fill<double>(C, T->xs, 2 * ns, 0.0);
  if (mode == 0) {
    if ((st = swm::state_in(C, swm::first_level(), in0, in1, in2, in3))) return st;
    if (T->tracers && (st = swm::tracer_in(C, swm::first_level(), in_tr))) return st;
  } else {   // 1, or 2: the initial state map alone
    const size_t nt = T->tracers ? (size_t)C->nt : 0;
    double* buf = nullptr;
    HIP_TRY(hipMalloc(&buf, (3 + nt) * ns * sizeof(double)));
    HIP_TRY(hipMemcpy(buf, in0, ns * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(buf + ns, in1, ns * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(buf + 2 * ns, in2, ns * sizeof(double), hipMemcpyHostToDevice));
    if (nt) {
      std::vector<double> tp(nt * ns);
      swm::tracer_to_planes(C, in_tr, tp.data());
      HIP_TRY(hipMemcpy(buf + 3 * ns, tp.data(), nt * ns * sizeof(double), hipMemcpyHostToDevice));
    }
    st = swm::init_tl(C, buf, buf + ns, buf + 2 * ns, nt ? buf + 3 * ns : nullptr);
    (void)hipStreamSynchronize(C->stream);
    (void)hipFree(buf);
    if (st) return st;
    // exact linearisation: xs, the initial solve's perturbation, is the
    // first step's warm start; else the solves do not depend on it
    if (!swm::exact_lin(C)) fill<double>(C, T->xs, 2 * ns, 0.0);
    if (mode == 2) {
      if ((st = swm::state_out(C, swm::first_level(), out_h, out_eta, out_div, out_u))) return st;
      return T->tracers ? swm::tracer_out(C, swm::first_level(), out_tr) : SWM_OK;
    }
  }
  const int nseg = (T->nsteps + T->every - 1) / T->every;
  for (int sg = 0; sg < nseg; ++sg) {
    const int len = std::min(T->every, T->nsteps - sg * T->every);
    if ((st = swm::recompute_segment(C, sg, len))) return st;
    for (int q = 0; q < len; ++q) {
      swm::set_step(C, sg * T->every + q + 1);
      if ((st = swm::tl_step(C, q))) return st;
    }
  }
  const int lev = swm::last_level(T->nsteps);
  if ((st = swm::state_out(C, lev, out_h, out_eta, out_div, out_u))) return st;
  if (T->tracers && (st = swm::tracer_out(C, lev, out_tr))) return st;
  // the model back at the window's end
  const int nck = (int)T->ckpt.size();
  if ((st = swm::restore(C, T->ckpt[nck - 1], T->ckpt_meta[2 * nck - 2], T->ckpt_meta[2 * nck - 1]))) return st;
  HIP_TRY(hipStreamSynchronize(C->stream));
  return SWM_OK;
}

swm_status swm_tl_run(swm_ctx* C, int mode, const double* in0, const double* in1, const double* in2,
                      const double* in3, double* out_h, double* out_eta, double* out_div, double* out_u) {
  if (!C || mode < 0 || mode > 2 || !in0 || !in1 || !in2 || (mode == 0 && !in3) || !out_h || !out_eta ||
      !out_div || !out_u) {
    swm::set_error("swm_tl_run: bad argument");
    return SWM_ERR_ARG;
  }
  return tl_run_impl(C, mode, in0, in1, in2, in3, nullptr, out_h, out_eta, out_div, out_u, nullptr);
}

swm_status swm_tl_run_tracer(swm_ctx* C, int mode, const double* in0, const double* in1, const double* in2,
                             const double* in3, const double* in_tr, double* out_h, double* out_eta,
                             double* out_div, double* out_u, double* out_tr) {
  if (!C || mode < 0 || mode > 2 || !in0 || !in1 || !in2 || (mode == 0 && !in3) || !in_tr || !out_h ||
      !out_eta || !out_div || !out_u || !out_tr || C->nt < 1) {
    swm::set_error("swm_tl_run_tracer: bad argument (a context with tracers, every array given)");
    return SWM_ERR_ARG;
  }
  return tl_run_impl(C, mode, in0, in1, in2, in3, in_tr, out_h, out_eta, out_div, out_u, out_tr);
}

// AD of the window: the adjoint of the state after the last step (lh, leta,
// ldiv, lu) back to the first step (mode 0: out the adjoint state (h, eta,
// div, u)) or to the control (mode 1: out0..2 the gradient with respect to
// zeta, delta, h; out3 unused).  With ltr / out_tr (tracer runs) the tracers'
// adjoints go in and out too.
static swm_status adj_run_impl(swm_ctx* C, int mode, const double* lh, const double* leta, const double* ldiv,
                               const double* lu, const double* ltr, double* out0, double* out1, double* out2,
                               double* out3, double* out_tr) {
  if (!C->tl || !C->tl->have_window) { swm::set_error("swm_adj_run: no window (swm_adj_forward first)"); return SWM_ERR_STATE; }
  HIP_TRY(hipSetDevice(C->device));
  TlAd* T = C->tl;
  const size_t ns = C->ns;
  swm_status st;
  if (swm::exact_lin(C) && (st = swm::exact_prepare(C))) return st;
  T->tracers = ltr != nullptr;
  swm::clear_state(C);
  HIP_TRY(hipMemsetAsync(T->lim_count, 0, sizeof(unsigned long long), C->stream));
  T->lim_steps = 0;
  if (T->tlim_count) HIP_TRY(hipMemsetAsync(T->tlim_count, 0, sizeof(unsigned long long), C->stream));
  T->tlim_steps = 0;
  fill<double>(C, T->xs, 2 * ns, 0.0);
  // mode 2: the initial state map alone, from the adjoint of its output
  const int lev_in = mode == 2 ? swm::first_level() : swm::last_level(T->nsteps);
  if ((st = swm::state_in(C, lev_in, lh, leta, ldiv, lu))) return st;
  if (T->tracers && (st = swm::tracer_in(C, lev_in, ltr))) return st;
  const int nseg = mode == 2 ? 0 : (T->nsteps + T->every - 1) / T->every;
  for (int sg = nseg - 1; sg >= 0; --sg) {
    const int len = std::min(T->every, T->nsteps - sg * T->every);
    if ((st = swm::recompute_segment(C, sg, len))) return st;
    for (int q = len - 1; q >= 0; --q) {
      swm::set_step(C, sg * T->every + q + 1);
      if ((st = swm::ad_step(C, q))) return st;
    }
  }
  if (mode == 0) {
    if ((st = swm::state_out(C, swm::first_level(), out0, out1, out2, out3))) return st;
    if (T->tracers && (st = swm::tracer_out(C, swm::first_level(), out_tr))) return st;
  } else {
    const size_t nt = T->tracers ? (size_t)C->nt : 0;
    double* buf = nullptr;
    HIP_TRY(hipMalloc(&buf, (3 + nt) * ns * sizeof(double)));
    st = swm::init_ad(C, buf, buf + ns, buf + 2 * ns, nt ? buf + 3 * ns : nullptr);
    if (!st) {
      (void)hipStreamSynchronize(C->stream);
      std::vector<double> tp(nt * ns);
      if (hipMemcpy(out0, buf, ns * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
          hipMemcpy(out1, buf + ns, ns * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
          hipMemcpy(out2, buf + 2 * ns, ns * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
          (nt && hipMemcpy(tp.data(), buf + 3 * ns, nt * ns * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)) {
        swm::set_error("swm_adj_run: copy of the gradient failed");
        st = SWM_ERR_HIP;
      } else if (nt) {
        swm::tracer_from_planes(C, tp.data(), out_tr);
      }
    }
    (void)hipFree(buf);
    if (st) return st;
  }
  const int nck = (int)T->ckpt.size();
  if ((st = swm::restore(C, T->ckpt[nck - 1], T->ckpt_meta[2 * nck - 2], T->ckpt_meta[2 * nck - 1]))) return st;
  HIP_TRY(hipStreamSynchronize(C->stream));
  return SWM_OK;
}

swm_status swm_adj_run(swm_ctx* C, int mode, const double* lh, const double* leta, const double* ldiv,
                       const double* lu, double* out0, double* out1, double* out2, double* out3) {
  if (!C || mode < 0 || mode > 2 || !lh || !leta || !ldiv || !lu || !out0 || !out1 || !out2 ||
      (mode == 0 && !out3)) {
    swm::set_error("swm_adj_run: bad argument");
    return SWM_ERR_ARG;
  }
  return adj_run_impl(C, mode, lh, leta, ldiv, lu, nullptr, out0, out1, out2, out3, nullptr);
}

swm_status swm_adj_run_tracer(swm_ctx* C, int mode, const double* lh, const double* leta, const double* ldiv,
                              const double* lu, const double* ltr, double* out0, double* out1, double* out2,
                              double* out3, double* out_tr) {
  if (!C || mode < 0 || mode > 2 || !lh || !leta || !ldiv || !lu || !ltr || !out0 || !out1 || !out2 ||
      (mode == 0 && !out3) || !out_tr || C->nt < 1) {
    swm::set_error("swm_adj_run_tracer: bad argument (a context with tracers, every array given)");
    return SWM_ERR_ARG;
  }
  return adj_run_impl(C, mode, lh, leta, ldiv, lu, ltr, out0, out1, out2, out3, out_tr);
}

swm_status swm_set_linearisation(swm_ctx* C, int exact) {
  if (!C || exact < 0 || exact > 1) { swm::set_error("swm_set_linearisation: bad argument"); return SWM_ERR_ARG; }
  C->lin_exact = exact;
  return SWM_OK;
}

swm_status swm_set_halo_shadow(swm_ctx* C, int enable) {
  if (!C || enable < 0 || enable > 1) { swm::set_error("swm_set_halo_shadow: bad argument"); return SWM_ERR_ARG; }
  C->halo_shadow = enable;
  C->epoch += 1;
  return SWM_OK;
}

swm_status swm_halo_shadow_stats(const swm_ctx* C, long long* checks, long long* mismatches) {
  if (!C || !checks || !mismatches) { swm::set_error("swm_halo_shadow_stats: bad argument"); return SWM_ERR_ARG; }
  *checks = C->shadow_checks;
  *mismatches = C->shadow_bad;
  return SWM_OK;
}

swm_status swm_restart_clock(swm_ctx* C) {
  if (!C || !C->initialized) { swm::set_error("swm_restart_clock: needs an initialised context"); return SWM_ERR_STATE; }
  HIP_TRY(hipSetDevice(C->device));
  swm_status st = latest_into_both_levels(C);
  if (st) return st;
  C->step_count = 0;
  C->nsol = 0;
  C->epoch += 1;
  pointers(C);
  HIP_TRY(hipStreamSynchronize(C->stream));
  return SWM_OK;
}

swm_status swm_adj_limiter_stats(swm_ctx* C, long long* below_one, long long* edges) {
  if (!C || !below_one || !edges) { swm::set_error("swm_adj_limiter_stats: bad argument"); return SWM_ERR_ARG; }
  if (!C->tl || !C->tl->have_window) {
    swm::set_error("swm_adj_limiter_stats: no window (swm_adj_forward first)");
    return SWM_ERR_STATE;
  }
  HIP_TRY(hipSetDevice(C->device));
  HIP_TRY(hipStreamSynchronize(C->stream));
  unsigned long long n = 0;
  HIP_TRY(hipMemcpy(&n, C->tl->lim_count, sizeof(n), hipMemcpyDeviceToHost));
  *below_one = (long long)n;
  // three edges per interior slot (the poles, in halo slots, own none)
  const long long per = 3LL * (C->iim - 2) * (C->jjm - 2) * 10;
  *edges = per * C->tl->lim_steps;
  return SWM_OK;
}

swm_status swm_adj_tracer_limiter_stats(swm_ctx* C, long long* below_one, long long* edges) {
  if (!C || !below_one || !edges) { swm::set_error("swm_adj_tracer_limiter_stats: bad argument"); return SWM_ERR_ARG; }
  if (!C->tl || !C->tl->have_window) {
    swm::set_error("swm_adj_tracer_limiter_stats: no window (swm_adj_forward first)");
    return SWM_ERR_STATE;
  }
  HIP_TRY(hipSetDevice(C->device));
  HIP_TRY(hipStreamSynchronize(C->stream));
  unsigned long long n = 0;
  if (C->tl->tlim_count) HIP_TRY(hipMemcpy(&n, C->tl->tlim_count, sizeof(n), hipMemcpyDeviceToHost));
  *below_one = (long long)n;
  const long long per = 3LL * (C->iim - 2) * (C->jjm - 2) * 10 * C->nt;
  *edges = per * C->tl->tlim_steps;
  return SWM_OK;
}

}  // extern "C"

#include "swm_ensemble.hip"
