// swm_ensemble.hip -- ensembles of member groups advanced together in HIP
// graphs, and the autotuner of their launch configuration (included by
// swm_device.hip).  No counterpart in the reference, which runs one model per
// process.
//
// A group is a context of nmem members on one grid (swm_create_members); an
// ensemble step advances group k by substeps[k] of its own steps, so groups
// at different resolutions keep the same model time (a level g+1 group takes
// two steps of half the length per level g step).  The launches of one or
// more ensemble steps are captured once per host-side phase of the groups
// (ABM3 level and weights, which solver buffer holds which iterate, the
// launch configuration) into an executable graph; later steps in the same
// phase replay it with one hipGraphLaunch and only update the host-side
// bookkeeping.  The mixed solve then needs a fixed number of correction
// cycles (swm_set_fixed_cycles), since no host-side test can run inside a
// graph; the residuals are tracked on the device instead.
//
// The autotuner times the candidate configurations (one stream or a stream
// per group inside the graph, ensemble steps per graph, workgroup sizes of the
// stencil and multigrid kernels) on the ensemble's own state, keeps the
// fastest, and restores the state and clocks it started from.

struct swm_ensemble {
  std::vector<swm_ctx*> g;
  std::vector<int> sub;
  int device = 0;
  hipStream_t s0 = nullptr;
  std::vector<hipStream_t> gs;
  std::vector<hipEvent_t> fork, join;
  swm_ensemble_choice cfg{0, 1, 256, 256, 0.0};
  struct Entry {
    std::vector<long long> key;
    hipGraphExec_t exec = nullptr;
    size_t nodes = 0;
    // the groups' solver buffers and flags after the captured steps
    std::vector<std::array<double*, 4>> bufs;
    std::vector<int> nsol;
    std::vector<int> vcycles;
    std::vector<long long> vtotal;   // V-cycles of the captured steps, per group
  };
  std::vector<Entry> cache;
  hipEvent_t ev_a = nullptr, ev_b = nullptr;
  double last_ms = 0.0;
  size_t last_nodes = 0;
};

namespace swm {
namespace {

// step counts that give the same launches: the first six differ (weights of
// the first steps, time levels), after that the levels repeat every six
long long step_class(int sc) { return sc < 6 ? sc : 6 + sc % 6; }

std::vector<long long> ens_key(const swm_ensemble* e, int K) {
  std::vector<long long> k = {K, e->cfg.concurrent, e->cfg.blk_step, e->cfg.blk_mg};
  for (const swm_ctx* C : e->g) {
    k.push_back(C->epoch);
    k.push_back(step_class(C->step_count));
    k.push_back(C->nsol);
    k.push_back((long long)(intptr_t)C->xout);
    k.push_back((long long)(intptr_t)C->xalt);
    k.push_back((long long)(intptr_t)C->xprev);
    k.push_back((long long)(intptr_t)C->xprev2);
  }
  return k;
}

void apply_cfg(swm_ensemble* e) {
  for (swm_ctx* C : e->g) {
    C->blk_step = e->cfg.blk_step;
    C->blk_mg = e->cfg.blk_mg;
  }
}

void drop_cache(swm_ensemble* e) {
  for (auto& en : e->cache)
    if (en.exec) (void)hipGraphExecDestroy(en.exec);
  e->cache.clear();
}

// capture K ensemble steps into a new cache entry; the groups' host-side
// state advances as if the steps had run
swm_status capture(swm_ensemble* e, int K, swm_ensemble::Entry& en) {
  const int ng = (int)e->g.size();
  std::vector<hipStream_t> saved(ng);
  for (int k = 0; k < ng; ++k) saved[k] = e->g[k]->stream;
  std::vector<long long> v0(ng);
  for (int k = 0; k < ng; ++k) v0[k] = e->g[k]->vcycles_total;
  HIP_TRY(hipStreamBeginCapture(e->s0, hipStreamCaptureModeThreadLocal));
  swm_status st = SWM_OK;
  hipError_t he = hipSuccess;
  for (int k = 0; k < ng && he == hipSuccess; ++k) {
    if (e->cfg.concurrent) {
      he = hipEventRecord(e->fork[k], e->s0);
      if (he == hipSuccess) he = hipStreamWaitEvent(e->gs[k], e->fork[k], 0);
      e->g[k]->stream = e->gs[k];
    } else {
      e->g[k]->stream = e->s0;
    }
  }
  for (int n = 0; n < K && !st && he == hipSuccess; ++n)
    for (int k = 0; k < ng && !st; ++k)
      for (int q = 0; q < e->sub[k] && !st; ++q) st = step_once(e->g[k]);
  if (e->cfg.concurrent)
    for (int k = 0; k < ng && he == hipSuccess; ++k) {
      he = hipEventRecord(e->join[k], e->gs[k]);
      if (he == hipSuccess) he = hipStreamWaitEvent(e->s0, e->join[k], 0);
    }
  hipGraph_t graph = nullptr;
  const hipError_t he_end = hipStreamEndCapture(e->s0, &graph);
  for (int k = 0; k < ng; ++k) e->g[k]->stream = saved[k];
  if (st) {
    if (graph) (void)hipGraphDestroy(graph);
    return st;
  }
  if (he != hipSuccess || he_end != hipSuccess || !graph) {
    if (graph) (void)hipGraphDestroy(graph);
    swm::set_error(std::string("ensemble capture failed: ") + hipGetErrorString(he != hipSuccess ? he : he_end));
    return SWM_ERR_HIP;
  }
  size_t nn = 0;
  (void)hipGraphGetNodes(graph, nullptr, &nn);
  const hipError_t hi = hipGraphInstantiate(&en.exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (hi != hipSuccess) {
    swm::set_error(std::string("ensemble graph instantiation failed: ") + hipGetErrorString(hi));
    return SWM_ERR_HIP;
  }
  en.nodes = nn;
  for (int k = 0; k < ng; ++k) en.vtotal.push_back(e->g[k]->vcycles_total - v0[k]);
  for (swm_ctx* C : e->g) {
    en.bufs.push_back({C->xout, C->xalt, C->xprev, C->xprev2});
    en.nsol.push_back(C->nsol);
    en.vcycles.push_back(C->last_vcycles);
  }
  return SWM_OK;
}

// the bookkeeping of K replayed ensemble steps
void advance_host(swm_ensemble* e, int K, const swm_ensemble::Entry& en) {
  for (size_t k = 0; k < e->g.size(); ++k) {
    swm_ctx* C = e->g[k];
    const int n = K * e->sub[k];
    for (int q = 0; q < n; ++q) C->nsecond = (long long)((double)C->nsecond + C->dt);   // as step_once
    C->step_count += n;
    pointers(C);
    weights(C);
    C->xout = en.bufs[k][0];
    C->xalt = en.bufs[k][1];
    C->xprev = en.bufs[k][2];
    C->xprev2 = en.bufs[k][3];
    C->nsol = en.nsol[k];
    C->last_vcycles = en.vcycles[k];
    C->solves += n;
    C->vcycles_total += en.vtotal[k];
  }
}

swm_status ens_steps(swm_ensemble* e, int nsteps) {
  apply_cfg(e);
  HIP_TRY(hipEventRecord(e->ev_a, e->s0));
  size_t nodes = 0;
  for (int n = 0; n < nsteps;) {
    const int K = nsteps - n >= e->cfg.steps_per_graph ? e->cfg.steps_per_graph : 1;
    const std::vector<long long> key = ens_key(e, K);
    int hit = -1;
    for (size_t c = 0; c < e->cache.size(); ++c)
      if (e->cache[c].key == key) { hit = (int)c; break; }
    if (hit < 0) {
      swm_ensemble::Entry en;
      en.key = key;
      swm_status st = capture(e, K, en);
      if (st) return st;
      e->cache.push_back(std::move(en));
      hit = (int)e->cache.size() - 1;
    } else {
      advance_host(e, K, e->cache[hit]);
    }
    HIP_TRY(hipGraphLaunch(e->cache[hit].exec, e->s0));
    nodes += e->cache[hit].nodes;
    n += K;
  }
  HIP_TRY(hipEventRecord(e->ev_b, e->s0));
  HIP_TRY(hipEventSynchronize(e->ev_b));
  float ms = 0.0f;
  HIP_TRY(hipEventElapsedTime(&ms, e->ev_a, e->ev_b));
  e->last_ms = nsteps > 0 ? ms / nsteps : 0.0;
  e->last_nodes = nodes;
  HIP_TRY(hipGetLastError());
  return SWM_OK;
}

// the groups' state and host-side clocks, to return to after tuning
struct EnsSnapshot {
  std::vector<void*> mem;
  struct Host {
    int step_count, last_vcycles;
    long long nsecond, solves, vcycles_total, checks_total;
    int nsol;
    double *xout, *xalt, *xprev, *xprev2;
  };
  std::vector<Host> host;
};

swm_status snapshot(swm_ensemble* e, EnsSnapshot& sn) {
  for (swm_ctx* C : e->g) {
    void* p = nullptr;
    const size_t bytes = C->mb * (size_t)C->nmem;
    HIP_TRY(hipMalloc(&p, bytes));
    sn.mem.push_back(p);
    HIP_TRY(hipMemcpyAsync(p, C->marena.base, bytes, hipMemcpyDeviceToDevice, e->s0));
    sn.host.push_back({C->step_count, C->last_vcycles, C->nsecond, C->solves, C->vcycles_total, C->checks_total,
                       C->nsol, C->xout, C->xalt, C->xprev, C->xprev2});
  }
  HIP_TRY(hipStreamSynchronize(e->s0));
  return SWM_OK;
}

swm_status restore(swm_ensemble* e, EnsSnapshot& sn) {
  for (size_t k = 0; k < e->g.size(); ++k) {
    swm_ctx* C = e->g[k];
    HIP_TRY(hipMemcpyAsync(C->marena.base, sn.mem[k], C->mb * (size_t)C->nmem, hipMemcpyDeviceToDevice, e->s0));
    const EnsSnapshot::Host& h = sn.host[k];
    C->step_count = h.step_count; C->last_vcycles = h.last_vcycles; C->nsecond = h.nsecond;
    C->solves = h.solves; C->vcycles_total = h.vcycles_total; C->checks_total = h.checks_total;
    C->nsol = h.nsol; C->xout = h.xout; C->xalt = h.xalt; C->xprev = h.xprev; C->xprev2 = h.xprev2;
    pointers(C);
    weights(C);
  }
  HIP_TRY(hipStreamSynchronize(e->s0));
  for (void* p : sn.mem) (void)hipFree(p);
  sn.mem.clear();
  return SWM_OK;
}

}  // namespace
}  // namespace swm

extern "C" {

swm_status swm_ensemble_create(swm_ctx* const* groups, const int* substeps, int ngroups, swm_ensemble** out) {
  if (!groups || !substeps || !out || ngroups < 1 || ngroups > 16) {
    swm::set_error("swm_ensemble_create: bad argument");
    return SWM_ERR_ARG;
  }
  for (int k = 0; k < ngroups; ++k) {
    const swm_ctx* C = groups[k];
    if (!C || substeps[k] < 1 || substeps[k] > 64 || C->device != groups[0]->device) {
      swm::set_error("swm_ensemble_create: every group needs a context on the same device and 1..64 substeps");
      return SWM_ERR_ARG;
    }
    for (int j = 0; j < k; ++j)
      if (groups[j] == C) { swm::set_error("swm_ensemble_create: a context appears twice"); return SWM_ERR_ARG; }
  }
  swm_ensemble* e = new swm_ensemble;
  e->device = groups[0]->device;
  for (int k = 0; k < ngroups; ++k) { e->g.push_back(groups[k]); e->sub.push_back(substeps[k]); }
  auto fail = [&](hipError_t he) {
    swm::set_error(std::string("swm_ensemble_create: ") + hipGetErrorString(he));
    swm_ensemble_destroy(e);
    return SWM_ERR_HIP;
  };
  hipError_t he;
  if ((he = hipSetDevice(e->device)) != hipSuccess) return fail(he);
  if ((he = hipStreamCreateWithFlags(&e->s0, hipStreamNonBlocking)) != hipSuccess) return fail(he);
  e->gs.assign(ngroups, nullptr);
  e->fork.assign(ngroups, nullptr);
  e->join.assign(ngroups, nullptr);
  for (int k = 0; k < ngroups; ++k) {
    if ((he = hipStreamCreateWithFlags(&e->gs[k], hipStreamNonBlocking)) != hipSuccess) return fail(he);
    if ((he = hipEventCreateWithFlags(&e->fork[k], hipEventDisableTiming)) != hipSuccess) return fail(he);
    if ((he = hipEventCreateWithFlags(&e->join[k], hipEventDisableTiming)) != hipSuccess) return fail(he);
  }
  if ((he = hipEventCreate(&e->ev_a)) != hipSuccess) return fail(he);
  if ((he = hipEventCreate(&e->ev_b)) != hipSuccess) return fail(he);
  *out = e;
  return SWM_OK;
}

void swm_ensemble_destroy(swm_ensemble* e) {
  if (!e) return;
  // teardown: errors here have no caller to report to
  (void)hipSetDevice(e->device);
  if (e->s0) (void)hipStreamSynchronize(e->s0);
  swm::drop_cache(e);
  for (hipStream_t s : e->gs)
    if (s) (void)hipStreamDestroy(s);
  for (hipEvent_t v : e->fork)
    if (v) (void)hipEventDestroy(v);
  for (hipEvent_t v : e->join)
    if (v) (void)hipEventDestroy(v);
  if (e->ev_a) (void)hipEventDestroy(e->ev_a);
  if (e->ev_b) (void)hipEventDestroy(e->ev_b);
  if (e->s0) (void)hipStreamDestroy(e->s0);
  delete e;
}

swm_status swm_ensemble_set(swm_ensemble* e, const swm_ensemble_choice* c) {
  auto bs = [](int b) { return b == 64 || b == 128 || b == 256; };
  if (!e || !c || (c->concurrent != 0 && c->concurrent != 1) || c->steps_per_graph < 1 || c->steps_per_graph > 12 ||
      !bs(c->blk_step) || !bs(c->blk_mg)) {
    swm::set_error("swm_ensemble_set: concurrent 0/1, 1..12 steps per graph, workgroup sizes 64, 128 or 256");
    return SWM_ERR_ARG;
  }
  e->cfg.concurrent = c->concurrent;
  e->cfg.steps_per_graph = c->steps_per_graph;
  e->cfg.blk_step = c->blk_step;
  e->cfg.blk_mg = c->blk_mg;
  return SWM_OK;
}

swm_status swm_ensemble_get(const swm_ensemble* e, swm_ensemble_choice* c) {
  if (!e || !c) { swm::set_error("swm_ensemble_get: bad argument"); return SWM_ERR_ARG; }
  *c = e->cfg;
  c->ms_per_step = e->last_ms;
  return SWM_OK;
}

swm_status swm_ensemble_step(swm_ensemble* e, int nsteps) {
  if (!e || nsteps < 0) { swm::set_error("swm_ensemble_step: bad argument"); return SWM_ERR_ARG; }
  for (const swm_ctx* C : e->g) {
    if (!C->initialized) { swm::set_error("swm_ensemble_step: a group is not initialised (every member)"); return SWM_ERR_STATE; }
    if (C->halo_shadow) {
      swm::set_error("swm_ensemble_step: the halo shadow check cannot run inside a captured graph");
      return SWM_ERR_STATE;
    }
    if (C->precision != SWM_MIXED || C->fixed_cycles < 1) {
      swm::set_error("swm_ensemble_step: every group runs the mixed step with a fixed schedule (swm_set_fixed_cycles)");
      return SWM_ERR_STATE;
    }
    if (C->profile) { swm::set_error("swm_ensemble_step: kernel profiling is not captured; switch it off"); return SWM_ERR_STATE; }
  }
  HIP_TRY(hipSetDevice(e->device));
  for (swm_ctx* C : e->g) HIP_TRY(hipStreamSynchronize(C->stream));
  return swm::ens_steps(e, nsteps);
}

swm_status swm_ensemble_tune(swm_ensemble* e, int rounds, swm_ensemble_choice* trials, int max_trials, int* ntrials) {
  if (!e || rounds < 1 || rounds > 100 || max_trials < 0 || (max_trials > 0 && !trials)) {
    swm::set_error("swm_ensemble_tune: bad argument");
    return SWM_ERR_ARG;
  }
  swm_status st = swm_ensemble_step(e, 0);   // the checks of swm_ensemble_step
  if (st) return st;
  swm::EnsSnapshot sn;
  if ((st = swm::snapshot(e, sn))) return st;
  // each trial: six ensemble steps (the period of the launch phases, and a
  // whole number of graphs for 1, 2, 3 or 6 steps per graph) to capture,
  // then rounds x six steps timed.  The first steps of a run differ (ABM3
  // start), so the trials start from step six at least.
  int nt = 0;
  const swm_ensemble_choice start = e->cfg;
  swm_ensemble_choice best = start;
  double best_ms = 1e300;
  auto trial = [&](swm_ensemble_choice c) -> swm_status {
    e->cfg = c;
    swm_status s2 = swm::ens_steps(e, 6);
    if (s2) return s2;
    double tot = 0.0;
    for (int r = 0; r < rounds; ++r) {
      if ((s2 = swm::ens_steps(e, 6))) return s2;
      tot += e->last_ms;
    }
    c.ms_per_step = tot / rounds;
    if (nt < max_trials) trials[nt] = c;
    ++nt;
    if (c.ms_per_step < best_ms) { best_ms = c.ms_per_step; best = c; }
    return SWM_OK;
  };
  const bool warm = std::all_of(e->g.begin(), e->g.end(), [](const swm_ctx* C) { return C->step_count >= 6; });
  if (!warm) {
    e->cfg = start;
    if ((st = swm::ens_steps(e, 6))) { (void)swm::restore(e, sn); return st; }
  }
  // schedule: one stream or a stream per group, ensemble steps per graph
  for (int conc = 0; conc < (e->g.size() > 1 ? 2 : 1) && !st; ++conc)
    for (int K : {1, 3, 6}) {
      swm_ensemble_choice c{conc, K, 256, 256, 0.0};
      if ((st = trial(c))) break;
    }
  // workgroup sizes of the stencil kernels, then of the multigrid kernels
  for (int b : {64, 128}) {
    if (st) break;
    swm_ensemble_choice c = best;
    c.blk_step = b;
    st = trial(c);
  }
  const int bstep = best.blk_step;
  for (int b : {64, 128}) {
    if (st) break;
    swm_ensemble_choice c = best;
    c.blk_step = bstep;
    c.blk_mg = b;
    st = trial(c);
  }
  e->cfg = best;
  e->cfg.ms_per_step = 0.0;
  swm::drop_cache(e);
  const swm_status rs = swm::restore(e, sn);
  if (ntrials) *ntrials = nt;
  if (st) return st;
  return rs;
}

}  // extern "C"
