// The fused MD step and its hipGraph capture, plus run-time knobs.
//
// Stream layout of one step (the reference driver's order of stages):
//   s0: NHC half 1 (+ wrap) --fork2--> cycle energies --fork--> reset F,
//       trap, ring springs (pair forces here unless they are Ewald), Ewald
//       phase tables, structure factors and reciprocal forces (into a side
//       buffer) --wait join, join2--> add the pair and exchange end-bead
//       forces, NHC half 2 (one kernel with massive thermostats)
// With massive thermostats NHC half 1 and the wrap are one kernel, and so are
// the force reset, trap and interior springs.
//   s1: (after fork) forward + backward recursions, end-bead exchange forces
//       into a side buffer --join--> (added into F on s0)
//   s2: (Ewald) real-space tile sums, forked from s0 right after NHC half 1
//       (positions only) --join2--> s0 then adds the real-space,
//       reciprocal and exchange forces
// The recursions only read cycle energies, so they overlap the pair forces;
// the end-bead exchange forces are added into F after both branches are done
// (the serial profiling order computes and adds them there in one kernel).
// The real-space grid goes first on s2: it fills the device while the short
// cycle-energy kernels (highest wave priority) run beside it, and the
// recursions then overlap the matrix-core k-space kernels, which leave them
// more of the compute units than the real-space waves do (UEG N=1024 P=32:
// 0.219 -> 0.199 ms per step against the real-space grid on s0 after the
// recursion fork).  The k-space kernels run on s0 after the bead-local
// forces, beside the end of the real-space grid, rather than after it on s2
// (0.1975 -> 0.192 ms with the phase tables moved, 0.185 -> 0.1815 ms with
// the structure factors and forces following them).
//
// Batches of independent replicas (md_pimd_run_batch_x) run the same stages
// with one launch per stage over a packed table of all replicas, captured
// into one hipGraph per batch step.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "md_hip_x.h"
#include "xmd_internal.h"

namespace xmd {

struct StepPlan {
  DevPimd *p = nullptr;
  int trap = 0;
  int pair_kind = -1;
  double dt = 0;
  int pbc = 0, image = 0;
  double L[3] = {0, 0, 0};
};

static int plan_of(md_pimd_t_x *pimd, const md_pimd_step_config_x *cfg, StepPlan &sp) {
  if (ensure_pimd_device(pimd)) return -1;
  sp.p = pimd_state(pimd, false);
  if (cfg->trap && cfg->trap != md_harmonic_trap_force_x) {
    fprintf(stderr, "xmd: trap callback has no device kernel\n");
    return -1;
  }
  sp.trap = cfg->trap != nullptr;
  sp.pair_kind = -1;
  if (cfg->pair) {
    sp.pair_kind = pair_kind_of(cfg->pair);
    if (sp.pair_kind < 0) {
      fprintf(stderr, "xmd: pair-force callback has no device kernel\n");
      return -1;
    }
  }
  sp.dt = cfg->dt;
  sp.pbc = cfg->pbc;
  sp.image = cfg->image;
  for (int d = 0; d < 3; ++d) sp.L[d] = box_L(pimd->sim, d);
  return 0;
}

static BatchItem item_of(const StepPlan &sp) {
  BatchItem it;
  it.p = sp.p;
  it.dt = sp.dt;
  it.pbc = sp.pbc;
  it.image = sp.image;
  it.trap = sp.trap;
  it.pair_kind = sp.pair_kind;
  for (int d = 0; d < 3; ++d) it.L[d] = sp.L[d];
  return it;
}

// A stream pair for one system's step: s0 carries the step, s1 the
// recursion branch (fork / join events between them).
struct Lane {
  hipStream_t s0 = nullptr, s1 = nullptr, s2 = nullptr;
  hipEvent_t fork = nullptr, join = nullptr, fork2 = nullptr, join2 = nullptr;
};

static Lane main_lane() {
  Runtime &R = rt();
  Lane L;
  L.s0 = R.s0;
  L.s1 = R.s1;
  L.s2 = R.s2;
  L.fork = R.fork;
  L.join = R.join;
  L.fork2 = R.fork2;
  L.join2 = R.join2;
  return L;
}

// ev: optional per-stage events (2 per stage: start, stop), may be null.
// serial: run every stage on s0 (true per-stage device times for profiling).
// fused (may be null: F complete on return): set (Fpart non-null) when the
// pair and exchange parts are left for NHC half 2 to add into F (massive
// thermostats), instead of a combine kernel.
static int enqueue_forces(StepPlan &sp, hipEvent_t *ev, bool serial, const Lane &R, ForceSums *fused) {
  DevPimd *p = sp.p;
  DevSim *d = p->ds;
  hipStream_t sx = serial ? R.s0 : R.s1;
  auto rec = [&](int k, hipStream_t st) -> int {
    if (ev) XMD_CHECK(hipEventRecord(ev[k], st));
    return 0;
  };
  // Reciprocal space (positions only) forks onto s2 first: the phase tables
  // and structure factors overlap the cycle energies.
  const int kfork =
      (!serial && sp.pair_kind >= 0) ? launch_pair_ewald_fork(p, sp.pair_kind, R.s0, R.s2, R.fork2, R.join2) : 0;
  if (kfork < 0) return -1;
  // Cycle energies (short, position-only); then the two recursions are
  // forked onto s1 before the real-space grid is queued, so their
  // workgroups are resident before it fills the device.
  if (rec(2, R.s0) || (serial ? launch_enk(p, sp.image, R.s0) : launch_enk_prep(p, sp.image, R.s0)) ||
      rec(3, R.s0))
    return -1;
  if (!serial) {
    XMD_CHECK(hipEventRecord(R.fork, R.s0));
    XMD_CHECK(hipStreamWaitEvent(R.s1, R.fork, 0));
  }
  if (rec(4, sx) || (serial ? launch_recursion(p, true, true, sx, ev ? ev[16] : nullptr) : launch_recursion_solve(p, sx)) ||
      rec(5, sx))
    return -1;
  // overlapped: the end-bead exchange forces follow the recursions on s1
  // into a side buffer, added into F after the join
  if (!serial && launch_exchange_force(p, sp.image, R.s1, true)) return -1;
  if (!serial) XMD_CHECK(hipEventRecord(R.join, R.s1));
  // bead-local and pair forces
  if (rec(6, R.s0)) return -1;
  if (d->massive) {
    if (launch_step_bead_forces(item_of(sp), R.s0)) return -1;
  } else {
    if (launch_reset_force(d, R.s0)) return -1;
    if (sp.trap && launch_trap_force(d, md_trap_frequency_x, md_trap_center_x, (double)p->P, R.s0)) return -1;
    if (launch_ring_springs(p, sp.image, sp.L, R.s0)) return -1;
  }
  if (rec(7, R.s0)) return -1;
  if (kfork && launch_pair_kspace(p, R.s0)) return -1;
  if (kfork) {
    if (rec(8, R.s0) || rec(9, R.s0) || rec(10, R.s0) || rec(11, R.s0)) return -1;
  } else if (sp.pair_kind >= 0) {
    if (rec(8, R.s0) || launch_pair_real(p, sp.pair_kind, R.s0) || rec(9, R.s0)) return -1;
    if (pair_has_recip(sp.pair_kind)) {
      if (rec(10, R.s0) || launch_pair_recip(p, sp.pair_kind, R.s0) || rec(11, R.s0)) return -1;
    } else if (rec(10, R.s0) || rec(11, R.s0)) {
      return -1;
    }
  } else if (rec(8, R.s0) || rec(9, R.s0) || rec(10, R.s0) || rec(11, R.s0)) {
    return -1;
  }
  if (!serial) {
    XMD_CHECK(hipStreamWaitEvent(R.s0, R.join, 0));
    if (kfork) {
      // one kernel adds the real-space, reciprocal and exchange forces
      XMD_CHECK(hipStreamWaitEvent(R.s0, R.join2, 0));
      if (fused && d->massive) {
        if (rec(12, R.s0) || force_sums(p, p->Fx, fused) || rec(13, R.s0)) return -1;
      } else if (rec(12, R.s0) || launch_combine_forces(p, p->Fx, R.s0) || rec(13, R.s0)) {
        return -1;
      }
    } else if (rec(12, R.s0) || launch_add_exchange_force(p, R.s0) || rec(13, R.s0)) {
      return -1;
    }
  } else if (rec(12, R.s0) || launch_exchange_force(p, sp.image, R.s0, false) || rec(13, R.s0)) {
    return -1;
  }
  return 0;
}

// head: NHC half 1 starts the step (false: the previous step's chained tail
// did it).  chain (massive thermostats): NHC half 2 and the next step's half
// 1 run as one kernel at the end of this step.
static int enqueue_step(StepPlan &sp, hipEvent_t *ev, bool serial, const Lane &R, bool head = true,
                        bool chain = false) {
  DevSim *d = sp.p->ds;
  if (ev) XMD_CHECK(hipEventRecord(ev[0], R.s0));
  if (!head) {
  } else if (d->massive) {
    if (launch_step_nhc(item_of(sp), 1, R.s0)) return -1;
  } else {
    if (launch_nhc(d, 1, sp.dt, R.s0)) return -1;
    if (sp.pbc && launch_pbc(d, sp.L, R.s0)) return -1;
  }
  if (ev) XMD_CHECK(hipEventRecord(ev[1], R.s0));
  ForceSums fused;
  if (enqueue_forces(sp, ev, serial, R, &fused)) return -1;
  if (ev) XMD_CHECK(hipEventRecord(ev[14], R.s0));
  if (d->massive ? launch_step_nhc(item_of(sp), chain ? 3 : 2, R.s0, fused.Fpart ? &fused : nullptr)
                 : launch_nhc(d, 2, sp.dt, R.s0))
    return -1;
  if (ev) XMD_CHECK(hipEventRecord(ev[15], R.s0));
  return 0;
}

// One step of a whole batch, the single-system step's stage order with one
// launch per stage over all replicas; the recursions fork onto s1 as in the
// single-system step.
struct Batch {
  BeadBatch *bead = nullptr;
  ExchBatch *exch = nullptr;
  PairBatch *pair = nullptr;
  ~Batch() {
    bead_batch_free(bead);
    exch_batch_free(exch);
    pair_batch_free(pair);
  }
};

// A batch's tables and executable step graph are kept after a call and reused
// by the next call on the same objects with the same configs, interaction
// globals and object state (temperature, exchange parameter, box), as long
// as no pair-kernel buffer was reallocated in between; anything else
// rebuilds them.
struct BatchCache {
  std::vector<md_pimd_t_x *> pimds;
  std::vector<md_pimd_step_config_x> cfgs;
  std::vector<double> key;
  unsigned epoch = 0;
  Batch *B = nullptr;
  hipGraphExec_t exec = nullptr;
};
static BatchCache g_bcache;

static std::vector<double> batch_key(md_pimd_t_x *const *pimds, int n) {
  std::vector<double> k = {md_trap_frequency_x, md_trap_center_x[0], md_trap_center_x[1], md_trap_center_x[2],
                           md_gaussian_strength_x, md_gaussian_range_x, md_coulomb_strength_x, md_coulomb_truc_x,
                           (double)md_coulomb_n_sum_x, (double)g_nhc_len, g_ewald_xcut, g_ewald_split_aL,
                           md_he_eps_x, md_he_A_x, md_he_alpha_x, md_he_C6_x, md_he_C8_x, md_he_C10_x, md_he_D_x,
                           md_he_rm_x};
  for (int i = 0; i < n; ++i) {
    k.push_back(pimds[i]->beta);
    k.push_back(pimds[i]->omegaP);
    k.push_back(pimds[i]->vi);
    k.push_back(pimds[i]->sim->T);
    for (int d = 0; d < 3; ++d) k.push_back(box_L(pimds[i]->sim, d));
  }
  return k;
}

static void batch_cache_drop() {
  if (g_bcache.exec) (void)hipGraphExecDestroy(g_bcache.exec);
  delete g_bcache.B;
  g_bcache = BatchCache();
}

static int enqueue_batch(Batch &B, const Lane &L) {
  if (bead_batch_nhc(B.bead, 1, L.s0) || exch_batch_enk(B.exch, L.s0)) return -1;
  XMD_CHECK(hipEventRecord(L.fork, L.s0));
  XMD_CHECK(hipStreamWaitEvent(L.s1, L.fork, 0));
  if (exch_batch_recursion(B.exch, L.s1)) return -1;
  XMD_CHECK(hipEventRecord(L.join, L.s1));
  if (bead_batch_forces(B.bead, L.s0) || pair_batch_forces(B.pair, L.s0)) return -1;
  XMD_CHECK(hipStreamWaitEvent(L.s0, L.join, 0));
  if (exch_batch_force(B.exch, L.s0) || bead_batch_nhc(B.bead, 2, L.s0)) return -1;
  return 0;
}

}  // namespace xmd

using namespace xmd;

extern "C" {

int md_set_nhc_length_x(int M) {
  if (M < 2 || M > MD_NHC_LENGTH_X || M % 2) return -1;
  g_nhc_len = M;
  return 0;
}

int md_get_nhc_length_x(void) { return g_nhc_len; }

int md_set_ewald_real_cutoff_x(md_num_t_x x_cut) {
  if (!(x_cut > 0)) return -1;
  g_ewald_xcut = x_cut;
  return 0;
}

int md_set_ewald_split_x(md_num_t_x alpha_L) {
  if (!(alpha_L == alpha_L)) return -1;
  g_ewald_split_aL = alpha_L;
  return 0;
}

int md_hip_synchronize_x(void) {
  if (ensure_runtime()) return -1;
  XMD_CHECK(hipStreamSynchronize(rt().s1));
  XMD_CHECK(hipStreamSynchronize(rt().s0));
  return recursion_status_all();
}

int md_pimd_inject_recursion_timeouts_x(md_pimd_t_x *pimd, int count) {
  if (count < 0 || ensure_pimd_device(pimd)) return -1;
  return inject_recursion_timeouts(pimd_state(pimd, false), count);
}

int md_set_recursion_frame_limit_x(double limit) {
  set_recursion_frame_limit(limit);
  return 0;
}

int md_pimd_get_backward_potentials_x(md_pimd_t_x *pimd, md_num_t_x *Vt) {
  DevPimd *p = pimd_state(pimd, false);
  if (!p || !p->back_dev) return -1;
  XMD_CHECK(hipStreamSynchronize(rt().s0));
  return backward_potentials(p, Vt);
}

int md_pimd_get_connection_matrix_x(md_pimd_t_x *pimd, md_num_t_x *G) {
  DevPimd *p = pimd_state(pimd, false);
  if (!p || !p->back_dev || !p->vbn_dev) return -1;
  return connection_matrix(p, G);
}

int md_pimd_compute_forces_x(md_pimd_t_x *pimd, const md_pimd_step_config_x *cfg) {
  StepPlan sp;
  if (plan_of(pimd, cfg, sp)) return -1;
  return enqueue_forces(sp, nullptr, false, main_lane(), nullptr);
}

int md_pimd_run_steps_x(md_pimd_t_x *pimd, const md_pimd_step_config_x *cfg, int nsteps) {
  StepPlan sp;
  if (plan_of(pimd, cfg, sp)) return -1;
  if (nsteps <= 0) return 0;
  const Lane L = main_lane();
  if (!cfg->use_graph || nsteps == 1) {
    for (int s = 0; s < nsteps; ++s)
      if (enqueue_step(sp, nullptr, false, L)) return -1;
    return 0;
  }
  // Every call captures the step afresh, so the graph holds this call's
  // parameters and buffers.  The step's lazily created buffers are made
  // first, on the host; when that allocated anything, or on the first call
  // on an object, the first step runs eagerly (so allocation and the
  // capture's kernel-only graph cannot interfere) and the capture runs on the
  // host while the device works on that step.  A call of at least
  // kLongCall graph steps captures kStepsPerGraph consecutive steps into one
  // graph, which saves most of the boundaries between graph launches
  // (0.2486 -> 0.2479 ms per step for UEG N=1024 P=32); shorter calls
  // capture one (the capture's host time would not be amortised); steps left
  // over run eagerly after the launches.  With massive thermostats the
  // captured steps chain NHC half 2 of one step and half 1 of the next into
  // one kernel (one launch and one pass over the thermostat state fewer per
  // step boundary inside the graph).  An
  // executable graph kept from the previous call on the same object is
  // updated in place when the new graph has its topology, which saves the
  // instantiation.
  constexpr int kStepsPerGraph = 4, kLongCall = 32;
  Runtime &R = rt();
  DevPimd *p = sp.p;
  if (autotune_if_needed(p, sp.pair_kind)) return -1;
  const int fresh = pair_prepare(p, sp.pair_kind);
  if (fresh < 0) return -1;
  const int eager = (fresh || !p->step_exec) ? 1 : 0;
  if (eager && enqueue_step(sp, nullptr, false, L)) return -1;
  const int per = nsteps - eager >= kLongCall ? kStepsPerGraph : 1, launches = (nsteps - eager) / per;
  hipGraph_t graph = nullptr;
  XMD_CHECK(hipStreamBeginCapture(R.s0, hipStreamCaptureModeGlobal));
  int rc = 0;
  const bool chain = p->ds->massive;
  for (int k = 0; k < per && !rc; ++k) rc = enqueue_step(sp, nullptr, false, L, k == 0 || !chain, chain && k + 1 < per);
  const hipError_t ec = hipStreamEndCapture(R.s0, &graph);
  if (rc || ec != hipSuccess || !graph) {
    fprintf(stderr, "xmd: step capture failed\n");
    return -1;
  }
  hipGraphExec_t exec = (hipGraphExec_t)p->step_exec;
  bool updated = false;
  if (exec) {
    hipGraphNode_t bad = nullptr;
    hipGraphExecUpdateResult res;
    updated = hipGraphExecUpdate(exec, graph, &bad, &res) == hipSuccess && res == hipGraphExecUpdateSuccess;
    if (!updated) {
      (void)hipGetLastError();
      XMD_CHECK(hipGraphExecDestroy(exec));
      exec = nullptr;
    }
  }
  if (!updated) XMD_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  p->step_exec = exec;
  XMD_CHECK(hipGraphDestroy(graph));
  for (int s = 0; s < launches; ++s) XMD_CHECK(hipGraphLaunch(exec, R.s0));
  for (int s = eager + launches * per; s < nsteps; ++s)
    if (enqueue_step(sp, nullptr, false, L)) return -1;
  XMD_CHECK(hipStreamSynchronize(R.s0));
  return recursion_status(sp.p);
}

int md_pimd_profile_steps_x(md_pimd_t_x *pimd, const md_pimd_step_config_x *cfg, int nsteps, int serial,
                            double *stage_ms) {
  StepPlan sp;
  if (plan_of(pimd, cfg, sp)) return -1;
  const int NE = 18;
  hipEvent_t ev[NE];
  for (int k = 0; k < NE; ++k) XMD_CHECK(hipEventCreate(&ev[k]));
  for (int k = 0; k < MD_PIMD_NSTAGES_X; ++k) stage_ms[k] = 0;
  // stage k <- events (a, b)
  const int pairs[8][2] = {{0, 1}, {2, 3}, {4, 5}, {6, 7}, {8, 9}, {10, 11}, {12, 13}, {14, 15}};
  for (int s = 0; s < nsteps; ++s) {
    if (serial) set_real_space_probe(ev[17]);
    const int bad = enqueue_step(sp, ev, serial != 0, main_lane());
    const bool probe = serial && real_space_probe_hit();
    set_real_space_probe(nullptr);
    if (bad) return -1;
    XMD_CHECK(hipStreamSynchronize(rt().s0));
    if (recursion_status(sp.p)) return -1;
    for (int k = 0; k < 8; ++k) {
      float ms = 0;
      XMD_CHECK(hipEventElapsedTime(&ms, ev[pairs[k][0]], ev[pairs[k][1]]));
      stage_ms[k] += ms;
    }
    float ms = 0;
    XMD_CHECK(hipEventElapsedTime(&ms, ev[0], ev[15]));
    stage_ms[8] += ms;
    if (serial) {
      XMD_CHECK(hipEventElapsedTime(&ms, ev[16], ev[5]));
      stage_ms[9] += ms;
      if (probe) XMD_CHECK(hipEventElapsedTime(&ms, ev[8], ev[17]));
      stage_ms[10] += probe ? ms : 0.0f;
    }
  }
  for (int k = 0; k < MD_PIMD_NSTAGES_X; ++k) stage_ms[k] /= nsteps;
  for (int k = 0; k < NE; ++k) XMD_CHECK(hipEventDestroy(ev[k]));
  return 0;
}

int md_pimd_run_batch_x(md_pimd_t_x *const *pimds, const md_pimd_step_config_x *cfgs, int n, int nsteps) {
  if (!pimds || !cfgs || n <= 0) return -1;
  if (nsteps <= 0) return 0;
  std::vector<BatchItem> items(n);
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < i; ++j)
      if (pimds[j] == pimds[i] || pimds[j]->sim == pimds[i]->sim) {
        fprintf(stderr, "xmd: a batch lists one system twice\n");
        return -1;
      }
    StepPlan sp;
    if (plan_of(pimds[i], &cfgs[i], sp)) return -1;
    if (autotune_if_needed(sp.p, sp.pair_kind)) return -1;
    items[i] = item_of(sp);
  }
  const Lane L = main_lane();
  Runtime &R = rt();
  const bool graph = cfgs[0].use_graph != 0;
  if (graph) {
    // the kept graph of the previous call, when nothing it depends on changed
    const std::vector<double> key = batch_key(pimds, n);
    bool same = g_bcache.exec && g_bcache.pimds.size() == (size_t)n && g_bcache.key == key &&
                g_bcache.epoch == pair_alloc_epoch();
    for (int i = 0; same && i < n; ++i)
      same = g_bcache.pimds[i] == pimds[i] && !memcmp(&g_bcache.cfgs[i], &cfgs[i], sizeof(md_pimd_step_config_x));
    if (same) {
      for (int s = 0; s < nsteps; ++s) XMD_CHECK(hipGraphLaunch(g_bcache.exec, R.s0));
      XMD_CHECK(hipStreamSynchronize(R.s0));
      return exch_batch_status(g_bcache.B->exch);
    }
    batch_cache_drop();
  }
  Batch *B = new Batch();
  B->bead = bead_batch_create(items);
  B->exch = exch_batch_create(items);
  B->pair = pair_batch_create(items);
  if (!B->bead || !B->exch || !B->pair) {
    delete B;
    return -1;
  }
  // The first step runs eagerly; the rest replay one captured graph of a
  // whole batch step (or run eagerly when the configs ask for no graph).
  if (enqueue_batch(*B, L)) {
    delete B;
    return -1;
  }
  XMD_CHECK(hipStreamSynchronize(R.s0));
  if (!graph) {
    int rc = 0;
    for (int s = 1; s < nsteps && !rc; ++s) rc = enqueue_batch(*B, L);
    if (!rc && hipStreamSynchronize(R.s0) != hipSuccess) rc = -1;
    if (!rc) rc = exch_batch_status(B->exch);
    delete B;
    return rc;
  }
  hipGraph_t g;
  hipGraphExec_t exec;
  XMD_CHECK(hipStreamBeginCapture(R.s0, hipStreamCaptureModeGlobal));
  int rc = enqueue_batch(*B, L);
  hipError_t ec = hipStreamEndCapture(R.s0, &g);
  if (rc || ec != hipSuccess) {
    fprintf(stderr, "xmd: batch capture failed\n");
    delete B;
    return -1;
  }
  XMD_CHECK(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0));
  XMD_CHECK(hipGraphDestroy(g));
  g_bcache.pimds.assign(pimds, pimds + n);
  g_bcache.cfgs.assign(cfgs, cfgs + n);
  g_bcache.key = batch_key(pimds, n);
  g_bcache.epoch = pair_alloc_epoch();
  g_bcache.B = B;
  g_bcache.exec = exec;
  for (int s = 1; s < nsteps; ++s) XMD_CHECK(hipGraphLaunch(exec, R.s0));
  XMD_CHECK(hipStreamSynchronize(R.s0));
  return exch_batch_status(B->exch);
}

}  // extern "C"
