// Internal declarations of the MI355X PIMD engine (not installed).
//
// Device state model
// ------------------
// The public structs keep the reference's host layout (AoS beads,
// particle-major).  Each md_simulation_t_x / md_pimd_t_x / md_stats_t_x gets a
// side record here that owns its device mirror.  The device copy is
// authoritative once a GPU entry point has touched the object and until
// md_simulation_sync_host_x(sim, 0) hands authority back to the host, which
// is the same contract as the reference's OpenCL build
// (md_util_x.c:481-551).
//
// Device layout (structure of arrays, FP64):
//   bead (l, j) (1-based) lives at slot s = (j-1)*N + (l-1)  ("slice-major"),
//   coordinate d of slot s at X[d*nb + s]; same for V and F.  One bead slice
//   is a contiguous run of N slots, which is what the pair kernels tile over.
//   A simulation without a PIMD object uses N = nb, P = 1 (identity order).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

extern "C" {
#include "md_pimd_x.h"
}

#define XMD_CHECK(call)                                                                    \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess) {                                                                \
      fprintf(stderr, "xmd: HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, __LINE__); \
      return -1;                                                                           \
    }                                                                                      \
  } while (0)

#define XMD_CHECK_NULL(call)                                                               \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess) {                                                                \
      fprintf(stderr, "xmd: HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, __LINE__); \
      return NULL;                                                                         \
    }                                                                                      \
  } while (0)

namespace xmd {

// ------------------------------------------------------------------ runtime
struct Runtime {
  bool ready = false;
  int device = 0;
  hipStream_t s0 = nullptr;   // main stream (all API work is ordered on it)
  hipStream_t s1 = nullptr;   // exchange branch of the fused step
  hipStream_t s2 = nullptr;   // reciprocal-space branch of the fused step
  hipEvent_t fork = nullptr, join = nullptr, fork2 = nullptr, join2 = nullptr;
  int num_cus = 0;
};
Runtime &rt();
int ensure_runtime();

// Run-time chain length of the NHC integrator (even, <= MD_NHC_LENGTH_X).
extern int g_nhc_len;

// -------------------------------------------------------------- device sim
struct DevSim {
  md_simulation_t_x *host = nullptr;
  int nb = 0;            // beads
  int N = 0, P = 1;      // layout (particles per slice, slices)
  bool dev_valid = false;
  double *X = nullptr, *V = nullptr, *F = nullptr, *M = nullptr;  // SoA, 3*nb / nb
  void *stage = nullptr;                                          // AoS staging (nb*80 B)
  // thermostat chains
  int Nf = 0;
  bool massive = false;   // one chain per degree of freedom, in device DOF order
  int Mlen = 8;
  double *th = nullptr, *vth = nullptr, *Q = nullptr;  // [k*Nf + chain]
  int *gf = nullptr;      // general mode: chain target DOF count (nhcs[i].f)
  int *gf0 = nullptr, *gf2 = nullptr;  // general mode: DOF range per chain (reference order)
  void *nstage = nullptr; // AoS chain staging (Nf*200 B)
  double T = 0;
  bool uniform_mass = true;
  double *red = nullptr;  // reduction scratch (partials + result)
  int red_len = 0;
};

struct DevPimd {
  md_pimd_t_x *host = nullptr;
  DevSim *ds = nullptr;
  int N = 0, P = 0;
  double *ENk = nullptr;    // reference row layout: E_n^(k) at (n-1)n/2 + k-1
  double *ET = nullptr;     // column layout: cycle of particles m+1..m+k at off(m) + k-1
  double *ENk2 = nullptr;
  double *Eint = nullptr;   // N
  double *Cpre = nullptr;   // chain prefix sums, double-double (2*(N+1))
  double *C2pre = nullptr;  // ENk2 link prefix sums, double-double (2*(N+1))
  double *VBN = nullptr;    // N+1
  double *Lf = nullptr;     // forward  L(n) = -beta V_B^(n), N+1
  double *Rb = nullptr;     // backward R(m) = -beta Vt^(m),  N+1
  double *Gm = nullptr;     // optional connection matrix (debug/tests), N*N
  double *work = nullptr;   // estimator scratch
  double *Fx = nullptr;     // exchange forces of the end beads, [slice P-1, slice 0][dim][particle] (6N)
  double *rec = nullptr;    // recursion state: values, accumulators, block coefficients
  int *recsync = nullptr;   // recursion sync words: 4*nblk unused, error word, injected-timeout count
  int *done = nullptr;      // finished-workgroup count of k_chain_scan (zero between launches)
  void *pair_bufs = nullptr;  // pair-kernel tile schedule, partial sums and k-space tables (pair_forces.hip)
  void *step_exec = nullptr;  // executable step graph of the last md_pimd_run_steps_x (hipGraphExec_t)
  void *sur = nullptr;        // Ewald surrogate attached for md_pimd_run_steps_surrogate_x (Surrogate)
  int sur_mode = 0;           // 0 exact reciprocal space, 1 surrogate, 2 judge (exact used, surrogate compared)
  bool enk_dev = false, enk2_dev = false, vbn_dev = false, back_dev = false;
  bool enk_host_stale = false;
  unsigned enk_version = 0, back_version = 0;
  int last_image = 0;
};

DevSim *sim_state(md_simulation_t_x *sim, bool create);
DevPimd *pimd_state(md_pimd_t_x *pimd, bool create);
void forget_sim(md_simulation_t_x *sim);
int set_sim_layout(md_simulation_t_x *sim, int N, int P);
int ensure_device(md_simulation_t_x *sim);          // upload if host is authoritative
int ensure_pimd_device(md_pimd_t_x *pimd);
int download_sim(DevSim *d);                         // device -> host mirror
int download_pimd(DevPimd *p);

// reductions: sum of n doubles on device into *out (host), deterministic
int device_sum(const double *src, int n, double *out);

inline double box_L(const md_simulation_t_x *sim, int d) {
  if (sim->box && sim->box->type == MD_ND_RECT_BOX_X && sim->box->box)
    return ((nd_rect_t_x *)sim->box->box)->L[d];
  return 0.0;
}

// ------------------------------------------------------------ kernel entry
// Launchers (enqueue on `st`; return 0 / -1).
int launch_reset_force(DevSim *d, hipStream_t st);
int launch_trap_force(DevSim *d, double w, const double *c, double invP, hipStream_t st);
int launch_pbc(DevSim *d, const double *L, hipStream_t st);
int launch_nhc(DevSim *d, int half, double h, hipStream_t st);
int launch_ring_springs(DevPimd *p, int image, const double *L, hipStream_t st);

// PK_EWALD_R2 (energies only): -erf(alpha r)/r of the minimum image, the
// term md_coulomb_3d_ewald_energy_R2_x adds to the R split.
enum PairKind { PK_GAUSS = 0, PK_PGAUSS = 1, PK_EWALD = 2, PK_EWALD_R = 3, PK_EWALD_NI = 4, PK_LJ = 5, PK_HE = 6,
                PK_EWALD_R2 = 7 };

// helium potential parameters (md_simulation_x.c:365-383), set by md_set_helium_parameters_x
extern "C" md_num_t_x md_he_eps_x, md_he_A_x, md_he_alpha_x, md_he_C6_x, md_he_C8_x, md_he_C10_x, md_he_D_x,
    md_he_rm_x;
int pair_kind_of(md_pair_force_t_x pf);
int pair_energy_kind_of(md_pair_energy_t_x pe);
int launch_pair_force(DevPimd *p, int kind, hipStream_t st);
// Pair and exchange forces computed in parts, summed into F: the real-space
// tile sums over P, then the reciprocal forces, then (Fx given) the stored
// exchange forces of the end beads (slices P-1 and 0).  Fpart == nullptr:
// nothing to add.
struct ForceSums {
  const double *Fpart = nullptr, *Frec = nullptr, *Fx = nullptr;
  int T = 0, N = 0, P = 0;
};

// F[k] plus the parts, in the order above (so the rounding is that of the
// unsplit kernels: k_reduce_tiles, k_recip_force_mfma, k_add_exch in turn).
__device__ __forceinline__ double combined_force(const double *__restrict__ F, const ForceSums &c, int n3, int k) {
  double s = 0;
  for (int q = 0; q < c.T; ++q) s += c.Fpart[(size_t)q * n3 + k];
  double f = F[k] + s / (double)c.P;
  f += c.Frec[k];
  if (c.Fx) {
    const int nb = c.N * c.P, d = k / nb, i = k - d * nb;
    if (i >= (c.P - 1) * c.N) f += c.Fx[d * c.N + (i - (c.P - 1) * c.N)];
    if (i < c.N) f += c.Fx[(3 + d) * c.N + i];
  }
  return f;
}

// Makes (host side, nothing launched) every buffer the step's pair kernels of
// this kind need; 1 when anything was allocated, 0 when all existed, -1 on
// error.
int pair_prepare(DevPimd *p, int kind);
// Count of device (re)allocations made by the pair kernels' buffers so far
// (a captured graph that names them stays valid while it is unchanged).
unsigned pair_alloc_epoch();
// Real-space Ewald launch choice autotuning (pair_forces.hip): with it on,
// tune the object's (N, P) bucket before first use (never during capture).
int autotune_if_needed(DevPimd *p, int kind);
// Ewald forces in parts: launch_pair_ewald_fork forks s2 from st (through
// `fork`), queues there the real-space tile kernel (per-tile partial forces;
// with the learned stand-in in use, then its kernels) and records `join`;
// launch_pair_kspace, called later in st's order, queues the phase tables,
// structure factors and reciprocal forces (into a side buffer) on st.  All of
// these read positions only.  After `join` (and the exchange branch)
// launch_combine_forces adds the real-space, reciprocal and, when Fx is
// given, exchange end-bead forces into F in the order of the unsplit
// kernels.  The fork returns 1 when it applies (Ewald with the table
// kernels), 0 when nothing was launched (use launch_pair_force), -1 on error.
int launch_pair_ewald_fork(DevPimd *p, int kind, hipStream_t st, hipStream_t s2, hipEvent_t fork, hipEvent_t join);
int launch_pair_kspace(DevPimd *p, hipStream_t st);
int launch_combine_forces(DevPimd *p, const double *Fx, hipStream_t st);
// The parts launch_combine_forces adds, for a fused consumer of F.
int force_sums(DevPimd *p, const double *Fx, ForceSums *out);
int launch_pair_real(DevPimd *p, int kind, hipStream_t st);
// Exact reciprocal-space Ewald terms of the step's plan (alpha, k set of
// md_coulomb_3d_ewald_force_x): forces (accelerations / P, [3][nb]) into Frec
// and each bead's share of its slice's reciprocal energy into Erec [nb];
// alpha_out (may be null) receives the split.  ewald_split_alpha: that split.
int launch_recip_labels(DevPimd *p, double *Frec, double *Erec, hipStream_t st, double *alpha_out);
double ewald_split_alpha(DevPimd *p);
// The same reciprocal terms restricted to the low shells |G| <= nlow / L_max
// (G = m / L): forces stored in Frec, per-bead energies in Erec.
int launch_recip_shells(DevPimd *p, int nlow, double *Frec, double *Erec, hipStream_t st);
// Reciprocal-space forces of the step from the attached surrogate (sur_mode
// 1) or exact with the surrogate's error against them recorded (sur_mode 2),
// into Frec ([3][nb], accelerations / P), on st (surrogate.hip).
int surrogate_kspace(DevPimd *p, double *Frec, hipStream_t st);
int launch_pair_recip(DevPimd *p, int kind, hipStream_t st);
bool pair_has_recip(int kind);
int pair_energy(DevPimd *p, int kind, double *out);

int launch_enk(DevPimd *p, int image, hipStream_t st);
// launch_enk plus the recursion set-up of both directions in two kernels;
// launch_recursion_solve then runs both recursions (together: launch_enk
// followed by launch_recursion(p, true, true)).
int launch_enk_prep(DevPimd *p, int image, hipStream_t st);
int launch_recursion_solve(DevPimd *p, hipStream_t st);
int init_lnn(DevPimd *p);  // ln n table at p->work[0..N], once per object
// mid (optional): recorded between the set-up and the persistent solve.
int launch_recursion(DevPimd *p, bool fwd, bool bwd, hipStream_t st, hipEvent_t mid = nullptr);
// stored: write the end-bead forces to p->Fx instead of adding them to F;
// launch_add_exchange_force adds them later (both on the caller's ordering).
int launch_exchange_force(DevPimd *p, int image, hipStream_t st, bool stored);
int launch_add_exchange_force(DevPimd *p, hipStream_t st);
int connection_matrix(DevPimd *p, double *host_G);
int backward_potentials(DevPimd *p, double *host_Vt);
int bead_sum(DevSim *d, int mode, const double *trap_c, double w, const double *L, double *out);
extern double g_ewald_xcut;
extern double g_ewald_split_aL;
// Error word of an object's persistent recursion solve: recursion_status
// reads and clears it (-1 when a hand-off timed out since the last read);
// recursion_status_all does so for every object with a recursion launched
// since its last read.  inject_recursion_timeouts: the next `count` forward
// solves of the object give up their first hand-off wait (test hook).
int recursion_status(DevPimd *p);
int recursion_status_all();
// Device address of the allocation-free recursion's scratch array for this
// device (once per device, at runtime set-up).
int noalloc_init(int device);
int inject_recursion_timeouts(DevPimd *p, int count);
// Range limit of the chain-frame block solve (test hook; <= 0: default).
void set_recursion_frame_limit(double limit);
// Serial stage profile: event recorded after the tabulated real-space Ewald
// kernel (nullptr: none), and whether a launch recorded it since it was set.
void set_real_space_probe(hipEvent_t e);
bool real_space_probe_hit();

// ------------------------------------------------------ batched replicas
// One launch per step stage covers every replica of a batch: each stage has
// a device table of per-replica arguments and a prefix table of per-replica
// block counts (batch_rep in xmd_device.h maps a block to its replica).
struct BatchItem {
  DevPimd *p = nullptr;
  double dt = 0;
  int pbc = 0, image = 0, trap = 0, pair_kind = -1;
  double L[3] = {0, 0, 0};
};

// Device copy of a host table plus its block prefix (grid = start[nrep]).
struct BatchTable {
  void *reps = nullptr;
  int *start = nullptr;
  int nrep = 0, total = 0;
};
int batch_table_make(BatchTable &t, const void *reps, size_t rep_bytes, const std::vector<int> &blocks);
void batch_table_free(BatchTable &t);

// Fused bead-local stages of one system's step (massive thermostats only).
// half 2 with cs (cs->Fpart set): F += the parts first (combined_force),
// stored, and used by the half-kick.  half 3: half 2 (with cs as for half 2)
// and then the next step's half 1 (+ wrap) in one kernel.
int launch_step_nhc(const BatchItem &it, int half, hipStream_t st, const ForceSums *cs = nullptr);
int launch_step_bead_forces(const BatchItem &it, hipStream_t st);

struct BeadBatch;  // integrator stages
BeadBatch *bead_batch_create(const std::vector<BatchItem> &items);
void bead_batch_free(BeadBatch *b);
int bead_batch_nhc(BeadBatch *b, int half, hipStream_t st);     // NHC half (+ wrap after half 1)
int bead_batch_forces(BeadBatch *b, hipStream_t st);            // reset, trap, interior springs

struct ExchBatch;  // cycle energies, recursions, exchange forces
ExchBatch *exch_batch_create(const std::vector<BatchItem> &items);
void exch_batch_free(ExchBatch *b);
int exch_batch_enk(ExchBatch *b, hipStream_t st);
int exch_batch_recursion(ExchBatch *b, hipStream_t st);
int exch_batch_force(ExchBatch *b, hipStream_t st);
int exch_batch_status(ExchBatch *b);

struct PairBatch;  // pair forces
PairBatch *pair_batch_create(const std::vector<BatchItem> &items);
void pair_batch_free(PairBatch *b);
int pair_batch_forces(PairBatch *b, hipStream_t st);

}  // namespace xmd
