// Device runtime and host<->device residency of the engine's objects.
//
// One HIP device, one main stream.  Objects are uploaded lazily on their
// first GPU use and stay resident; md_simulation_sync_host_x /
// md_pimd_sync_host_x copy the device state back into the reference-layout
// host mirrors (read_only = 0 also returns authority to the host, as in
// md_util_x.c:481-551 / md_pimd_x.c:271-328).
#include <cmath>
#include <cstring>
#include <unordered_map>

#include "xmd_internal.h"

namespace xmd {

static Runtime g_rt;
int g_nhc_len = MD_NHC_LENGTH_X;

Runtime &rt() { return g_rt; }

int ensure_runtime() {
  if (g_rt.ready) return 0;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    fprintf(stderr, "xmd: no HIP device available; the engine has no CPU path\n");
    return -1;
  }
  XMD_CHECK(hipSetDevice(g_rt.device));
  XMD_CHECK(hipStreamCreateWithFlags(&g_rt.s0, hipStreamNonBlocking));
  XMD_CHECK(hipStreamCreateWithFlags(&g_rt.s1, hipStreamNonBlocking));
  XMD_CHECK(hipStreamCreateWithFlags(&g_rt.s2, hipStreamNonBlocking));
  XMD_CHECK(hipEventCreateWithFlags(&g_rt.fork, hipEventDisableTiming));
  XMD_CHECK(hipEventCreateWithFlags(&g_rt.join, hipEventDisableTiming));
  XMD_CHECK(hipEventCreateWithFlags(&g_rt.fork2, hipEventDisableTiming));
  XMD_CHECK(hipEventCreateWithFlags(&g_rt.join2, hipEventDisableTiming));
  hipDeviceProp_t prop;
  XMD_CHECK(hipGetDeviceProperties(&prop, g_rt.device));
  g_rt.num_cus = prop.multiProcessorCount;
  if (noalloc_init(g_rt.device)) return -1;
  g_rt.ready = true;
  return 0;
}

static std::unordered_map<const void *, DevSim *> g_sims;
static std::unordered_map<const void *, DevPimd *> g_pimds;

DevSim *sim_state(md_simulation_t_x *sim, bool create) {
  auto it = g_sims.find(sim);
  if (it != g_sims.end()) return it->second;
  if (!create) return nullptr;
  DevSim *d = new DevSim();
  d->host = sim;
  d->nb = sim->N;
  d->N = sim->N;
  d->P = 1;
  g_sims[sim] = d;
  return d;
}

DevPimd *pimd_state(md_pimd_t_x *pimd, bool create) {
  auto it = g_pimds.find(pimd);
  if (it != g_pimds.end()) return it->second;
  if (!create) return nullptr;
  DevPimd *p = new DevPimd();
  p->host = pimd;
  p->N = pimd->N;
  p->P = pimd->P;
  p->ds = sim_state(pimd->sim, true);
  g_pimds[pimd] = p;
  return p;
}

int set_sim_layout(md_simulation_t_x *sim, int N, int P) {
  DevSim *d = sim_state(sim, true);
  if (N * P != sim->N) return -1;
  if (d->dev_valid && (d->N != N || d->P != P)) {
    if (!md_simulation_sync_host_x(sim, 0)) return -1;
  }
  d->N = N;
  d->P = P;
  return 0;
}

// ------------------------------------------------------------- kernels

// Reference bead index b = l*P + j (0-based) -> device slot j*N + l.
__device__ __forceinline__ int slot_of(int b, int N, int P) { return (b % P) * N + b / P; }

__global__ void k_beads_to_device(const md_particle_t_x *__restrict__ in, double *X, double *V, double *F,
                                  double *M, int nb, int N, int P) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const md_particle_t_x p = in[b];
  int s = slot_of(b, N, P);
  M[s] = p.m;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    X[d * nb + s] = p.x[d];
    V[d * nb + s] = p.v[d];
    F[d * nb + s] = p.f[d];
  }
}

__global__ void k_beads_to_host(md_particle_t_x *__restrict__ out, const double *X, const double *V,
                                const double *F, const double *M, int nb, int N, int P) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  int s = slot_of(b, N, P);
  md_particle_t_x p;
  p.m = M[s];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    p.x[d] = X[d * nb + s];
    p.v[d] = V[d * nb + s];
    p.f[d] = F[d * nb + s];
  }
  out[b] = p;
}

// Chains: massive mode stores chain q (= reference degree of freedom q) at
// device DOF index d*nb + slot(q/3); general mode keeps chain order.
__global__ void k_chains_to_device(const md_nhc_t_x *__restrict__ in, double *th, double *vth, double *Q,
                                   int Nf, int nb, int N, int P, int massive, int M) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Nf) return;
  int e = massive ? (i % 3) * nb + slot_of(i / 3, N, P) : i;
  for (int k = 0; k < M; ++k) {
    th[k * Nf + e] = in[i].theta[k];
    vth[k * Nf + e] = in[i].vtheta[k];
    Q[k * Nf + e] = in[i].Q[k];
  }
}

__global__ void k_chains_to_host(md_nhc_t_x *__restrict__ out, const double *th, const double *vth,
                                 const double *Q, int Nf, int nb, int N, int P, int massive, int M) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Nf) return;
  int e = massive ? (i % 3) * nb + slot_of(i / 3, N, P) : i;
  for (int k = 0; k < M; ++k) {
    out[i].theta[k] = th[k * Nf + e];
    out[i].vtheta[k] = vth[k * Nf + e];
    out[i].Q[k] = Q[k * Nf + e];
  }
}

// Deterministic sum: one block, fixed per-thread strides, tree in LDS.
__global__ __launch_bounds__(1024) void k_sum(const double *__restrict__ a, int n, double *out) {
  __shared__ double red[1024];
  double s = 0;
  for (int i = threadIdx.x; i < n; i += 1024) s += a[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

int batch_table_make(BatchTable &t, const void *reps, size_t rep_bytes, const std::vector<int> &blocks) {
  t.nrep = (int)blocks.size();
  std::vector<int> start(t.nrep + 1, 0);
  for (int r = 0; r < t.nrep; ++r) start[r + 1] = start[r] + blocks[r];
  t.total = start[t.nrep];
  XMD_CHECK(hipMalloc(&t.reps, rep_bytes * (t.nrep > 0 ? t.nrep : 1)));
  XMD_CHECK(hipMalloc(&t.start, start.size() * sizeof(int)));
  if (t.nrep > 0) XMD_CHECK(hipMemcpy(t.reps, reps, rep_bytes * t.nrep, hipMemcpyHostToDevice));
  XMD_CHECK(hipMemcpy(t.start, start.data(), start.size() * sizeof(int), hipMemcpyHostToDevice));
  return 0;
}

void batch_table_free(BatchTable &t) {
  if (t.reps) (void)hipFree(t.reps);
  if (t.start) (void)hipFree(t.start);
  t = BatchTable();
}

int device_sum(const double *src, int n, double *out) {
  static double *d_out = nullptr;
  if (!d_out) XMD_CHECK(hipMalloc(&d_out, sizeof(double)));
  hipLaunchKernelGGL(k_sum, dim3(1), dim3(1024), 0, rt().s0, src, n, d_out);
  XMD_CHECK(hipGetLastError());
  XMD_CHECK(hipMemcpyAsync(out, d_out, sizeof(double), hipMemcpyDeviceToHost, rt().s0));
  XMD_CHECK(hipStreamSynchronize(rt().s0));
  return 0;
}

// -------------------------------------------------------- upload/download

static int alloc_sim(DevSim *d) {
  md_simulation_t_x *s = d->host;
  if (d->X && d->nb == s->N && d->Nf == s->Nf) return 0;
  d->nb = s->N;
  d->Nf = s->Nf;
  size_t nb = (size_t)d->nb;
  XMD_CHECK(hipMalloc(&d->X, 3 * nb * sizeof(double)));
  XMD_CHECK(hipMalloc(&d->V, 3 * nb * sizeof(double)));
  XMD_CHECK(hipMalloc(&d->F, 3 * nb * sizeof(double)));
  XMD_CHECK(hipMalloc(&d->M, nb * sizeof(double)));
  XMD_CHECK(hipMalloc(&d->stage, nb * sizeof(md_particle_t_x)));
  XMD_CHECK(hipMalloc(&d->th, (size_t)MD_NHC_LENGTH_X * d->Nf * sizeof(double)));
  XMD_CHECK(hipMalloc(&d->vth, (size_t)MD_NHC_LENGTH_X * d->Nf * sizeof(double)));
  XMD_CHECK(hipMalloc(&d->Q, (size_t)MD_NHC_LENGTH_X * d->Nf * sizeof(double)));
  XMD_CHECK(hipMalloc(&d->gf, d->Nf * sizeof(int)));
  XMD_CHECK(hipMalloc(&d->gf0, d->Nf * sizeof(int)));
  XMD_CHECK(hipMalloc(&d->gf2, d->Nf * sizeof(int)));
  XMD_CHECK(hipMalloc(&d->nstage, (size_t)d->Nf * sizeof(md_nhc_t_x)));
  d->red_len = 4096;
  XMD_CHECK(hipMalloc(&d->red, d->red_len * sizeof(double)));
  return 0;
}

int ensure_device(md_simulation_t_x *sim) {
  if (ensure_runtime()) return -1;
  DevSim *d = sim_state(sim, true);
  if (d->dev_valid) return 0;
  if (alloc_sim(d)) return -1;
  hipStream_t st = rt().s0;
  int nb = d->nb;
  XMD_CHECK(hipMemcpyAsync(d->stage, sim->particles, nb * sizeof(md_particle_t_x), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_beads_to_device, dim3((nb + 255) / 256), dim3(256), 0, st,
                     (const md_particle_t_x *)d->stage, d->X, d->V, d->F, d->M, nb, d->N, d->P);
  XMD_CHECK(hipGetLastError());
  // chains: detect massive thermostatting (one chain per degree of freedom)
  int Nf = sim->Nf;
  bool massive = (Nf == 3 * nb);
  std::vector<int> f(Nf), f0(Nf), f2(Nf);
  int acc = 0;
  for (int i = 0; i < Nf; ++i) {
    f[i] = sim->nhcs[i].f;
    f0[i] = acc;
    f2[i] = (i == Nf - 1) ? 3 * nb - acc : f[i];
    acc += f[i];
    if (f[i] != 1) massive = false;
  }
  d->massive = massive;
  d->Mlen = g_nhc_len;
  d->T = sim->T;
  XMD_CHECK(hipMemcpyAsync(d->gf, f.data(), Nf * sizeof(int), hipMemcpyHostToDevice, st));
  XMD_CHECK(hipMemcpyAsync(d->gf0, f0.data(), Nf * sizeof(int), hipMemcpyHostToDevice, st));
  XMD_CHECK(hipMemcpyAsync(d->gf2, f2.data(), Nf * sizeof(int), hipMemcpyHostToDevice, st));
  XMD_CHECK(hipMemcpyAsync(d->nstage, sim->nhcs, Nf * sizeof(md_nhc_t_x), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_chains_to_device, dim3((Nf + 255) / 256), dim3(256), 0, st, (const md_nhc_t_x *)d->nstage,
                     d->th, d->vth, d->Q, Nf, nb, d->N, d->P, massive ? 1 : 0, MD_NHC_LENGTH_X);
  XMD_CHECK(hipGetLastError());
  // uniform-mass check per slice (the Ewald structure-factor path needs it)
  d->uniform_mass = true;
  for (int b = 1; b < nb && d->uniform_mass; ++b)
    if (sim->particles[b].m != sim->particles[0].m) d->uniform_mass = false;
  // the staging copies must complete before the host buffers may change
  XMD_CHECK(hipStreamSynchronize(st));
  d->dev_valid = true;
  return 0;
}

int download_sim(DevSim *d) {
  md_simulation_t_x *sim = d->host;
  hipStream_t st = rt().s0;
  int nb = d->nb, Nf = d->Nf;
  hipLaunchKernelGGL(k_beads_to_host, dim3((nb + 255) / 256), dim3(256), 0, st, (md_particle_t_x *)d->stage,
                     d->X, d->V, d->F, d->M, nb, d->N, d->P);
  XMD_CHECK(hipGetLastError());
  XMD_CHECK(hipMemcpyAsync(sim->particles, d->stage, nb * sizeof(md_particle_t_x), hipMemcpyDeviceToHost, st));
  // chain slots beyond the integrated length keep their host values
  XMD_CHECK(hipMemcpyAsync(d->nstage, sim->nhcs, Nf * sizeof(md_nhc_t_x), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_chains_to_host, dim3((Nf + 255) / 256), dim3(256), 0, st, (md_nhc_t_x *)d->nstage, d->th,
                     d->vth, d->Q, Nf, nb, d->N, d->P, d->massive ? 1 : 0, MD_NHC_LENGTH_X);
  XMD_CHECK(hipGetLastError());
  XMD_CHECK(hipMemcpyAsync(sim->nhcs, d->nstage, Nf * sizeof(md_nhc_t_x), hipMemcpyDeviceToHost, st));
  XMD_CHECK(hipStreamSynchronize(st));
  return 0;
}

static int alloc_pimd(DevPimd *p) {
  if (p->ENk) return 0;
  size_t N = (size_t)p->N;
  size_t tri = N * (N + 1) / 2;
  XMD_CHECK(hipMalloc(&p->ENk, tri * sizeof(double)));
  XMD_CHECK(hipMalloc(&p->ET, tri * sizeof(double)));
  XMD_CHECK(hipMalloc(&p->ENk2, tri * sizeof(double)));
  XMD_CHECK(hipMalloc(&p->Eint, N * sizeof(double)));
  XMD_CHECK(hipMalloc(&p->Cpre, 2 * (N + 1) * sizeof(double)));
  XMD_CHECK(hipMalloc(&p->C2pre, 2 * (N + 1) * sizeof(double)));
  XMD_CHECK(hipMalloc(&p->VBN, (N + 1) * sizeof(double)));
  XMD_CHECK(hipMalloc(&p->Lf, (N + 1) * sizeof(double)));
  XMD_CHECK(hipMalloc(&p->Rb, (N + 1) * sizeof(double)));
  XMD_CHECK(hipMalloc(&p->work, 8 * (N + 1) * sizeof(double) + 64 * 1024 * sizeof(double)));
  XMD_CHECK(hipMalloc(&p->Fx, 6 * N * sizeof(double)));
  // recursion state: values and accumulators 10(N+1), then per direction and
  // 64-row block the in-block (64*64) and previous-block (64*64)
  // coefficients, then the chain offsets (65) per direction and block
  const size_t nblk = (N + 63) / 64;
  XMD_CHECK(hipMalloc(&p->rec, (10 * (size_t)(N + 1) + 2 * nblk * (64 * 64 + 64 * 64 + 65)) * sizeof(double)));
  XMD_CHECK(hipMalloc(&p->recsync, (4 * nblk + 2) * sizeof(int)));
  XMD_CHECK(hipMemset(p->recsync, 0, (4 * nblk + 2) * sizeof(int)));
  XMD_CHECK(hipMalloc(&p->done, sizeof(int)));
  XMD_CHECK(hipMemset(p->done, 0, sizeof(int)));
  // ln n, n = 0..N, in the first N+1 doubles of the scratch (constant)
  return init_lnn(p);
}

int ensure_pimd_device(md_pimd_t_x *pimd) {
  DevPimd *p = pimd_state(pimd, true);
  if (ensure_runtime()) return -1;
  if (pimd->P < 2) {
    fprintf(stderr, "xmd: PIMD objects need P >= 2 beads\n");
    return -1;
  }
  if (set_sim_layout(pimd->sim, pimd->N, pimd->P)) return -1;
  if (ensure_device(pimd->sim)) return -1;
  if (alloc_pimd(p)) return -1;
  return 0;
}

int download_pimd(DevPimd *p) {
  md_pimd_t_x *h = p->host;
  hipStream_t st = rt().s0;
  size_t tri = (size_t)p->N * (p->N + 1) / 2;
  if (p->enk_dev) {
    XMD_CHECK(hipMemcpyAsync(h->ENk[0], p->ENk, tri * sizeof(double), hipMemcpyDeviceToHost, st));
    XMD_CHECK(hipMemcpyAsync(h->Eint, p->Eint, p->N * sizeof(double), hipMemcpyDeviceToHost, st));
  }
  if (p->enk2_dev) XMD_CHECK(hipMemcpyAsync(h->ENk2[0], p->ENk2, tri * sizeof(double), hipMemcpyDeviceToHost, st));
  if (p->vbn_dev) XMD_CHECK(hipMemcpyAsync(h->VBN, p->VBN, (p->N + 1) * sizeof(double), hipMemcpyDeviceToHost, st));
  XMD_CHECK(hipStreamSynchronize(st));
  return 0;
}

}  // namespace xmd

using namespace xmd;

extern "C" {

// md_util_x.c:136-351: seeds the generator and brings up the device.
int md_init_setup_x(FILE *out, void *config) {
  md_init_config_x *c = (md_init_config_x *)config;
  md_set_seed_x(c->seed);
  if (ensure_runtime()) return -1;
  if (out) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, rt().device) == hipSuccess)
      fprintf(out, "Device: %s (%s), %d CUs\n", prop.name, prop.gcnArchName, prop.multiProcessorCount);
    fprintf(out, "Initialization completed\n");
  }
  srand(c->seed);
  return 0;
}

md_simulation_t_x *md_simulation_sync_host_x(md_simulation_t_x *sim, int read_only) {
  DevSim *d = sim_state(sim, false);
  if (!d || !d->dev_valid) return sim;
  if (download_sim(d)) return NULL;
  if (!read_only) d->dev_valid = false;
  return sim;
}

md_pimd_t_x *md_pimd_sync_host_x(md_pimd_t_x *pimd, int read_only) {
  DevPimd *p = pimd_state(pimd, false);
  if (!p) return pimd;
  if (download_pimd(p)) return NULL;
  if (!read_only) p->enk_dev = p->enk2_dev = p->vbn_dev = p->back_dev = false;
  return pimd;
}

// md_pimd_x.c:330-513 (host mirrors as in the reference, device mirrors lazily).
md_pimd_t_x *md_alloc_pimd_x(md_simulation_t_x *sim, int N, int P, md_num_t_x vi, md_num_t_x T) {
  if (N <= 0 || P <= 0 || N * P != sim->N) return NULL;
  md_pimd_t_x *p = (md_pimd_t_x *)calloc(1, sizeof(md_pimd_t_x));
  if (!p) return NULL;
  p->sim = sim;
  p->N = N;
  p->P = P;
  p->vi = vi;
  sim->T = T;
  p->beta = 1.0 / (MD_kB_X * T);
  p->omegaP = sqrt((md_num_t_x)P) / (p->beta * MD_hBar_X);
  p->VBN = (md_num_t_x *)calloc(N + 1, sizeof(md_num_t_x));
  p->minE = (md_num_t_x *)calloc(N + 1, sizeof(md_num_t_x));
  p->Eint = (md_num_t_x *)calloc(N, sizeof(md_num_t_x));
  p->Evir = (md_num_t_x *)calloc(N + 1, sizeof(md_num_t_x));
  p->ENk_count = N * (N + 1) / 2;
  md_num_t_x ***tri[3] = {&p->ENk, &p->ENk2, &p->ENk3};
  for (int t = 0; t < 3; ++t) {
    md_num_t_x **rows = (md_num_t_x **)malloc(sizeof(md_num_t_x *) * N);
    md_num_t_x *base = (md_num_t_x *)calloc((size_t)p->ENk_count, sizeof(md_num_t_x));
    if (!rows || !base) return NULL;
    for (int n = 1, off = 0; n <= N; off += n, ++n) rows[n - 1] = base + off;
    *tri[t] = rows;
  }
  p->pcount_in = N * (N + 1) / 2;
  p->pcount_ex = N * (N - 1) / 2;
  p->pair_in = (md_index_pair_t_x *)malloc(sizeof(md_index_pair_t_x) * p->pcount_in);
  p->pair_ex = (md_index_pair_t_x *)malloc(sizeof(md_index_pair_t_x) * (p->pcount_ex > 0 ? p->pcount_ex : 1));
  p->Eindices = (int *)malloc(sizeof(int) * (size_t)N * N);
  p->pbW = (int *)calloc((size_t)N * P * MD_DIMENSION_X, sizeof(int));
  if (!p->VBN || !p->minE || !p->Eint || !p->Evir || !p->pair_in || !p->pair_ex || !p->Eindices || !p->pbW)
    return NULL;
  for (int i = 0, a = 0, b = 0; i < N; ++i)
    for (int j = 0; j <= i; ++j) {
      p->pair_in[a].i = i;
      p->pair_in[a].j = j;
      p->Eindices[i * N + j] = a++;
      if (j < i) {
        p->pair_ex[b].i = i;
        p->pair_ex[b++].j = j;
      }
    }
  p->cut_Pj = -1;
  p->mc_par = -1;
  p->pbc = 0;
  p->worm_index = -1;
  p->old_worm_index = -1;
  if (set_sim_layout(sim, N, P)) return NULL;
  pimd_state(p, true);
  return p;
}

void md_pimd_change_temperature_x(md_pimd_t_x *pimd, md_num_t_x T) {
  pimd->sim->T = T;
  pimd->beta = 1.0 / (MD_kB_X * T);
  pimd->omegaP = sqrt((md_num_t_x)pimd->P) / (pimd->beta * MD_hBar_X);
  DevSim *d = sim_state(pimd->sim, false);
  if (d) d->T = T;
}

}  // extern "C"
