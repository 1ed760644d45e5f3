// Integrator and per-bead kernels: Nose-Hoover-chain velocity Verlet halves
// (md_simulation_x.c:599-792), force reset (:421-472), periodic wrap
// (:794-856), harmonic trap per bead (md_pimd_x.c:1297-1391), interior ring
// springs (md_pimd_x.c:2690-2698) and the kinetic / chain-energy reductions
// (md_statistics_x.c:960-983).
//
// All of these are HBM-streaming kernels: one thread per degree of freedom
// (massive thermostats) or per chain, coalesced over the slice-major SoA.
#include <cmath>

#include "xmd_device.h"
#include "xmd_internal.h"

namespace xmd {

__device__ __forceinline__ int slot_of_ref(int b, int N, int P) { return (b % P) * N + b / P; }

__device__ __forceinline__ int img_count(double x, double L) {
  int c = (int)(fabs(x) / L);
  if (x > L) return -c;
  if (x < 0) return c + 1;
  return 0;
}

// md_simulation_x.c:70-81
__device__ __forceinline__ double min_image(double a, double L) {
  double s = a + L / 2;
  s += L * img_count(s, L);
  return s - L / 2;
}

__global__ void k_fill(double *a, double v, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

__global__ void k_trap(const double *__restrict__ X, double *__restrict__ F, int nb, double w, double c0, double c1,
                       double c2, double invP_unused, double P) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * nb) return;
  int d = i / nb;
  double c = d == 0 ? c0 : (d == 1 ? c1 : c2);
  double f = 0;
  f += -w * w * (X[i] - c);
  F[i] += f / P;
}

__global__ void k_pbc(double *__restrict__ X, int nb, double L0, double L1, double L2) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * nb) return;
  int d = i / nb;
  double L = d == 0 ? L0 : (d == 1 ? L1 : L2);
  double x = X[i];
  if (x < 0)
    x += L;
  else if (x > L)
    x -= L;
  X[i] = x;
}

// Interior springs of every ring polymer, beads j = 2..P-1 (1-based).
__global__ void k_ring_springs(const double *__restrict__ X, const double *__restrict__ M, double *__restrict__ F,
                               int N, int P, double om, int image, double L0, double L1, double L2) {
  int nb = N * P;
  int s = blockIdx.x * blockDim.x + threadIdx.x + N;  // skip slice 1
  if (s >= nb - N) return;                            // and slice P
  double m = M[s];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double L = d == 0 ? L0 : (d == 1 ? L1 : L2);
    double x = X[d * nb + s];
    double a = x - X[d * nb + s + N];
    double c = x - X[d * nb + s - N];
    if (image) {
      a = min_image(a, L);
      c = min_image(c, L);
    }
    double g = 0;
    g += m * om * om * a;
    g += m * om * om * c;
    F[d * nb + s] += -g / m;
  }
}

// ---------------------------------------------------------------- NHC

// One chain per degree of freedom.  Thread e owns DOF e (device order), its
// velocity, position, force and chain slots th/vth/Q[k*ndof + e].
template <int HALF>
__device__ __forceinline__ void nhc_massive_body(double *__restrict__ X, double *__restrict__ V, double f,
                                                 const double *__restrict__ M,
                                                 double *__restrict__ th, double *__restrict__ vth,
                                                 const double *__restrict__ Qa, int nb, int Mlen, double T, double h,
                                                 int e) {
  int ndof = 3 * nb;
  double m = M[e % nb];
  double v = V[e];
  double t[MD_NHC_LENGTH_X], vt[MD_NHC_LENGTH_X], Q[MD_NHC_LENGTH_X], a[MD_NHC_LENGTH_X];
#pragma unroll
  for (int k = 0; k < MD_NHC_LENGTH_X; ++k)
    if (k < Mlen) {
      t[k] = th[k * ndof + e];
      vt[k] = vth[k * ndof + e];
      Q[k] = Qa[k * ndof + e];
    }
  const int M2 = Mlen / 2;
  auto chain_force = [&]() {
    a[0] = (m * v * v - 1 * T) / Q[0];
#pragma unroll
    for (int k = 1; k < MD_NHC_LENGTH_X; ++k)
      if (k < Mlen) a[k] = (Q[k - 1] * vt[k - 1] * vt[k - 1] - T) / Q[k];
  };
  auto odd_update = [&]() {
#pragma unroll
    for (int j = 1; j <= MD_NHC_LENGTH_X / 2; ++j)
      if (j <= M2) {
        double up = (j == M2) ? 0.0 : vt[2 * j];
        vt[2 * j - 1] = vt[2 * j - 1] * exp(-0.5 * h * up) + 0.5 * h * a[2 * j - 1] * exp(-0.25 * h * up);
      }
  };
  if (HALF == 1) {
    chain_force();
    v = v * exp(-0.5 * h * vt[0]) + 0.5 * h * f * exp(-0.25 * h * vt[0]);
#pragma unroll
    for (int j = 1; j <= MD_NHC_LENGTH_X / 2; ++j)
      if (j <= M2) t[2 * j - 2] = t[2 * j - 2] + h * vt[2 * j - 2] / 2;
    odd_update();
    X[e] = X[e] + h * v;
#pragma unroll
    for (int j = 1; j <= MD_NHC_LENGTH_X / 2; ++j)
      if (j <= M2) t[2 * j - 1] = t[2 * j - 1] + h * vt[2 * j - 1];
    chain_force();
#pragma unroll
    for (int j = 1; j <= MD_NHC_LENGTH_X / 2; ++j)
      if (j <= M2)
        vt[2 * j - 2] = vt[2 * j - 2] * exp(-h * vt[2 * j - 1]) + h * a[2 * j - 2] * exp(-0.5 * h * vt[2 * j - 1]);
  } else {
    v = v * exp(-0.5 * h * vt[0]) + 0.5 * h * f * exp(-0.25 * h * vt[0]);
#pragma unroll
    for (int j = 1; j <= MD_NHC_LENGTH_X / 2; ++j)
      if (j <= M2) t[2 * j - 2] = t[2 * j - 2] + h * vt[2 * j - 2] / 2;
    chain_force();
    odd_update();
  }
  V[e] = v;
#pragma unroll
  for (int k = 0; k < MD_NHC_LENGTH_X; ++k)
    if (k < Mlen) {
      th[k * ndof + e] = t[k];
      vth[k * ndof + e] = vt[k];
    }
}

template <int HALF>
__global__ void k_nhc_massive(double *__restrict__ X, double *__restrict__ V, const double *__restrict__ F,
                              const double *__restrict__ M, double *__restrict__ th, double *__restrict__ vth,
                              const double *__restrict__ Qa, int nb, int Mlen, double T, double h) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 3 * nb) return;
  nhc_massive_body<HALF>(X, V, F[e], M, th, vth, Qa, nb, Mlen, T, h, e);
}

// General grouping: thread per chain, its DOF range [f0, f0+f2) in the
// reference's degree-of-freedom order (bead-major, then dimension).
template <int HALF>
__global__ void k_nhc_groups(double *__restrict__ X, double *__restrict__ V, const double *__restrict__ F,
                             const double *__restrict__ M, double *__restrict__ th, double *__restrict__ vth,
                             const double *__restrict__ Qa, const int *__restrict__ gf, const int *__restrict__ gf0,
                             const int *__restrict__ gf2, int Nf, int nb, int N, int P, int Mlen, double T,
                             double h) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Nf) return;
  const int f0 = gf0[i], f2 = gf2[i], fcnt = gf[i];
  double t[MD_NHC_LENGTH_X], vt[MD_NHC_LENGTH_X], Q[MD_NHC_LENGTH_X], a[MD_NHC_LENGTH_X];
  for (int k = 0; k < Mlen; ++k) {
    t[k] = th[k * Nf + i];
    vt[k] = vth[k * Nf + i];
    Q[k] = Qa[k * Nf + i];
  }
  const int M2 = Mlen / 2;
  auto idx = [&](int q) { return (q % 3) * nb + slot_of_ref(q / 3, N, P); };
  auto chain_force = [&]() {
    double s = 0;
    for (int j = 0; j < f2; ++j) {
      int e = idx(f0 + j);
      double v = V[e];
      s += M[e % nb] * v * v;
    }
    a[0] = (s - fcnt * T) / Q[0];
    for (int k = 1; k < Mlen; ++k) a[k] = (Q[k - 1] * vt[k - 1] * vt[k - 1] - T) / Q[k];
  };
  auto kick = [&]() {
    double e1 = exp(-0.5 * h * vt[0]), e2 = exp(-0.25 * h * vt[0]);
    for (int j = 0; j < f2; ++j) {
      int e = idx(f0 + j);
      V[e] = V[e] * e1 + 0.5 * h * F[e] * e2;
    }
  };
  auto odd_update = [&]() {
    for (int j = 1; j <= M2; ++j) {
      double up = (j == M2) ? 0.0 : vt[2 * j];
      vt[2 * j - 1] = vt[2 * j - 1] * exp(-0.5 * h * up) + 0.5 * h * a[2 * j - 1] * exp(-0.25 * h * up);
    }
  };
  if (HALF == 1) {
    chain_force();
    kick();
    for (int j = 1; j <= M2; ++j) t[2 * j - 2] = t[2 * j - 2] + h * vt[2 * j - 2] / 2;
    odd_update();
    for (int j = 0; j < f2; ++j) {
      int e = idx(f0 + j);
      X[e] = X[e] + h * V[e];
    }
    for (int j = 1; j <= M2; ++j) t[2 * j - 1] = t[2 * j - 1] + h * vt[2 * j - 1];
    chain_force();
    for (int j = 1; j <= M2; ++j)
      vt[2 * j - 2] = vt[2 * j - 2] * exp(-h * vt[2 * j - 1]) + h * a[2 * j - 2] * exp(-0.5 * h * vt[2 * j - 1]);
  } else {
    kick();
    for (int j = 1; j <= M2; ++j) t[2 * j - 2] = t[2 * j - 2] + h * vt[2 * j - 2] / 2;
    chain_force();
    odd_update();
  }
  for (int k = 0; k < Mlen; ++k) {
    th[k * Nf + i] = t[k];
    vth[k * Nf + i] = vt[k];
  }
}

// Reductions into per-block partials (summed deterministically afterwards).
__global__ __launch_bounds__(256) void k_kinetic(const double *__restrict__ V, const double *__restrict__ M, int nb,
                                                 double *__restrict__ part) {
  __shared__ double red[256];
  double s = 0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < nb; i += gridDim.x * 256) {
    double m = M[i];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      double v = V[d * nb + i];
      s += 0.5 * m * v * v;
    }
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_nhc_energy(const double *__restrict__ th, const double *__restrict__ vth,
                                                    const double *__restrict__ Qa, const int *__restrict__ gf,
                                                    int Nf, int massive, int Mlen, double T,
                                                    double *__restrict__ part) {
  __shared__ double red[256];
  double s = 0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < Nf; i += gridDim.x * 256) {
    int f = massive ? 1 : gf[i];
    for (int k = 0; k < Mlen; ++k) {
      double vt = vth[k * Nf + i];
      s += 0.5 * Qa[k * Nf + i] * vt * vt;
    }
    s += f * T * th[i];
    for (int k = 1; k < Mlen; ++k) s += T * th[k * Nf + i];
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Trap energy per bead (md_pimd_x.c:1518-1522), virial sums (:3914-3920, 3937-3954).
__global__ __launch_bounds__(256) void k_bead_sums(const double *__restrict__ X, const double *__restrict__ F,
                                                   const double *__restrict__ M, int N, int P, int mode,
                                                   double w, double c0, double c1, double c2, double L0,
                                                   double L1, double L2, double *__restrict__ part) {
  __shared__ double red[256];
  const int nb = N * P;
  double s = 0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < nb; i += gridDim.x * 256) {
    double m = M[i];
    if (mode == 0) {  // trap energy / P
      double dx = X[i] - c0, dy = X[nb + i] - c1, dz = X[2 * nb + i] - c2;
      double d = sqrt(dx * dx + dy * dy + dz * dz);
      s += 0.5 * m * w * w * d * d / P;
    } else if (mode == 1) {  // plain virial: -(x - L/2) f m
      s += -(X[i] - L0 / 2) * F[i] * m;
      s += -(X[nb + i] - L1 / 2) * F[nb + i] * m;
      s += -(X[2 * nb + i] - L2 / 2) * F[2 * nb + i] * m;
    } else {  // centroid virial: -(x - centroid) f m
      int l = i % N;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        double c = 0;
        for (int j = 0; j < P; ++j) c += X[d * nb + j * N + l] / P;
        s += -(X[d * nb + i] - c) * F[d * nb + i] * m;
      }
    }
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w2 = 128; w2 > 0; w2 >>= 1) {
    if ((int)threadIdx.x < w2) red[threadIdx.x] += red[threadIdx.x + w2];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// ------------------------------------------------- batched replicas
//
// Per-replica arguments of the bead-local stages; every thread owns one
// degree of freedom e = d * nb + s of one replica, with the same arithmetic
// as the single-system kernels above.
struct BeadRep {
  double *X, *V, *F, *th, *vth;
  const double *M, *Q;
  int nb, N, P, Mlen;
  double T, h;
  int pbc, image, trap;
  double L[3], w, c[3], om;
  ForceSums cs;  // half 2: parts added into F first (cs.Fpart set)
};

// NHC half-kick (+ drift for HALF 1), then the box wrap after HALF 1, for
// degree of freedom e < 3 nb.  HALF 3: half 2 of one step and half 1 of the
// next in one pass (the force does not change between them; every chain is
// its own degree of freedom's, so the arithmetic is that of the two kernels).
template <int HALF>
__device__ __forceinline__ void nhc_wrap_body(const BeadRep &R, int e) {
  double f;
  if (HALF != 1 && R.cs.Fpart) {
    f = combined_force(R.F, R.cs, 3 * R.nb, e);
    R.F[e] = f;
  } else {
    f = R.F[e];
  }
  if (HALF != 1) nhc_massive_body<2>(R.X, R.V, f, R.M, R.th, R.vth, R.Q, R.nb, R.Mlen, R.T, R.h, e);
  if (HALF != 2) nhc_massive_body<1>(R.X, R.V, f, R.M, R.th, R.vth, R.Q, R.nb, R.Mlen, R.T, R.h, e);
  if (HALF != 2 && R.pbc) {
    const int d = e / R.nb;
    const double L = R.L[d];
    double x = R.X[e];
    if (x < 0)
      x += L;
    else if (x > L)
      x -= L;
    R.X[e] = x;
  }
}

template <int HALF>
__global__ __launch_bounds__(256) void kb_nhc(const BeadRep *__restrict__ reps, const int *__restrict__ start,
                                              int nrep) {
  const int r = batch_rep(start, nrep, blockIdx.x);
  const BeadRep &R = reps[r];
  const int e = (blockIdx.x - start[r]) * 256 + threadIdx.x;
  if (e >= 3 * R.nb) return;
  nhc_wrap_body<HALF>(R, e);
}

template <int HALF>
__global__ __launch_bounds__(256) void k_nhc_wrap(BeadRep R) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= 3 * R.nb) return;
  nhc_wrap_body<HALF>(R, e);
}

// Force reset, trap (md_pimd_x.c:1370-1389) and interior ring springs of
// degree of freedom i < 3 nb: the arithmetic of k_fill, k_trap and
// k_ring_springs in one pass.
__device__ __forceinline__ void bead_forces_body(const BeadRep &R, int i) {
  const int nb = R.nb;
  const int d = i / nb, s = i - d * nb;
  double Fv = 0.0;
  if (R.trap) {
    double f = 0;
    f += -R.w * R.w * (R.X[i] - R.c[d]);
    Fv += f / R.P;
  }
  if (s >= R.N && s < nb - R.N) {
    const double m = R.M[s], om = R.om;
    const double x = R.X[i];
    double a = x - R.X[i + R.N];
    double c = x - R.X[i - R.N];
    if (R.image) {
      a = min_image(a, R.L[d]);
      c = min_image(c, R.L[d]);
    }
    double g = 0;
    g += m * om * om * a;
    g += m * om * om * c;
    Fv += -g / m;
  }
  R.F[i] = Fv;
}

__global__ __launch_bounds__(256) void kb_bead_forces(const BeadRep *__restrict__ reps, const int *__restrict__ start,
                                                      int nrep) {
  const int r = batch_rep(start, nrep, blockIdx.x);
  const BeadRep &R = reps[r];
  const int i = (blockIdx.x - start[r]) * 256 + threadIdx.x;
  if (i >= 3 * R.nb) return;
  bead_forces_body(R, i);
}

__global__ __launch_bounds__(256) void k_bead_forces(BeadRep R) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= 3 * R.nb) return;
  bead_forces_body(R, i);
}

static BeadRep bead_rep(const BatchItem &it) {
  DevPimd *p = it.p;
  DevSim *d = p->ds;
  BeadRep R;
  R.X = d->X;
  R.V = d->V;
  R.F = d->F;
  R.th = d->th;
  R.vth = d->vth;
  R.M = d->M;
  R.Q = d->Q;
  R.nb = d->nb;
  R.N = p->N;
  R.P = p->P;
  R.Mlen = g_nhc_len;
  R.T = d->host->T;
  R.h = it.dt;
  R.pbc = it.pbc;
  R.image = it.image;
  R.trap = it.trap;
  for (int k = 0; k < 3; ++k) {
    R.L[k] = it.L[k];
    R.c[k] = md_trap_center_x[k];
  }
  R.w = md_trap_frequency_x;
  R.om = p->host->omegaP;
  return R;
}

// Fused bead-local stages of one system's step (massive thermostats):
// NHC half (+ wrap after half 1), or reset + trap + interior springs.
int launch_step_nhc(const BatchItem &it, int half, hipStream_t st, const ForceSums *cs) {
  BeadRep R = bead_rep(it);
  if (half != 1 && cs) R.cs = *cs;
  const dim3 g((3 * R.nb + 255) / 256);
  if (half == 1)
    hipLaunchKernelGGL(k_nhc_wrap<1>, g, dim3(256), 0, st, R);
  else if (half == 2)
    hipLaunchKernelGGL(k_nhc_wrap<2>, g, dim3(256), 0, st, R);
  else
    hipLaunchKernelGGL(k_nhc_wrap<3>, g, dim3(256), 0, st, R);
  XMD_CHECK(hipGetLastError());
  return 0;
}

int launch_step_bead_forces(const BatchItem &it, hipStream_t st) {
  const BeadRep R = bead_rep(it);
  hipLaunchKernelGGL(k_bead_forces, dim3((3 * R.nb + 255) / 256), dim3(256), 0, st, R);
  XMD_CHECK(hipGetLastError());
  return 0;
}

struct BeadBatch {
  BatchTable t;
};

BeadBatch *bead_batch_create(const std::vector<BatchItem> &items) {
  std::vector<BeadRep> reps;
  std::vector<int> blocks;
  for (const BatchItem &it : items) {
    if (!it.p->ds->massive) {
      fprintf(stderr, "xmd: batched steps need massive thermostats (one chain per degree of freedom)\n");
      return nullptr;
    }
    reps.push_back(bead_rep(it));
    blocks.push_back((3 * it.p->ds->nb + 255) / 256);
  }
  BeadBatch *b = new BeadBatch();
  if (batch_table_make(b->t, reps.data(), sizeof(BeadRep), blocks)) {
    delete b;
    return nullptr;
  }
  return b;
}

void bead_batch_free(BeadBatch *b) {
  if (!b) return;
  batch_table_free(b->t);
  delete b;
}

int bead_batch_nhc(BeadBatch *b, int half, hipStream_t st) {
  const BeadRep *reps = (const BeadRep *)b->t.reps;
  if (half == 1)
    hipLaunchKernelGGL(kb_nhc<1>, dim3(b->t.total), dim3(256), 0, st, reps, b->t.start, b->t.nrep);
  else
    hipLaunchKernelGGL(kb_nhc<2>, dim3(b->t.total), dim3(256), 0, st, reps, b->t.start, b->t.nrep);
  XMD_CHECK(hipGetLastError());
  return 0;
}

int bead_batch_forces(BeadBatch *b, hipStream_t st) {
  hipLaunchKernelGGL(kb_bead_forces, dim3(b->t.total), dim3(256), 0, st, (const BeadRep *)b->t.reps, b->t.start,
                     b->t.nrep);
  XMD_CHECK(hipGetLastError());
  return 0;
}

// ------------------------------------------------------------ launchers

static inline unsigned blocks_for(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

int launch_reset_force(DevSim *d, hipStream_t st) {
  hipLaunchKernelGGL(k_fill, dim3(blocks_for(3L * d->nb, 256)), dim3(256), 0, st, d->F, 0.0, 3 * d->nb);
  XMD_CHECK(hipGetLastError());
  return 0;
}

int launch_trap_force(DevSim *d, double w, const double *c, double P, hipStream_t st) {
  hipLaunchKernelGGL(k_trap, dim3(blocks_for(3L * d->nb, 256)), dim3(256), 0, st, d->X, d->F, d->nb, w, c[0], c[1],
                     c[2], 0.0, P);
  XMD_CHECK(hipGetLastError());
  return 0;
}

int launch_pbc(DevSim *d, const double *L, hipStream_t st) {
  hipLaunchKernelGGL(k_pbc, dim3(blocks_for(3L * d->nb, 256)), dim3(256), 0, st, d->X, d->nb, L[0], L[1], L[2]);
  XMD_CHECK(hipGetLastError());
  return 0;
}

int launch_ring_springs(DevPimd *p, int image, const double *L, hipStream_t st) {
  if (p->P < 3) return 0;
  DevSim *d = p->ds;
  long n = (long)p->N * (p->P - 2);
  hipLaunchKernelGGL(k_ring_springs, dim3(blocks_for(n, 256)), dim3(256), 0, st, d->X, d->M, d->F, p->N, p->P,
                     p->host->omegaP, image, L[0], L[1], L[2]);
  XMD_CHECK(hipGetLastError());
  return 0;
}

int launch_nhc(DevSim *d, int half, double h, hipStream_t st) {
  double T = d->host->T;
  int Mlen = g_nhc_len;
  if (d->massive) {
    long n = 3L * d->nb;
    if (half == 1)
      hipLaunchKernelGGL(k_nhc_massive<1>, dim3(blocks_for(n, 256)), dim3(256), 0, st, d->X, d->V, d->F, d->M, d->th,
                         d->vth, d->Q, d->nb, Mlen, T, h);
    else
      hipLaunchKernelGGL(k_nhc_massive<2>, dim3(blocks_for(n, 256)), dim3(256), 0, st, d->X, d->V, d->F, d->M, d->th,
                         d->vth, d->Q, d->nb, Mlen, T, h);
  } else {
    if (half == 1)
      hipLaunchKernelGGL(k_nhc_groups<1>, dim3(blocks_for(d->Nf, 128)), dim3(128), 0, st, d->X, d->V, d->F, d->M,
                         d->th, d->vth, d->Q, d->gf, d->gf0, d->gf2, d->Nf, d->nb, d->N, d->P, Mlen, T, h);
    else
      hipLaunchKernelGGL(k_nhc_groups<2>, dim3(blocks_for(d->Nf, 128)), dim3(128), 0, st, d->X, d->V, d->F, d->M,
                         d->th, d->vth, d->Q, d->gf, d->gf0, d->gf2, d->Nf, d->nb, d->N, d->P, Mlen, T, h);
  }
  XMD_CHECK(hipGetLastError());
  return 0;
}

static int partial_sum(DevSim *d, int nblk, double *out) { return device_sum(d->red, nblk, out); }

int bead_sum(DevSim *d, int mode, const double *trap_c, double w, const double *L, double *out) {
  int nblk = (int)std::min<long>(blocks_for(d->nb, 256), 1024);
  hipLaunchKernelGGL(k_bead_sums, dim3(nblk), dim3(256), 0, rt().s0, d->X, d->F, d->M, d->N, d->P, mode, w,
                     trap_c ? trap_c[0] : 0.0, trap_c ? trap_c[1] : 0.0, trap_c ? trap_c[2] : 0.0, L[0], L[1],
                     L[2], d->red);
  XMD_CHECK(hipGetLastError());
  return partial_sum(d, nblk, out);
}

}  // namespace xmd

using namespace xmd;

extern "C" {

int md_reset_force_x(md_simulation_t_x *sim) {
  if (ensure_device(sim)) return -1;
  return launch_reset_force(sim_state(sim, false), rt().s0);
}

int md_update_nhc_VV3_1_x(md_simulation_t_x *sim, md_num_t_x h) {
  if (ensure_device(sim)) return -1;
  return launch_nhc(sim_state(sim, false), 1, h, rt().s0);
}

int md_update_nhc_VV3_2_x(md_simulation_t_x *sim, md_num_t_x h) {
  if (ensure_device(sim)) return -1;
  return launch_nhc(sim_state(sim, false), 2, h, rt().s0);
}

int md_periodic_boundary_x(md_simulation_t_x *sim) {
  if (!sim->box || sim->box->type != MD_ND_RECT_BOX_X) return 0;
  if (ensure_device(sim)) return -1;
  double L[3] = {box_L(sim, 0), box_L(sim, 1), box_L(sim, 2)};
  return launch_pbc(sim_state(sim, false), L, rt().s0);
}

// Chain forces of chain i on the host mirror (md_simulation_x.c:599-613);
// a diagnostic helper, not used by the per-step path.
void md_calc_nhc_force_x(md_simulation_t_x *sim, int f0, int f2, int i, md_num_t_x *res) {
  md_simulation_sync_host_x(sim, 1);
  double s = 0;
  for (int j = 0; j < f2; ++j) {
    int q = f0 + j;
    const md_particle_t_x *p = &sim->particles[q / 3];
    s += p->m * p->v[q % 3] * p->v[q % 3];
  }
  res[0] = (s - sim->nhcs[i].f * MD_kB_X * sim->T) / sim->nhcs[i].Q[0];
  for (int j = 1; j < g_nhc_len; ++j)
    res[j] = (sim->nhcs[i].Q[j - 1] * sim->nhcs[i].vtheta[j - 1] * sim->nhcs[i].vtheta[j - 1] - MD_kB_X * sim->T) /
             sim->nhcs[i].Q[j];
}

int md_pimd_calc_trap_force_x(md_pimd_t_x *pimd, md_trap_force_t_x tf) {
  if (tf != md_harmonic_trap_force_x) {
    fprintf(stderr, "xmd: trap callback has no device kernel\n");
    return -1;
  }
  if (ensure_pimd_device(pimd)) return -1;
  return launch_trap_force(sim_state(pimd->sim, false), md_trap_frequency_x, md_trap_center_x, (double)pimd->P,
                           rt().s0);
}

int md_pimd_calc_trap_energy_x(md_pimd_t_x *pimd, md_trap_energy_t_x te, md_stats_t_x *stats) {
  if (te != md_harmonic_trap_energy_x) {
    fprintf(stderr, "xmd: trap energy callback has no device kernel\n");
    return -1;
  }
  if (ensure_pimd_device(pimd)) return -1;
  double L[3] = {0, 0, 0}, e = 0;
  if (bead_sum(sim_state(pimd->sim, false), 0, md_trap_center_x, md_trap_frequency_x, L, &e)) return -1;
  stats->es[stats->N - 1] += e;
  return 0;
}

int md_pimd_calc_virial_energy_x(md_pimd_t_x *pimd, md_stats_t_x *stats) {
  if (ensure_pimd_device(pimd)) return -1;
  md_simulation_t_x *sim = pimd->sim;
  double L[3] = {box_L(sim, 0), box_L(sim, 1), box_L(sim, 2)}, e = 0;
  if (bead_sum(sim_state(sim, false), 1, nullptr, 0, L, &e)) return -1;
  stats->es[stats->N - 1] += e / 2.0;
  return 0;
}

int md_pimd_calc_centroid_virial_energy_x(md_pimd_t_x *pimd, md_stats_t_x *stats, int fc) {
  if (ensure_pimd_device(pimd)) return -1;
  double L[3] = {0, 0, 0}, e = 0;
  if (bead_sum(sim_state(pimd->sim, false), 2, nullptr, 0, L, &e)) return -1;
  stats->es[stats->N - 1] += e / 2.0;
  stats->es[stats->N - 1] += (MD_DIMENSION_X * pimd->N - fc / ((md_num_t_x)pimd->P)) / (2 * pimd->beta);
  return 0;
}

int md_calc_kinetic_x(md_simulation_t_x *sim, md_stats_t_x *stats) {
  if (ensure_device(sim)) return -1;
  DevSim *d = sim_state(sim, false);
  int nblk = (int)std::min<long>(blocks_for(d->nb, 256), 1024);
  hipLaunchKernelGGL(k_kinetic, dim3(nblk), dim3(256), 0, rt().s0, d->V, d->M, d->nb, d->red);
  XMD_CHECK(hipGetLastError());
  double e = 0;
  if (partial_sum(d, nblk, &e)) return -1;
  stats->es[stats->N - 1] += e;
  return 0;
}

// md_statistics_x.c:855-958 (nocl 946-955): sum m v^2 over the total
// degree-of-freedom count of the thermostat groups, into Ts[N-1].
int md_calc_temperature_x(md_simulation_t_x *sim, md_stats_t_x *stats) {
  long f = 0;
  for (int i = 0; i < sim->Nf; ++i) f += sim->fs[i];
  if (f <= 0 || !stats->Ts) return -1;
  if (ensure_device(sim)) return -1;
  DevSim *d = sim_state(sim, false);
  int nblk = (int)std::min<long>(blocks_for(d->nb, 256), 1024);
  hipLaunchKernelGGL(k_kinetic, dim3(nblk), dim3(256), 0, rt().s0, d->V, d->M, d->nb, d->red);
  XMD_CHECK(hipGetLastError());
  double e = 0;  // sum m v^2 / 2
  if (partial_sum(d, nblk, &e)) return -1;
  stats->Ts[stats->N - 1] += 2.0 * e / (double)f;
  return 0;
}

int md_calc_nhc_kinetic_x(md_simulation_t_x *sim, md_stats_t_x *stats) {
  if (ensure_device(sim)) return -1;
  DevSim *d = sim_state(sim, false);
  int nblk = (int)std::min<long>(blocks_for(d->Nf, 256), 1024);
  hipLaunchKernelGGL(k_nhc_energy, dim3(nblk), dim3(256), 0, rt().s0, d->th, d->vth, d->Q, d->gf, d->Nf,
                     d->massive ? 1 : 0, g_nhc_len, sim->T, d->red);
  XMD_CHECK(hipGetLastError());
  double e = 0;
  if (partial_sum(d, nblk, &e)) return -1;
  stats->es[stats->N - 1] += e;
  return 0;
}

int md_periodic_boundary_count_x(md_simulation_t_x *sim, int *count) {
  if (!md_simulation_sync_host_x(sim, 0)) return -1;
  for (int l = 0; l < sim->N; ++l)
    for (int i = 0; i < MD_DIMENSION_X; ++i) {
      double L = box_L(sim, i);
      count[MD_DIMENSION_X * l + i] = md_periodic_image_count_x(sim->particles[l].x[i], L);
      sim->particles[l].x[i] += count[MD_DIMENSION_X * l + i] * L;
    }
  return 0;
}

int md_periodic_boundary_recover_x(md_simulation_t_x *sim, int *count) {
  if (!md_simulation_sync_host_x(sim, 0)) return -1;
  for (int l = 0; l < sim->N; ++l)
    for (int i = 0; i < MD_DIMENSION_X; ++i) sim->particles[l].x[i] -= count[MD_DIMENSION_X * l + i] * box_L(sim, i);
  return 0;
}

int md_simulation_rescale_velocity_x(md_simulation_t_x *sim, md_num_t_x T) {
  if (!md_simulation_sync_host_x(sim, 0)) return -1;
  double s = 0;
  for (int i = 0; i < sim->N; ++i)
    for (int j = 0; j < MD_DIMENSION_X; ++j) s += sim->particles[i].m * sim->particles[i].v[j] * sim->particles[i].v[j];
  double lam = sqrt(sim->fc * MD_kB_X * T / s);
  for (int i = 0; i < sim->N; ++i)
    for (int j = 0; j < MD_DIMENSION_X; ++j) sim->particles[i].v[j] *= lam;
  return 0;
}

}  // extern "C"
