// Per-sample observables on device-resident beads: density distribution
// (md_pimd_x.c:1826-1967), pair correlation (md_pimd_x.c:2974-3040) and the
// structure factor S(q) (md_pimd_x.c:4958-5000), nocl semantics.
//
// Histograms: the kernels count samples per bin with integer atomics (exact,
// order independent); the host then adds the reference's increment (1/P or
// 1/norm) once per count into stats->fxs[stats->N-1][bin].  Every bin only
// ever receives that one increment value, so the result equals the
// reference's sequential accumulation bit for bit.
#include <cmath>

#include "xmd_device.h"
#include "xmd_internal.h"

namespace xmd {

__device__ __forceinline__ double obs_dist(double dx, double dy, double dz, const double *L, int image) {
  if (image) {
    dx = dv_min_image(dx, L[0]);
    dy = dv_min_image(dy, L[1]);
    dz = dv_min_image(dz, L[2]);
  }
  return sqrt(dx * dx + dy * dy + dz * dz);
}

struct ObsGeo {
  double L[3];
  double c[3];
  double incre;
  int points, image;
};

// md_pimd_x.c:1947-1962: every bead's distance to the center.
__global__ void k_density_hist(const double *__restrict__ X, int nb, ObsGeo g, unsigned long long *__restrict__ cnt) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nb) return;
  const double r = obs_dist(X[s] - g.c[0], X[nb + s] - g.c[1], X[2 * nb + s] - g.c[2], g.L, g.image);
  const int b = (int)(r / g.incre);
  if (b < g.points) atomicAdd(&cnt[b], 1ull);
}

// md_pimd_x.c:3019-3036: bead 1 of particle l against bead pi of every j != l.
__global__ void k_paircorr_hist(const double *__restrict__ X, int N, int nb, int pi, ObsGeo g,
                                unsigned long long *__restrict__ cnt) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N * N) return;
  const int l = t / N, j = t - l * N;
  if (l == j) return;
  const int sa = l, sb = (pi - 1) * N + j;  // slice-major slots of beads (l, 1) and (j, pi)
  const double r = obs_dist(X[sa] - X[sb], X[nb + sa] - X[nb + sb], X[2 * nb + sa] - X[2 * nb + sb], g.L, g.image);
  const int b = (int)(r / g.incre);
  if (b < g.points) atomicAdd(&cnt[b], 1ull);
}

// md_pimd_x.c:4967-4996: one workgroup per q.  res(q) = sum_d [C1 Cpi + S1 Spi] / N / 3,
// C1 = sum_l cos(q x_{l,1,d}), Cpi over bead pi, S likewise with sin.
__global__ __launch_bounds__(256) void k_sk_structure(const double *__restrict__ X, int N, int nb, int pi, double q0,
                                                      double qincre, double *__restrict__ out) {
  __shared__ double red[4][256];
  const int j = blockIdx.x;
  const double q = q0 + j * qincre;
  double res = 0;
  for (int d = 0; d < 3; ++d) {
    double c1 = 0, c2 = 0, s1 = 0, s2 = 0;
    for (int l = threadIdx.x; l < N; l += blockDim.x) {
      double sn, cs;
      sincos(q * X[d * nb + l], &sn, &cs);
      c1 += cs;
      s1 += sn;
      sincos(q * X[d * nb + (pi - 1) * N + l], &sn, &cs);
      c2 += cs;
      s2 += sn;
    }
    red[0][threadIdx.x] = c1;
    red[1][threadIdx.x] = c2;
    red[2][threadIdx.x] = s1;
    red[3][threadIdx.x] = s2;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w)
        for (int k = 0; k < 4; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
      __syncthreads();
    }
    res += red[0][0] * red[1][0] / N;
    res += red[2][0] * red[3][0] / N;
    __syncthreads();
  }
  if (threadIdx.x == 0) out[j] = res / 3;
}

static int obs_geo(md_simulation_t_x *sim, ObsGeo &g) {
  for (int d = 0; d < 3; ++d) g.L[d] = box_L(sim, d);
  return 0;
}

// counts -> stats->fxs[N-1][b] += incr, once per count
static int add_counts(md_stats_t_x *stats, const unsigned long long *dcnt, int points, double incr) {
  std::vector<unsigned long long> h(points);
  XMD_CHECK(hipMemcpy(h.data(), dcnt, points * sizeof(unsigned long long), hipMemcpyDeviceToHost));
  double *fx = stats->fxs[stats->N - 1];
  for (int b = 0; b < points; ++b)
    for (unsigned long long c = 0; c < h[b]; ++c) fx[b] += incr;
  return 0;
}

static unsigned long long *obs_counts(int points) {
  static unsigned long long *buf = nullptr;
  static int cap = 0;
  if (points > cap) {
    if (buf) (void)hipFree(buf);
    buf = nullptr;
    if (hipMalloc(&buf, points * sizeof(unsigned long long)) != hipSuccess) return nullptr;
    cap = points;
  }
  return buf;
}

}  // namespace xmd

using namespace xmd;

extern "C" {

int md_pimd_calc_density_distribution_x(md_pimd_t_x *pimd, md_stats_t_x *stats, md_num_t_x rmax, int image,
                                        md_num_t_x *center) {
  if (!pimd || !stats || stats->points <= 0 || !center) return -1;
  if (ensure_pimd_device(pimd)) return -1;
  DevPimd *p = pimd_state(pimd, false);
  ObsGeo g;
  obs_geo(pimd->sim, g);
  for (int d = 0; d < 3; ++d) g.c[d] = center[d];
  g.incre = rmax / stats->points;
  g.points = stats->points;
  g.image = image;
  unsigned long long *cnt = obs_counts(g.points);
  if (!cnt) return -1;
  hipStream_t st = rt().s0;
  XMD_CHECK(hipMemsetAsync(cnt, 0, g.points * sizeof(unsigned long long), st));
  const int nb = p->ds->nb;
  hipLaunchKernelGGL(k_density_hist, dim3((nb + 255) / 256), dim3(256), 0, st, p->ds->X, nb, g, cnt);
  XMD_CHECK(hipGetLastError());
  XMD_CHECK(hipStreamSynchronize(st));
  return add_counts(stats, cnt, g.points, 1.0 / pimd->P);
}

int md_pimd_calc_pair_correlation_x(md_pimd_t_x *pimd, md_stats_t_x *stats, md_num_t_x rmax, md_num_t_x norm,
                                    int pi, int image) {
  if (!pimd || !stats || stats->points <= 0 || pi < 1 || pi > pimd->P) return -1;
  if (ensure_pimd_device(pimd)) return -1;
  DevPimd *p = pimd_state(pimd, false);
  ObsGeo g;
  obs_geo(pimd->sim, g);
  g.incre = rmax / stats->points;
  g.points = stats->points;
  g.image = image;
  unsigned long long *cnt = obs_counts(g.points);
  if (!cnt) return -1;
  hipStream_t st = rt().s0;
  XMD_CHECK(hipMemsetAsync(cnt, 0, g.points * sizeof(unsigned long long), st));
  const int N = p->N;
  hipLaunchKernelGGL(k_paircorr_hist, dim3((N * N + 255) / 256), dim3(256), 0, st, p->ds->X, N, p->ds->nb, pi, g,
                     cnt);
  XMD_CHECK(hipGetLastError());
  XMD_CHECK(hipStreamSynchronize(st));
  return add_counts(stats, cnt, g.points, 1.0 / norm);
}

int md_pimd_calc_Sk_structure_x(md_pimd_t_x *pimd, md_stats_t_x *stats, md_num_t_x q0, md_num_t_x qincre, int pi) {
  if (!pimd || !stats || stats->points <= 0 || pi < 1 || pi > pimd->P) return -1;
  if (ensure_pimd_device(pimd)) return -1;
  DevPimd *p = pimd_state(pimd, false);
  hipStream_t st = rt().s0;
  double *out = p->ds->red;  // at least 256 doubles
  std::vector<double> h(stats->points);
  for (int j0 = 0; j0 < stats->points; j0 += 256) {
    const int n = std::min(256, stats->points - j0);
    hipLaunchKernelGGL(k_sk_structure, dim3(n), dim3(256), 0, st, p->ds->X, p->N, p->ds->nb, pi, q0 + j0 * qincre,
                       qincre, out);
    XMD_CHECK(hipGetLastError());
    XMD_CHECK(hipMemcpyAsync(h.data() + j0, out, n * sizeof(double), hipMemcpyDeviceToHost, st));
    XMD_CHECK(hipStreamSynchronize(st));
  }
  for (int j = 0; j < stats->points; ++j) stats->fxs[stats->N - 1][j] += h[j];
  return 0;
}

}  // extern "C"
