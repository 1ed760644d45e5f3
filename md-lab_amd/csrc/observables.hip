// Per-sample observables on device-resident beads, nocl semantics:
// density distribution (md_pimd_x.c:1826-1967), imaginary-time correlation
// of x (ITCF, md_pimd_x.c:2723-2872), pair correlations (one and two systems,
// in-plane, and 2-D |dx|,|dy| histograms; md_pimd_x.c:2974-3280,4869-4956)
// and the structure factors S(q) on axes and on spherical shells
// (md_pimd_x.c:4958-5450, md_statistics_x.c:104-161).
//
// Histograms: the kernels count samples per bin with integer atomics (exact,
// order independent); the host then adds the reference's increment (1/P or
// 1/norm) once per count into stats->fxs[stats->N-1][bin].  Every bin only
// ever receives that one increment value, so the result equals the
// reference's sequential accumulation bit for bit.
#include <algorithm>
#include <cmath>
#include <vector>

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

// Bead sets of a pair histogram: bead slots slot0 + l (l < N) of one
// system, in the slice-major layout X[d * nb + slot].
struct BeadSet {
  const double *X;
  int N, nb, slot0;
};

enum HistMode { HM_RADIAL = 0, HM_AXES2D = 1, HM_ITCF = 2 };

struct PairHist {
  BeadSet a, b;
  int mode;
  int same;   // a and b are one system: skip l == j
  int two;    // ITCF of two systems: count (i1, i2) and (i2, i1)
  int image;
  double L[3];
  double incre, rmin;
  int n, points;  // n = (int)sqrt(points) for the 2-D histograms
};

__device__ __forceinline__ int clamp_bin(int i1, int i2, int n, int points) {
  int idx = i1 * n + i2;
  if (idx < 0) idx = 0;
  if (idx >= points) idx = points - 1;
  return idx;
}

// Every (l of a, j of b):
//  HM_RADIAL (md_pimd_x.c:3019-3036, 3125-3136): bin (int)(r / incre), dropped beyond the last bin;
//  HM_AXES2D (md_pimd_x.c:4889-4909): bins (int)(|dx| / incre), (int)(|dy| / incre), clamped;
//  HM_ITCF   (md_pimd_x.c:2731-2746, 2851-2870): bins of (x - rmin) / incre along dimension 0, clamped.
__global__ __launch_bounds__(256) void k_pair_hist(PairHist h, unsigned long long *__restrict__ cnt) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)h.a.N * h.b.N) return;
  const int l = (int)(t / h.b.N), j = (int)(t - (long)l * h.b.N);
  if (h.same && l == j) return;
  const int sa = h.a.slot0 + l, sb = h.b.slot0 + j;
  const double *XA = h.a.X, *XB = h.b.X;
  const int na = h.a.nb, nb = h.b.nb;
  if (h.mode == HM_RADIAL) {
    const double r =
        obs_dist(XA[sa] - XB[sb], XA[na + sa] - XB[nb + sb], XA[2 * na + sa] - XB[2 * nb + sb], h.L, h.image);
    const int bin = (int)(r / h.incre);
    if (bin < h.points) atomicAdd(&cnt[bin], 1ull);
  } else if (h.mode == HM_AXES2D) {
    double dx = XA[sa] - XB[sb], dy = XA[na + sa] - XB[nb + sb];
    if (h.image) {
      dx = dv_min_image(dx, h.L[0]);
      dy = dv_min_image(dy, h.L[1]);
    }
    const int i1 = (int)(fabs(dx) / h.incre), i2 = (int)(fabs(dy) / h.incre);
    atomicAdd(&cnt[clamp_bin(i1, i2, h.n, h.points)], 1ull);
  } else {
    const int i1 = (int)((XA[sa] - h.rmin) / h.incre), i2 = (int)((XB[sb] - h.rmin) / h.incre);
    atomicAdd(&cnt[clamp_bin(i1, i2, h.n, h.points)], 1ull);
    if (h.two) atomicAdd(&cnt[clamp_bin(i2, i1, h.n, h.points)], 1ull);
  }
}

// md_pimd_x.c:3248-3278: in-plane distance of beads l != j of one slice,
// both inside the window box = {x0, x1, y0, y1}.
__global__ __launch_bounds__(256) void k_paircorr_box2d(const double *__restrict__ X, int N, int nb, int P, double x0,
                                                        double x1, double y0, double y1, double incre, int points,
                                                        unsigned long long *__restrict__ cnt) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)P * N * N) return;
  const int slice = (int)(t / ((long)N * N));
  const int rem = (int)(t - (long)slice * N * N);
  const int l = rem / N, j = rem - l * N;
  if (l == j) return;
  const int sa = slice * N + l, sb = slice * N + j;
  const double ax = X[sa], ay = X[nb + sa], bx = X[sb], by = X[nb + sb];
  if (ax < x0 || ax > x1 || ay < y0 || ay > y1) return;
  if (bx < x0 || bx > x1 || by < y0 || by > y1) return;
  const double dx = ax - bx, dy = ay - by;
  const double r = sqrt(dx * dx + dy * dy);
  const int bin = (int)(r / incre);
  if (bin < points) atomicAdd(&cnt[bin], 1ull);
}

// S(q) of two bead sets (each the union of up to two BeadSets):
//   res(q) = sum [C1 C2 + S1 S2] / norm,  C = sum cos(q . x), S = sum sin(q . x),
// on axes (q along each dimension in turn, the three terms averaged;
// md_pimd_x.c:4967-4996, 5052-5083, 5152-5203) or for one q vector per point
// (md_pimd_x.c:5268-5340, 5344-5402, 5406-5450).  One workgroup per q point.
struct SkArgs {
  BeadSet s1[2], s2[2];
  int n1, n2;
  double norm;
  int sphere;
  double q0, qincre;
  const double *qis;  // sphere: 3 per point
};

__device__ __forceinline__ void sk_accum(const BeadSet &b, double qx, double qy, double qz, int axis, double &c,
                                         double &s) {
  for (int l = threadIdx.x; l < b.N; l += blockDim.x) {
    const int k = b.slot0 + l;
    const double t = axis >= 0 ? qx * b.X[axis * b.nb + k]
                               : qx * b.X[k] + qy * b.X[b.nb + k] + qz * b.X[2 * b.nb + k];
    double sn, cs;
    sincos(t, &sn, &cs);
    c += cs;
    s += sn;
  }
}

__global__ __launch_bounds__(256) void k_sk(SkArgs a, int j0, double *__restrict__ out) {
  __shared__ double red[4][256];
  const int j = j0 + blockIdx.x;
  double res = 0;
  const int nterm = a.sphere ? 1 : 3;
  for (int d = 0; d < nterm; ++d) {
    double qx, qy = 0, qz = 0;
    int axis;
    if (a.sphere) {
      qx = a.qis[3 * j];
      qy = a.qis[3 * j + 1];
      qz = a.qis[3 * j + 2];
      axis = -1;
    } else {
      qx = a.q0 + j * a.qincre;
      axis = d;
    }
    double c1 = 0, s1 = 0, c2 = 0, s2 = 0;
    for (int k = 0; k < a.n1; ++k) sk_accum(a.s1[k], qx, qy, qz, axis, c1, s1);
    for (int k = 0; k < a.n2; ++k) sk_accum(a.s2[k], qx, qy, qz, axis, c2, s2);
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
    res += red[0][0] * red[1][0] / a.norm;
    res += red[2][0] * red[3][0] / a.norm;
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = a.sphere ? res : res / 3;
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
    cap = 0;
    if (hipMalloc(&buf, points * sizeof(unsigned long long)) != hipSuccess) return nullptr;
    cap = points;
  }
  return buf;
}

// Device scratch of n doubles for S(q) outputs and q vectors (grown on demand, kept).
static double *obs_doubles(int slot, int n) {
  static double *buf[2] = {nullptr, nullptr};
  static int cap[2] = {0, 0};
  if (n > cap[slot]) {
    if (buf[slot]) (void)hipFree(buf[slot]);
    buf[slot] = nullptr;
    cap[slot] = 0;
    if (hipMalloc(&buf[slot], n * sizeof(double)) != hipSuccess) return nullptr;
    cap[slot] = n;
  }
  return buf[slot];
}

// Bead slice `bead` (1-based) of every particle of a system.
static BeadSet bead_set(DevPimd *p, int bead) {
  BeadSet b;
  b.X = p->ds->X;
  b.N = p->N;
  b.nb = p->ds->nb;
  b.slot0 = (bead - 1) * p->N;
  return b;
}

static DevPimd *obs_pimd(md_pimd_t_x *pimd) {
  if (!pimd || ensure_pimd_device(pimd)) return nullptr;
  return pimd_state(pimd, false);
}

static int run_pair_hist(PairHist &h, md_stats_t_x *stats, double incr) {
  unsigned long long *cnt = obs_counts(h.points);
  if (!cnt) return -1;
  hipStream_t st = rt().s0;
  XMD_CHECK(hipMemsetAsync(cnt, 0, h.points * sizeof(unsigned long long), st));
  const long work = (long)h.a.N * h.b.N;
  if (work > 0) {
    hipLaunchKernelGGL(k_pair_hist, dim3((unsigned)((work + 255) / 256)), dim3(256), 0, st, h, cnt);
    XMD_CHECK(hipGetLastError());
  }
  XMD_CHECK(hipStreamSynchronize(st));
  return add_counts(stats, cnt, h.points, incr);
}

static PairHist pair_hist(DevPimd *a, int bead_a, DevPimd *b, int bead_b, md_simulation_t_x *box_sim, int mode,
                          int points) {
  PairHist h;
  h.a = bead_set(a, bead_a);
  h.b = bead_set(b, bead_b);
  h.mode = mode;
  h.same = 0;
  h.two = 0;
  h.image = 0;
  for (int d = 0; d < 3; ++d) h.L[d] = box_L(box_sim, d);
  h.incre = 0;
  h.rmin = 0;
  h.n = (int)sqrt((double)points);
  h.points = points;
  return h;
}

// S(q) for the points j < npts of stats, added into stats->fxs[N-1][j].
static int run_sk(SkArgs &a, md_stats_t_x *stats, int npts, const double *qis_host) {
  if (npts <= 0) return 0;
  hipStream_t st = rt().s0;
  double *out = obs_doubles(0, npts);
  if (!out) return -1;
  a.qis = nullptr;
  if (a.sphere) {
    double *q = obs_doubles(1, 3 * npts);
    if (!q) return -1;
    XMD_CHECK(hipMemcpyAsync(q, qis_host, 3 * npts * sizeof(double), hipMemcpyHostToDevice, st));
    a.qis = q;
  }
  hipLaunchKernelGGL(k_sk, dim3(npts), dim3(256), 0, st, a, 0, out);
  XMD_CHECK(hipGetLastError());
  std::vector<double> h(npts);
  XMD_CHECK(hipMemcpyAsync(h.data(), out, npts * sizeof(double), hipMemcpyDeviceToHost, st));
  XMD_CHECK(hipStreamSynchronize(st));
  for (int j = 0; j < npts; ++j) stats->fxs[stats->N - 1][j] += h[j];
  return 0;
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
  DevPimd *p = obs_pimd(pimd);
  if (!p || !stats || stats->points <= 0 || pi < 1 || pi > pimd->P) return -1;
  PairHist h = pair_hist(p, 1, p, pi, pimd->sim, HM_RADIAL, stats->points);
  h.same = 1;
  h.image = image;
  h.incre = rmax / stats->points;
  return run_pair_hist(h, stats, 1.0 / norm);
}

int md_pimd_calc_pair_correlation_2_x(md_pimd_t_x *pimd, md_pimd_t_x *pimd2, md_stats_t_x *stats, md_num_t_x rmax,
                                      md_num_t_x norm, int pi, int image) {
  if (!pimd || !pimd2 || pimd->P != pimd2->P) return -1;
  DevPimd *a = obs_pimd(pimd), *b = obs_pimd(pimd2);
  if (!a || !b || !stats || stats->points <= 0 || pi < 1 || pi > pimd->P) return -1;
  PairHist h = pair_hist(a, 1, b, pi, pimd->sim, HM_RADIAL, stats->points);
  h.image = image;
  h.incre = rmax / stats->points;
  return run_pair_hist(h, stats, 1.0 / norm);
}

int md_pimd_calc_pair_correlation_2d_x(md_pimd_t_x *pimd, md_stats_t_x *stats, md_num_t_x rmax, md_num_t_x norm,
                                       md_num_t_x *box) {
  DevPimd *p = obs_pimd(pimd);
  if (!p || !stats || stats->points <= 0 || !box) return -1;
  unsigned long long *cnt = obs_counts(stats->points);
  if (!cnt) return -1;
  hipStream_t st = rt().s0;
  XMD_CHECK(hipMemsetAsync(cnt, 0, stats->points * sizeof(unsigned long long), st));
  const long work = (long)p->P * p->N * p->N;
  hipLaunchKernelGGL(k_paircorr_box2d, dim3((unsigned)((work + 255) / 256)), dim3(256), 0, st, p->ds->X, p->N,
                     p->ds->nb, p->P, box[0], box[1], box[2], box[3], rmax / stats->points, stats->points, cnt);
  XMD_CHECK(hipGetLastError());
  XMD_CHECK(hipStreamSynchronize(st));
  return add_counts(stats, cnt, stats->points, 1.0 / norm);
}

int md_pimd_calc_2d_pair_correlation_x(md_pimd_t_x *pimd, md_stats_t_x *stats, md_num_t_x rmax, md_num_t_x norm,
                                       int pi, int image) {
  DevPimd *p = obs_pimd(pimd);
  if (!p || !stats || stats->points <= 0 || pi < 1 || pi > pimd->P) return -1;
  PairHist h = pair_hist(p, 1, p, pi, pimd->sim, HM_AXES2D, stats->points);
  h.same = 1;
  h.image = image;
  h.incre = rmax / h.n;
  return run_pair_hist(h, stats, 1.0 / norm);
}

int md_pimd_calc_2d_pair_correlation_2_x(md_pimd_t_x *pimd, md_pimd_t_x *pimd2, md_stats_t_x *stats,
                                         md_num_t_x rmax, md_num_t_x norm, int pi, int image) {
  if (!pimd || !pimd2 || pimd->P != pimd2->P) return -1;
  DevPimd *a = obs_pimd(pimd), *b = obs_pimd(pimd2);
  if (!a || !b || !stats || stats->points <= 0 || pi < 1 || pi > pimd->P) return -1;
  PairHist h = pair_hist(a, 1, b, pi, pimd->sim, HM_AXES2D, stats->points);
  h.image = image;
  h.incre = rmax / h.n;
  return run_pair_hist(h, stats, 1.0 / norm);
}

int md_pimd_calc_ITCF_x(md_pimd_t_x *pimd, md_stats_t_x *stats, md_num_t_x rmin, md_num_t_x rmax, int pi) {
  DevPimd *p = obs_pimd(pimd);
  if (!p || !stats || stats->points <= 0 || pi < 1 || pi > pimd->P) return -1;
  PairHist h = pair_hist(p, 1, p, pi, pimd->sim, HM_ITCF, stats->points);
  h.rmin = rmin;
  h.incre = (rmax - rmin) / h.n;
  return run_pair_hist(h, stats, 1.0 / pimd->P);
}

int md_pimd_calc_ITCF_2_x(md_pimd_t_x *pimd, md_pimd_t_x *pimd2, md_stats_t_x *stats, md_num_t_x rmin,
                          md_num_t_x rmax, int pi) {
  if (!pimd || !pimd2 || pimd->P != pimd2->P) return -1;
  DevPimd *a = obs_pimd(pimd), *b = obs_pimd(pimd2);
  if (!a || !b || !stats || stats->points <= 0 || pi < 1 || pi > pimd->P) return -1;
  PairHist h = pair_hist(a, 1, b, pi, pimd->sim, HM_ITCF, stats->points);
  h.two = 1;
  h.rmin = rmin;
  h.incre = (rmax - rmin) / h.n;
  return run_pair_hist(h, stats, 1.0 / pimd->P);
}

int md_pimd_calc_Sk_structure_x(md_pimd_t_x *pimd, md_stats_t_x *stats, md_num_t_x q0, md_num_t_x qincre, int pi) {
  DevPimd *p = obs_pimd(pimd);
  if (!p || !stats || stats->points <= 0 || pi < 1 || pi > pimd->P) return -1;
  SkArgs a{};
  a.s1[0] = bead_set(p, 1);
  a.s2[0] = bead_set(p, pi);
  a.n1 = a.n2 = 1;
  a.norm = p->N;
  a.q0 = q0;
  a.qincre = qincre;
  return run_sk(a, stats, stats->points, nullptr);
}

int md_pimd_calc_Sk_structure_2_x(md_pimd_t_x *pimd, md_pimd_t_x *pimd2, md_stats_t_x *stats, md_num_t_x q0,
                                  md_num_t_x qincre, int pi) {
  if (!pimd || !pimd2 || pimd->P != pimd2->P) return -1;
  DevPimd *p = obs_pimd(pimd), *p2 = obs_pimd(pimd2);
  if (!p || !p2 || !stats || stats->points <= 0 || pi < 1 || pi > pimd->P) return -1;
  SkArgs a{};
  a.s1[0] = bead_set(p, 1);
  a.s2[0] = bead_set(p2, pi);
  a.n1 = a.n2 = 1;
  a.norm = p->N;
  a.q0 = q0;
  a.qincre = qincre;
  return run_sk(a, stats, stats->points, nullptr);
}

int md_pimd_calc_Sk_structure_full_2_x(md_pimd_t_x *pimd, md_pimd_t_x *pimd2, md_stats_t_x *stats, md_num_t_x q0,
                                       md_num_t_x qincre, int pi) {
  if (!pimd || !pimd2 || pimd->P != pimd2->P) return -1;
  DevPimd *p = obs_pimd(pimd), *p2 = obs_pimd(pimd2);
  if (!p || !p2 || !stats || stats->points <= 0 || pi < 1 || pi > pimd->P) return -1;
  SkArgs a{};
  a.s1[0] = bead_set(p, 1);
  a.s1[1] = bead_set(p2, 1);
  a.s2[0] = bead_set(p, pi);
  a.s2[1] = bead_set(p2, pi);
  a.n1 = a.n2 = 2;
  a.norm = p->N + p2->N;
  a.q0 = q0;
  a.qincre = qincre;
  return run_sk(a, stats, stats->points, nullptr);
}

// Spherical variants: the second bead slice is (pi - 1) % P + 1 and the
// points are the first min(sk_sphere->N, stats->points) shells.
int md_pimd_calc_sphere_Sk_structure_x(md_pimd_t_x *pimd, md_stats_t_x *stats, int pi, md_sk_sphere_t_x *sk_sphere) {
  DevPimd *p = obs_pimd(pimd);
  if (!p || !stats || !sk_sphere || pi < 1) return -1;
  const int b2 = (pi - 1) % pimd->P + 1;
  SkArgs a{};
  a.s1[0] = bead_set(p, 1);
  a.s2[0] = bead_set(p, b2);
  a.n1 = a.n2 = 1;
  a.norm = p->N;
  a.sphere = 1;
  return run_sk(a, stats, std::min(sk_sphere->N, stats->points), sk_sphere->qis);
}

int md_pimd_calc_sphere_Sk_structure_full_2_x(md_pimd_t_x *pimd, md_pimd_t_x *pimd2, md_stats_t_x *stats, int pi,
                                              md_sk_sphere_t_x *sk_sphere) {
  if (!pimd || !pimd2 || pimd->P != pimd2->P) return -1;
  DevPimd *p = obs_pimd(pimd), *p2 = obs_pimd(pimd2);
  if (!p || !p2 || !stats || !sk_sphere || pi < 1) return -1;
  const int b2 = (pi - 1) % pimd->P + 1;
  SkArgs a{};
  a.s1[0] = bead_set(p, 1);
  a.s1[1] = bead_set(p2, 1);
  a.s2[0] = bead_set(p, b2);
  a.s2[1] = bead_set(p2, b2);
  a.n1 = a.n2 = 2;
  a.norm = p->N + p2->N;
  a.sphere = 1;
  return run_sk(a, stats, std::min(sk_sphere->N, stats->points), sk_sphere->qis);
}

int md_pimd_calc_sphere_Sk_structure_3_x(md_pimd_t_x *pimd, md_pimd_t_x *pimd2, md_pimd_t_x *pimd3,
                                         md_stats_t_x *stats, int pi, md_sk_sphere_t_x *sk_sphere) {
  if (!pimd || !pimd2 || !pimd3 || pimd->P != pimd2->P || pimd->P != pimd3->P) return -1;
  DevPimd *p = obs_pimd(pimd), *p2 = obs_pimd(pimd2), *p3 = obs_pimd(pimd3);
  if (!p || !p2 || !p3 || !stats || !sk_sphere || pi < 1) return -1;
  const int b2 = (pi - 1) % pimd->P + 1;
  SkArgs a{};
  a.s1[0] = bead_set(p, 1);
  a.s1[1] = bead_set(p2, 1);
  a.s2[0] = bead_set(p3, b2);
  a.n1 = 2;
  a.n2 = 1;
  a.norm = sqrt((double)((p->N + p2->N) * p3->N));
  a.sphere = 1;
  return run_sk(a, stats, std::min(sk_sphere->N, stats->points), sk_sphere->qis);
}

}  // extern "C"
