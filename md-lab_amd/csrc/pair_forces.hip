// Per-bead-slice pair forces and energies (md_pimd_x.c:1527-1824) for the
// Gaussian, periodic Gaussian and 3-D Ewald interactions
// (md_simulation_x.c:166-346, md_statistics_x.c:340-396).
//
// Kernel shapes (gfx950, wave64):
//  * k_pair<MODE>: grid (ceil(N/64), P), 256 threads.  The 64 lanes of every
//    wave own 64 beads i of one slice; the four waves split the partner list
//    (stride 4) of an LDS-staged partner tile, so every partner read is an
//    LDS broadcast.  Partials are combined through LDS and written once.
//    Result per bead: f_i += sum_{j != i} F(x_i, x_j) / m_{max(i,j)} / P,
//    i.e. the reference's l > l' loop with Newton's third law unrolled into
//    ordered pairs (the reaction is divided by the higher-index bead's mass
//    exactly as at md_pimd_x.c:1669-1673).
//  * Ewald real space uses the reference's own image set -- the cube
//    |n_d| <= n_sum around the raw separation p1 - p2 (md_simulation_x.c:282)
//    -- pruned to images within r_c = 6/alpha (erfc(6) = 2e-17; the dropped
//    terms are below 1e-16 of a typical pair force).
//  * Ewald reciprocal space uses per-slice structure factors
//    S(k) = sum_l exp(i k.x_l) over the reference's k cube
//    (|m_d| <= n_sum, k != 0) folded to a half space:
//      f_i = sum_{k in half} 4 pi A_k g_k Im(exp(i k.x_i) S(k)*) / m,
//    A_k = q exp(-pi^2 G^2/alpha^2) / (pi V G^2), which equals the
//    reference's per-pair sum (md_simulation_x.c:266-281) because
//    sum_{l != i} sin(k.(x_i - x_l)) = Im(exp(i k.x_i) S(k)*).
//    k_sk_partial builds S for 64-bead chunks (tables of exp(i m theta) in
//    LDS), k_sk_reduce sums the chunks, k_recip_force applies them with the
//    wave's S(k) read through scalar loads.
#include <cmath>
#include <vector>

#include "xmd_internal.h"

namespace xmd {

constexpr int TILE = 512;

__device__ __forceinline__ int pf_img_count(double x, double L) {
  int c = (int)(fabs(x) / L);
  if (x > L) return -c;
  if (x < 0) return c + 1;
  return 0;
}

__device__ __forceinline__ double pf_min_image(double a, double L) {
  double s = a + L / 2;
  s += L * pf_img_count(s, L);
  return s - L / 2;
}

struct PairParams {
  double L[3];
  double g, s2;        // Gaussian: strength, range^2
  double q, alpha;     // Ewald: charge product, splitting
  double rc2;          // real-space cutoff^2
  int nsum;            // Ewald cube half-width
  double P;            // beads (pair forces are divided by P)
};

// Real-space Ewald contribution of one raw separation (p1 - p2) to the
// force (accumulated into f) and/or the energy.  skip_min / only_min select
// the reference's R and NI splits (md_simulation_x.c:295-346).
template <bool FORCE, bool ENERGY>
__device__ __forceinline__ void ewald_real(double dx, double dy, double dz, const PairParams &pp, int skip_min,
                                           int only_min, double invm, double &fx, double &fy, double &fz,
                                           double &e) {
  const double al = pp.alpha, q = pp.q;
  const double c_gauss = q * 2 * al / sqrt(M_PI);
  if (only_min) {
    double rx = pf_min_image(dx, pp.L[0]), ry = pf_min_image(dy, pp.L[1]), rz = pf_min_image(dz, pp.L[2]);
    double R = sqrt(rx * rx + ry * ry + rz * rz);
    if (FORCE) {
      double w = c_gauss * exp(-al * al * R * R) / (R * R) + q * erfc(al * R) / (R * R * R);
      fx += rx * w * invm;
      fy += ry * w * invm;
      fz += rz * w * invm;
    }
    if (ENERGY) e += q * erfc(al * R) / R;
    return;
  }
  const double rc = sqrt(pp.rc2);
  double mx = 0, my = 0, mz = 0;
  if (skip_min) {
    mx = fabs(pf_min_image(-dx, pp.L[0]));
    my = fabs(pf_min_image(-dy, pp.L[1]));
    mz = fabs(pf_min_image(-dz, pp.L[2]));
  }
  const int n = pp.nsum;
  int ax = max(-n, (int)ceil((dx - rc) / pp.L[0])), bx = min(n, (int)floor((dx + rc) / pp.L[0]));
  int ay = max(-n, (int)ceil((dy - rc) / pp.L[1])), by = min(n, (int)floor((dy + rc) / pp.L[1]));
  int az = max(-n, (int)ceil((dz - rc) / pp.L[2])), bz = min(n, (int)floor((dz + rc) / pp.L[2]));
  for (int i = ax; i <= bx; ++i) {
    double rx = dx - i * pp.L[0];
    double rx2 = rx * rx;
    if (rx2 > pp.rc2) continue;
    for (int j = ay; j <= by; ++j) {
      double ry = dy - j * pp.L[1];
      double rxy2 = rx2 + ry * ry;
      if (rxy2 > pp.rc2) continue;
      for (int k = az; k <= bz; ++k) {
        double rz = dz - k * pp.L[2];
        double R2 = rxy2 + rz * rz;
        if (R2 > pp.rc2) continue;
        if (skip_min && fabs(fabs(rx) - mx) < 1e-4 && fabs(fabs(ry) - my) < 1e-4 && fabs(fabs(rz) - mz) < 1e-4)
          continue;
        double R = sqrt(R2);
        double ec = erfc(al * R);
        if (FORCE) {
          double w = c_gauss * exp(-al * al * R * R) / (R * R) + q * ec / (R * R * R);
          fx += rx * w * invm;
          fy += ry * w * invm;
          fz += rz * w * invm;
        }
        if (ENERGY) e += q * ec / R;
      }
    }
  }
}

// MODE: PK_GAUSS, PK_PGAUSS, PK_EWALD (real part), PK_EWALD_R (real part,
// minimum image skipped), PK_EWALD_NI (minimum image only).  ENERGY: write
// per-block energy partials instead of forces.
template <int MODE, bool ENERGY>
__global__ __launch_bounds__(256) void k_pair(const double *__restrict__ X, const double *__restrict__ M,
                                              double *__restrict__ F, int N, int nb, PairParams pp,
                                              double *__restrict__ epart) {
  __shared__ double sx[TILE], sy[TILE], sz[TILE], sm[TILE];
  __shared__ double red[4][4][64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int slice = blockIdx.y;
  const int base = slice * N;
  const int i = blockIdx.x * 64 + lane;
  const bool act = i < N;
  double xi = 0, yi = 0, zi = 0, mi = 1;
  if (act) {
    xi = X[base + i];
    yi = X[nb + base + i];
    zi = X[2 * nb + base + i];
    mi = M[base + i];
  }
  double fx = 0, fy = 0, fz = 0, e = 0;
  for (int t0 = 0; t0 < N; t0 += TILE) {
    const int cnt = min(TILE, N - t0);
    __syncthreads();
    for (int q = threadIdx.x; q < cnt; q += 256) {
      sx[q] = X[base + t0 + q];
      sy[q] = X[nb + base + t0 + q];
      sz[q] = X[2 * nb + base + t0 + q];
      sm[q] = M[base + t0 + q];
    }
    __syncthreads();
    if (act) {
      for (int q = wv; q < cnt; q += 4) {
        const int j = t0 + q;
        if (j == i) continue;
        const double invm = 1.0 / (j > i ? sm[q] : mi);
        double dx = xi - sx[q], dy = yi - sy[q], dz = zi - sz[q];
        if (MODE == PK_GAUSS || MODE == PK_PGAUSS) {
          if (MODE == PK_PGAUSS) {
            dx = pf_min_image(dx, pp.L[0]);
            dy = pf_min_image(dy, pp.L[1]);
            dz = pf_min_image(dz, pp.L[2]);
          }
          double d = sqrt(dx * dx + dy * dy + dz * dz);
          double w = (pp.g / (M_PI * pp.s2)) * exp(-d * d / pp.s2);
          if (ENERGY) {
            e += w;
          } else {
            fx += w * (2 * dx) / pp.s2 * invm;
            fy += w * (2 * dy) / pp.s2 * invm;
            fz += w * (2 * dz) / pp.s2 * invm;
          }
        } else {
          ewald_real<!ENERGY, ENERGY>(dx, dy, dz, pp, MODE == PK_EWALD_R, MODE == PK_EWALD_NI, invm, fx, fy, fz, e);
        }
      }
    }
  }
  red[0][wv][lane] = fx;
  red[1][wv][lane] = fy;
  red[2][wv][lane] = fz;
  red[3][wv][lane] = e;
  __syncthreads();
  if (ENERGY) {
    if (threadIdx.x < 64) {
      double s = red[3][0][lane] + red[3][1][lane] + red[3][2][lane] + red[3][3][lane];
      for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
      if (lane == 0) epart[blockIdx.y * gridDim.x + blockIdx.x] = s;
    }
    return;
  }
  if (wv == 0 && act) {
    double gx = red[0][0][lane] + red[0][1][lane] + red[0][2][lane] + red[0][3][lane];
    double gy = red[1][0][lane] + red[1][1][lane] + red[1][2][lane] + red[1][3][lane];
    double gz = red[2][0][lane] + red[2][1][lane] + red[2][2][lane] + red[2][3][lane];
    F[base + i] += gx / pp.P;
    F[nb + base + i] += gy / pp.P;
    F[2 * nb + base + i] += gz / pp.P;
  }
}

// ------------------------------------------------------ reciprocal space

// k-vector bookkeeping: columns (mx, my) of the half space, mx in [0, n],
// my in [-n, n], keeping mx > 0 or (mx == 0, my >= 0); each column carries
// the 2n+1 values mz in [-n, n] (column (0,0) only uses mz > 0; its other
// slots have zero weight).
struct KSpace {
  int n = -1, ncol = 0, nk = 0;
  double L[3] = {0, 0, 0}, alpha = 0, q = 0;
  double *coef = nullptr;  // [ncol*(2n+1)]: 4 pi A_k (force) ; energy uses A_k = coef / (4 pi)
  int *cols = nullptr;     // [2*ncol]: mx, my
  double *S = nullptr;     // [P][nk] complex (re, im)
  double *Spart = nullptr; // [P][nchunk][nk] complex
  size_t S_cap = 0, Spart_cap = 0;
};

static KSpace g_ks;

static int ensure_kspace(const double *L, double alpha, double q, int n, int P, int nchunk) {
  KSpace &K = g_ks;
  bool same = K.n == n && K.alpha == alpha && K.q == q && K.L[0] == L[0] && K.L[1] == L[1] && K.L[2] == L[2];
  if (!same) {
    std::vector<int> cols;
    for (int mx = 0; mx <= n; ++mx)
      for (int my = -n; my <= n; ++my)
        if (mx > 0 || my >= 0) {
          cols.push_back(mx);
          cols.push_back(my);
        }
    int ncol = (int)cols.size() / 2, nz = 2 * n + 1;
    std::vector<double> coef((size_t)ncol * nz, 0.0);
    double V = L[0] * L[1] * L[2];
    for (int c = 0; c < ncol; ++c)
      for (int t = 0; t < nz; ++t) {
        int mx = cols[2 * c], my = cols[2 * c + 1], mz = t - n;
        if (mx == 0 && my == 0 && mz <= 0) continue;
        double gx = mx / L[0], gy = my / L[1], gz = mz / L[2];
        double G2 = gx * gx + gy * gy + gz * gz;
        double A = q * exp(-M_PI * M_PI * G2 / (alpha * alpha)) / (G2 * V * M_PI);
        coef[(size_t)c * nz + t] = 4 * M_PI * A;
      }
    if (K.coef) {
      (void)hipFree(K.coef);
      (void)hipFree(K.cols);
    }
    XMD_CHECK(hipMalloc(&K.coef, coef.size() * sizeof(double)));
    XMD_CHECK(hipMalloc(&K.cols, cols.size() * sizeof(int)));
    XMD_CHECK(hipMemcpy(K.coef, coef.data(), coef.size() * sizeof(double), hipMemcpyHostToDevice));
    XMD_CHECK(hipMemcpy(K.cols, cols.data(), cols.size() * sizeof(int), hipMemcpyHostToDevice));
    K.n = n;
    K.ncol = ncol;
    K.nk = ncol * nz;
    K.alpha = alpha;
    K.q = q;
    for (int d = 0; d < 3; ++d) K.L[d] = L[d];
  }
  size_t need_S = (size_t)P * K.nk * 2, need_part = (size_t)P * nchunk * K.nk * 2;
  if (need_S > K.S_cap) {
    if (K.S) (void)hipFree(K.S);
    XMD_CHECK(hipMalloc(&K.S, need_S * sizeof(double)));
    K.S_cap = need_S;
  }
  if (need_part > K.Spart_cap) {
    if (K.Spart) (void)hipFree(K.Spart);
    XMD_CHECK(hipMalloc(&K.Spart, need_part * sizeof(double)));
    K.Spart_cap = need_part;
  }
  return 0;
}

constexpr int SK_CHUNK = 64;

// Work item = (column, half of the mz range); tables of exp(i m theta_d) for
// the chunk's beads live in LDS: ex[b][0..n], ey[b][-n..n], ez[b][-n..n].
__global__ __launch_bounds__(320) void k_sk_partial(const double *__restrict__ X, int N, int nb, int n,
                                                    const int *__restrict__ cols, int ncol, double L0, double L1,
                                                    double L2, double *__restrict__ Spart, int nchunk) {
  extern __shared__ double tab[];
  const int nz = 2 * n + 1, rowx = n + 1;
  const int per_bead = 2 * (rowx + 2 * nz);
  const int slice = blockIdx.y, chunk = blockIdx.x;
  const int b0 = chunk * SK_CHUNK;
  const int nbeads = min(SK_CHUNK, N - b0);
  // tables
  for (int w = threadIdx.x; w < nbeads * 3; w += blockDim.x) {
    int b = w / 3, d = w % 3;
    double L = d == 0 ? L0 : (d == 1 ? L1 : L2);
    double x = X[d * nb + slice * N + b0 + b];
    double th = 2 * M_PI * x / L;
    double *t = tab + b * per_bead + (d == 0 ? 0 : (d == 1 ? 2 * rowx : 2 * (rowx + nz)));
    int lo = d == 0 ? 0 : -n;
    for (int m = lo; m <= n; ++m) {
      double sn, cs;
      sincos(m * th, &sn, &cs);
      t[2 * (m - lo)] = cs;
      t[2 * (m - lo) + 1] = sn;
    }
  }
  __syncthreads();
  const int nk = ncol * nz;
  for (int item = threadIdx.x; item < 2 * ncol; item += blockDim.x) {
    const int col = item >> 1, half = item & 1;
    const int mx = cols[2 * col], my = cols[2 * col + 1];
    const int zlo = half ? 1 : -n, zhi = half ? n : 0;
    double acc_re[17], acc_im[17];
#pragma unroll
    for (int t = 0; t < 17; ++t) acc_re[t] = acc_im[t] = 0;
    for (int b = 0; b < nbeads; ++b) {
      const double *t = tab + b * per_bead;
      double axr = t[2 * mx], axi = t[2 * mx + 1];
      double ayr = t[2 * rowx + 2 * (my + n)], ayi = t[2 * rowx + 2 * (my + n) + 1];
      double er = axr * ayr - axi * ayi, ei = axr * ayi + axi * ayr;
      const double *tz = t + 2 * (rowx + nz);
#pragma unroll
      for (int u = 0; u < 17; ++u) {
        int mz = zlo + u;
        if (mz > zhi) break;
        double zr = tz[2 * (mz + n)], zi = tz[2 * (mz + n) + 1];
        acc_re[u] += er * zr - ei * zi;
        acc_im[u] += er * zi + ei * zr;
      }
    }
    double *out = Spart + ((size_t)(slice * nchunk + chunk) * nk) * 2;
#pragma unroll
    for (int u = 0; u < 17; ++u) {
      int mz = zlo + u;
      if (mz > zhi) break;
      int kk = col * nz + (mz + n);
      out[2 * kk] = acc_re[u];
      out[2 * kk + 1] = acc_im[u];
    }
  }
}

__global__ void k_sk_reduce(const double *__restrict__ Spart, double *__restrict__ S, int nk, int nchunk, int P) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nk * P) return;
  int slice = t / nk, kk = t % nk;
  double re = 0, im = 0;
  for (int c = 0; c < nchunk; ++c) {
    const double *p = Spart + ((size_t)(slice * nchunk + c) * nk + kk) * 2;
    re += p[0];
    im += p[1];
  }
  S[2 * (size_t)t] = re;
  S[2 * (size_t)t + 1] = im;
}

// Thread = bead; the four waves of a block share 64 beads and split the
// columns.  exp(i k.x) advances along mz by repeated multiplication.
__global__ __launch_bounds__(256) void k_recip_force(const double *__restrict__ X, const double *__restrict__ M,
                                                     double *__restrict__ F, int N, int nb, int n,
                                                     const int *__restrict__ cols, int ncol,
                                                     const double *__restrict__ coef, const double *__restrict__ S,
                                                     double L0, double L1, double L2, double P) {
  __shared__ double red[3][4][64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int slice = blockIdx.y;
  const int i = blockIdx.x * 64 + lane;
  const bool act = i < N;
  const int nz = 2 * n + 1, nk = ncol * nz;
  const int s = slice * N + (act ? i : 0);
  const double tx = 2 * M_PI * X[s] / L0, ty = 2 * M_PI * X[nb + s] / L1, tz = 2 * M_PI * X[2 * nb + s] / L2;
  double z1r, z1i, znr, zni;
  sincos(tz, &z1i, &z1r);
  sincos(-n * tz, &zni, &znr);
  const double *Ss = S + (size_t)slice * nk * 2;
  const int c0 = (ncol * wv) / 4, c1 = (ncol * (wv + 1)) / 4;
  double fx = 0, fy = 0, fz = 0;
  for (int c = c0; c < c1; ++c) {
    const int mx = cols[2 * c], my = cols[2 * c + 1];
    double axr, axi, ayr, ayi;
    sincos(mx * tx, &axi, &axr);
    sincos(my * ty, &ayi, &ayr);
    double xyr = axr * ayr - axi * ayi, xyi = axr * ayi + axi * ayr;
    double er = xyr * znr - xyi * zni, ei = xyr * zni + xyi * znr;
    const double *Sc = Ss + (size_t)c * nz * 2;
    const double *Kc = coef + (size_t)c * nz;
    double acc_x = 0, acc_z = 0;
    for (int t = 0; t < nz; ++t) {
      const double sr = Sc[2 * t], si = Sc[2 * t + 1];
      const double w = Kc[t] * (ei * sr - er * si);
      acc_x += w;
      acc_z += w * (t - n);
      double nr = er * z1r - ei * z1i;
      ei = er * z1i + ei * z1r;
      er = nr;
    }
    fx += acc_x * mx;
    fy += acc_x * my;
    fz += acc_z;
  }
  red[0][wv][lane] = fx;
  red[1][wv][lane] = fy;
  red[2][wv][lane] = fz;
  __syncthreads();
  if (wv == 0 && act) {
    double gx = red[0][0][lane] + red[0][1][lane] + red[0][2][lane] + red[0][3][lane];
    double gy = red[1][0][lane] + red[1][1][lane] + red[1][2][lane] + red[1][3][lane];
    double gz = red[2][0][lane] + red[2][1][lane] + red[2][2][lane] + red[2][3][lane];
    const double m = M[s];
    F[s] += gx / L0 / m / P;
    F[nb + s] += gy / L1 / m / P;
    F[2 * nb + s] += gz / L2 / m / P;
  }
}

// Reciprocal energy of every slice: sum_k A_k (|S_k|^2 - N), into partials.
__global__ __launch_bounds__(256) void k_recip_energy(const double *__restrict__ S, const double *__restrict__ coef,
                                                      int nk, int P, int N, double *__restrict__ part) {
  __shared__ double red[256];
  double s = 0;
  for (int t = blockIdx.x * 256 + threadIdx.x; t < nk * P; t += gridDim.x * 256) {
    int kk = t % nk;
    double re = S[2 * (size_t)t], im = S[2 * (size_t)t + 1];
    s += coef[kk] / (4 * M_PI) * (re * re + im * im - N);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// ------------------------------------------------------------ launchers

int pair_kind_of(md_pair_force_t_x pf) {
  if (pf == md_gaussian_force_x) return PK_GAUSS;
  if (pf == md_periodic_gaussian_force_x) return PK_PGAUSS;
  if (pf == md_coulomb_3d_ewald_force_x) return PK_EWALD;
  if (pf == md_coulomb_3d_ewald_force_R_x) return PK_EWALD_R;
  if (pf == md_coulomb_3d_ewald_force_NI_x) return PK_EWALD_NI;
  return -1;
}

int pair_energy_kind_of(md_pair_energy_t_x pe) {
  if (pe == md_gaussian_energy_x) return PK_GAUSS;
  if (pe == md_periodic_gaussian_energy_x) return PK_PGAUSS;
  if (pe == md_coulomb_3d_ewald_energy_x) return PK_EWALD;
  return -1;
}

// Real-space cutoff: erfc(alpha r_c) = erfc(6) ~ 2e-17.
double g_ewald_xcut = 6.0;

static PairParams make_params(DevPimd *p) {
  PairParams pp;
  md_simulation_t_x *sim = p->host->sim;
  for (int d = 0; d < 3; ++d) pp.L[d] = box_L(sim, d);
  pp.g = md_gaussian_strength_x;
  pp.s2 = md_gaussian_range_x * md_gaussian_range_x;
  pp.q = md_coulomb_strength_x;
  pp.alpha = md_coulomb_truc_x;
  double rc = g_ewald_xcut / md_coulomb_truc_x;
  pp.rc2 = rc * rc;
  pp.nsum = md_coulomb_n_sum_x;
  pp.P = (double)p->P;
  return pp;
}

template <int MODE, bool ENERGY>
static int launch_pair_mode(DevPimd *p, const PairParams &pp, hipStream_t st, double *epart) {
  DevSim *d = p->ds;
  dim3 grid((p->N + 63) / 64, p->P);
  hipLaunchKernelGGL((k_pair<MODE, ENERGY>), grid, dim3(256), 0, st, d->X, d->M, d->F, p->N, d->nb, pp, epart);
  XMD_CHECK(hipGetLastError());
  return 0;
}

static int launch_recip(DevPimd *p, const PairParams &pp, hipStream_t st, bool force) {
  DevSim *d = p->ds;
  int n = pp.nsum;
  if (n > 16) {
    fprintf(stderr, "xmd: Ewald n_sum > 16 is not supported by the structure-factor kernels\n");
    return -1;
  }
  int nchunk = (p->N + SK_CHUNK - 1) / SK_CHUNK;
  if (ensure_kspace(pp.L, pp.alpha, pp.q, n, p->P, nchunk)) return -1;
  KSpace &K = g_ks;
  int nz = 2 * n + 1;
  size_t lds = (size_t)SK_CHUNK * 2 * ((n + 1) + 2 * nz) * sizeof(double);
  hipLaunchKernelGGL(k_sk_partial, dim3(nchunk, p->P), dim3(320), lds, st, d->X, p->N, d->nb, n, K.cols, K.ncol,
                     pp.L[0], pp.L[1], pp.L[2], K.Spart, nchunk);
  XMD_CHECK(hipGetLastError());
  int tot = K.nk * p->P;
  hipLaunchKernelGGL(k_sk_reduce, dim3((tot + 255) / 256), dim3(256), 0, st, K.Spart, K.S, K.nk, nchunk, p->P);
  XMD_CHECK(hipGetLastError());
  if (force) {
    hipLaunchKernelGGL(k_recip_force, dim3((p->N + 63) / 64, p->P), dim3(256), 0, st, d->X, d->M, d->F, p->N, d->nb,
                       n, K.cols, K.ncol, K.coef, K.S, pp.L[0], pp.L[1], pp.L[2], pp.P);
    XMD_CHECK(hipGetLastError());
  }
  return 0;
}

int launch_pair_real(DevPimd *p, int kind, hipStream_t st) {
  PairParams pp = make_params(p);
  switch (kind) {
    case PK_GAUSS:
      return launch_pair_mode<PK_GAUSS, false>(p, pp, st, nullptr);
    case PK_PGAUSS:
      return launch_pair_mode<PK_PGAUSS, false>(p, pp, st, nullptr);
    case PK_EWALD:
    case PK_EWALD_R:
    case PK_EWALD_NI:
      if (!p->ds->uniform_mass) {
        fprintf(stderr, "xmd: the Ewald kernels need equal bead masses\n");
        return -1;
      }
      if (kind == PK_EWALD) return launch_pair_mode<PK_EWALD, false>(p, pp, st, nullptr);
      if (kind == PK_EWALD_R) return launch_pair_mode<PK_EWALD_R, false>(p, pp, st, nullptr);
      return launch_pair_mode<PK_EWALD_NI, false>(p, pp, st, nullptr);
    default:
      return -1;
  }
}

bool pair_has_recip(int kind) { return kind == PK_EWALD || kind == PK_EWALD_R; }

int launch_pair_recip(DevPimd *p, hipStream_t st) {
  PairParams pp = make_params(p);
  return launch_recip(p, pp, st, true);
}

int launch_pair_force(DevPimd *p, int kind, hipStream_t st) {
  if (launch_pair_real(p, kind, st)) return -1;
  if (pair_has_recip(kind)) return launch_pair_recip(p, st);
  return 0;
}

int pair_energy(DevPimd *p, int kind, double *out) {
  PairParams pp = make_params(p);
  DevSim *d = p->ds;
  hipStream_t st = rt().s0;
  int nblk = ((p->N + 63) / 64) * p->P;
  if (nblk > d->red_len) {
    fprintf(stderr, "xmd: reduction scratch too small\n");
    return -1;
  }
  double e_pairs = 0;
  if (kind == PK_GAUSS) {
    if (launch_pair_mode<PK_GAUSS, true>(p, pp, st, d->red)) return -1;
  } else if (kind == PK_PGAUSS) {
    if (launch_pair_mode<PK_PGAUSS, true>(p, pp, st, d->red)) return -1;
  } else if (kind == PK_EWALD) {
    if (launch_pair_mode<PK_EWALD, true>(p, pp, st, d->red)) return -1;
  } else {
    return -1;
  }
  if (device_sum(d->red, nblk, &e_pairs)) return -1;
  double e = e_pairs / 2;  // ordered pairs counted twice
  if (kind == PK_EWALD) {
    if (launch_recip(p, pp, st, false)) return -1;
    KSpace &K = g_ks;
    int nb2 = 256;
    hipLaunchKernelGGL(k_recip_energy, dim3(nb2), dim3(256), 0, st, K.S, K.coef, K.nk, p->P, p->N, d->red);
    XMD_CHECK(hipGetLastError());
    double er = 0;
    if (device_sum(d->red, nb2, &er)) return -1;
    double V = pp.L[0] * pp.L[1] * pp.L[2];
    double bg = -pp.q * M_PI / (pp.alpha * pp.alpha * V) * 0.5 * p->N * (p->N - 1) * p->P;
    e += er + bg;
  }
  *out = e / p->P;
  return 0;
}

}  // namespace xmd

using namespace xmd;

extern "C" {

int md_pimd_calc_pair_force_x(md_pimd_t_x *pimd, md_pair_force_t_x pf) {
  int kind = pair_kind_of(pf);
  if (kind < 0) {
    fprintf(stderr, "xmd: pair-force callback has no device kernel\n");
    return -1;
  }
  if (ensure_pimd_device(pimd)) return -1;
  return launch_pair_force(pimd_state(pimd, false), kind, rt().s0);
}

int md_pimd_calc_pair_energy_x(md_pimd_t_x *pimd, md_pair_energy_t_x pe, md_stats_t_x *stats) {
  int kind = pair_energy_kind_of(pe);
  if (kind < 0) {
    fprintf(stderr, "xmd: pair-energy callback has no device kernel\n");
    return -1;
  }
  if (ensure_pimd_device(pimd)) return -1;
  double e = 0;
  if (pair_energy(pimd_state(pimd, false), kind, &e)) return -1;
  stats->es[stats->N - 1] += e;
  return 0;
}

}  // extern "C"
