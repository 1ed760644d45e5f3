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

#include "xmd_device.h"
#include "xmd_internal.h"

namespace xmd {

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

// (1 + 2x) exp(x^2) erfc(x) on 0 <= x <= 6.5 as a degree-20 Chebyshev series
// in s(y), y = (x - 3)/(x + 3) (the transform of Shepherd & Laframboise,
// Math. Comp. 36 (1981) 249); coefficients fitted here at 40 digits,
// max relative error 4.4e-16 in FP64 evaluation.
__device__ __constant__ double kErfcxCheb[21] = {
    1.2007410605728304,     0.065198450961816964,   -0.092409466117412367,  0.034405477918768981,
    -0.0076296298533925291, 0.0010403531055794199,  -6.3272734583294028e-05, -4.6011890721509966e-06,
    1.0119739574626279e-06, 9.9431600651180912e-09, -1.3295432865350941e-08, -9.4590207011356631e-12,
    1.9114449190078913e-10, 3.6405902241118161e-12, -2.8882712550806537e-12, -1.780186683919635e-13,
    3.9543737314097065e-14, 5.4923737447575301e-15, -3.4794118944385234e-16, -1.0306316042149898e-16,
    -1.3876025693966222e-17};
constexpr double kErfcxY1 = 0.36842105263157893;  // y at x = 6.5

// erfc(x) given e = exp(-x^2), 0 <= x <= 6.5.  One reciprocal serves both
// the transform y = (x-3)/(x+3) and the final division by (1 + 2x).
__device__ __forceinline__ double erfc_cheb(double x, double e) {
  const double d1 = x + 3.0, d2 = 1.0 + 2.0 * x;
  const double a = d1 * d2;
  double r = __builtin_amdgcn_rcp(a);
  r = fma(fma(-a, r, 1.0), r, r);
  r = fma(fma(-a, r, 1.0), r, r);
  const double y = (x - 3.0) * (d2 * r);
  const double sc = 2.0 / (kErfcxY1 + 1.0), sh = (kErfcxY1 - 1.0) / (kErfcxY1 + 1.0);
  const double t = sc * y - sh, t2 = 2.0 * t;
  double b1 = 0.0, b2 = 0.0;
#pragma unroll
  for (int k = 20; k >= 1; --k) {
    const double b0 = fma(t2, b1, kErfcxCheb[k] - b2);
    b2 = b1;
    b1 = b0;
  }
  const double c = fma(t, b1, kErfcxCheb[0] - b2);
  return e * (c * (d1 * r));
}

__device__ __forceinline__ double erfc_from_exp(double x, double e) {
  if (x > 6.5) return erfc(x);
  return erfc_cheb(x, e);
}

// One real-space image r (= p1 - image of p2) at distance R: force factor
// (q 2 alpha exp(-alpha^2 R^2)/(sqrt(pi) R^2) + q erfc(alpha R)/R^3) as at
// md_simulation_x.c:288-289, energy q erfc(alpha R)/R (md_statistics_x.c:393).
// INRANGE: alpha R <= 6.5 is guaranteed by the caller (images inside the
// cutoff 6/alpha).
template <bool FORCE, bool ENERGY, bool INRANGE = false>
__device__ __forceinline__ void ewald_image(double rx, double ry, double rz, double R2, const PairParams &pp,
                                            double &gx, double &gy, double &gz, double &e) {
  const double invR = rsqrt(R2);
  const double R = R2 * invR;
  const double x = pp.alpha * R;
  const double ex = exp(-x * x);
  const double ec = INRANGE ? erfc_cheb(x, ex) : erfc_from_exp(x, ex);
  if (FORCE) {
    const double i2 = invR * invR;
    const double w = i2 * ((pp.q * 2 * pp.alpha / sqrt(M_PI)) * ex + pp.q * ec * invR);
    gx += rx * w;
    gy += ry * w;
    gz += rz * w;
  }
  if (ENERGY) e += pp.q * ec * invR;
}

// Images of one pair within r_c when r_c < sqrt(2) L: every such image takes
// per dimension one of the three nearest periodic copies a (minimum image),
// b = a -+ L, c = a +- L, and at most one dimension takes c.  Candidates are
// numbered 0..19 (0-7: {a,b}^3, 8-11/12-15/16-19: c in x/y/z); each lane
// builds the bit mask of its candidates inside r_c and the wave walks the
// set bits, so a step costs the largest per-lane image count rather than
// all candidates any lane needs.  skip_min drops (a,a,a) (the R split),
// only_min keeps only it (the NI split).
template <bool FORCE, bool ENERGY>
__device__ __forceinline__ void ewald_real_near(double dx, double dy, double dz, const PairParams &pp, int skip_min,
                                                int only_min, double &gx, double &gy, double &gz, double &e) {
  const double ax = pf_min_image(dx, pp.L[0]), ay = pf_min_image(dy, pp.L[1]), az = pf_min_image(dz, pp.L[2]);
  if (only_min) {
    ewald_image<FORCE, ENERGY>(ax, ay, az, ax * ax + ay * ay + az * az, pp, gx, gy, gz, e);
    return;
  }
  const double bx = ax - copysign(pp.L[0], ax), by = ay - copysign(pp.L[1], ay), bz = az - copysign(pp.L[2], az);
  const double cx = ax + copysign(pp.L[0], ax), cy = ay + copysign(pp.L[1], ay), cz = az + copysign(pp.L[2], az);
  const double ax2 = ax * ax, ay2 = ay * ay, az2 = az * az;
  const double bx2 = bx * bx, by2 = by * by, bz2 = bz * bz;
  const double cx2 = cx * cx, cy2 = cy * cy, cz2 = cz * cz;
  const double rc2 = pp.rc2;
  unsigned mask = 0;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const double R2 = ((q & 1) ? bx2 : ax2) + ((q & 2) ? by2 : ay2) + ((q & 4) ? bz2 : az2);
    mask |= (R2 <= rc2 ? 1u : 0u) << q;
  }
  if (skip_min) mask &= ~1u;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const double yx = (u & 1) ? by2 : ay2, zx = (u & 2) ? bz2 : az2;  // c in x
    const double xy = (u & 1) ? bx2 : ax2, zy = (u & 2) ? bz2 : az2;  // c in y
    const double xz = (u & 1) ? bx2 : ax2, yz = (u & 2) ? by2 : ay2;  // c in z
    mask |= (cx2 + yx + zx <= rc2 ? 1u : 0u) << (8 + u);
    mask |= (xy + cy2 + zy <= rc2 ? 1u : 0u) << (12 + u);
    mask |= (xz + yz + cz2 <= rc2 ? 1u : 0u) << (16 + u);
  }
  while (mask) {
    const int q = __builtin_ctz(mask);
    mask &= mask - 1;
    const int cdim = q < 8 ? -1 : (q - 8) >> 2;
    const int u = q < 8 ? q : ((q - 8) & 3);
    const int xb = u & 1;
    const int yb = cdim == 0 ? (u & 1) : ((u >> 1) & 1);
    const int zb = cdim < 0 ? ((u >> 2) & 1) : ((u >> 1) & 1);
    const double rx = cdim == 0 ? cx : (xb ? bx : ax);
    const double ry = cdim == 1 ? cy : (yb ? by : ay);
    const double rz = cdim == 2 ? cz : (zb ? bz : az);
    ewald_image<FORCE, ENERGY, true>(rx, ry, rz, rx * rx + ry * ry + rz * rz, pp, gx, gy, gz, e);
  }
}

// Force (raw, not divided by mass) on bead i from bead j and/or pair energy.
template <int MODE, bool ENERGY, bool NEAR>
__device__ __forceinline__ void pair_term(double dx, double dy, double dz, const PairParams &pp, double &gx,
                                          double &gy, double &gz, double &e) {
  if (MODE == PK_GAUSS || MODE == PK_PGAUSS) {
    if (MODE == PK_PGAUSS) {
      dx = pf_min_image(dx, pp.L[0]);
      dy = pf_min_image(dy, pp.L[1]);
      dz = pf_min_image(dz, pp.L[2]);
    }
    const double d = sqrt(dx * dx + dy * dy + dz * dz);
    const double w = (pp.g / (M_PI * pp.s2)) * exp(-d * d / pp.s2);
    if (ENERGY) {
      e += w;
    } else {
      gx += w * (2 * dx) / pp.s2;
      gy += w * (2 * dy) / pp.s2;
      gz += w * (2 * dz) / pp.s2;
    }
  } else if (NEAR) {
    ewald_real_near<!ENERGY, ENERGY>(dx, dy, dz, pp, MODE == PK_EWALD_R, MODE == PK_EWALD_NI, gx, gy, gz, e);
  } else {
    ewald_real<!ENERGY, ENERGY>(dx, dy, dz, pp, MODE == PK_EWALD_R, MODE == PK_EWALD_NI, 1.0, gx, gy, gz, e);
  }
}

// Symmetric tile-pair kernel.  One 64-lane workgroup per (slice, I, J),
// I <= J (64-bead tiles of one slice).  Lane t owns bead i = 64 I + t; the
// J tile sits in LDS together with its reaction accumulators.  At rotation
// step s lane t meets partner (t + s) mod 64, so within a step every lane
// updates a different reaction slot and no atomics are needed.  Each pair
// is evaluated once (Newton's third law), divided by the mass of the
// higher-index bead as in md_pimd_x.c:1669-1673.
//
// Output: per-bead partial sums in Fpart[(d*nb + bead)*T + slot], where
// slot is the other tile of the pair, so every slot is written exactly once
// and k_reduce_tiles sums them in a fixed order (deterministic forces).
template <int MODE, bool ENERGY, bool NEAR>
__global__ __launch_bounds__(64) void k_pair_tiles(const double *__restrict__ X, const double *__restrict__ M,
                                                   double *__restrict__ Fpart, double *__restrict__ Epart,
                                                   const int2 *__restrict__ tiles, int N, int nb, int T, int ntp,
                                                   PairParams pp) {
  __shared__ double jx[64], jy[64], jz[64], jm[64], rx[64], ry[64], rz[64];
  const int item = blockIdx.x;
  const int slice = item / ntp;
  const int2 IJ = tiles[item - slice * ntp];
  const int I = IJ.x, J = IJ.y;
  const int t = threadIdx.x;
  const int base = slice * N;
  const int i = I * 64 + t, jl = J * 64 + t;
  const bool act = i < N;
  double xi = 0, yi = 0, zi = 0, mi = 1;
  if (act) {
    xi = X[base + i];
    yi = X[nb + base + i];
    zi = X[2 * nb + base + i];
    mi = M[base + i];
  }
  const bool jact = jl < N;
  jx[t] = jact ? X[base + jl] : 0.0;
  jy[t] = jact ? X[nb + base + jl] : 0.0;
  jz[t] = jact ? X[2 * nb + base + jl] : 0.0;
  jm[t] = jact ? M[base + jl] : 1.0;
  rx[t] = 0;
  ry[t] = 0;
  rz[t] = 0;
  __syncthreads();
  const bool diag = I == J;
  const int s_lo = diag ? 1 : 0, s_hi = diag ? 32 : 63;
  double fx = 0, fy = 0, fz = 0, e = 0;
  for (int s = s_lo; s <= s_hi; ++s) {
    const int j = (t + s) & 63;
    const int jg = J * 64 + j;
    const bool ok = act && jg < N && !(diag && s == 32 && t >= 32);
    if (ok) {
      double gx = 0, gy = 0, gz = 0;
      pair_term<MODE, ENERGY, NEAR>(xi - jx[j], yi - jy[j], zi - jz[j], pp, gx, gy, gz, e);
      if (!ENERGY) {
        const double invm = 1.0 / (jg > i ? jm[j] : mi);
        fx += gx * invm;
        fy += gy * invm;
        fz += gz * invm;
        rx[j] -= gx * invm;
        ry[j] -= gy * invm;
        rz[j] -= gz * invm;
      }
    }
  }
  __syncthreads();
  if (ENERGY) {
    e = wave_sum(e);
    if (t == 0) Epart[item] = e;
    return;
  }
  auto put = [&](int bead, int slot, double vx, double vy, double vz) {
    const size_t s0 = (size_t)(base + bead) * T + slot;
    Fpart[s0] = vx;
    Fpart[(size_t)nb * T + s0] = vy;
    Fpart[2 * (size_t)nb * T + s0] = vz;
  };
  if (diag) {
    if (act) put(i, I, fx + rx[t], fy + ry[t], fz + rz[t]);
  } else {
    if (act) put(i, J, fx, fy, fz);
    if (jact) put(jl, I, rx[t], ry[t], rz[t]);
  }
}

// F[d*nb + s] += sum_slot Fpart / P, slots in a fixed order.
__global__ void k_reduce_tiles(const double *__restrict__ Fpart, double *__restrict__ F, int n3, int T, double P) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n3) return;
  const double *src = Fpart + (size_t)k * T;
  double s = 0;
  for (int q = 0; q < T; ++q) s += src[q];
  F[k] += s / P;
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

// Tile-pair schedule and partial-sum scratch (grown on demand, kept).
struct TileBufs {
  int N = -1, T = 0, ntp = 0;
  int2 *tiles = nullptr;
  double *Fpart = nullptr, *Epart = nullptr;
  size_t fcap = 0, ecap = 0;
};
static TileBufs g_tb;

static int ensure_tiles(int N, int P) {
  TileBufs &B = g_tb;
  if (B.N != N) {
    int T = (N + 63) / 64;
    std::vector<int2> tl;
    for (int I = 0; I < T; ++I)
      for (int J = I; J < T; ++J) tl.push_back(make_int2(I, J));
    if (B.tiles) (void)hipFree(B.tiles);
    XMD_CHECK(hipMalloc(&B.tiles, tl.size() * sizeof(int2)));
    XMD_CHECK(hipMemcpy(B.tiles, tl.data(), tl.size() * sizeof(int2), hipMemcpyHostToDevice));
    B.N = N;
    B.T = T;
    B.ntp = (int)tl.size();
  }
  size_t fneed = (size_t)3 * N * P * B.T, eneed = (size_t)P * B.ntp;
  if (fneed > B.fcap) {
    if (B.Fpart) (void)hipFree(B.Fpart);
    XMD_CHECK(hipMalloc(&B.Fpart, fneed * sizeof(double)));
    B.fcap = fneed;
  }
  if (eneed > B.ecap) {
    if (B.Epart) (void)hipFree(B.Epart);
    XMD_CHECK(hipMalloc(&B.Epart, eneed * sizeof(double)));
    B.ecap = eneed;
  }
  return 0;
}

// The near-image enumeration covers every image within r_c when
// r_c < sqrt(2) L and the reference's image cube (|n| <= n_sum around the raw
// separation) holds all of them, i.e. n_sum >= 3 for wrapped positions.
static int near_ok(const PairParams &pp) {
  double Lmin = fmin(pp.L[0], fmin(pp.L[1], pp.L[2]));
  // (the image kernel's erfc series also needs alpha r_c <= 6.5)
  return (pp.nsum >= 3 && pp.rc2 < 2.0 * Lmin * Lmin && pp.alpha * sqrt(pp.rc2) <= 6.5) ? 1 : 0;
}

template <int MODE, bool ENERGY>
static int launch_pair_mode(DevPimd *p, const PairParams &pp, hipStream_t st, double *esum) {
  DevSim *d = p->ds;
  if (ensure_tiles(p->N, p->P)) return -1;
  TileBufs &B = g_tb;
  const bool near = (MODE == PK_EWALD_NI) || near_ok(pp);
  const int items = p->P * B.ntp;
  if (near)
    hipLaunchKernelGGL((k_pair_tiles<MODE, ENERGY, true>), dim3(items), dim3(64), 0, st, d->X, d->M, B.Fpart,
                       B.Epart, B.tiles, p->N, d->nb, B.T, B.ntp, pp);
  else
    hipLaunchKernelGGL((k_pair_tiles<MODE, ENERGY, false>), dim3(items), dim3(64), 0, st, d->X, d->M, B.Fpart,
                       B.Epart, B.tiles, p->N, d->nb, B.T, B.ntp, pp);
  XMD_CHECK(hipGetLastError());
  if (ENERGY) return device_sum(B.Epart, items, esum);
  const int n3 = 3 * d->nb;
  hipLaunchKernelGGL(k_reduce_tiles, dim3((n3 + 255) / 256), dim3(256), 0, st, B.Fpart, d->F, n3, B.T, pp.P);
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
  double e = 0;  // every unordered pair once
  if (kind == PK_GAUSS) {
    if (launch_pair_mode<PK_GAUSS, true>(p, pp, st, &e)) return -1;
  } else if (kind == PK_PGAUSS) {
    if (launch_pair_mode<PK_PGAUSS, true>(p, pp, st, &e)) return -1;
  } else if (kind == PK_EWALD) {
    if (launch_pair_mode<PK_EWALD, true>(p, pp, st, &e)) return -1;
  } else {
    return -1;
  }
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
