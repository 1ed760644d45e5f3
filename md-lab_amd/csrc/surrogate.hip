// Learned stand-in for the reciprocal-space part of the Ewald sum.
//
// Per bead slice (one configuration of N particles in the periodic box) the
// reciprocal energy of the engine's Ewald split is written as a sum of
// per-particle terms, E_rec = sum_i e_i, e_i = sum_k A_k (Re(e^{ik.x_i}
// S_k*) - 1) -- an exact split of sum_k A_k (|S_k|^2 - N) that the FP64
// reciprocal kernel emits alongside its forces (the training labels) -- and
// a small network predicts e_i from the neighbourhood of particle i in its
// slice:
//
//   descriptors D_i (FP32): 64 radial functions
//       G_f = sum_l exp(-eta_r (r_il - f dmu_r)^2) fc(r_il),  f = 0..63,
//     and 32 angular invariants from 8 channels of moment tensors
//       w_c(r) = exp(-eta_a (r - (c + 1/2) dmu_a)^2) fc(r),
//       M0_c = sum_l w_c,  M1_c = sum_l w_c u_il,  M2_c = sum_l w_c u_il u_il^T
//     (u_il the unit vector from i to its minimum image of l), whose
//       M0^2 +- |M1|^2       = sum_{l,m} w w (1 +- cos theta_lim)
//       M0^2 +- 2|M1|^2 + |M2|_F^2 = sum_{l,m} w w (1 +- cos theta_lim)^2
//     are the angular sums of Behler-Parrinello type (lambda = +-1, zeta =
//     1, 2; the l = m terms included) in O(neighbours) instead of
//     O(neighbours^2); fc(r) = (cos(pi r / r_c) + 1) / 2 below r_c;
//   network (inference in BF16 with FP32 accumulation on the matrix cores):
//       D -> (D - mean) / std -> tanh(W1 . + b1) -> tanh(W2 . + b2) -> w3 . + b3,
//     96 -> 128 -> 128 -> 1, output scaled back to Hartree;
//   forces: -dE/dx by the chain rule through the network (same kernel,
//     BF16 operands, FP32 accumulation) and the descriptors (FP32): each
//     bead computes the gradient of its own e_i with respect to every
//     neighbour separation, adds it to its own force and leaves the reaction
//     in a per-(bead, neighbour slot) buffer, which a second pass gathers
//     through the sorted neighbour lists (deterministic, no atomics).
//
// Training (FP32 master weights, FP32 kernels): forward, backward to the
// weights, a fixed-order reduction of the per-block weight gradients, and
// Adam; the BF16 copies used by inference are refreshed after every update.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "md_hip_x.h"
#include "xmd_device.h"
#include "xmd_internal.h"

namespace xmd {
namespace sur {

constexpr int NR = 64, NCH = 8, NA = 4 * NCH, ND = NR + NA, NMOM = 10 * NCH, H = 128;
constexpr int OFF_W1 = 0, OFF_B1 = ND * H, OFF_W2 = OFF_B1 + H, OFF_B2 = OFF_W2 + H * H, OFF_W3 = OFF_B2 + H,
              OFF_B3 = OFF_W3 + H, NPAR = OFF_B3 + 1;
constexpr int MAXNB = 192;  // neighbour-list capacity per bead
constexpr int TB = 8;       // beads per training workgroup
constexpr int GB = 256;     // beads per weight-gradient workgroup

struct Geo {
  double L0, L1, L2;
  double rc2;          // cutoff^2 (double: the neighbour test)
  float rc, pirc;      // cutoff, pi / cutoff
  float eta_r, dmu_r;  // radial Gaussians
  float eta_a, dmu_a;  // angular channels
  int N, nb;
};

__device__ __forceinline__ double mimg(double d, double L) { return d - L * rint(d / L); }

// Separation r = x_l - x_i (minimum image) of bead slots si, sl.
__device__ __forceinline__ void separation(const double *__restrict__ X, const Geo &g, int si, int sl, float &rx,
                                           float &ry, float &rz) {
  rx = (float)mimg(X[sl] - X[si], g.L0);
  ry = (float)mimg(X[g.nb + sl] - X[g.nb + si], g.L1);
  rz = (float)mimg(X[2 * g.nb + sl] - X[2 * g.nb + si], g.L2);
}

// ------------------------------------------------------------ neighbours

// Partners of every bead within r_c in its own slice, in increasing index
// order (the force pass finds a bead's slot in its partner's list by
// bisection).  One workgroup per (slice, NLB beads): the slice's fractional
// coordinates in LDS (FP32, wrapped to [0, 1)), one wave per bead at a time,
// its lanes testing 64 partners per round; a ballot and the lane's rank in it
// place the hits in order.  The test is the same FP32 arithmetic for (i, l)
// and (l, i) (negated differences), so the lists are symmetric.  cnt >
// MAXNB raises *overflow (the host reports it).  The ballots are kept too:
// bm[s][w] has bit k set when bead 64 w + k is a partner of s, and pc[s][w]
// counts the partners below 64 w, so the slot of i in l's list is
// pc[l][i/64] + popcount(bm[l][i/64] below bit i%64) without a search.
constexpr int NLB = 16;  // beads per neighbour-list workgroup
__global__ __launch_bounds__(256) void k_sur_nlist(const double *__restrict__ X, Geo g, int nw, int *__restrict__ nbr,
                                                   int *__restrict__ cnt, unsigned long long *__restrict__ bm,
                                                   int *__restrict__ pc, int *__restrict__ overflow) {
  extern __shared__ float sp[];  // [3][N]
  const int slice = blockIdx.y, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int base = slice * g.N;
  const double iL[3] = {1.0 / g.L0, 1.0 / g.L1, 1.0 / g.L2};
  for (int t = tid; t < 3 * g.N; t += 256) {
    const int d = t / g.N, l = t - d * g.N;
    const double u = X[(size_t)d * g.nb + base + l] * iL[d];
    sp[t] = (float)(u - floor(u));
  }
  __syncthreads();
  const float L0 = (float)g.L0, L1 = (float)g.L1, L2 = (float)g.L2, rc2 = (float)g.rc2;
  for (int k = wv; k < NLB; k += 4) {
    const int i = blockIdx.x * NLB + k;
    if (i >= g.N) break;
    const float xi = sp[i], yi = sp[g.N + i], zi = sp[2 * g.N + i];
    int *out = nbr + (size_t)(base + i) * MAXNB;
    int c = 0;
    for (int t0 = 0; t0 < g.N; t0 += 64) {
      const int l = t0 + lane;
      bool hit = false;
      if (l < g.N && l != i) {
        float dx = sp[l] - xi, dy = sp[g.N + l] - yi, dz = sp[2 * g.N + l] - zi;
        dx = (dx - rintf(dx)) * L0;
        dy = (dy - rintf(dy)) * L1;
        dz = (dz - rintf(dz)) * L2;
        hit = dx * dx + dy * dy + dz * dz < rc2;
      }
      const unsigned long long m = __ballot(hit);
      const int pos = c + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
      if (hit && pos < MAXNB) out[pos] = l;
      if (lane == 0) {
        bm[(size_t)(base + i) * nw + (t0 >> 6)] = m;
        pc[(size_t)(base + i) * nw + (t0 >> 6)] = c;
      }
      c += __popcll(m);
    }
    if (lane == 0) {
      cnt[base + i] = min(c, MAXNB);
      if (c > MAXNB) atomicMax(overflow, c);
    }
  }
}

// ------------------------------------------------------------ descriptors

// The descriptors are FP32 quantities: the cutoff function and the basis
// exps use the hardware approximations (__cosf, __sinf, __expf: relative
// error ~1e-6 over the arguments met here, |x| < 30), not the correctly
// rounded library calls, which cost several times the instructions.
__device__ __forceinline__ float cutf(const Geo &g, float r) { return 0.5f * (__cosf(g.pirc * r) + 1.0f); }
__device__ __forceinline__ float dcutf(const Geo &g, float r) { return -0.5f * g.pirc * __sinf(g.pirc * r); }

// Eight consecutive radial Gaussians e_k = exp(-eta_r (r - (f0 + k) dmu_r)^2),
// k = 0..7.  With eta_r dmu_r^2 = 1/8 and rho = r / dmu_r, e_k = exp(-(f0 + k
// - rho)^2 / 8) and e_{k+1} = e_k R_k, R_k = exp(-(2 (f0 + k - rho) + 1) / 8),
// R_{k+1} = R_k e^{-1/4}: two exps and products (relative error ~1e-6; a
// start far from r may underflow to 0, where every member is below e^-28 of
// the peak).
__device__ __forceinline__ void radial_gauss8(const Geo &g, float r, int f0, float *e) {
  const float d0 = f0 - r / g.dmu_r, q = 0.77880078307f;  // e^{-1/4}
  float v = __expf(-0.125f * d0 * d0), R = __expf(-0.125f * (2.f * d0 + 1.f));
  e[0] = v;
#pragma unroll
  for (int k = 1; k < 8; ++k) {
    v *= R;
    R *= q;
    e[k] = v;
  }
}

// Geometry of one pair (bead s, partner slot sl): distance r, unit vector u
// from s to the minimum image of sl, cutoff value fc (and its derivative)
// and whether it counts (r < r_c).  In the descriptor and force kernels each
// of a bead's SUB lanes computes the geometry of one of SUB consecutive
// partners and the lanes then share it (pair_from), instead of every lane
// computing every partner's.
struct PairGeo {
  float r, ux, uy, uz, fc, dfc;
  int in;
};
__device__ __forceinline__ PairGeo pair_geo(const double *__restrict__ X, const Geo &g, int s, int sl, bool valid) {
  PairGeo p;
  float rx, ry, rz;
  separation(X, g, s, sl, rx, ry, rz);
  p.r = sqrtf(rx * rx + ry * ry + rz * rz);
  p.in = valid && p.r < g.rc && p.r > 0.f;
  const float ir = p.in ? 1.0f / p.r : 0.f;
  p.ux = rx * ir;
  p.uy = ry * ir;
  p.uz = rz * ir;
  p.fc = cutf(g, p.r);
  p.dfc = dcutf(g, p.r);
  return p;
}
__device__ __forceinline__ PairGeo pair_from(const PairGeo &p, int src) {
  PairGeo q;
  q.r = __shfl(p.r, src, 64);
  q.ux = __shfl(p.ux, src, 64);
  q.uy = __shfl(p.uy, src, 64);
  q.uz = __shfl(p.uz, src, 64);
  q.fc = __shfl(p.fc, src, 64);
  q.dfc = __shfl(p.dfc, src, 64);
  q.in = __shfl(p.in, src, 64);
  return q;
}

// Sum over the SUB lanes of a bead (consecutive lanes), every lane left
// holding the total (fixed butterfly order: deterministic).
constexpr int SUB = 8;  // lanes per bead in the descriptor and force kernels
// DPP steps: swap neighbours (quad_perm 1,0,3,2), swap pairs (quad_perm
// 2,3,0,1), then lane i of each group of eight takes lane 7 - i
// (row_half_mirror); every lane ends with the same bits (the partial sums
// meet in the same pairings on every lane, and addition commutes exactly).
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, true));
}
static_assert(SUB == 8, "the DPP sum covers groups of eight lanes");
__device__ __forceinline__ float sub_sum(float v) {
  v += dpp_f<0xB1>(v);
  v += dpp_f<0x4E>(v);
  v += dpp_f<0x141>(v);
  return v;
}

// SUB = 8 lanes per bead, each owning eight radial functions (f = 8 sub ..
// 8 sub + 7) and one angular channel (c = sub), over all of the bead's
// neighbours: raw descriptors D [nb][ND] and moments Mom [nb][NMOM] (per
// channel M0, M1 x y z, M2 xx yy zz xy xz yz).
static_assert(NR == 8 * SUB && NCH == SUB, "one lane per channel, eight radial functions per lane");
__global__ __launch_bounds__(256) void k_sur_desc(const double *__restrict__ X, Geo g, const int *__restrict__ nbr,
                                                  const int *__restrict__ cnt, float *__restrict__ D,
                                                  float *__restrict__ Mom) {
  const int gt = blockIdx.x * 256 + threadIdx.x;
  const int s = gt / SUB, sub = gt % SUB;
  if (s >= g.nb) return;
  const int base = (s / g.N) * g.N, f0 = 8 * sub;
  const float mu_a = (sub + 0.5f) * g.dmu_a;
  float G[8], m[10];
#pragma unroll
  for (int k = 0; k < 8; ++k) G[k] = 0.f;
#pragma unroll
  for (int k = 0; k < 10; ++k) m[k] = 0.f;
  const int n = cnt[s];
  const int *lst = nbr + (size_t)s * MAXNB;
  const int grp = (threadIdx.x & 63) & ~(SUB - 1);
  for (int q0 = 0; q0 < n; q0 += SUB) {
    const int qm = min(q0 + sub, n - 1);
    const PairGeo mine = pair_geo(X, g, s, base + lst[qm], q0 + sub < n);
#pragma unroll
    for (int j = 0; j < SUB; ++j) {
      const PairGeo p = pair_from(mine, grp + j);
      if (!p.in) continue;
      float e[8];
      radial_gauss8(g, p.r, f0, e);
#pragma unroll
      for (int k = 0; k < 8; ++k) G[k] += e[k] * p.fc;
      const float d = p.r - mu_a;
      const float w = __expf(-g.eta_a * d * d) * p.fc;
      const float ux = p.ux, uy = p.uy, uz = p.uz;
      m[0] += w;
      m[1] += w * ux;
      m[2] += w * uy;
      m[3] += w * uz;
      m[4] += w * ux * ux;
      m[5] += w * uy * uy;
      m[6] += w * uz * uz;
      m[7] += w * ux * uy;
      m[8] += w * ux * uz;
      m[9] += w * uy * uz;
    }
  }
  float *Ds = D + (size_t)s * ND;
#pragma unroll
  for (int k = 0; k < 8; ++k) Ds[f0 + k] = G[k];
  const float a0 = m[0] * m[0], a1 = m[1] * m[1] + m[2] * m[2] + m[3] * m[3];
  const float a2 = m[4] * m[4] + m[5] * m[5] + m[6] * m[6] + 2.f * (m[7] * m[7] + m[8] * m[8] + m[9] * m[9]);
  Ds[NR + 4 * sub + 0] = a0 + a1;
  Ds[NR + 4 * sub + 1] = a0 - a1;
  Ds[NR + 4 * sub + 2] = a0 + 2.f * a1 + a2;
  Ds[NR + 4 * sub + 3] = a0 - 2.f * a1 + a2;
  float *Ms = Mom + (size_t)s * NMOM + 10 * sub;
#pragma unroll
  for (int k = 0; k < 10; ++k) Ms[k] = m[k];
}

// Round to nearest even BF16 (finite inputs).
__device__ __forceinline__ unsigned short to_bf16(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__host__ __device__ inline float bf16_to_f(unsigned short h) {
#ifdef __HIP_DEVICE_COMPILE__
  return __uint_as_float((unsigned)h << 16);
#else
  unsigned u = (unsigned)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
#endif
}

// Normalised descriptors: FP32 (training) and BF16 (inference).
__global__ void k_sur_norm(const float *__restrict__ D, const float *__restrict__ mean,
                           const float *__restrict__ istd, int n, float *__restrict__ Dn,
                           unsigned short *__restrict__ Db) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int f = t % ND;
  const float v = (D[t] - mean[f]) * istd[f];
  Dn[t] = v;
  Db[t] = to_bf16(v);
}

// ------------------------------------------------ inference (matrix cores)

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4f mfma16(v4s a, v4s b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ v4s lds4(const unsigned short *p) {
  const uint2 u = *(const uint2 *)p;  // 4 consecutive bf16, 8-byte aligned
  v4s r;
  r[0] = (short)(u.x & 0xffffu);
  r[1] = (short)(u.x >> 16);
  r[2] = (short)(u.y & 0xffffu);
  r[3] = (short)(u.y >> 16);
  return r;
}

struct MlpArgs {
  const unsigned short *Db;                    // [nb][ND] normalised descriptors, BF16
  const unsigned short *w1t, *w1, *w2t, *w2;   // BF16: W1^T [H][ND], W1 [ND][H], W2^T [H][H], W2 [H][H]
  const float *b1, *b2, *w3, *b3;              // FP32
  const float *istd;                           // [ND]
  float ymu, ysd;                              // e = (y + b3) ysd + ymu
  float *eps;                                  // [nb] predicted e_i (Hartree)
  float *dD;                                   // [nb][ND] de_i / dD_i (raw descriptors)
  int nb;
};

// One wave per tile of 16 beads; per tile three GEMM layers forward and two
// backward (to the descriptors), V_MFMA_F32_16X16X16_BF16 with the operand
// layout A[i][k]: lane i + 16 (k / 4), B[k][j]: lane j + 16 (k / 4) (four
// consecutive k per lane), D[i][j]: lane j + 16 (i / 4), register i % 4.
// Layer outputs go through a per-wave LDS tile (BF16, [16][128]) to become
// the next layer's A operand.  One tile per wave serves every layer: a layer
// loads its whole A operand into registers before it writes its output over
// the tile, and a wave's LDS accesses take effect in program order, so the
// waves need no barriers after the weights are staged.
constexpr int MW = 8;  // waves per workgroup (two per SIMD, sharing the staged weights)
__global__ __launch_bounds__(64 * MW) void k_sur_mlp(MlpArgs a) {
  extern __shared__ unsigned short lds[];
  unsigned short *sw1t = lds, *sw1 = sw1t + H * ND, *sw2t = sw1 + ND * H, *sw2 = sw2t + H * H;
  unsigned short *scr = sw2 + H * H;  // [MW][16][H]
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  for (int t = tid; t < H * ND / 4; t += 64 * MW) {
    ((uint2 *)sw1t)[t] = ((const uint2 *)a.w1t)[t];
    ((uint2 *)sw1)[t] = ((const uint2 *)a.w1)[t];
  }
  for (int t = tid; t < H * H / 4; t += 64 * MW) {
    ((uint2 *)sw2t)[t] = ((const uint2 *)a.w2t)[t];
    ((uint2 *)sw2)[t] = ((const uint2 *)a.w2)[t];
  }
  __syncthreads();
  unsigned short *s0 = scr + wv * 16 * H;
  const int ri = lane & 15, kq = 4 * (lane >> 4);  // A/B operand row/column and k offset
  const int ntile = (a.nb + 15) / 16;
  const int rounds = (ntile + gridDim.x * MW - 1) / (gridDim.x * MW);
  for (int rd = 0; rd < rounds; ++rd) {
    const int tile = (rd * gridDim.x + blockIdx.x) * MW + wv;
    const bool on = tile < ntile;
    const int b0 = tile * 16;
    // layer 1: A = descriptors of the tile's 16 beads
    v4s xa[ND / 16];
    {
      const int bead = b0 + ri;
      const bool ok = on && bead < a.nb;
#pragma unroll
      for (int ks = 0; ks < ND / 16; ++ks)
        xa[ks] = ok ? lds4(a.Db + (size_t)bead * ND + 16 * ks + kq) : v4s{0, 0, 0, 0};
    }
    float dh1[8][4], dh2[8][4];
#pragma unroll
    for (int ct = 0; ct < 8; ++ct) {
      v4f c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ks = 0; ks < ND / 16; ++ks) c = mfma16(xa[ks], lds4(sw1t + (16 * ct + ri) * ND + 16 * ks + kq), c);
      const int j = 16 * ct + ri;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const float h = tanhf(c[v] + a.b1[j]);
        dh1[ct][v] = 1.f - h * h;
        s0[(kq + v) * H + j] = to_bf16(h);
      }
    }
    // layer 2 and the output
    float y[4] = {0.f, 0.f, 0.f, 0.f};
    {
      v4s ha[H / 16];
#pragma unroll
      for (int ks = 0; ks < H / 16; ++ks) ha[ks] = lds4(s0 + ri * H + 16 * ks + kq);
#pragma unroll
      for (int ct = 0; ct < 8; ++ct) {
        v4f c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < H / 16; ++ks) c = mfma16(ha[ks], lds4(sw2t + (16 * ct + ri) * H + 16 * ks + kq), c);
        const int j = 16 * ct + ri;
        const float w3 = a.w3[j];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const float h = tanhf(c[v] + a.b2[j]);
          dh2[ct][v] = 1.f - h * h;
          y[v] += w3 * h;
          // delta2 = w3 (1 - h2^2), the next A operand
          s0[(kq + v) * H + j] = to_bf16(w3 * dh2[ct][v]);
        }
      }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
      for (int o = 1; o < 16; o <<= 1) y[v] += __shfl_xor(y[v], o, 64);
    if (on && ri == 0)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int bead = b0 + kq + v;
        if (bead < a.nb) a.eps[bead] = (y[v] + a.b3[0]) * a.ysd + a.ymu;
      }
    // delta1 = (delta2 W2^T) (1 - h1^2)
    {
      v4s da[H / 16];
#pragma unroll
      for (int ks = 0; ks < H / 16; ++ks) da[ks] = lds4(s0 + ri * H + 16 * ks + kq);
#pragma unroll
      for (int ct = 0; ct < 8; ++ct) {
        v4f c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < H / 16; ++ks) c = mfma16(da[ks], lds4(sw2 + (16 * ct + ri) * H + 16 * ks + kq), c);
        const int j = 16 * ct + ri;
#pragma unroll
        for (int v = 0; v < 4; ++v) s0[(kq + v) * H + j] = to_bf16(c[v] * dh1[ct][v]);
      }
    }
    // dy/dDn = delta1 W1^T, then back to raw descriptors and Hartree
    {
      v4s da[H / 16];
#pragma unroll
      for (int ks = 0; ks < H / 16; ++ks) da[ks] = lds4(s0 + ri * H + 16 * ks + kq);
#pragma unroll
      for (int ct = 0; ct < ND / 16; ++ct) {
        v4f c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < H / 16; ++ks) c = mfma16(da[ks], lds4(sw1 + (16 * ct + ri) * H + 16 * ks + kq), c);
        const int f = 16 * ct + ri;
        const float sc = a.istd[f] * a.ysd;
        if (on)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int bead = b0 + kq + v;
            if (bead < a.nb) a.dD[(size_t)bead * ND + f] = c[v] * sc;
          }
      }
    }
  }
}

constexpr size_t MLP_LDS = (size_t)(2 * H * ND + 2 * H * H + MW * 16 * H) * sizeof(unsigned short);
static_assert(MLP_LDS <= 160 * 1024, "staged weights and per-wave tiles fit the LDS");

// ------------------------------------------------------------ forces

// Pass 1, SUB = 8 lanes per bead i (lane sub: radial functions 8 sub ..
// 8 sub + 7 and angular channel sub, as in the descriptors): zeta = de_i /
// d(moments) from de_i / dD_i, then for every neighbour l the gradient of e_i
// with respect to the separation r = x_l - x_i,
//   grad = (sum_f gam_f g_f'(r) + sum_c w_c'(r) (z0 + z1.u + u.Z2.u)) u
//          + sum_c w_c(r) [(z1 - (z1.u) u) + 2 (Z2 u - (u.Z2.u) u)] / r,
// each lane's share added across the bead's lanes; added to bead i's force
// (F_i = -de_i/dx_i = +grad) and stored as the reaction on l at i's slot in
// l's list (from the list bitmaps of k_sur_nlist), V[l][slot], so that each
// bead's reactions lie contiguous for the second pass.
__global__ __launch_bounds__(256) void k_sur_force1(const double *__restrict__ X, Geo g, int nw,
                                                    const int *__restrict__ nbr, const int *__restrict__ cnt,
                                                    const unsigned long long *__restrict__ bm,
                                                    const int *__restrict__ pc, const float *__restrict__ dD,
                                                    const float *__restrict__ Mom, double *__restrict__ Fown,
                                                    float *__restrict__ V) {
  const int gt = blockIdx.x * 256 + threadIdx.x;
  const int s0 = gt / SUB, sub = gt % SUB;
  // lanes past the last bead follow bead nb-1 (their loads stay in bounds and
  // the shuffles see full groups) but store nothing
  const bool act = s0 < g.nb;
  const int s = act ? s0 : g.nb - 1;
  const int base = (s / g.N) * g.N, me = s - base, f0 = 8 * sub;
  const float mu_a = (sub + 0.5f) * g.dmu_a;
  const float *dd = dD + (size_t)s * ND;
  float gam[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) gam[k] = dd[f0 + k];
  float z[10];
  {
    const float d0 = dd[NR + 4 * sub], d1 = dd[NR + 4 * sub + 1], d2 = dd[NR + 4 * sub + 2], d3 = dd[NR + 4 * sub + 3];
    const float *mc = Mom + (size_t)s * NMOM + 10 * sub;
    z[0] = 2.f * mc[0] * (d0 + d1 + d2 + d3);
    const float k1 = 2.f * (d0 - d1 + 2.f * d2 - 2.f * d3);
#pragma unroll
    for (int k = 1; k < 4; ++k) z[k] = k1 * mc[k];
    const float k2 = 2.f * (d2 + d3);
#pragma unroll
    for (int k = 4; k < 10; ++k) z[k] = k2 * mc[k];
  }
  double fx = 0, fy = 0, fz = 0;
  const int n = cnt[s];
  const int *lst = nbr + (size_t)s * MAXNB;
  const int grp = (threadIdx.x & 63) & ~(SUB - 1);
  for (int q0 = 0; q0 < n; q0 += SUB) {
    const int qm = min(q0 + sub, n - 1);
    const int slm = base + lst[qm];
    const PairGeo mine = pair_geo(X, g, s, slm, q0 + sub < n);
    const size_t wi = (size_t)slm * nw + (me >> 6);
    const int slot = pc[wi] + __popcll(bm[wi] & ((1ull << (me & 63)) - 1ull));
#pragma unroll
    for (int j = 0; j < SUB; ++j) {
      if (q0 + j >= n) break;
      const PairGeo p = pair_from(mine, grp + j);
      float gx = 0.f, gy = 0.f, gz = 0.f;
      if (p.in) {
        const float r = p.r, fc = p.fc, dfc = p.dfc, ux = p.ux, uy = p.uy, uz = p.uz, ir = 1.0f / r;
        float e[8];
        radial_gauss8(g, r, f0, e);
        float rad = 0.f;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          const float d = r - (f0 + k) * g.dmu_r;
          rad += gam[k] * e[k] * (dfc - 2.f * g.eta_r * d * fc);
        }
        const float d = r - mu_a;
        const float ea = __expf(-g.eta_a * d * d);
        const float w = ea * fc, dw = ea * (dfc - 2.f * g.eta_a * d * fc);
        const float t1 = z[1] * ux + z[2] * uy + z[3] * uz;
        // Z2 u with the full symmetric tensor (xx yy zz xy xz yz)
        const float qx = z[4] * ux + z[7] * uy + z[8] * uz;
        const float qy = z[7] * ux + z[5] * uy + z[9] * uz;
        const float qz = z[8] * ux + z[9] * uy + z[6] * uz;
        const float t2 = qx * ux + qy * uy + qz * uz;
        rad += dw * (z[0] + t1 + t2);
        const float wr = w * ir;
        gx = rad * ux + wr * ((z[1] - t1 * ux) + 2.f * (qx - t2 * ux));
        gy = rad * uy + wr * ((z[2] - t1 * uy) + 2.f * (qy - t2 * uy));
        gz = rad * uz + wr * ((z[3] - t1 * uz) + 2.f * (qz - t2 * uz));
      }
      gx = sub_sum(gx);
      gy = sub_sum(gy);
      gz = sub_sum(gz);
      fx += gx;
      fy += gy;
      fz += gz;
      // lane j owns pair q0 + j: its reaction goes to the partner's slot
      if (act && sub == j && slot < MAXNB) {
        float *v = V + ((size_t)slm * MAXNB + slot) * 3;
        v[0] = gx;
        v[1] = gy;
        v[2] = gz;
      }
    }
  }
  if (act && sub == 0) {
    Fown[s] = fx;
    Fown[g.nb + s] = fy;
    Fown[2 * g.nb + s] = fz;
  }
}

// Pass 2, one thread per bead l: F_l = own part - the reactions left in its
// slots (in slot order), as an acceleration of the slice energy's 1/P share:
// out[d][s] = F / (m P), or added to out when add is set.
__global__ __launch_bounds__(128) void k_sur_force2(Geo g, const int *__restrict__ cnt, const double *__restrict__ Fown,
                                                    const float *__restrict__ V, const double *__restrict__ M,
                                                    double Pd, int add, double *__restrict__ out) {
  const int s = blockIdx.x * 128 + threadIdx.x;
  if (s >= g.nb) return;
  double fx = Fown[s], fy = Fown[g.nb + s], fz = Fown[2 * g.nb + s];
  const int n = cnt[s];
  const float *v = V + (size_t)s * MAXNB * 3;
  for (int q = 0; q < n; ++q) {
    fx -= v[3 * q];
    fy -= v[3 * q + 1];
    fz -= v[3 * q + 2];
  }
  const double w = 1.0 / (M[s] * Pd);
  if (add) {
    out[s] += fx * w;
    out[g.nb + s] += fy * w;
    out[2 * g.nb + s] += fz * w;
  } else {
    out[s] = fx * w;
    out[g.nb + s] = fy * w;
    out[2 * g.nb + s] = fz * w;
  }
}

// ------------------------------------------------------------ training

// Forward pass, FP32: TB beads per workgroup, one thread per hidden unit.
// H1, H2 [nb][H] activations, Y [nb] normalised outputs.
__global__ __launch_bounds__(H) void k_sur_tr_fwd(const float *__restrict__ Dn, const float *__restrict__ par, int nb,
                                                  float *__restrict__ H1, float *__restrict__ H2,
                                                  float *__restrict__ Y) {
  __shared__ float xin[TB][ND], h1[TB][H], red[TB][H];
  const int j = threadIdx.x, b0 = blockIdx.x * TB;
  for (int t = j; t < TB * ND; t += H) {
    const int b = t / ND, k = t - b * ND;
    xin[b][k] = b0 + b < nb ? Dn[(size_t)(b0 + b) * ND + k] : 0.f;
  }
  __syncthreads();
  float z[TB];
#pragma unroll
  for (int b = 0; b < TB; ++b) z[b] = par[OFF_B1 + j];
  for (int k = 0; k < ND; ++k) {
    const float w = par[OFF_W1 + k * H + j];
#pragma unroll
    for (int b = 0; b < TB; ++b) z[b] = fmaf(xin[b][k], w, z[b]);
  }
#pragma unroll
  for (int b = 0; b < TB; ++b) {
    h1[b][j] = tanhf(z[b]);
    if (b0 + b < nb) H1[(size_t)(b0 + b) * H + j] = h1[b][j];
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < TB; ++b) z[b] = par[OFF_B2 + j];
  for (int k = 0; k < H; ++k) {
    const float w = par[OFF_W2 + k * H + j];
#pragma unroll
    for (int b = 0; b < TB; ++b) z[b] = fmaf(h1[b][k], w, z[b]);
  }
  const float w3 = par[OFF_W3 + j];
#pragma unroll
  for (int b = 0; b < TB; ++b) {
    const float h = tanhf(z[b]);
    if (b0 + b < nb) H2[(size_t)(b0 + b) * H + j] = h;
    red[b][j] = w3 * h;
  }
  __syncthreads();
  for (int w = H / 2; w > 0; w >>= 1) {
    if (j < w)
#pragma unroll
      for (int b = 0; b < TB; ++b) red[b][j] += red[b][j + w];
    __syncthreads();
  }
  if (j < TB && b0 + j < nb) Y[b0 + j] = red[j][0] + par[OFF_B3];
}

// Backward pass, FP32: loss (1/n) sum (y - t)^2; dY = 2 (y - t) / n,
// delta2 = dY w3 (1 - h2^2), delta1 = (W2 delta2) (1 - h1^2) (W2T = W2
// transposed, [out][in], FP32).  Per-workgroup loss partial into lpart.
__global__ __launch_bounds__(H) void k_sur_tr_bwd(const float *__restrict__ par, const float *__restrict__ W2T,
                                                  const float *__restrict__ T, int nb, float invn,
                                                  const float *__restrict__ H1, const float *__restrict__ H2,
                                                  const float *__restrict__ Y, float *__restrict__ DY,
                                                  float *__restrict__ D1, float *__restrict__ D2,
                                                  float *__restrict__ lpart) {
  __shared__ float d2[TB][H], dy[TB];
  const int j = threadIdx.x, b0 = blockIdx.x * TB;
  if (j < TB) {
    const int b = b0 + j;
    const float e = b < nb ? Y[b] - T[b] : 0.f;
    dy[j] = 2.f * e * invn;
    if (b < nb) DY[b] = dy[j];
  }
  __syncthreads();
  if (j == 0) {
    float l = 0.f;
    for (int b = 0; b < TB; ++b)
      if (b0 + b < nb) {
        const float e = Y[b0 + b] - T[b0 + b];
        l += e * e;
      }
    lpart[blockIdx.x] = l;
  }
  const float w3 = par[OFF_W3 + j];
#pragma unroll
  for (int b = 0; b < TB; ++b) {
    const float h = b0 + b < nb ? H2[(size_t)(b0 + b) * H + j] : 0.f;
    d2[b][j] = dy[b] * w3 * (1.f - h * h);
    if (b0 + b < nb) D2[(size_t)(b0 + b) * H + j] = d2[b][j];
  }
  __syncthreads();
  // thread j = hidden unit k of layer 1: sum over layer-2 units jj of W2[j][jj] d2[jj]
  float acc[TB];
#pragma unroll
  for (int b = 0; b < TB; ++b) acc[b] = 0.f;
  for (int jj = 0; jj < H; ++jj) {
    const float w = W2T[jj * H + j];  // W2[j][jj]
#pragma unroll
    for (int b = 0; b < TB; ++b) acc[b] = fmaf(w, d2[b][jj], acc[b]);
  }
#pragma unroll
  for (int b = 0; b < TB; ++b)
    if (b0 + b < nb) {
      const float h = H1[(size_t)(b0 + b) * H + j];
      D1[(size_t)(b0 + b) * H + j] = acc[b] * (1.f - h * h);
    }
}

// Weight-gradient partials over GB beads per workgroup: blockIdx.y = 0:
// dW1[k][j] = sum_b Dn[b][k] D1[b][j] and db1; = 1: dW2[k][j] = sum_b
// H1[b][k] D2[b][j], db2, dw3[j] = sum_b DY[b] H2[b][j], db3.  One thread per
// j; part[block][NPAR] (entries a block does not own stay untouched).
__global__ __launch_bounds__(H) void k_sur_tr_grad(const float *__restrict__ Dn, const float *__restrict__ H1,
                                                   const float *__restrict__ H2, const float *__restrict__ DY,
                                                   const float *__restrict__ D1, const float *__restrict__ D2, int nb,
                                                   float *__restrict__ part) {
  __shared__ float xr[16][H];
  const int j = threadIdx.x, b0 = blockIdx.x * GB;
  const int b1 = min(nb, b0 + GB);
  float *pp = part + (size_t)blockIdx.x * NPAR;
  if (blockIdx.y == 0) {
    float g[ND], gb = 0.f;
#pragma unroll
    for (int k = 0; k < ND; ++k) g[k] = 0.f;
    for (int c0 = b0; c0 < b1; c0 += 16) {
      __syncthreads();
      for (int t = j; t < 16 * ND; t += H) {
        const int b = t / ND, k = t - b * ND;
        xr[b][k] = c0 + b < b1 ? Dn[(size_t)(c0 + b) * ND + k] : 0.f;
      }
      __syncthreads();
      for (int b = 0; b < 16 && c0 + b < b1; ++b) {
        const float d = D1[(size_t)(c0 + b) * H + j];
        gb += d;
#pragma unroll
        for (int k = 0; k < ND; ++k) g[k] = fmaf(xr[b][k], d, g[k]);
      }
    }
#pragma unroll
    for (int k = 0; k < ND; ++k) pp[OFF_W1 + k * H + j] = g[k];
    pp[OFF_B1 + j] = gb;
  } else {
    float g[H], gb = 0.f, g3 = 0.f, gb3 = 0.f;
#pragma unroll
    for (int k = 0; k < H; ++k) g[k] = 0.f;
    for (int c0 = b0; c0 < b1; c0 += 16) {
      __syncthreads();
      for (int t = j; t < 16 * H; t += H) {
        const int b = t / H, k = t - b * H;
        xr[b][k] = c0 + b < b1 ? H1[(size_t)(c0 + b) * H + k] : 0.f;
      }
      __syncthreads();
      for (int b = 0; b < 16 && c0 + b < b1; ++b) {
        const float d = D2[(size_t)(c0 + b) * H + j];
        gb += d;
        const float dy = DY[c0 + b];
        g3 = fmaf(dy, H2[(size_t)(c0 + b) * H + j], g3);
        gb3 += dy;
#pragma unroll
        for (int k = 0; k < H; ++k) g[k] = fmaf(xr[b][k], d, g[k]);
      }
    }
#pragma unroll
    for (int k = 0; k < H; ++k) pp[OFF_W2 + k * H + j] = g[k];
    pp[OFF_B2 + j] = gb;
    pp[OFF_W3 + j] = g3;
    if (j == 0) pp[OFF_B3] = gb3;
  }
}

// Sum of the partials over the workgroups, in order.
__global__ void k_sur_reduce(const float *__restrict__ part, int nblk, float *__restrict__ grad) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= NPAR) return;
  float s = 0.f;
  for (int b = 0; b < nblk; ++b) s += part[(size_t)b * NPAR + t];
  grad[t] = s;
}

// Adam (Kingma & Ba 2015), bias-corrected step size lr_t.
__global__ void k_sur_adam(float *__restrict__ par, float *__restrict__ m, float *__restrict__ v,
                           const float *__restrict__ grad, float lr_t, float b1, float b2, float eps) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= NPAR) return;
  const float gr = grad[t];
  const float mt = b1 * m[t] + (1.f - b1) * gr;
  const float vt = b2 * v[t] + (1.f - b2) * gr * gr;
  m[t] = mt;
  v[t] = vt;
  par[t] -= lr_t * mt / (sqrtf(vt) + eps);
}

// BF16 copies (both orientations) and the FP32 W2 transpose.
__global__ void k_sur_convert(const float *__restrict__ par, unsigned short *__restrict__ w1t,
                              unsigned short *__restrict__ w1, unsigned short *__restrict__ w2t,
                              unsigned short *__restrict__ w2, float *__restrict__ W2T) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < ND * H) {
    const int k = t / H, j = t - k * H;
    const unsigned short h = to_bf16(par[OFF_W1 + t]);
    w1[t] = h;
    w1t[j * ND + k] = h;
  }
  if (t < H * H) {
    const int k = t / H, j = t - k * H;
    const float x = par[OFF_W2 + t];
    const unsigned short h = to_bf16(x);
    w2[t] = h;
    w2t[j * H + k] = h;
    W2T[j * H + k] = x;
  }
}

__global__ void k_sur_zero(int *__restrict__ v) { *v = 0; }

// Targets t_i = (e_i - el_i - ymu) / ysd from the FP64 labels e and the
// exact low-shell part el (null: none).
__global__ void k_sur_targets(const double *__restrict__ E, const double *__restrict__ El, int n, double ymu,
                              double ysd, float *__restrict__ T) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) T[t] = (float)((E[t] - (El ? El[t] : 0.0) - ymu) / ysd);
}

// Error sums of a prediction against the exact terms: [0..2] sum (dF_d)^2,
// [3..5] sum F_d^2 (exact), per workgroup; per-slice energy differences
// sum_i (e_pred - e_exact) into dE[slice] (one workgroup per slice), the
// prediction being the network's eps plus the exact low-shell el (null:
// none).
__global__ __launch_bounds__(256) void k_sur_err(const double *__restrict__ Fs, const double *__restrict__ Fx,
                                                 const float *__restrict__ eps, const double *__restrict__ El,
                                                 const double *__restrict__ Ex, int N, int nb,
                                                 double *__restrict__ fpart, double *__restrict__ dE,
                                                 double *__restrict__ Eex) {
  __shared__ double red[8][256];
  const int slice = blockIdx.x, t = threadIdx.x;
  double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = t; i < N; i += 256) {
    const int s = slice * N + i;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double x = Fx[d * nb + s], e = Fs[d * nb + s] - x;
      a[d] += e * e;
      a[3 + d] += x * x;
    }
    a[6] += (double)eps[s] + (El ? El[s] : 0.0) - Ex[s];
    a[7] += Ex[s];
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) red[k][t] = a[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w)
#pragma unroll
      for (int k = 0; k < 8; ++k) red[k][t] += red[k][t + w];
    __syncthreads();
  }
  if (t == 0) {
#pragma unroll
    for (int k = 0; k < 6; ++k) fpart[slice * 6 + k] = red[k][0];
    dE[slice] = red[6][0];
    Eex[slice] = red[7][0];
  }
}

}  // namespace sur
}  // namespace xmd

// ================================================================ host side

namespace xmd {
using namespace sur;

struct Surrogate {
  double rc = 8.0;
  int nlow = 0;  // low reciprocal shells |m| <= nlow kept exact (0: none)
  float *par = nullptr, *am = nullptr, *av = nullptr, *grad = nullptr, *W2T = nullptr;
  unsigned short *w1t = nullptr, *w1 = nullptr, *w2t = nullptr, *w2 = nullptr;
  float *mean = nullptr, *istd = nullptr;
  double ymu = 0, ysd = 1;
  bool normed = false;
  int adam_t = 0;
  // per-size buffers (beads, slices)
  int cap = 0, cap_sl = 0, cap_gb = 0;
  size_t cap_bm = 0;  // list bitmap words (nb * ceil(N / 64))
  int *nbr = nullptr, *cnt = nullptr, *ovf = nullptr, *pc = nullptr;
  unsigned long long *bm = nullptr;
  float *D = nullptr, *Mom = nullptr, *Dn = nullptr, *dD = nullptr, *eps = nullptr, *V = nullptr;
  unsigned short *Db = nullptr;
  double *Fown = nullptr, *Fs = nullptr, *Fx = nullptr, *Ex = nullptr, *El = nullptr;
  float *H1 = nullptr, *H2 = nullptr, *Y = nullptr, *DY = nullptr, *D1 = nullptr, *D2 = nullptr, *T = nullptr;
  float *lpart = nullptr, *gpart = nullptr;
  double *fpart = nullptr, *dE = nullptr, *Eex = nullptr;
  double alpha = 0;  // Ewald split of the last labels
};

static Geo sur_geo(const Surrogate *S, DevPimd *p) {
  Geo g;
  md_simulation_t_x *sim = p->host->sim;
  g.L0 = box_L(sim, 0);
  g.L1 = box_L(sim, 1);
  g.L2 = box_L(sim, 2);
  g.rc = (float)S->rc;
  g.rc2 = S->rc * S->rc;
  g.pirc = (float)(M_PI / S->rc);
  g.dmu_r = (float)(S->rc / NR);
  // Gaussian width two spacings: eta_r dmu_r^2 = 1/8 (radial_gauss relies on it)
  g.eta_r = 1.0f / (2.0f * (2.0f * g.dmu_r) * (2.0f * g.dmu_r));
  g.dmu_a = (float)(S->rc / NCH);
  g.eta_a = 1.0f / (2.0f * g.dmu_a * g.dmu_a);
  g.N = p->N;
  g.nb = p->N * p->P;
  return g;
}

template <class T>
static int dalloc(T *&ptr, size_t n) {
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  XMD_CHECK(hipMalloc(&ptr, (n ? n : 1) * sizeof(T)));
  return 0;
}

static int sur_ensure(Surrogate *S, DevPimd *p) {
  const int nb = p->N * p->P;
  if (2.0 * S->rc > std::min(box_L(p->host->sim, 0), std::min(box_L(p->host->sim, 1), box_L(p->host->sim, 2)))) {
    fprintf(stderr, "xmd: surrogate cutoff %.3g exceeds half the box\n", S->rc);
    return -1;
  }
  if (p->N > 4096) {  // the neighbour-list kernel holds a slice's coordinates in LDS
    fprintf(stderr, "xmd: the surrogate supports N <= 4096 particles per slice (got %d)\n", p->N);
    return -1;
  }
  const size_t nwb = (size_t)nb * ((p->N + 63) / 64);
  if (nwb > S->cap_bm) {
    if (dalloc(S->bm, nwb) || dalloc(S->pc, nwb)) return -1;
    S->cap_bm = nwb;
  }
  if (nb > S->cap) {
    if (dalloc(S->nbr, (size_t)nb * MAXNB) || dalloc(S->cnt, nb) || dalloc(S->D, (size_t)nb * ND) ||
        dalloc(S->Mom, (size_t)nb * NMOM) || dalloc(S->Dn, (size_t)nb * ND) || dalloc(S->Db, (size_t)nb * ND) ||
        dalloc(S->dD, (size_t)nb * ND) || dalloc(S->eps, nb) || dalloc(S->V, (size_t)nb * MAXNB * 3) ||
        dalloc(S->Fown, 3 * (size_t)nb) || dalloc(S->Fs, 3 * (size_t)nb) || dalloc(S->Fx, 3 * (size_t)nb) ||
        dalloc(S->Ex, nb) || dalloc(S->El, nb) || dalloc(S->H1, (size_t)nb * H) || dalloc(S->H2, (size_t)nb * H) ||
        dalloc(S->Y, nb) ||
        dalloc(S->DY, nb) || dalloc(S->D1, (size_t)nb * H) || dalloc(S->D2, (size_t)nb * H) || dalloc(S->T, nb) ||
        dalloc(S->lpart, (nb + TB - 1) / TB))
      return -1;
    S->cap = nb;
  }
  const int ngb = (nb + GB - 1) / GB;
  if (ngb > S->cap_gb) {
    if (dalloc(S->gpart, (size_t)ngb * NPAR)) return -1;
    XMD_CHECK(hipMemset(S->gpart, 0, (size_t)ngb * NPAR * sizeof(float)));
    S->cap_gb = ngb;
  }
  if (p->P > S->cap_sl) {
    if (dalloc(S->fpart, 6 * (size_t)p->P) || dalloc(S->dE, p->P) || dalloc(S->Eex, p->P)) return -1;
    S->cap_sl = p->P;
  }
  return 0;
}

static int sur_convert(Surrogate *S, hipStream_t st) {
  hipLaunchKernelGGL(k_sur_convert, dim3((H * H + 255) / 256), dim3(256), 0, st, S->par, S->w1t, S->w1, S->w2t, S->w2,
                     S->W2T);
  XMD_CHECK(hipGetLastError());
  return 0;
}

Surrogate *sur_create(unsigned seed, double rc) {
  Surrogate *S = new Surrogate();
  S->rc = rc;
  std::vector<float> h(NPAR, 0.f);
  // Glorot-uniform weights from a xorshift stream, zero biases
  unsigned long long st = 0x9E3779B97F4A7C15ull ^ (unsigned long long)seed;
  auto uni = [&]() {
    st ^= st << 13;
    st ^= st >> 7;
    st ^= st << 17;
    return (double)(st >> 11) * (1.0 / 9007199254740992.0);
  };
  const double a1 = sqrt(6.0 / (ND + H)), a2 = sqrt(6.0 / (H + H)), a3 = sqrt(6.0 / (H + 1));
  for (int t = 0; t < ND * H; ++t) h[OFF_W1 + t] = (float)((2 * uni() - 1) * a1);
  for (int t = 0; t < H * H; ++t) h[OFF_W2 + t] = (float)((2 * uni() - 1) * a2);
  for (int t = 0; t < H; ++t) h[OFF_W3 + t] = (float)((2 * uni() - 1) * a3);
  std::vector<float> one(ND, 1.f), zero(ND, 0.f);
  if (dalloc(S->par, NPAR) || dalloc(S->am, NPAR) || dalloc(S->av, NPAR) || dalloc(S->grad, NPAR) ||
      dalloc(S->W2T, H * H) || dalloc(S->w1t, ND * H) || dalloc(S->w1, ND * H) || dalloc(S->w2t, H * H) ||
      dalloc(S->w2, H * H) || dalloc(S->mean, ND) || dalloc(S->istd, ND) ||
      hipMemcpy(S->par, h.data(), NPAR * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(S->am, 0, NPAR * sizeof(float)) != hipSuccess ||
      hipMemset(S->av, 0, NPAR * sizeof(float)) != hipSuccess ||
      hipMemcpy(S->mean, zero.data(), ND * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(S->istd, one.data(), ND * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      sur_convert(S, rt().s0) || hipStreamSynchronize(rt().s0) != hipSuccess) {
    delete S;
    return nullptr;
  }
  return S;
}

// Neighbour lists and raw descriptors of every bead (and the normalised
// copies when the normalisation is set).
static int sur_features(Surrogate *S, DevPimd *p, hipStream_t st) {
  const Geo g = sur_geo(S, p);
  hipLaunchKernelGGL(k_sur_zero, dim3(1), dim3(1), 0, st, S->ovf);  // (a kernel: this may be captured)
  hipLaunchKernelGGL(k_sur_nlist, dim3((p->N + NLB - 1) / NLB, p->P), dim3(256), 3 * p->N * sizeof(float), st,
                     p->ds->X, g, (p->N + 63) / 64, S->nbr, S->cnt, S->bm, S->pc, S->ovf);
  hipLaunchKernelGGL(k_sur_desc, dim3((g.nb * SUB + 255) / 256), dim3(256), 0, st, p->ds->X, g, S->nbr, S->cnt, S->D,
                     S->Mom);
  const int n = g.nb * ND;
  hipLaunchKernelGGL(k_sur_norm, dim3((n + 255) / 256), dim3(256), 0, st, S->D, S->mean, S->istd, n, S->Dn, S->Db);
  XMD_CHECK(hipGetLastError());
  return 0;
}

// Predicted per-bead energies (eps) and forces (accelerations of the 1/P
// share, [3][nb]) stored into out, or added to it when add is set.
static int sur_infer(Surrogate *S, DevPimd *p, double *out, hipStream_t st, int add = 0) {
  const Geo g = sur_geo(S, p);
  MlpArgs a;
  a.Db = S->Db;
  a.w1t = S->w1t;
  a.w1 = S->w1;
  a.w2t = S->w2t;
  a.w2 = S->w2;
  a.b1 = S->par + OFF_B1;
  a.b2 = S->par + OFF_B2;
  a.w3 = S->par + OFF_W3;
  a.b3 = S->par + OFF_B3;
  a.istd = S->istd;
  a.ymu = (float)S->ymu;
  a.ysd = (float)S->ysd;
  a.eps = S->eps;
  a.dD = S->dD;
  a.nb = g.nb;
  const int ntile = (g.nb + 15) / 16;
  const int grid = std::max(1, std::min(rt().num_cus, (ntile + MW - 1) / MW));
  hipLaunchKernelGGL(k_sur_mlp, dim3(grid), dim3(64 * MW), MLP_LDS, st, a);
  hipLaunchKernelGGL(k_sur_force1, dim3((g.nb * SUB + 255) / 256), dim3(256), 0, st, p->ds->X, g, (p->N + 63) / 64,
                     S->nbr, S->cnt, S->bm, S->pc, S->dD, S->Mom, S->Fown, S->V);
  hipLaunchKernelGGL(k_sur_force2, dim3((g.nb + 127) / 128), dim3(128), 0, st, g, S->cnt, S->Fown, S->V, p->ds->M,
                     (double)p->P, add, out);
  XMD_CHECK(hipGetLastError());
  return 0;
}

static int sur_overflow(Surrogate *S) {
  int v = 0;
  XMD_CHECK(hipMemcpy(&v, S->ovf, sizeof(int), hipMemcpyDeviceToHost));
  if (v) {
    fprintf(stderr, "xmd: surrogate neighbour list overflow (%d > %d partners within r_c)\n", v, MAXNB);
    return -1;
  }
  return 0;
}

// Feature and label normalisation from the current configuration.
static int sur_normalize(Surrogate *S, DevPimd *p) {
  const int nb = p->N * p->P;
  XMD_CHECK(hipStreamSynchronize(rt().s0));
  std::vector<float> D((size_t)nb * ND);
  std::vector<double> E(nb), El(nb, 0.0);
  XMD_CHECK(hipMemcpy(D.data(), S->D, D.size() * sizeof(float), hipMemcpyDeviceToHost));
  XMD_CHECK(hipMemcpy(E.data(), S->Ex, nb * sizeof(double), hipMemcpyDeviceToHost));
  if (S->nlow) XMD_CHECK(hipMemcpy(El.data(), S->El, nb * sizeof(double), hipMemcpyDeviceToHost));
  for (int b = 0; b < nb; ++b) E[b] -= El[b];  // the network's targets
  std::vector<float> mean(ND), istd(ND);
  for (int f = 0; f < ND; ++f) {
    double s = 0, s2 = 0;
    for (int b = 0; b < nb; ++b) {
      const double x = D[(size_t)b * ND + f];
      s += x;
      s2 += x * x;
    }
    const double m = s / nb, var = std::max(0.0, s2 / nb - m * m);
    mean[f] = (float)m;
    istd[f] = (float)(var > 1e-20 ? 1.0 / sqrt(var) : 1.0);
  }
  double s = 0, s2 = 0;
  for (double e : E) {
    s += e;
    s2 += e * e;
  }
  S->ymu = s / nb;
  S->ysd = sqrt(std::max(1e-30, s2 / nb - S->ymu * S->ymu));
  XMD_CHECK(hipMemcpy(S->mean, mean.data(), ND * sizeof(float), hipMemcpyHostToDevice));
  XMD_CHECK(hipMemcpy(S->istd, istd.data(), ND * sizeof(float), hipMemcpyHostToDevice));
  S->normed = true;
  const int n = nb * ND;
  hipLaunchKernelGGL(k_sur_norm, dim3((n + 255) / 256), dim3(256), 0, rt().s0, S->D, S->mean, S->istd, n, S->Dn,
                     S->Db);
  XMD_CHECK(hipGetLastError());
  return 0;
}

// Weight gradient of the loss (1/n) sum_i (y_i - t_i)^2 on the current
// features and targets S->T; *loss (may be null) receives the loss.
static int sur_gradient(Surrogate *S, int nb, double *loss, hipStream_t st) {
  const int nt = (nb + TB - 1) / TB, ng = (nb + GB - 1) / GB;
  hipLaunchKernelGGL(k_sur_tr_fwd, dim3(nt), dim3(H), 0, st, S->Dn, S->par, nb, S->H1, S->H2, S->Y);
  hipLaunchKernelGGL(k_sur_tr_bwd, dim3(nt), dim3(H), 0, st, S->par, S->W2T, S->T, nb, 1.0f / nb, S->H1, S->H2, S->Y,
                     S->DY, S->D1, S->D2, S->lpart);
  hipLaunchKernelGGL(k_sur_tr_grad, dim3(ng, 2), dim3(H), 0, st, S->Dn, S->H1, S->H2, S->DY, S->D1, S->D2, nb,
                     S->gpart);
  hipLaunchKernelGGL(k_sur_reduce, dim3((NPAR + 255) / 256), dim3(256), 0, st, S->gpart, ng, S->grad);
  XMD_CHECK(hipGetLastError());
  if (loss) {
    std::vector<float> lp(nt);
    XMD_CHECK(hipStreamSynchronize(st));
    XMD_CHECK(hipMemcpy(lp.data(), S->lpart, nt * sizeof(float), hipMemcpyDeviceToHost));
    double l = 0;
    for (float x : lp) l += x;
    *loss = l / nb;
  }
  return 0;
}

// Exact labels Fx, Ex; with low shells kept exact, also their part: forces
// stored in Fs (the base of a prediction), energies in El.
static int sur_labels(Surrogate *S, DevPimd *p, hipStream_t st) {
  if (launch_recip_labels(p, S->Fx, S->Ex, st, &S->alpha)) return -1;
  return S->nlow ? launch_recip_shells(p, S->nlow, S->Fs, S->El, st) : 0;
}

int sur_prepare(Surrogate *S, DevPimd *p) {
  if (!S->ovf && dalloc(S->ovf, 1)) return -1;
  return sur_ensure(S, p);
}

int sur_train(Surrogate *S, DevPimd *p, double lr, int iters, double *loss) {
  if (sur_prepare(S, p)) return -1;
  hipStream_t st = rt().s0;
  const int nb = p->N * p->P;
  if (sur_labels(S, p, st) || sur_features(S, p, st)) return -1;
  if (!S->normed && sur_normalize(S, p)) return -1;
  hipLaunchKernelGGL(k_sur_targets, dim3((nb + 255) / 256), dim3(256), 0, st, S->Ex, S->nlow ? S->El : nullptr, nb,
                     S->ymu, S->ysd, S->T);
  XMD_CHECK(hipGetLastError());
  const float b1 = 0.9f, b2 = 0.999f;
  for (int it = 0; it < iters; ++it) {
    if (sur_gradient(S, nb, it == iters - 1 ? loss : nullptr, st)) return -1;
    ++S->adam_t;
    const double lr_t = lr * sqrt(1.0 - pow(b2, S->adam_t)) / (1.0 - pow(b1, S->adam_t));
    hipLaunchKernelGGL(k_sur_adam, dim3((NPAR + 255) / 256), dim3(256), 0, st, S->par, S->am, S->av, S->grad,
                       (float)lr_t, b1, b2, 1e-8f);
    XMD_CHECK(hipGetLastError());
  }
  if (sur_convert(S, st)) return -1;
  return sur_overflow(S);
}

// Surrogate vs exact on the current configuration (both evaluated here):
// out[0] mean over slices of (E_pred - E_exact) / N, out[1] the largest
// |E_pred - E_exact| / N of a slice (Hartree per particle), out[2..4] RMS
// force error per component (Hartree / bohr, the slice energy's force),
// out[5..7] RMS exact reciprocal force per component, out[8] overall
// relative RMS force error, out[9] mean exact reciprocal energy per particle.
int sur_check(Surrogate *S, DevPimd *p, double *out) {
  if (sur_prepare(S, p)) return -1;
  hipStream_t st = rt().s0;
  if (sur_labels(S, p, st) || sur_features(S, p, st) || sur_infer(S, p, S->Fs, st, S->nlow > 0)) return -1;
  hipLaunchKernelGGL(k_sur_err, dim3(p->P), dim3(256), 0, st, S->Fs, S->Fx, S->eps, S->nlow ? S->El : nullptr, S->Ex,
                     p->N, p->N * p->P, S->fpart, S->dE, S->Eex);
  XMD_CHECK(hipGetLastError());
  XMD_CHECK(hipStreamSynchronize(st));
  std::vector<double> fp(6 * p->P), de(p->P), ee(p->P);
  XMD_CHECK(hipMemcpy(fp.data(), S->fpart, fp.size() * sizeof(double), hipMemcpyDeviceToHost));
  XMD_CHECK(hipMemcpy(de.data(), S->dE, p->P * sizeof(double), hipMemcpyDeviceToHost));
  XMD_CHECK(hipMemcpy(ee.data(), S->Eex, p->P * sizeof(double), hipMemcpyDeviceToHost));
  double f[6] = {0, 0, 0, 0, 0, 0}, dm = 0, dmax = 0, em = 0;
  for (int j = 0; j < p->P; ++j) {
    for (int k = 0; k < 6; ++k) f[k] += fp[6 * j + k];
    dm += de[j];
    dmax = std::max(dmax, fabs(de[j]));
    em += ee[j];
  }
  const double nb = (double)p->N * p->P, scale = (double)p->P;  // accelerations of the 1/P share -> slice forces
  out[0] = dm / nb;
  out[1] = dmax / p->N;
  for (int d = 0; d < 3; ++d) {
    out[2 + d] = sqrt(f[d] / nb) * scale;
    out[5 + d] = sqrt(f[3 + d] / nb) * scale;
  }
  out[8] = sqrt((f[0] + f[1] + f[2]) / (f[3] + f[4] + f[5]));
  out[9] = em / nb;
  return sur_overflow(S);
}

// The step's reciprocal-space stage with a surrogate attached (see
// md_pimd_run_steps_surrogate_x): mode 1 the surrogate's forces (the exact
// low shells plus the network); mode 2 the exact FP64 terms (forces used by
// the step, per-bead energies kept) and the surrogate's prediction beside
// them, with the error sums left in fpart / dE.
int surrogate_kspace(DevPimd *p, double *Frec, hipStream_t st) {
  Surrogate *S = (Surrogate *)p->sur;
  if (!S) return -1;
  const int add = S->nlow > 0;
  if (p->sur_mode == 1) {
    if (add && launch_recip_shells(p, S->nlow, Frec, S->El, st)) return -1;
    return (sur_features(S, p, st) || sur_infer(S, p, Frec, st, add)) ? -1 : 0;
  }
  if (launch_recip_labels(p, Frec, S->Ex, st, &S->alpha) ||
      (add && launch_recip_shells(p, S->nlow, S->Fs, S->El, st)) || sur_features(S, p, st) ||
      sur_infer(S, p, S->Fs, st, add))
    return -1;
  hipLaunchKernelGGL(k_sur_err, dim3(p->P), dim3(256), 0, st, S->Fs, Frec, S->eps, add ? S->El : nullptr, S->Ex, p->N,
                     p->N * p->P, S->fpart, S->dE, S->Eex);
  XMD_CHECK(hipGetLastError());
  return 0;
}

// Error sums of the last judge step (after a synchronisation): relative RMS
// force error and the mean and largest per-slice energy error per particle.
static int sur_judge_read(Surrogate *S, DevPimd *p, double *frel, double *emean, double *emax) {
  std::vector<double> fp(6 * p->P), de(p->P);
  XMD_CHECK(hipMemcpy(fp.data(), S->fpart, fp.size() * sizeof(double), hipMemcpyDeviceToHost));
  XMD_CHECK(hipMemcpy(de.data(), S->dE, p->P * sizeof(double), hipMemcpyDeviceToHost));
  double num = 0, den = 0, m = 0, mx = 0;
  for (int j = 0; j < p->P; ++j) {
    num += fp[6 * j] + fp[6 * j + 1] + fp[6 * j + 2];
    den += fp[6 * j + 3] + fp[6 * j + 4] + fp[6 * j + 5];
    m += de[j];
    mx = std::max(mx, fabs(de[j]) / p->N);
  }
  *frel = den > 0 ? sqrt(num / den) : 0.0;
  *emean = m / ((double)p->N * p->P);
  *emax = mx;
  return 0;
}

}  // namespace xmd

using namespace xmd;

struct md_ewald_surrogate_x {
  Surrogate *s;
};

namespace {
// device slot (slice-major) <-> reference bead index (particle-major)
inline int slot_to_bead(int s, int N, int P) { return (s % N) * P + s / N; }

int sur_pimd(md_pimd_t_x *pimd, DevPimd *&p) {
  if (ensure_pimd_device(pimd)) return -1;
  p = pimd_state(pimd, false);
  return p ? 0 : -1;
}
}  // namespace

extern "C" {

md_ewald_surrogate_x *md_alloc_ewald_surrogate_x(unsigned seed, md_num_t_x r_cut) {
  if (!(r_cut > 0) || ensure_runtime()) return nullptr;
  Surrogate *S = sur_create(seed, r_cut);
  if (!S) return nullptr;
  md_ewald_surrogate_x *h = new md_ewald_surrogate_x;
  h->s = S;
  return h;
}

int md_ewald_surrogate_set_exact_shells_x(md_ewald_surrogate_x *h, int nlow) {
  if (!h || nlow < 0) return -1;
  if (nlow != h->s->nlow) h->s->normed = false;  // the targets change
  h->s->nlow = nlow;
  return 0;
}

int md_ewald_surrogate_train_x(md_ewald_surrogate_x *h, md_pimd_t_x *pimd, md_num_t_x lr, int iters,
                               md_num_t_x *loss) {
  DevPimd *p;
  if (!h || iters < 1 || sur_pimd(pimd, p)) return -1;
  double l = 0;
  if (sur_train(h->s, p, lr, iters, &l)) return -1;
  if (loss) *loss = l;
  return 0;
}

int md_ewald_surrogate_check_x(md_ewald_surrogate_x *h, md_pimd_t_x *pimd, md_num_t_x *report) {
  DevPimd *p;
  if (!h || !report || sur_pimd(pimd, p)) return -1;
  return sur_check(h->s, p, report);
}

int md_ewald_surrogate_params_x(md_ewald_surrogate_x *h, float *get, const float *set) {
  if (!h) return -1;
  Surrogate *S = h->s;
  XMD_CHECK(hipStreamSynchronize(rt().s0));
  if (get) XMD_CHECK(hipMemcpy(get, S->par, NPAR * sizeof(float), hipMemcpyDeviceToHost));
  if (set) {
    XMD_CHECK(hipMemcpy(S->par, set, NPAR * sizeof(float), hipMemcpyHostToDevice));
    if (sur_convert(S, rt().s0)) return -1;
    XMD_CHECK(hipStreamSynchronize(rt().s0));
  }
  return NPAR;
}

int md_ewald_surrogate_normalization_x(md_ewald_surrogate_x *h, md_num_t_x *get, const md_num_t_x *set) {
  if (!h) return -1;
  Surrogate *S = h->s;
  std::vector<float> m(ND), is(ND);
  XMD_CHECK(hipStreamSynchronize(rt().s0));
  if (get) {
    XMD_CHECK(hipMemcpy(m.data(), S->mean, ND * sizeof(float), hipMemcpyDeviceToHost));
    XMD_CHECK(hipMemcpy(is.data(), S->istd, ND * sizeof(float), hipMemcpyDeviceToHost));
    for (int f = 0; f < ND; ++f) {
      get[f] = m[f];
      get[ND + f] = is[f];
    }
    get[2 * ND] = S->ymu;
    get[2 * ND + 1] = S->ysd;
  }
  if (set) {
    for (int f = 0; f < ND; ++f) {
      m[f] = (float)set[f];
      is[f] = (float)set[ND + f];
    }
    XMD_CHECK(hipMemcpy(S->mean, m.data(), ND * sizeof(float), hipMemcpyHostToDevice));
    XMD_CHECK(hipMemcpy(S->istd, is.data(), ND * sizeof(float), hipMemcpyHostToDevice));
    S->ymu = set[2 * ND];
    S->ysd = set[2 * ND + 1];
    S->normed = true;
  }
  return 2 * ND + 2;
}

// Surrogate prediction on the current configuration, in reference bead
// order (b = (l-1) P + (j-1)): raw descriptors desc[b][96] (FP32), per-bead
// energies e_bead[b] and slice-energy forces force[b][3] = -dE_slice/dx
// (Hartree / bohr); any output may be NULL.
int md_ewald_surrogate_eval_x(md_ewald_surrogate_x *h, md_pimd_t_x *pimd, float *desc, md_num_t_x *e_bead,
                              md_num_t_x *force) {
  DevPimd *p;
  if (!h || sur_pimd(pimd, p)) return -1;
  Surrogate *S = h->s;
  if (sur_prepare(S, p) || sur_features(S, p, rt().s0) || sur_infer(S, p, S->Fs, rt().s0)) return -1;
  XMD_CHECK(hipStreamSynchronize(rt().s0));
  if (sur_overflow(S)) return -1;
  const int N = p->N, P = p->P, nb = N * P;
  if (desc) {
    std::vector<float> D((size_t)nb * ND);
    XMD_CHECK(hipMemcpy(D.data(), S->D, D.size() * sizeof(float), hipMemcpyDeviceToHost));
    for (int s = 0; s < nb; ++s)
      memcpy(desc + (size_t)slot_to_bead(s, N, P) * ND, D.data() + (size_t)s * ND, ND * sizeof(float));
  }
  if (e_bead) {
    std::vector<float> e(nb);
    XMD_CHECK(hipMemcpy(e.data(), S->eps, nb * sizeof(float), hipMemcpyDeviceToHost));
    for (int s = 0; s < nb; ++s) e_bead[slot_to_bead(s, N, P)] = e[s];
  }
  if (force) {
    std::vector<double> F(3 * (size_t)nb), M(nb);
    XMD_CHECK(hipMemcpy(F.data(), S->Fs, F.size() * sizeof(double), hipMemcpyDeviceToHost));
    XMD_CHECK(hipMemcpy(M.data(), p->ds->M, nb * sizeof(double), hipMemcpyDeviceToHost));
    for (int s = 0; s < nb; ++s)
      for (int d = 0; d < 3; ++d) force[3 * slot_to_bead(s, N, P) + d] = F[d * (size_t)nb + s] * M[s] * P;
  }
  return 0;
}

// Exact labels of the current configuration (reference bead order): each
// bead's share of its slice's reciprocal energy and the slice-energy forces,
// less the exact low-shell part when low shells are kept (the network's
// targets).
int md_ewald_surrogate_labels_x(md_ewald_surrogate_x *h, md_pimd_t_x *pimd, md_num_t_x *e_bead, md_num_t_x *force) {
  DevPimd *p;
  if (!h || sur_pimd(pimd, p)) return -1;
  Surrogate *S = h->s;
  if (sur_prepare(S, p) || sur_labels(S, p, rt().s0)) return -1;
  XMD_CHECK(hipStreamSynchronize(rt().s0));
  const int N = p->N, P = p->P, nb = N * P;
  std::vector<double> E(nb), F(3 * (size_t)nb), M(nb), El(nb, 0.0), Fl(3 * (size_t)nb, 0.0);
  XMD_CHECK(hipMemcpy(E.data(), S->Ex, nb * sizeof(double), hipMemcpyDeviceToHost));
  XMD_CHECK(hipMemcpy(F.data(), S->Fx, F.size() * sizeof(double), hipMemcpyDeviceToHost));
  XMD_CHECK(hipMemcpy(M.data(), p->ds->M, nb * sizeof(double), hipMemcpyDeviceToHost));
  if (S->nlow) {
    XMD_CHECK(hipMemcpy(El.data(), S->El, nb * sizeof(double), hipMemcpyDeviceToHost));
    XMD_CHECK(hipMemcpy(Fl.data(), S->Fs, Fl.size() * sizeof(double), hipMemcpyDeviceToHost));
  }
  for (int s = 0; s < nb; ++s) {
    const int b = slot_to_bead(s, N, P);
    if (e_bead) e_bead[b] = E[s] - El[s];
    if (force)
      for (int d = 0; d < 3; ++d) force[3 * b + d] = (F[d * (size_t)nb + s] - Fl[d * (size_t)nb + s]) * M[s] * P;
  }
  return 0;
}

// Gradient of the training loss (1/n) sum_b (y_b - t_b)^2, t_b = (e_b -
// ymu) / ysd, with respect to the FP32 parameters at the current
// configuration, for given per-bead targets e (reference bead order); the
// normalisation must be set.  Nothing is updated.
int md_ewald_surrogate_gradient_x(md_ewald_surrogate_x *h, md_pimd_t_x *pimd, const md_num_t_x *e_bead, float *grad,
                                  md_num_t_x *loss) {
  DevPimd *p;
  if (!h || !e_bead || !grad || sur_pimd(pimd, p)) return -1;
  Surrogate *S = h->s;
  if (!S->normed) return -1;
  const int N = p->N, P = p->P, nb = N * P;
  if (sur_prepare(S, p) || sur_features(S, p, rt().s0)) return -1;
  std::vector<float> t(nb);
  for (int s = 0; s < nb; ++s) t[s] = (float)((e_bead[slot_to_bead(s, N, P)] - S->ymu) / S->ysd);
  XMD_CHECK(hipStreamSynchronize(rt().s0));
  XMD_CHECK(hipMemcpy(S->T, t.data(), nb * sizeof(float), hipMemcpyHostToDevice));
  double l = 0;
  if (sur_gradient(S, nb, &l, rt().s0)) return -1;
  XMD_CHECK(hipMemcpy(grad, S->grad, NPAR * sizeof(float), hipMemcpyDeviceToHost));
  if (loss) *loss = l;
  return sur_overflow(S);
}

// nsteps exact MD steps, each followed by `iters` Adam steps on the exact
// labels of the configuration it reached (every bead slice is one training
// configuration).  report[0] / [1]: the loss of the first / last update.
int md_pimd_train_surrogate_steps_x(md_pimd_t_x *pimd, const md_pimd_step_config_x *cfg, md_ewald_surrogate_x *h,
                                    int nsteps, md_num_t_x lr, int iters, md_num_t_x *report) {
  DevPimd *p;
  if (!h || !cfg || nsteps < 1 || iters < 1 || sur_pimd(pimd, p)) return -1;
  Surrogate *S = h->s;
  md_pimd_step_config_x c1 = *cfg;
  c1.use_graph = 0;
  p->sur_mode = 0;
  double first = 0, last = 0;
  for (int k = 0; k < nsteps; ++k) {
    if (md_pimd_run_steps_x(pimd, &c1, 1)) return -1;
    double l = 0;
    if (sur_train(S, p, lr, iters, &l)) return -1;
    if (k == 0) first = l;
    last = l;
  }
  if (md_hip_synchronize_x()) return -1;
  if (report) {
    report[0] = first;
    report[1] = last;
  }
  return 0;
}

// nsteps MD steps with the surrogate standing in for the reciprocal-space
// Ewald term (see md_hip_x.h): surrogate steps replayed as graphs, every
// check_every-th step a judge step, exact fallback with online training
// when a judge exceeds either threshold.
int md_pimd_run_steps_surrogate_x(md_pimd_t_x *pimd, const md_pimd_step_config_x *cfg, md_ewald_surrogate_x *h,
                                  int nsteps, int check_every, md_num_t_x f_threshold, md_num_t_x e_threshold,
                                  md_num_t_x lr, md_num_t_x *report) {
  DevPimd *p;
  if (!h || !cfg || nsteps < 0 || check_every < 1 || sur_pimd(pimd, p)) return -1;
  Surrogate *S = h->s;
  if (!S->normed) {
    fprintf(stderr, "xmd: the surrogate has not been trained (no normalisation)\n");
    return -1;
  }
  if (sur_prepare(S, p)) return -1;
  md_pimd_step_config_x c1 = *cfg;
  c1.use_graph = 0;
  p->sur = S;
  int done = 0, nsur = 0, njudge = 0, nexact = 0, nfb = 0, rc = 0;
  double fsum = 0, fmax = 0, esum = 0, emax = 0;
  while (done < nsteps && !rc) {
    const int run = std::min(check_every - 1, nsteps - done);
    if (run > 0) {
      p->sur_mode = 1;
      rc = md_pimd_run_steps_x(pimd, cfg, run);
      nsur += run;
      done += run;
      if (rc || done >= nsteps) break;
    }
    p->sur_mode = 2;  // judge: exact terms drive the step, the surrogate is compared
    if ((rc = md_pimd_run_steps_x(pimd, &c1, 1)) != 0) break;
    if ((rc = md_hip_synchronize_x()) != 0) break;
    ++njudge;
    ++done;
    double frel, em, emx;
    if ((rc = sur_judge_read(S, p, &frel, &em, &emx)) != 0) break;
    fsum += frel;
    fmax = std::max(fmax, frel);
    esum += em;
    emax = std::max(emax, emx);
    if (frel > f_threshold || emx > e_threshold) {
      // fallback: exact steps, each followed by one online training update
      ++nfb;
      p->sur_mode = 0;
      const int k = std::min(check_every, nsteps - done);
      for (int i = 0; i < k && !rc; ++i) {
        rc = md_pimd_run_steps_x(pimd, &c1, 1);
        if (!rc && lr > 0) rc = sur_train(S, p, lr, 1, nullptr);
      }
      nexact += k;
      done += k;
    }
  }
  p->sur_mode = 0;
  p->sur = nullptr;
  if (!rc) rc = md_hip_synchronize_x();
  if (!rc) rc = sur_overflow(S);
  if (report) {
    report[0] = nsur;
    report[1] = njudge;
    report[2] = nexact;
    report[3] = nfb;
    report[4] = njudge ? fsum / njudge : 0.0;
    report[5] = fmax;
    report[6] = njudge ? esum / njudge : 0.0;
    report[7] = emax;
  }
  return rc;
}

}  // extern "C"
