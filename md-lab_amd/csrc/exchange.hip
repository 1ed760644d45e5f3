// Bosonic / fictitious-particle exchange on the GPU: cycle energies E_n^(k)
// (md_pimd_x.c:1969-2210), the forward V_B recursion (:814-953), the
// backward potentials, connection probabilities and exchange forces
// (:2369-2721), and the estimator recursions (:1111-1295, 3959-4556,
// 4652-4705).
//
// Formulation used throughout.  With particles 1..N and
//   c(n, m) = (n-m-1) ln xi - beta E_n^(n-m) - ln n       (0 <= m < n <= N)
// (only n - m = 1 is present when xi = 0), the reference's recursions are
//   forward   L(n) = LSE_{m<n} [c(n,m) + L(m)],  L(0) = 0,  V_B^(n) = -L(n)/beta
//   backward  R(m) = LSE_{n>m} [c(n,m) + R(n)],  R(N) = 0,  Vt^(m)  = -R(m)/beta
//   G(l, j)   = exp(c(l, j-1) + L(j-1) + R(l) - L(N))  for j <= l,
//               1 - exp(L(l) + R(l) - L(N))            for j = l+1.
// These are algebraically identical to md_pimd_xminE_x / md_pimd_fill_VB_x /
// md_pimd_fast_xminE_x / md_pimd_fast_fill_force_VB_x; the shift of each
// log-sum-exp is the running maximum instead of a separate minimum pass.
//
// Cycle energies come from double-double prefix sums of the open-chain
// energies C_i = Eint_i + s(x_{i,P}, x_{i+1,1}):
//   E_n^(k) = sum_{i=n-k+1}^{n-1} C_i + Eint_n + s(x_{n,P}, x_{n-k+1,1}),
// the closed form of the reference's recurrence (md_pimd_x.c:2150-2206), so
// every (n, k) is independent and the fill is one parallel pass.  Two copies
// are kept: the reference's row layout (E_n^(.) contiguous) for the forward
// pass and the host mirror, and a column layout (cycles sharing their first
// particle contiguous) for the backward pass and the exchange-force columns.
//
// The recursions are solved as a blocked right-looking triangular solve in
// log-sum-exp form (see the block-solve section below): both directions in
// one persistent launch, 64 positions per block.
#include <cmath>

#include "xmd_device.h"
#include "xmd_internal.h"

namespace xmd {

constexpr double NEG = -1.0e300;  // "no term" marker (exp(NEG - x) == 0)
constexpr int RB = 64;             // recursion block (rows per diagonal block)
constexpr int RT = 1024;           // estimator workgroup
constexpr double LSE_SKIP = 40.0;  // log-sum-exp terms below max - 40 are < 4.3e-18 relative

__device__ __forceinline__ size_t tri_row(int n) { return (size_t)(n - 1) * n / 2; }  // row n (1-based)
__device__ __forceinline__ size_t col_off(int m, int N) { return (size_t)m * N - (size_t)m * (m - 1) / 2; }

struct Geo {
  double L0, L1, L2;
  double iL0, iL1, iL2;
  int image;
};

// Minimum image by rounding, for squared distances only: at an exact
// +-L/2 tie it may pick the other sign than md_minimum_image_x, which
// leaves d^2 unchanged.
__device__ __forceinline__ double comp_sq(double a, double L, double iL, int image) {
  return image ? fma(-L, rint(a * iL), a) : a;
}

// s(a, b) = 1/2 m_a w^2 d^2 (energy only; md_pimd_x.c:2120-2124).
__device__ __forceinline__ double spring_e(const double *__restrict__ X, int nb, int sa, int sb, double m, double om,
                                           const Geo &g) {
  const double dx = comp_sq(X[sa] - X[sb], g.L0, g.iL0, g.image);
  const double dy = comp_sq(X[nb + sa] - X[nb + sb], g.L1, g.iL1, g.image);
  const double dz = comp_sq(X[2 * nb + sa] - X[2 * nb + sb], g.L2, g.iL2, g.image);
  // distance, then its square, in the reference's order of operations: on
  // ill-conditioned exchange inputs the forces follow ENk's last bits
  const double d = sqrt(dx * dx + dy * dy + dz * dz);
  return 0.5 * m * om * om * d * d;
}

__device__ __forceinline__ double comp(double a, double L, int image) { return image ? dv_min_image(a, L) : a; }

// s(a, b) = 1/2 m_a w^2 d^2 with d the (minimum-image) distance, as in
// md_pimd_x.c:2120-2124.
__device__ __forceinline__ double spring(const double *X, int nb, int sa, int sb, double m, double om, Geo g) {
  double dx = comp(X[sa] - X[sb], g.L0, g.image);
  double dy = comp(X[nb + sa] - X[nb + sb], g.L1, g.image);
  double dz = comp(X[2 * nb + sa] - X[2 * nb + sb], g.L2, g.image);
  double d = sqrt(dx * dx + dy * dy + dz * dz);
  return 0.5 * m * om * om * d * d;
}

// 1/2 m_a w^2 (x_a - x_b).(x_a - x_c), md_pimd_x.c:4376-4420.
__device__ __forceinline__ double e2(const double *X, const double *M, int nb, int sa, int sb, int sc, double om,
                                     Geo g) {
  double r = 0;
  double m = M[sa];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double L = d == 0 ? g.L0 : (d == 1 ? g.L1 : g.L2);
    double u = comp(X[d * nb + sa] - X[d * nb + sb], L, g.image);
    double v = comp(X[d * nb + sa] - X[d * nb + sc], L, g.image);
    r += 0.5 * m * om * om * u * v;
  }
  return r;
}

// ------------------------------------------------------------ E_n^(k)

// Particle l (0-based): Eint_l, E_l^(1), link C_l.  A block holds 64
// particles (lanes); its four waves split the P - 1 interior springs and
// the partial sums are added in a fixed order.
__device__ __forceinline__ void chain_body(int bx, const double *__restrict__ X, const double *__restrict__ M, int N,
                                           int P, double om, Geo g, double *__restrict__ Eint,
                                           double *__restrict__ Clink, double *__restrict__ ENk,
                                           double *__restrict__ ET) {
  __shared__ double part[4][64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int l = bx * 64 + lane;
  const int nb = N * P;
  double e = 0;
  if (l < N) {
    const int j0 = ((P - 1) * wv) / 4, j1 = ((P - 1) * (wv + 1)) / 4;
    for (int j = j0; j < j1; ++j) e += spring_e(X, nb, j * N + l, (j + 1) * N + l, M[j * N + l], om, g);
  }
  part[wv][lane] = e;
  __syncthreads();
  if (wv == 0 && l < N) {
    e = ((part[0][lane] + part[1][lane]) + part[2][lane]) + part[3][lane];
    Eint[l] = e;
    const int sP = (P - 1) * N + l;
    const double e1 = e + spring_e(X, nb, sP, l, M[sP], om, g);
    ENk[tri_row(l + 1)] = e1;
    ET[col_off(l, N)] = e1;
    Clink[l] = (l < N - 1) ? e + spring_e(X, nb, sP, l + 1, M[sP], om, g) : 0.0;
  }
}

__global__ __launch_bounds__(256) void k_chain(const double *__restrict__ X, const double *__restrict__ M, int N,
                                               int P, double om, Geo g, double *__restrict__ Eint,
                                               double *__restrict__ Clink, double *__restrict__ ENk,
                                               double *__restrict__ ET) {
  chain_body(blockIdx.x, X, M, N, P, om, g, Eint, Clink, ENk, ET);
}

// Exclusive double-double prefix sum of in[0..n-1] into pre[2*i] (hi, lo)
// by one workgroup of NT threads.
template <int NT = 256>
__device__ __forceinline__ void scan_dd_body(const double *__restrict__ in, int n, double *__restrict__ pre) {
  __shared__ double sh[NT], sl[NT];
  const int t = threadIdx.x;
  const int seg = (n + NT - 1) / NT;
  const int a = min(n, t * seg), b = min(n, a + seg);
  dd tot{0, 0};
  for (int i = a; i < b; ++i) tot = dd_add(tot, dd{in[i], 0});
  sh[t] = tot.hi;
  sl[t] = tot.lo;
  __syncthreads();
  for (int w = 1; w < NT; w <<= 1) {
    dd v{sh[t], sl[t]};
    if (t >= w) v = dd_add(v, dd{sh[t - w], sl[t - w]});
    __syncthreads();
    sh[t] = v.hi;
    sl[t] = v.lo;
    __syncthreads();
  }
  dd run = t > 0 ? dd{sh[t - 1], sl[t - 1]} : dd{0, 0};
  for (int i = a; i < b; ++i) {
    pre[2 * i] = run.hi;
    pre[2 * i + 1] = run.lo;
    run = dd_add(run, dd{in[i], 0});
  }
}

__global__ __launch_bounds__(256) void k_scan_dd(const double *__restrict__ in, int n, double *__restrict__ pre) {
  scan_dd_body(in, n, pre);
}

// In the concurrent step the cycle energies, the recursion set-up, the
// recursion and the end-bead exchange forces form a chain of dependent
// kernels while the pair-force grids fill the compute units beside them: the
// waves of these kernels take the highest issue priority (measured 0.267 ->
// 0.249 ms per step for UEG N=1024 P=32 for the short kernels; for the
// persistent recursion 0.235 -> 0.226 ms once the faster k-space kernels had
// put it on the longest chain).
__device__ __forceinline__ void chain_priority() { __builtin_amdgcn_s_setprio(3); }

// k_chain and k_scan_dd in one launch: the last workgroup to finish its
// particles (counted in *done, which it resets) scans the links.
__global__ __launch_bounds__(256) void k_chain_scan(const double *__restrict__ X, const double *__restrict__ M, int N,
                                                    int P, double om, Geo g, double *__restrict__ Eint,
                                                    double *__restrict__ Clink, double *__restrict__ ENk,
                                                    double *__restrict__ ET, double *__restrict__ pre,
                                                    int *__restrict__ done) {
  __shared__ int last;
  chain_priority();
  chain_body(blockIdx.x, X, M, N, P, om, g, Eint, Clink, ENk, ET);
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0)
    last = __hip_atomic_fetch_add(done, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == (int)gridDim.x - 1;
  __syncthreads();
  if (!last) return;
  __threadfence();
  scan_dd_body(Clink, N, pre);
  if (threadIdx.x == 0) __hip_atomic_store(done, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Row n' (0-based) of the cycle-energy triangle, k >= 2, both layouts.
__device__ __forceinline__ void enk_fill_body(int r, const double *__restrict__ X, const double *__restrict__ M,
                                              int N, int P, double om, Geo g, const double *__restrict__ Eint,
                                              const double *__restrict__ pre, double *__restrict__ ENk,
                                              double *__restrict__ ET) {
  // r: 0-based last particle of the cycle
  const int nb = N * P;
  const int sP = (P - 1) * N + r;
  const double mP = M[sP], ein = Eint[r];
  const dd ar{pre[2 * r], pre[2 * r + 1]};
  for (int k = 2 + threadIdx.x; k <= r + 1; k += blockDim.x) {
    const int m0 = r - k + 1;  // 0-based first particle
    double chain = dd_sub_to_double(ar, dd{pre[2 * m0], pre[2 * m0 + 1]});
    double E = chain + ein + spring_e(X, nb, sP, m0, mP, om, g);
    ENk[tri_row(r + 1) + k - 1] = E;
    ET[col_off(m0, N) + k - 1] = E;
  }
}

__global__ __launch_bounds__(256) void k_enk_fill(const double *__restrict__ X, const double *__restrict__ M, int N,
                                                  int P, double om, Geo g, const double *__restrict__ Eint,
                                                  const double *__restrict__ pre, double *__restrict__ ENk,
                                                  double *__restrict__ ET) {
  enk_fill_body(blockIdx.x, X, M, N, P, om, g, Eint, pre, ENk, ET);
}

// Positions, masses and prefix sums from which a cycle energy is computed
// directly, with the arithmetic of k_chain (k = 1) and k_enk_fill (k >= 2).
struct EnkArgs {
  const double *X, *M, *Eint, *pre;
  double om;
  Geo g;
  int N, P;
};

// E of the cycle of particles m0 .. r (0-based, m0 <= r).
__device__ __forceinline__ double E_cycle(const EnkArgs &e, int m0, int r) {
  const int nb = e.N * e.P;
  const int sP = (e.P - 1) * e.N + r;
  const double mP = e.M[sP], ein = e.Eint[r];
  if (m0 == r) return ein + spring_e(e.X, nb, sP, r, mP, e.om, e.g);
  const double chain = dd_sub_to_double(dd{e.pre[2 * r], e.pre[2 * r + 1]}, dd{e.pre[2 * m0], e.pre[2 * m0 + 1]});
  return chain + ein + spring_e(e.X, nb, sP, m0, mP, e.om, e.g);
}

// Host-provided ENk (row layout) -> column layout.
__global__ void k_enk_transpose(const double *__restrict__ ENk, double *__restrict__ ET, int N) {
  const int r = blockIdx.x;
  for (int k = 1 + threadIdx.x; k <= r + 1; k += blockDim.x) {
    int m0 = r - k + 1;
    ET[col_off(m0, N) + k - 1] = ENk[tri_row(r + 1) + k - 1];
  }
}

// ------------------------------------------------------------ recursion

struct RecArgs {
  const double *ENk;  // row layout
  const double *ET;   // column layout
  const double *lnn;  // ln n, n = 0..N (ln 0 unused)
  double *Lf;         // forward output, N+1
  double *Rb;         // backward output, N+1
  double *VBN;        // V_B^(n) = -L(n)/beta
  double *val2;       // [2][N+1] solved values per direction, position order (hand-off: NaN until solved)
  double *accM, *accS;  // [2][N+1] accumulators of column 0 (log-sum-exp pairs)
  double *finM, *finS;  // [2][N+1] row blocks' accumulators over blocks 0..G-2 (hand-off: NaN until folded)
  double *dg;         // [2][nblk][RB][RB] in-block coefficients [r][q] (0 unless r < q), chain frame, exponentiated
  double *la;         // [2][nblk][RB][RB] the block's rows on the previous block, likewise
  double *ech;        // [2][nblk][RB+1] chain offsets E(0..RB-1) and the step Delta into the next block
  int *sync;          // [4 nblk] unused, error word, injected-timeout count
  int N, nblk;
  double beta, lnxi;
  int xi0;            // xi == 0: only k = 1 terms
  double big;         // upper end of the chain frame's accepted range
};

// Position p = 0..N along the solve order; p' < p are solved earlier.
//   forward:  value L(p),      coef(p, p') = c(n = p,      m = p')
//   backward: value R(N - p),  coef(p, p') = c(n = N - p', m = N - p)
// In both directions k = p - p'.
template <bool FWD>
__device__ __forceinline__ double rec_coef(const RecArgs &a, int p, int pp) {
  const int k = p - pp;
  if (a.xi0 && k != 1) return NEG;
  double E, ln_n;
  if (FWD) {
    E = a.ENk[tri_row(p) + k - 1];
    ln_n = a.lnn[p];
  } else {
    E = a.ET[col_off(a.N - p, a.N) + k - 1];
    ln_n = a.lnn[a.N - pp];
  }
  return (k - 1) * a.lnxi - a.beta * E - ln_n;
}

// rec_coef with E computed from positions and prefix sums (the value the
// tables hold) instead of read from them.
template <bool FWD>
__device__ __forceinline__ double rec_coef_direct(const RecArgs &a, const EnkArgs &e, int p, int pp) {
  const int k = p - pp;
  if (a.xi0 && k != 1) return NEG;
  const double E = FWD ? E_cycle(e, pp, p - 1) : E_cycle(e, a.N - p, a.N - pp - 1);
  const double ln_n = FWD ? a.lnn[p] : a.lnn[a.N - pp];
  return (k - 1) * a.lnxi - a.beta * E - ln_n;
}

// ---- blocked right-looking solve (positions in blocks of RB rows) ----
//
// Direction dir (0 forward, 1 backward) keeps, per position p, the solved
// value val[p] and a running log-sum-exp (M, S) of the terms coef(p, p') +
// val[p'] over solved positions p' outside p's own block.
//   k_rec_prep   (one launch; in the fused step its work is part of
//                k_enk_prep) accumulators from position 0, the hand-off
//                buffers' NaN marks, and the chain-frame coefficients of
//                every diagonal block and of every block on its predecessor,
//                so the serial kernels read them with independent, coalesced
//                loads;
//   k_rec_persist (one launch) per direction a solver workgroup that solves
//                block b and folds it into block b+1's rows (look-ahead),
//                and one updater workgroup per later row block that folds
//                each finished block into its rows.
// Every row of block B thus receives column 0 (prep), blocks 0..B-2 (its
// updater) and block B-1 (the solver's look-ahead).
// Lanes are rows throughout in the updaters: log-sum-exps run along each
// lane, so no cross-lane reductions are needed.
constexpr int RSQ = RB * RB;

// coef(p, p') read through the other layout (column p' of the forward
// problem is contiguous in ET, of the backward problem in ENk), so that
// lanes p, p+1, ... at fixed p' load consecutive doubles.
template <bool FWD>
__device__ __forceinline__ double rec_coef_col(const RecArgs &a, int p, int pp) {
  const int k = p - pp;
  if (a.xi0 && k != 1) return NEG;
  double E, ln_n;
  if (FWD) {
    E = a.ET[col_off(pp, a.N) + k - 1];
    ln_n = a.lnn[p];
  } else {
    E = a.ENk[tri_row(a.N - pp) + k - 1];
    ln_n = a.lnn[a.N - pp];
  }
  return (k - 1) * a.lnxi - a.beta * E - ln_n;
}

// rec_coef without branches (ok false: NEG), so that unrolled callers issue
// their loads together.
template <bool FWD>
__device__ __forceinline__ double rec_coef_nb(const RecArgs &a, int p, int pp, bool ok) {
  const int k = p - pp;
  ok = ok && !(a.xi0 && k != 1);
  const size_t idx = ok ? (FWD ? tri_row(p) : col_off(a.N - p, a.N)) + k - 1 : 0;
  const double E = FWD ? a.ENk[idx] : a.ET[idx];
  const double ln_n = a.lnn[ok ? (FWD ? p : a.N - pp) : 0];
  return ok ? (k - 1) * a.lnxi - a.beta * E - ln_n : NEG;
}

// rec_coef_col for the 16 columns pp0 .. pp0+15 of row p, all loads issued
// before any is used (clamped indices instead of branches, which would make
// each load wait for the one before); act false: every coefficient NEG.
template <bool FWD>
__device__ __forceinline__ void rec_coef_col16(const RecArgs &a, int p, int pp0, bool act, double *c) {
  double E[16], ln[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int pp = pp0 + j, k = p - pp;
    const bool ok = act && !(a.xi0 && k != 1);
    const size_t idx = ok ? (FWD ? col_off(pp, a.N) : tri_row(a.N - pp)) + k - 1 : 0;
    E[j] = FWD ? a.ET[idx] : a.ENk[idx];
    ln[j] = a.lnn[ok ? (FWD ? p : a.N - pp) : 0];
  }
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int pp = pp0 + j, k = p - pp;
    const bool ok = act && !(a.xi0 && k != 1);
    c[j] = ok ? (k - 1) * a.lnxi - a.beta * E[j] - ln[j] : NEG;
  }
}

// DIRECT: coefficients from positions and prefix sums (rec_coef_direct, e
// used) instead of the cycle-energy tables, so the set-up may run beside the
// kernel that fills them.
template <bool FWD, bool DIRECT>
struct RecCoef {
  const RecArgs &a;
  const EnkArgs &e;
  __device__ double operator()(int p, int pp) const {
    return DIRECT ? rec_coef_direct<FWD>(a, e, p, pp) : rec_coef<FWD>(a, p, pp);
  }
  __device__ double col(int p, int pp) const {
    return DIRECT ? rec_coef_direct<FWD>(a, e, p, pp) : rec_coef_col<FWD>(a, p, pp);
  }
};

// Value of position 0 and the accumulators of column 0 (t = position).
template <bool FWD, bool DIRECT>
__device__ __forceinline__ void rec_prep_acc(const RecArgs &a, int dir, int t, const EnkArgs &e) {
  const RecCoef<FWD, DIRECT> coef{a, e};
  const int N = a.N;
  if (t > N) return;
  double *val = a.val2 + (size_t)dir * (N + 1);
  double *AM = a.accM + (size_t)dir * (N + 1), *AS = a.accS + (size_t)dir * (N + 1);
  a.finM[(size_t)dir * (N + 1) + t] = __builtin_nan("");
  a.finS[(size_t)dir * (N + 1) + t] = __builtin_nan("");
  if (t == 0) {
    val[0] = 0.0;
    if (FWD) {
      a.Lf[0] = 0.0;
      a.VBN[0] = 0.0;
    } else {
      a.Rb[N] = 0.0;
    }
  } else {
    val[t] = __builtin_nan("");
    const double c = coef(t, 0);
    AM[t] = c > NEG ? c : NEG;
    AS[t] = c > NEG ? 1.0 : 0.0;
  }
}

// Chain offsets of row block b into E[0 .. RB], lane i of the calling wave
// computing E[i]: E(0) = 0, E(i) = E(i-1) + coef(p0+i, p0+i-1) over the
// block's rows (constant past them), and E[RB] = Delta = E(63) +
// coef(p0+64, p0+63), the k = 1 step into the next block (0 for the last
// block).  Every workgroup that needs block b's offsets computes them with
// this code, in the same order, so they agree bit for bit.
template <bool FWD, bool DIRECT>
__device__ __forceinline__ void rec_chain_offsets(const RecArgs &a, int b, const EnkArgs &e, double *E) {
  const RecCoef<FWD, DIRECT> coef{a, e};
  const int lane = threadIdx.x & 63;
  const int p0 = 1 + b * RB, nrows = min(RB, a.N - p0 + 1);
  double d = (lane >= 1 && lane < nrows) ? coef(p0 + lane, p0 + lane - 1) : 0.0;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const double t = __shfl_up(d, o, 64);
    if (lane >= o) d += t;
  }
  E[lane] = d;
  if (lane == 63) E[RB] = b + 1 < a.nblk ? d + coef(p0 + RB, p0 + RB - 1) : 0.0;
}

// Coefficients of row block b in the frame of its k = 1 chain, one part of
// REC_PREP_PARTS (every workgroup of the block computes the offsets, part 0
// stores them):
//   in-block    A(q, r)  = exp(c(p0+q, p0+r) + E(r) - E(q)),   r < q, stored [r][q]
//   look-ahead  F(q, r)  = exp(c(p0+q, p0-RB+r) + E'(r) - Delta' - E(q))
// with E' and Delta' those of block b-1 (F: block b's rows on block b-1's
// columns, stored [r][q]).  Terms without a coefficient are 0; an exponent
// past the FP64 range gives inf, which the solver's range vote catches.
constexpr int REC_PREP_PARTS = 4;

template <bool FWD, bool DIRECT>
__device__ __forceinline__ void rec_prep_block(const RecArgs &a, int dir, int b, int part, const EnkArgs &e,
                                               double (*E)[RB + 1]) {
  const RecCoef<FWD, DIRECT> coef{a, e};
  const int tid = threadIdx.x, wv = tid >> 6;
  if (wv == 0) rec_chain_offsets<FWD, DIRECT>(a, b, e, E[0]);
  if (wv == 1 && b > 0) rec_chain_offsets<FWD, DIRECT>(a, b - 1, e, E[1]);
  __syncthreads();
  const int p0 = 1 + b * RB, nrows = min(RB, a.N - p0 + 1);
  if (part == 0 && tid <= RB) a.ech[((size_t)dir * a.nblk + b) * (RB + 1) + tid] = E[0][tid];
  double *dg = a.dg + ((size_t)dir * a.nblk + b) * RSQ;
  for (int w = part * 256 + tid; w < RSQ; w += REC_PREP_PARTS * 256) {
    const int r = w / RB, q = w - r * RB;
    const double c = (r < q && q < nrows) ? coef(p0 + q, p0 + r) : NEG;
    dg[w] = c > NEG ? exp(c + E[0][r] - E[0][q]) : 0.0;
  }
  double *la = a.la + ((size_t)dir * a.nblk + b) * (RB * RB);
  for (int w = part * 256 + tid; w < RB * RB; w += REC_PREP_PARTS * 256) {
    const int r = w / RB, q = w - r * RB;
    const double c = (b > 0 && q < nrows) ? coef.col(p0 + q, p0 - RB + r) : NEG;
    la[w] = c > NEG ? exp(c + E[1][r] - E[1][RB] - E[0][q]) : 0.0;
  }
}

// Workgroups of one direction's set-up: [0, rec_prep_acc_wgs) the values
// and accumulators (thread = position), then REC_PREP_PARTS per row block
// for the coefficients.  The error word at sync[4 nblk] keeps a timeout of
// any earlier solve until the host has read it (recursion_status clears it
// there).
__host__ __device__ inline int rec_prep_acc_wgs(int N) { return (N + 1 + 255) / 256; }
__host__ __device__ inline int rec_prep_wgs(int N, int nblk) { return rec_prep_acc_wgs(N) + nblk * REC_PREP_PARTS; }

template <bool DIRECT>
__device__ __forceinline__ void rec_prep_wg(const RecArgs &a, int dir, int wg, const EnkArgs &e,
                                            double (*E)[RB + 1]) {
  const int nacc = rec_prep_acc_wgs(a.N);
  if (wg < nacc) {
    const int t = wg * 256 + threadIdx.x;
    if (dir == 0)
      rec_prep_acc<true, DIRECT>(a, 0, t, e);
    else
      rec_prep_acc<false, DIRECT>(a, 1, t, e);
  } else {
    const int w = wg - nacc;
    if (dir == 0)
      rec_prep_block<true, DIRECT>(a, 0, w / REC_PREP_PARTS, w % REC_PREP_PARTS, e, E);
    else
      rec_prep_block<false, DIRECT>(a, 1, w / REC_PREP_PARTS, w % REC_PREP_PARTS, e, E);
  }
}

// grid (rec_prep_wgs, directions)
__global__ __launch_bounds__(256) void k_rec_prep(RecArgs a, int dir0) {
  __shared__ double E[2][RB + 1];
  rec_prep_wg<false>(a, dir0 + blockIdx.y, blockIdx.x, EnkArgs{}, E);
}

// k_enk_fill (workgroups 0 .. N-1) and the recursion set-up of both
// directions (the next 2 nprepb workgroups, coefficients computed directly)
// in one launch.
__global__ __launch_bounds__(256) void k_enk_prep(EnkArgs e, double *__restrict__ ENk, double *__restrict__ ET,
                                                  RecArgs a, int nprepb) {
  __shared__ double E[2][RB + 1];
  chain_priority();
  const int bx = blockIdx.x;
  if (bx < e.N) {
    enk_fill_body(bx, e.X, e.M, e.N, e.P, e.om, e.g, e.Eint, e.pre, ENk, ET);
    return;
  }
  const int w = bx - e.N, dir = w / nprepb;
  rec_prep_wg<true>(a, dir, w - dir * nprepb, e, E);
}

// (M, S) <- log-sum-exp merge with (m, s).
__device__ __forceinline__ void lse_pair(double &M, double &S, double m, double s) {
  if (!(m > NEG)) return;
  if (m > M) {
    S = S * exp(M - m) + s;
    M = m;
  } else if (m > M - LSE_SKIP) {
    S += s * exp(m - M);
  }
}

// Two-pass log-sum-exp of 16 terms held by this lane.
__device__ __forceinline__ void lse16(const double *t, double &M, double &S) {
  M = NEG;
#pragma unroll
  for (int j = 0; j < 16; ++j) M = fmax(M, t[j]);
  S = 0.0;
  if (M > NEG) {
#pragma unroll
    for (int j = 0; j < 16; ++j) S += (t[j] > M - LSE_SKIP) ? exp(t[j] - M) : 0.0;
  }
}

// ---- one persistent launch for the whole solve ----
//
// Per direction one solver workgroup (blocks 0, 1, 2, ... in order, with
// the look-ahead fold) and one updater workgroup per row block G >= 2 (folds
// blocks 0 .. G-2 into its 64 rows as soon as each is solved).  All of the
// launch's workgroups are resident from the start, so no block waits for a
// new dispatch while other kernels fill the device.  The hand-offs carry
// their own readiness: the set-up marks val[] and fin[MS][] with NaN, the
// solver writes block b's values into val[], an updater its rows' finished
// accumulators into fin[MS][], both with device-scope relaxed stores (to the
// coherence point, past the writer's L2), and the reader polls the values
// themselves with device-scope loads until none is NaN.  No flag, so no
// cache-wide write-back or invalidate per hand-off.  Every poll is bounded;
// a timeout raises the error word (checked by the host at synchronisation
// points) and lets all polls fall through, so the grid always drains.
constexpr int REC_SPIN_MAX = 1 << 22;

__device__ __forceinline__ void st_handoff(double *p, double v) {
  __hip_atomic_store((unsigned long long *)p, (unsigned long long)__double_as_longlong(v), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ double ld_handoff(const double *p) {
  return __longlong_as_double(
      (long long)__hip_atomic_load((unsigned long long *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

// Lanes [0, n) of the calling wave read p[lane] (and q[lane] if q) until no
// value is NaN; at most `spin` rounds (wave-uniform; spin < 0 gives up at
// once: the injected timeout of the test hook), then the error word is
// raised and the values are whatever was read last.  While the last value
// is still NaN a round reads only that one (a single request per wave: many
// workgroups poll the same lines).
__device__ __forceinline__ void rec_poll(const double *p, const double *q, int n, int *err, int spin, double &pv,
                                         double &qv) {
  const int lane = threadIdx.x & 63;
  const double *last = q ? q + n - 1 : p + n - 1;
  for (int it = 0;; ++it) {
    const double probe = ld_handoff(last);
    if (spin >= 0 && probe == probe) {
      pv = lane < n ? ld_handoff(p + lane) : 0.0;
      qv = (q && lane < n) ? ld_handoff(q + lane) : 0.0;
      if (!__any(pv != pv || qv != qv)) return;
    }
    if (__builtin_amdgcn_readfirstlane(__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) != 0)
      return;
    if (it >= spin) {
      if (lane == 0) __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return;
    }
    __builtin_amdgcn_s_sleep(2);
  }
}

// Folds blocks 0 .. G-2 into row block G's accumulators (lane = row, the
// four waves split each block's 64 columns, 16 each), in block order,
// starting from column 0's (M, S); wave 0 returns the result.  POLL (an
// updater): each wave polls its 16 columns' values itself; otherwise this
// workgroup solved the blocks (its barriers order those stores before these
// plain loads).  The next block's coefficients are in flight while a block
// is folded; the four partials of a block merge in wave 0 through
// double-buffered LDS (one barrier per block).
template <bool FWD, bool POLL>
__device__ __forceinline__ void rec_fold_rows(const RecArgs &a, int dir, int G, double (*pm)[4][RB],
                                              double (*ps)[4][RB], double &Ma, double &Sa) {
  const int N = a.N;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  int *err = a.sync + 4 * a.nblk;
  const double *val = a.val2 + (size_t)dir * (N + 1);
  const double *AM = a.accM + (size_t)dir * (N + 1), *AS = a.accS + (size_t)dir * (N + 1);
  const int p = 1 + G * RB + lane;
  const bool act = p <= N;
  Ma = NEG;
  Sa = 0.0;
  if (wv == 0 && act) {
    Ma = AM[p];
    Sa = AS[p];
  }
  double c[16];
  rec_coef_col16<FWD>(a, p, 1 + 16 * wv, act, c);
  for (int b = 0; b <= G - 2; ++b) {
    const int pprev = 1 + b * RB;
    double v, unused;
    if (POLL)
      rec_poll(val + pprev + 16 * wv, nullptr, 16, err, REC_SPIN_MAX, v, unused);
    else
      v = lane < 16 ? val[pprev + 16 * wv + lane] : 0.0;
    double cn[16];
    rec_coef_col16<FWD>(a, p, pprev + RB + 16 * wv, act && b + 1 <= G - 2, cn);
    double t[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const double vj = readlane_d(v, j);
      t[j] = c[j] > NEG ? c[j] + vj : NEG;
    }
    double M, S;
    lse16(t, M, S);
    const int buf = b & 1;
    pm[buf][wv][lane] = M;
    ps[buf][wv][lane] = S;
    __syncthreads();
    if (wv == 0) {
#pragma unroll
      for (int w = 0; w < 4; ++w) lse_pair(Ma, Sa, pm[buf][w][lane], ps[buf][w][lane]);
    }
#pragma unroll
    for (int j = 0; j < 16; ++j) c[j] = cn[j];
  }
}

template <bool FWD>
__device__ __forceinline__ void rec_updater(const RecArgs &a, int dir, int G, double (*pm)[4][RB],
                                            double (*ps)[4][RB]) {
  double Ma, Sa;
  rec_fold_rows<FWD, true>(a, dir, G, pm, ps, Ma, Sa);
  const int p = 1 + G * RB + (threadIdx.x & 63);
  if (threadIdx.x < 64 && p <= a.N) {
    st_handoff(a.finM + (size_t)dir * (a.N + 1) + p, Ma);
    st_handoff(a.finS + (size_t)dir * (a.N + 1) + p, Sa);
  }
}

struct RecLds {
  double dg[2][RSQ];            // in-block coefficients [r][q] (ping-pong)
  double sg[RB];                // per-row shifts of an exact block solve
  double wl[RB];                // W of the block solved last
  double lv[RB];                // its values L
  double tacc[4][RB];           // look-ahead partial sums of the four waves
  double nm[RB], ns[RB];        // next block's accumulators, loaded by wave 1
  double pm[2][4][RB], ps[2][4][RB];  // fold partials (updater, or a self-fold in SOLO)
};

// LDS written by some lanes of a wave and read by others of the same wave:
// the fence keeps the compiler from forwarding or reordering across it.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Upper end of the chain frame's range (RecArgs::big, lowered only by a
// test hook): W up to e^575, so no sum of non-negative terms overflows, and
// a term that underflows is below e^-133 of its row's W (>= 1).
constexpr double REC_BIG = 1e250;

// Exact solve of row block b by wave 0 alone, without the chain frame:
// coefficients from the tables, per-row shifts sigma_q = max(M_q, max_{r<q}
// c(q,r) + sigma_r) by a (max,+) scan (no transcendental on its chain), the
// scaled exps a(q,r) = exp(c(q,r) + sigma_r - sigma_q) and the substitution
// W_q = b_q + sum_{r<q} a(q,r) W_r, W_q = exp(L(q) - sigma_q) (every term
// obeys c(q,r) + L(r) <= L(q) <= sigma_q + (rows) ln N).  Returns L of the
// lane's row (lane = row, accumulators (M, Sm)).
template <bool FWD>
__device__ __forceinline__ double rec_block_exact(const RecArgs &a, int p0, int nrows, double M, double Sm,
                                                  double *dg, double *sg) {
  const int lane = threadIdx.x & 63;
  // lane = row q, element [r][q] = c(q, r) for r < q (NEG otherwise)
#pragma unroll 16
  for (int r = 0; r < RB - 1; ++r) dg[r * RB + lane] = rec_coef_nb<FWD>(a, p0 + lane, p0 + r, r < lane && lane < nrows);
  double sig = lane < nrows ? M : NEG;
  for (int r = 0; r < RB - 1; ++r) sig = fmax(sig, dg[r * RB + lane] + readlane_d(sig, r));
  sg[lane] = lane < nrows ? sig : NEG;
  wave_lds_sync();
  for (int r = 0; r < RB - 1; ++r) {
    const double x = dg[r * RB + lane];
    dg[r * RB + lane] = x > NEG ? exp(x + sg[r] - sig) : 0.0;
  }
  double W = 0.0;
  if (lane < nrows && M > NEG && M > sig - 745.0) W = Sm * exp(M - sig);
  for (int r = 0; r < RB - 1; ++r) W = fma(dg[r * RB + lane], readlane_d(W, r), W);
  return sig + log(W);
}

// The solver walks the blocks in order.  Block b is solved in the frame of
// its k = 1 chain: row q's unknown is W_q = exp(L(q) - sigma - E(q)), with
// sigma = L(p0) (the block's first row has no in-block predecessor, so its
// accumulator is its value) and E the chain offsets of the set-up.  Since
// L(q) >= L(q-1) + c(q, q-1), L(q) >= sigma + E(q): every W_q >= 1.  The
// frame's coefficients A and look-ahead weights F come exponentiated from the
// set-up, so per block:
//   all waves: issue the loads of block b+1's coefficients;
//   wave 0: right-hand sides W_q = S_q exp(M_q - sigma - E(q)) (one exp),
//           forward substitution W_q += sum_{r<q} A(q, r) W_r (one FMA per
//           step), L(q) = sigma + E(q) + ln W_q, publish the block;
//   wave 1 meanwhile: poll block b+1's accumulators from its updater;
//   all waves: look-ahead sums Y(q') = sum_r F(q', r) W_r over a quarter of
//           the columns each (block b+1's row q' = lane);
//   wave 0: merge sigma + Delta + E'(q') + ln Y(q') into block b+1's rows
//           (E', Delta: block b+1's offsets and the step into it, so this is
//           the log-sum-exp of c(q', r) + L(r) over block b).
// A block whose W leaves [0, REC_BIG] (a row more than ~575 above its
// chain, or an overflowed coefficient) fails a wave vote; wave 0 then solves
// it with rec_block_exact and folds it into block b+1 from its values, as it
// also does for a row whose look-ahead sum left the range.
// SOLO: no updater workgroups; the solver folds blocks 0 .. b-1 into block
// b+1 itself (same order and arithmetic as an updater), so it waits on no
// other workgroup.
template <bool FWD, bool SOLO = false>
__device__ __forceinline__ void rec_solver(const RecArgs &a, int dir, RecLds &S) {
  const int N = a.N;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  int *err = a.sync + 4 * a.nblk;
  // Test hook: while sync[4 nblk + 1] (set by the host, never by a normal
  // run) is positive, a forward solve takes one from it and its first poll
  // of an updater (wave 1's) gives up at once, as a timed-out poll would.
  int spin = REC_SPIN_MAX;
  if (!SOLO && FWD && tid == 64) {
    int *inject = a.sync + 4 * a.nblk + 1;
    if (__hip_atomic_load(inject, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > 0 &&
        __hip_atomic_fetch_add(inject, -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > 0)
      spin = -1;
  }
  if (wv == 1) spin = __builtin_amdgcn_readfirstlane(spin);
  double *val = a.val2 + (size_t)dir * (N + 1);
  const double *AM = a.accM + (size_t)dir * (N + 1), *AS = a.accS + (size_t)dir * (N + 1);
  const double *FM = a.finM + (size_t)dir * (N + 1), *FS = a.finS + (size_t)dir * (N + 1);
  const double *DGd = a.dg + (size_t)dir * a.nblk * RSQ;
  const double *LAd = a.la + (size_t)dir * a.nblk * (RB * RB);
  const double *ECd = a.ech + (size_t)dir * a.nblk * (RB + 1);
  constexpr int DPT = RSQ / 256;           // in-block coefficients per thread
  constexpr int LPW = RB / 4;              // look-ahead columns per lane
  // block 0: coefficients, accumulators (column 0 only), chain offsets
  for (int w = tid; w < RSQ; w += 256) S.dg[0][w] = DGd[w];
  double M = NEG, Sm = 0.0, E = 0.0, Dl = 0.0;
  if (wv == 0) {
    const int nr0 = min(RB, N);
    M = lane < nr0 ? AM[1 + lane] : NEG;
    Sm = lane < nr0 ? AS[1 + lane] : 0.0;
    E = ECd[lane];
    Dl = ECd[RB];
  }
  for (int b = 0; b < a.nblk; ++b) {
    const int cur = b & 1;
    const int p0 = 1 + b * RB, nrows = min(RB, N - p0 + 1);
    const bool more = b + 1 < a.nblk;
    const int p1 = p0 + RB, nrows1 = more ? min(RB, N - p1 + 1) : 0;
    __syncthreads();
    // block b+1's coefficients, in flight during this block's solve
    double dgr[DPT], lar[LPW], En = 0.0, Dn = 0.0;
    if (more) {
#pragma unroll
      for (int k = 0; k < DPT; ++k) {
        const int w = tid + 256 * k;
        dgr[k] = DGd[(size_t)(b + 1) * RSQ + w];
      }
#pragma unroll
      for (int k = 0; k < LPW; ++k) lar[k] = LAd[(size_t)(b + 1) * (RB * RB) + (wv + 4 * k) * RB + lane];
      if (wv == 0) {
        En = ECd[(size_t)(b + 1) * (RB + 1) + lane];
        Dn = ECd[(size_t)(b + 1) * (RB + 1) + RB];
      }
    }
    double *dg = S.dg[cur];
    double sigma = 0.0;
    bool fast = true;
    if (wv == 0) {
      sigma = readlane_d(M, 0) + log(readlane_d(Sm, 0));
      double W = lane < nrows ? Sm * exp(M - sigma - E) : 0.0;
      // column r of dg holds A(q, r) at row q = lane (0 unless r < q)
#pragma unroll
      for (int r = 0; r < RB - 1; ++r) W = fma(dg[r * RB + lane], readlane_d(W, r), W);
      fast = __all(W <= a.big);  // NaN fails too
      const double v = fast ? sigma + E + log(W) : rec_block_exact<FWD>(a, p0, nrows, M, Sm, dg, S.sg);
      S.wl[lane] = W;
      S.lv[lane] = v;
      if (lane < nrows) {
        const int p = p0 + lane;
        st_handoff(val + p, v);  // the updaters poll these
        if (FWD) {
          a.Lf[p] = v;
          a.VBN[p] = -v / a.beta;
        } else {
          a.Rb[N - p] = v;
        }
      }
    } else if (wv == 1 && !SOLO && more) {
      // block b+1's accumulators: column 0 (set-up) for block 1, blocks
      // 0 .. b-1 from its updater otherwise
      double m = NEG, sv = 0.0;
      if (b + 1 >= 2) {
        rec_poll(FM + p1, FS + p1, nrows1, err, spin, m, sv);
        spin = REC_SPIN_MAX;
      } else if (lane < nrows1) {
        m = AM[p1 + lane];
        sv = AS[p1 + lane];
      }
      S.nm[lane] = lane < nrows1 ? m : NEG;
      S.ns[lane] = lane < nrows1 ? sv : 0.0;
    }
    __syncthreads();
    if (!more) break;
    {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < LPW; ++k) acc = fma(lar[k], S.wl[wv + 4 * k], acc);
      S.tacc[wv][lane] = acc;
    }
#pragma unroll
    for (int k = 0; k < DPT; ++k) {
      const int w = tid + 256 * k;
      S.dg[cur ^ 1][w] = dgr[k];
    }
    // SOLO: block b+1's accumulators over blocks 0 .. b-1, folded here
    double Mf = NEG, Sf = 0.0;
    if (SOLO && b + 1 >= 2) rec_fold_rows<FWD, false>(a, dir, b + 1, S.pm, S.ps, Mf, Sf);
    __syncthreads();
    if (wv == 0) {
      double Mn, Sn;
      if (SOLO) {
        Mn = b + 1 >= 2 ? Mf : (lane < nrows1 ? AM[p1 + lane] : NEG);
        Sn = b + 1 >= 2 ? Sf : (lane < nrows1 ? AS[p1 + lane] : 0.0);
      } else {
        Mn = S.nm[lane];
        Sn = S.ns[lane];
      }
      const double Y = (S.tacc[0][lane] + S.tacc[1][lane]) + (S.tacc[2][lane] + S.tacc[3][lane]);
      if (lane < nrows1) {
        if (fast && Y <= a.big) {
          if (Y > 0.0) lse_pair(Mn, Sn, sigma + Dl + En + log(Y), 1.0);
        } else {
          // block b folded into row q' from its values (block b is full)
          const int p = p1 + lane;
          double m = NEG, s = 0.0;
          for (int r0 = 0; r0 < RB; r0 += 16) {
            double c[16], t[16], m16, s16;
            rec_coef_col16<FWD>(a, p, p0 + r0, true, c);
#pragma unroll
            for (int j = 0; j < 16; ++j) t[j] = c[j] > NEG ? c[j] + S.lv[r0 + j] : NEG;
            lse16(t, m16, s16);
            lse_pair(m, s, m16, s16);
          }
          lse_pair(Mn, Sn, m, s);
        }
      }
      M = Mn;
      Sm = Sn;
      E = En;
      Dl = Dn;
    }
  }
}

__global__ __launch_bounds__(256) void k_rec_persist(RecArgs a, int dir0) {
  __shared__ RecLds S;
  chain_priority();
  const int per = 1 + max(0, a.nblk - 2);
  const int dir = dir0 + blockIdx.x / per, role = blockIdx.x % per;
  if (role == 0) {
    if (dir == 0)
      rec_solver<true>(a, 0, S);
    else
      rec_solver<false>(a, 1, S);
  } else {
    if (dir == 0)
      rec_updater<true>(a, 0, role + 1, S.pm, S.ps);
    else
      rec_updater<false>(a, 1, role + 1, S.pm, S.ps);
  }
}

__global__ void k_lnn(double *lnn, int N) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= N) lnn[i] = i > 0 ? log((double)i) : 0.0;
}

__global__ void k_vbn_to_L(const double *VBN, double *Lf, int N, double beta) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= N) Lf[i] = -beta * VBN[i];
}

// --------------------------------------------------------- exchange force

// Block per particle l (1-based l = blockIdx.x + 1): bead P against every
// bead 1 (row l of G) and bead 1 against every bead P (column l of G).
// Fx null: the forces are added into F.  Otherwise they are stored in Fx
// ([slice P-1, slice 0][dimension][particle]) for k_add_exch to add later.
__device__ __forceinline__ void exch_force_body(int bx, const double *__restrict__ X, const double *__restrict__ M,
                                                double *__restrict__ F, int N, int P, double om, Geo g,
                                                const RecArgs &a, double *__restrict__ Fx = nullptr) {
  __shared__ double red[8][4];
  const int l = bx + 1;
  const int nb = N * P;
  const int sP = (P - 1) * N + (l - 1), s1 = l - 1;
  const double LN = a.Lf[N];
  const double Rl = a.Rb[l], Llm1 = a.Lf[l - 1];
  const double xP[3] = {X[sP], X[nb + sP], X[2 * nb + sP]};
  const double x1[3] = {X[s1], X[nb + s1], X[2 * nb + s1]};
  const double Ls[3] = {g.L0, g.L1, g.L2};
  double gs = 0, gc[3] = {0, 0, 0}, hs = 0, hc[3] = {0, 0, 0};
  // row l: partners j = 1..l+1 (bead 1 of particle j)
  for (int j = 1 + threadIdx.x; j <= l + 1; j += blockDim.x) {
    double G;
    if (j <= l) {
      const int k = l - j + 1;
      if (a.xi0 && k != 1) continue;
      const double c = (k - 1) * a.lnxi - a.beta * a.ENk[tri_row(l) + k - 1] - a.lnn[l];
      G = exp(c + a.Lf[j - 1] + Rl - LN);
    } else {
      if (l == N) continue;
      G = 1 - exp(a.Lf[l] + Rl - LN);
    }
    const int sj = j - 1;
    gs += G;
#pragma unroll
    for (int d = 0; d < 3; ++d) gc[d] += G * comp(xP[d] - X[d * nb + sj], Ls[d], g.image);
  }
  // column l: partners j' = l-1 .. N (bead P of particle j')
  for (int jp = max(1, l - 1) + threadIdx.x; jp <= N; jp += blockDim.x) {
    double G;
    if (jp >= l) {
      const int k = jp - l + 1;
      if (a.xi0 && k != 1) continue;
      const double c = (k - 1) * a.lnxi - a.beta * a.ET[col_off(l - 1, N) + k - 1] - a.lnn[jp];
      G = exp(c + Llm1 + a.Rb[jp] - LN);
    } else {
      G = 1 - exp(a.Lf[l - 1] + a.Rb[l - 1] - LN);
    }
    const int sj = (P - 1) * N + jp - 1;
    hs += G;
#pragma unroll
    for (int d = 0; d < 3; ++d) hc[d] += G * comp(x1[d] - X[d * nb + sj], Ls[d], g.image);
  }
  // sums over the workgroup: butterfly within each wave, then the four wave
  // totals in a fixed order
  double v[8] = {gs, gc[0], gc[1], gc[2], hs, hc[0], hc[1], hc[2]};
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] += __shfl_xor(v[q], o, 64);
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int q = 0; q < 8; ++q) red[q][wv] = v[q];
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < 8; ++q) red[q][0] = ((red[q][0] + red[q][1]) + red[q][2]) + red[q][3];
    const int sPm = (P - 2) * N + (l - 1), s2 = N + (l - 1);
    const double mP = M[sP], m1 = M[s1];
    for (int d = 0; d < 3; ++d) {
      double aP = comp(xP[d] - X[d * nb + sPm], Ls[d], g.image);
      double gradP = (mP * om * om) * (red[0][0] * aP) + (mP * om * om) * red[1 + d][0];
      double a1 = comp(x1[d] - X[d * nb + s2], Ls[d], g.image);
      double grad1 = (m1 * om * om) * (red[4][0] * a1) + (m1 * om * om) * red[5 + d][0];
      if (Fx) {
        Fx[d * N + (l - 1)] = -gradP / mP;
        Fx[(3 + d) * N + (l - 1)] = -grad1 / m1;
      } else {
        F[d * nb + sP] += -gradP / mP;
        F[d * nb + s1] += -grad1 / m1;
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_exch_force(const double *__restrict__ X, const double *__restrict__ M,
                                                    double *__restrict__ F, int N, int P, double om, Geo g,
                                                    RecArgs a, double *__restrict__ Fx) {
  chain_priority();
  exch_force_body(blockIdx.x, X, M, F, N, P, om, g, a, Fx);
}

// F of the end beads (slices P-1 and 0, P >= 2) += the stored exchange forces.
__global__ __launch_bounds__(256) void k_add_exch(double *__restrict__ F, const double *__restrict__ Fx, int N, int P) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * N) return;
  const int d = t / N, l = t - d * N;
  const size_t nb = (size_t)N * P;
  F[d * nb + (size_t)(P - 1) * N + l] += Fx[t];
  F[d * nb + l] += Fx[3 * N + t];
}

// Optional full connection matrix (tests): G[(l-1)*N + (j-1)].
__global__ void k_connection_matrix(RecArgs a, double *__restrict__ G) {
  const int N = a.N;
  const int l = blockIdx.x + 1;
  const double LN = a.Lf[N];
  for (int j = 1 + threadIdx.x; j <= N; j += blockDim.x) {
    double v = 0;
    if (j == l + 1) {
      v = 1 - exp(a.Lf[l] + a.Rb[l] - LN);
    } else if (j <= l) {
      const int k = l - j + 1;
      if (!(a.xi0 && k != 1)) {
        const double c = (k - 1) * a.lnxi - a.beta * a.ENk[tri_row(l) + k - 1] - a.lnn[l];
        v = exp(c + a.Lf[j - 1] + a.Rb[l] - LN);
      }
    }
    G[(size_t)(l - 1) * N + (j - 1)] = v;
  }
}

// ------------------------------------------------------------ ENk2

// Link terms: nocl orientation b_i = e2((i,1),(i+1,P),(i+1,1)); forward
// orientation a_i = e2((i+1,1),(i,P),(i,1))  (0-based particles).
__global__ void k_chain2(const double *__restrict__ X, const double *__restrict__ M, int N, int P, double om, Geo g,
                         int forward, double *__restrict__ link) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int nb = N * P;
  if (i == N - 1) {
    link[i] = 0;
    return;
  }
  const int last = (P - 1) * N;
  link[i] = forward ? e2(X, M, nb, i + 1, last + i, i, om, g) : e2(X, M, nb, i, last + i + 1, i + 1, om, g);
}

__global__ __launch_bounds__(256) void k_enk2_fill(const double *__restrict__ X, const double *__restrict__ M, int N,
                                                   int P, double om, Geo g, int forward,
                                                   const double *__restrict__ pre, double *__restrict__ ENk2) {
  const int r = blockIdx.x;
  const int nb = N * P, last = (P - 1) * N;
  const dd ar{pre[2 * r], pre[2 * r + 1]};
  for (int k = 1 + threadIdx.x; k <= r + 1; k += blockDim.x) {
    const int m0 = r - k + 1;
    double v;
    if (k == 1) {
      v = e2(X, M, nb, r, last + r, r, om, g);
    } else {
      double links = dd_sub_to_double(ar, dd{pre[2 * m0], pre[2 * m0 + 1]});
      double close = forward ? e2(X, M, nb, m0, last + r, r, om, g) : e2(X, M, nb, r, last + m0, m0, om, g);
      v = links + close;
    }
    ENk2[tri_row(r + 1) + k - 1] = v;
  }
}

// ------------------------------------------------------------ estimators

// res(n) = sum_{m<n} w(n,m) (res(m) - B(n,m)) / sum_{m<n} w(n,m) with
// w(n,m) = exp(c(n,m) + L(m) - L(n)), B = ENk or ENk2
// (md_pimd_x.c:1273-1288, 4294-4319).  Linear recursion: blocked like the
// forward pass, the diagonal is a 64-step forward substitution.
__global__ __launch_bounds__(RT) void k_estimator(RecArgs a, const double *__restrict__ Bm, double *__restrict__ out) {
  extern __shared__ double lds[];
  const int N = a.N;
  double *wd = lds;                       // RB*(RB-1)/2 weights
  double *bd = wd + RB * (RB - 1) / 2;    // RB*(RB-1)/2 matrix values
  double *ao = bd + RB * (RB - 1) / 2;    // RB
  double *wo = ao + RB;                   // RB
  double *res = wo + RB;                  // N+1
  double *Ls = res + N + 1;               // N+1
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  for (int p = tid; p <= N; p += RT) Ls[p] = a.Lf[p];
  if (tid == 0) res[0] = 0.0;
  __syncthreads();
  for (int p0 = 1; p0 <= N; p0 += RB) {
    const int nrows = min(RB, N - p0 + 1);
    for (int r = wv; r < nrows; r += RT / 64) {
      const int n = p0 + r;
      double acc = 0, sw = 0;
      for (int m = lane; m < p0; m += 64) {
        double c = rec_coef<true>(a, n, m);
        if (c > NEG) {
          double w = exp(c + Ls[m] - Ls[n]);
          acc += w * (res[m] - Bm[tri_row(n) + (n - m) - 1]);
          sw += w;
        }
      }
      acc = wave_sum(acc);
      sw = wave_sum(sw);
      if (lane == 0) {
        ao[r] = acc;
        wo[r] = sw;
      }
    }
    for (int w = tid; w < RB * (RB - 1) / 2; w += RT) {
      int q = (int)((1.0 + sqrt(1.0 + 8.0 * w)) / 2.0);
      while (q * (q - 1) / 2 > w) --q;
      while ((q + 1) * q / 2 <= w) ++q;
      int r = w - q * (q - 1) / 2;
      double wt = 0, bv = 0;
      if (q < nrows) {
        const int n = p0 + q, m = p0 + r;
        double c = rec_coef<true>(a, n, m);
        if (c > NEG) wt = exp(c + Ls[m] - Ls[n]);
        bv = Bm[tri_row(n) + (n - m) - 1];
      }
      wd[w] = wt;
      bd[w] = bv;
    }
    __syncthreads();
    if (wv == 0) {
      double acc = lane < nrows ? ao[lane] : 0.0;
      double sw = lane < nrows ? wo[lane] : 0.0;
      for (int r = 0; r < nrows; ++r) {
        const double rr = readlane_d(acc, r) / readlane_d(sw, r);
        if (lane > r && lane < nrows) {
          const int w = lane * (lane - 1) / 2 + r;
          acc += wd[w] * (rr - bd[w]);
          sw += wd[w];
        }
      }
      if (lane < nrows) res[p0 + lane] = acc / sw;
    }
    __syncthreads();
  }
  if (tid == 0) *out = res[N];
}

// -1/2 sum_{j>=2} (x_{l,j} - x_{l,1}) . f_{l,j} m  partials (md_pimd_x.c:4332-4345)
__global__ __launch_bounds__(256) void k_vir_term(const double *__restrict__ X, const double *__restrict__ F,
                                                  const double *__restrict__ M, int N, int P, Geo g,
                                                  double *__restrict__ part) {
  __shared__ double red[256];
  const int nb = N * P;
  double s = 0;
  for (int i = N + blockIdx.x * 256 + threadIdx.x; i < nb; i += gridDim.x * 256) {
    const int l = i % N;
    const double m = M[i];
    s += -comp(X[i] - X[l], g.L0, g.image) * F[i] * m;
    s += -comp(X[nb + i] - X[nb + l], g.L1, g.image) * F[nb + i] * m;
    s += -comp(X[2 * nb + i] - X[2 * nb + l], g.L2, g.image) * F[2 * nb + i] * m;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Box-centred virial sum_{all beads} (x - L/2) . f m (md_pimd_x.c:5508-5515).
__global__ __launch_bounds__(256) void k_box_virial(const double *__restrict__ X, const double *__restrict__ F,
                                                    const double *__restrict__ M, int nb, Geo g,
                                                    double *__restrict__ part) {
  __shared__ double red[256];
  double s = 0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < nb; i += gridDim.x * 256) {
    const double m = M[i];
    s += (X[i] - g.L0 / 2.0) * F[i] * m;
    s += (X[nb + i] - g.L1 / 2.0) * F[nb + i] * m;
    s += (X[2 * nb + i] - g.L2 / 2.0) * F[2 * nb + i] * m;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Sign-tracked forward recursion at |xi'| (md_pimd_x.c:4675-4701): one
// workgroup, one n at a time (a per-sample estimator, not per step).
// One workgroup per xi' of a grid: out[b] = sign * exp(-beta V'(N) + beta V_B(N)).
__global__ __launch_bounds__(256) void k_vi_sign(RecArgs a, const double *__restrict__ vi2s,
                                                 double *__restrict__ outs) {
  const double vi2 = vi2s[blockIdx.x];
  double *out = outs + blockIdx.x;
  __shared__ double red[256];
  extern __shared__ double lds[];
  const int N = a.N;
  double *V2 = lds, *sg = lds + N + 1;
  const double av = fabs(vi2), sv = vi2 < 0 ? -1.0 : 1.0;
  const double lv = av > 0 ? log(av) : 0.0;
  if (threadIdx.x == 0) {
    V2[0] = 0;
    sg[0] = 1;
  }
  __syncthreads();
  for (int n = 1; n <= N; ++n) {
    double mn = 1e10;
    if (vi2 == 0.0) {
      mn = a.beta * (a.ENk[tri_row(n)] + V2[n - 1]);
    } else {
      for (int k = 1 + threadIdx.x; k <= n; k += 256) mn = fmin(mn, -(k - 1) * lv + a.beta * (a.ENk[tri_row(n) + k - 1] + V2[n - k]));
      red[threadIdx.x] = mn;
      __syncthreads();
      for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
      }
      mn = red[0];
      __syncthreads();
    }
    double s = 0;
    for (int k = 1 + threadIdx.x; k <= n; k += 256) {
      if (vi2 == 0.0 && k != 1) continue;
      double sk = ((k - 1) % 2 == 1) ? sv : 1.0;
      double arg = (vi2 == 0.0 ? 0.0 : (k - 1) * lv) - a.beta * (a.ENk[tri_row(n) + k - 1] + V2[n - k]) + mn;
      s += sg[n - k] * sk * exp(arg);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      double t = red[0];
      sg[n] = t >= 0 ? 1 : -1;
      V2[n] = (mn - log(fabs(t)) + log((double)n)) / a.beta;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sg[N] * exp(-a.beta * V2[N] - a.Lf[N]);
}

// ------------------------------------------------- batched replicas
//
// One launch per stage over all replicas of a batch; the per-replica work of
// each kernel is the single-system kernel's body with the same block
// decomposition (virtual block index = global block - start[replica]).
struct ExRep {
  const double *X, *M;
  double *F;
  int N, P;
  double om;
  Geo g;
  double *Eint, *clink, *Cpre, *ENk, *ET;
  RecArgs a;
  int prep_blocks;  // blocks per direction of the recursion set-up
};

__global__ __launch_bounds__(256) void kb_chain(const ExRep *__restrict__ reps, const int *__restrict__ start,
                                                int nrep) {
  const int r = batch_rep(start, nrep, blockIdx.x);
  const ExRep &R = reps[r];
  chain_body(blockIdx.x - start[r], R.X, R.M, R.N, R.P, R.om, R.g, R.Eint, R.clink, R.ENk, R.ET);
}

// one 1024-thread block per replica
__global__ __launch_bounds__(256) void kb_scan_dd(const ExRep *__restrict__ reps) {
  const ExRep &R = reps[blockIdx.x];
  scan_dd_body(R.clink, R.N, R.Cpre);
}

__global__ __launch_bounds__(256) void kb_enk_fill(const ExRep *__restrict__ reps, const int *__restrict__ start,
                                                   int nrep) {
  const int r = batch_rep(start, nrep, blockIdx.x);
  const ExRep &R = reps[r];
  enk_fill_body(blockIdx.x - start[r], R.X, R.M, R.N, R.P, R.om, R.g, R.Eint, R.Cpre, R.ENk, R.ET);
}

// both directions: blocks [0, prep_blocks) forward, [prep_blocks, 2 prep_blocks) backward
__global__ __launch_bounds__(256) void kb_rec_prep(const ExRep *__restrict__ reps, const int *__restrict__ start,
                                                   int nrep) {
  const int r = batch_rep(start, nrep, blockIdx.x);
  __shared__ double E[2][RB + 1];
  const ExRep &R = reps[r];
  const int lb = blockIdx.x - start[r];
  const int dir = lb >= R.prep_blocks ? 1 : 0, bx = lb - dir * R.prep_blocks;
  rec_prep_wg<false>(R.a, dir, bx, EnkArgs{}, E);
}

// one workgroup per (replica, direction), each solving its whole recursion
// alone (no hand-off between workgroups)
__global__ __launch_bounds__(256) void kb_rec_solo(const ExRep *__restrict__ reps) {
  __shared__ RecLds S;
  const ExRep &R = reps[blockIdx.x >> 1];
  if ((blockIdx.x & 1) == 0)
    rec_solver<true, true>(R.a, 0, S);
  else
    rec_solver<false, true>(R.a, 1, S);
}

__global__ __launch_bounds__(256) void kb_exch_force(const ExRep *__restrict__ reps, const int *__restrict__ start,
                                                     int nrep) {
  const int r = batch_rep(start, nrep, blockIdx.x);
  const ExRep &R = reps[r];
  exch_force_body(blockIdx.x - start[r], R.X, R.M, R.F, R.N, R.P, R.om, R.g, R.a);
}

// ------------------------------------------------------------ launchers

static Geo geo_of(DevPimd *p, int image) {
  md_simulation_t_x *sim = p->host->sim;
  Geo g;
  g.L0 = box_L(sim, 0);
  g.L1 = box_L(sim, 1);
  g.L2 = box_L(sim, 2);
  g.iL0 = g.L0 != 0 ? 1.0 / g.L0 : 0.0;
  g.iL1 = g.L1 != 0 ? 1.0 / g.L1 : 0.0;
  g.iL2 = g.L2 != 0 ? 1.0 / g.L2 : 0.0;
  g.image = image;
  return g;
}

static double g_rec_frame_limit = REC_BIG;

void set_recursion_frame_limit(double limit) { g_rec_frame_limit = limit > 0.0 ? limit : REC_BIG; }

static RecArgs rec_args(DevPimd *p) {
  RecArgs a;
  a.ENk = p->ENk;
  a.ET = p->ET;
  a.lnn = p->work;  // first N+1 entries of the scratch hold ln n
  a.Lf = p->Lf;
  a.Rb = p->Rb;
  a.VBN = p->VBN;
  a.val2 = p->rec;
  a.accM = p->rec + 2 * (p->N + 1);
  a.accS = p->rec + 4 * (p->N + 1);
  a.finM = p->rec + 6 * (p->N + 1);
  a.finS = p->rec + 8 * (p->N + 1);
  a.nblk = (p->N + RB - 1) / RB;
  a.dg = p->rec + 10 * (p->N + 1);
  a.la = a.dg + 2 * (size_t)a.nblk * RSQ;
  a.ech = a.la + 2 * (size_t)a.nblk * (RB * RB);
  a.big = g_rec_frame_limit;
  a.sync = p->recsync;
  a.N = p->N;
  a.beta = p->host->beta;
  a.xi0 = p->host->vi == 0.0;
  a.lnxi = a.xi0 ? 0.0 : log(p->host->vi);
  return a;
}

int init_lnn(DevPimd *p) {
  hipLaunchKernelGGL(k_lnn, dim3((p->N + 256) / 256), dim3(256), 0, rt().s0, p->work, p->N);
  XMD_CHECK(hipGetLastError());
  return 0;
}

int launch_enk(DevPimd *p, int image, hipStream_t st) {
  DevSim *d = p->ds;
  const int N = p->N, P = p->P;
  Geo g = geo_of(p, image);
  double om = p->host->omegaP;
  double *clink = p->work + (N + 1);
  hipLaunchKernelGGL(k_chain, dim3((N + 63) / 64), dim3(256), 0, st, d->X, d->M, N, P, om, g, p->Eint, clink,
                     p->ENk, p->ET);
  XMD_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_scan_dd, dim3(1), dim3(256), 0, st, clink, N, p->Cpre);
  XMD_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_enk_fill, dim3(N), dim3(256), 0, st, d->X, d->M, N, P, om, g, p->Eint, p->Cpre, p->ENk, p->ET);
  XMD_CHECK(hipGetLastError());
  p->enk_dev = true;
  p->last_image = image;
  p->enk_version++;
  return 0;
}


// Objects with a recursion launched since their error word was last read
// (each listed once; objects live until the process ends).
static std::vector<DevPimd *> g_rec_pending;

static void rec_launched(DevPimd *p) {
  for (DevPimd *q : g_rec_pending)
    if (q == p) return;
  g_rec_pending.push_back(p);
}

// Error word of the persistent solve (a timed-out hand-off in any solve since
// the last read): read after a synchronisation, then cleared here on the
// host, so a timeout in any step of a multi-step call is reported once.
int recursion_status(DevPimd *p) {
  if (!p || !p->recsync) return 0;
  for (size_t i = 0; i < g_rec_pending.size(); ++i)
    if (g_rec_pending[i] == p) {
      g_rec_pending.erase(g_rec_pending.begin() + i);
      break;
    }
  const int nblk = (p->N + RB - 1) / RB;
  int err = 0;
  XMD_CHECK(hipMemcpy(&err, p->recsync + 4 * nblk, sizeof(int), hipMemcpyDeviceToHost));
  if (err) {
    fprintf(stderr, "xmd: exchange recursion hand-off timed out\n");
    XMD_CHECK(hipMemset(p->recsync + 4 * nblk, 0, sizeof(int)));
    return -1;
  }
  return 0;
}

// recursion_status of every object with a recursion launched since its last
// read (after a device-wide synchronisation).
static int noalloc_status();

int recursion_status_all() {
  int rc = noalloc_status();
  while (!g_rec_pending.empty())
    if (recursion_status(g_rec_pending.back())) rc = -1;
  return rc;
}

int inject_recursion_timeouts(DevPimd *p, int count) {
  const int nblk = (p->N + RB - 1) / RB;
  XMD_CHECK(hipStreamSynchronize(rt().s0));
  XMD_CHECK(hipMemcpy(p->recsync + 4 * nblk + 1, &count, sizeof(int), hipMemcpyHostToDevice));
  return 0;
}

int launch_enk_prep(DevPimd *p, int image, hipStream_t st) {
  DevSim *d = p->ds;
  const int N = p->N, P = p->P;
  EnkArgs e;
  e.X = d->X;
  e.M = d->M;
  e.Eint = p->Eint;
  e.pre = p->Cpre;
  e.om = p->host->omegaP;
  e.g = geo_of(p, image);
  e.N = N;
  e.P = P;
  double *clink = p->work + (N + 1);
  hipLaunchKernelGGL(k_chain_scan, dim3((N + 63) / 64), dim3(256), 0, st, d->X, d->M, N, P, e.om, e.g, p->Eint, clink,
                     p->ENk, p->ET, p->Cpre, p->done);
  XMD_CHECK(hipGetLastError());
  RecArgs a = rec_args(p);
  const int nprepb = rec_prep_wgs(N, a.nblk);
  hipLaunchKernelGGL(k_enk_prep, dim3(N + 2 * nprepb), dim3(256), 0, st, e, p->ENk, p->ET, a, nprepb);
  XMD_CHECK(hipGetLastError());
  p->enk_dev = true;
  p->last_image = image;
  p->enk_version++;
  return 0;
}

int launch_recursion_solve(DevPimd *p, hipStream_t st) {
  RecArgs a = rec_args(p);
  const int per = 1 + std::max(0, a.nblk - 2);
  hipLaunchKernelGGL(k_rec_persist, dim3(2 * per), dim3(256), 0, st, a, 0);
  XMD_CHECK(hipGetLastError());
  rec_launched(p);
  p->vbn_dev = true;
  p->back_dev = true;
  p->back_version = p->enk_version;
  return 0;
}

int launch_recursion(DevPimd *p, bool fwd, bool bwd, hipStream_t st, hipEvent_t mid) {
  const int ndir = (fwd ? 1 : 0) + (bwd ? 1 : 0);
  if (ndir == 0) return 0;
  const int dir0 = fwd ? 0 : 1;
  const int N = p->N;
  RecArgs a = rec_args(p);
  hipLaunchKernelGGL(k_rec_prep, dim3(rec_prep_wgs(N, a.nblk), ndir), dim3(256), 0, st, a, dir0);
  XMD_CHECK(hipGetLastError());
  if (mid) XMD_CHECK(hipEventRecord(mid, st));
  const int per = 1 + std::max(0, a.nblk - 2);
  hipLaunchKernelGGL(k_rec_persist, dim3(ndir * per), dim3(256), 0, st, a, dir0);
  XMD_CHECK(hipGetLastError());
  rec_launched(p);
  if (fwd) p->vbn_dev = true;
  if (bwd) {
    p->back_dev = true;
    p->back_version = p->enk_version;
  }
  return 0;
}

int launch_exchange_force(DevPimd *p, int image, hipStream_t st, bool stored) {
  DevSim *d = p->ds;
  Geo g = geo_of(p, image);
  RecArgs a = rec_args(p);
  hipLaunchKernelGGL(k_exch_force, dim3(p->N), dim3(256), 0, st, d->X, d->M, d->F, p->N, p->P, p->host->omegaP, g, a,
                     stored ? p->Fx : nullptr);
  XMD_CHECK(hipGetLastError());
  return 0;
}

int launch_add_exchange_force(DevPimd *p, hipStream_t st) {
  hipLaunchKernelGGL(k_add_exch, dim3((3 * p->N + 255) / 256), dim3(256), 0, st, p->ds->F, p->Fx, p->N, p->P);
  XMD_CHECK(hipGetLastError());
  return 0;
}

// ------------------------------------------------ allocation-free V_B
//
// md_pimd_fill_VB_noalloc_x runs the recursion of md_pimd_fill_VB_x (the same
// set-up and persistent-solve kernels) on cycle energies the caller already
// holds in device memory, with all of its working memory in the file-scope
// array `scratch` below instead of buffers the library allocates.  Layout
// (doubles): the column copy of the cycle energies, ln n, the forward and
// backward values, the solve state (values, accumulators, hand-off values,
// block coefficients of both directions), then the sync words as ints.  They
// are placed so that the error word and the injected-timeout count sit at
// fixed positions for every N (the 4 nblk unused words of a smaller N end
// where those of N_max end); the host clears the error word when it reports
// it.
constexpr int NA_NMAX = 1024;
constexpr int NA_NBLK = (NA_NMAX + RB - 1) / RB;
constexpr size_t NA_ET = 0;
constexpr size_t NA_LNN = NA_ET + (size_t)NA_NMAX * (NA_NMAX + 1) / 2;
constexpr size_t NA_LF = NA_LNN + (NA_NMAX + 1);
constexpr size_t NA_RB = NA_LF + (NA_NMAX + 1);
constexpr size_t NA_REC = NA_RB + (NA_NMAX + 1);
constexpr size_t NA_SYNC = NA_REC + 10 * (size_t)(NA_NMAX + 1) + 2 * (size_t)NA_NBLK * (2 * RSQ + RB + 1);
constexpr int NA_NSYNC = 4 * NA_NBLK + 2;  // flags of N_max, error word, injected-timeout count (unused here)
constexpr size_t NA_DOUBLES = NA_SYNC + (NA_NSYNC + 1) / 2;
static_assert(NA_DOUBLES * sizeof(double) < 7u * 1024 * 1024, "scratch stays below 7 MiB");

static __device__ __align__(256) double scratch[NA_DOUBLES];

// Device address of `scratch` per device (the host-side name is not a device
// address), taken once per device by noalloc_init at runtime set-up.
constexpr int NA_MAXDEV = 64;
static double *g_scratch[NA_MAXDEV];
static bool g_noalloc_pending = false;  // a solve was queued since the last status read

int noalloc_init(int device) {
  if (device < 0 || device >= NA_MAXDEV) return -1;
  if (g_scratch[device]) return 0;
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    fprintf(stderr, "xmd: hipGetSymbolAddress(scratch) failed: %s\n", hipGetErrorString(e));
    return -1;
  }
  g_scratch[device] = (double *)ptr;
  return 0;
}

static int *noalloc_err_word(double *base) { return (int *)(base + NA_SYNC) + 4 * NA_NBLK; }

// ln n (n = 0..N) into the scratch's table.
__global__ void k_noalloc_lnn(double *__restrict__ lnn, int N) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= N) lnn[i] = i > 0 ? log((double)i) : 0.0;
}

// Vt^(m) = -R(m) / beta.
__global__ void k_noalloc_vt(const double *__restrict__ Rb, double *__restrict__ Vt, int N, double beta) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= N) Vt[i] = -Rb[i] / beta;
}

static int noalloc_fill_vb(int N, double beta, double vi, const double *ENk, double *VBN, double *Vt) {
  Runtime &R = rt();
  double *base = g_scratch[R.device];
  const int nblk = (N + RB - 1) / RB;
  RecArgs a;
  a.ENk = ENk;
  a.ET = base + NA_ET;
  a.lnn = base + NA_LNN;
  a.Lf = base + NA_LF;
  a.Rb = base + NA_RB;
  a.VBN = VBN;
  a.val2 = base + NA_REC;
  a.accM = a.val2 + 2 * (N + 1);
  a.accS = a.val2 + 4 * (N + 1);
  a.finM = a.val2 + 6 * (N + 1);
  a.finS = a.val2 + 8 * (N + 1);
  a.dg = a.val2 + 10 * (N + 1);
  a.la = a.dg + 2 * (size_t)nblk * RSQ;
  a.ech = a.la + 2 * (size_t)nblk * (RB * RB);
  a.big = g_rec_frame_limit;
  a.sync = (int *)(base + NA_SYNC) + 4 * (NA_NBLK - nblk);  // error word at the fixed slot 4 NA_NBLK
  a.N = N;
  a.nblk = nblk;
  a.beta = beta;
  a.xi0 = vi == 0.0;
  a.lnxi = a.xi0 ? 0.0 : log(vi);
  const hipStream_t st = R.s0;
  hipLaunchKernelGGL(k_noalloc_lnn, dim3((N + 256) / 256), dim3(256), 0, st, base + NA_LNN, N);
  hipLaunchKernelGGL(k_enk_transpose, dim3(N), dim3(256), 0, st, ENk, base + NA_ET, N);
  const int ndir = Vt ? 2 : 1;
  hipLaunchKernelGGL(k_rec_prep, dim3(rec_prep_wgs(N, nblk), ndir), dim3(256), 0, st, a, 0);
  const int per = 1 + std::max(0, nblk - 2);
  hipLaunchKernelGGL(k_rec_persist, dim3(ndir * per), dim3(256), 0, st, a, 0);
  if (Vt) hipLaunchKernelGGL(k_noalloc_vt, dim3((N + 256) / 256), dim3(256), 0, st, base + NA_RB, Vt, N, beta);
  XMD_CHECK(hipGetLastError());
  g_noalloc_pending = true;
  return 0;
}

// Error word of the allocation-free solves queued since the last read
// (call after a synchronisation); cleared once reported.
static int noalloc_status() {
  if (!g_noalloc_pending) return 0;
  g_noalloc_pending = false;
  int *err = noalloc_err_word(g_scratch[rt().device]);
  int v = 0;
  XMD_CHECK(hipMemcpy(&v, err, sizeof(int), hipMemcpyDeviceToHost));
  if (v) {
    fprintf(stderr, "xmd: exchange recursion hand-off timed out (allocation-free solve)\n");
    const int zero = 0;
    XMD_CHECK(hipMemcpy(err, &zero, sizeof(int), hipMemcpyHostToDevice));
    return -1;
  }
  return 0;
}

struct ExchBatch {
  std::vector<DevPimd *> ps;
  std::vector<int> image;
  BatchTable chain, enk, prep, force;
  ExRep *reps = nullptr;  // device table shared by the per-replica-grid kernels
  int nrep = 0;
};

ExchBatch *exch_batch_create(const std::vector<BatchItem> &items) {
  ExchBatch *b = new ExchBatch();
  std::vector<ExRep> reps;
  std::vector<int> bc, be, bp, bf;
  for (const BatchItem &it : items) {
    DevPimd *p = it.p;
    DevSim *d = p->ds;
    ExRep R;
    R.X = d->X;
    R.M = d->M;
    R.F = d->F;
    R.N = p->N;
    R.P = p->P;
    R.om = p->host->omegaP;
    R.g = geo_of(p, it.image);
    R.Eint = p->Eint;
    R.clink = p->work + (p->N + 1);
    R.Cpre = p->Cpre;
    R.ENk = p->ENk;
    R.ET = p->ET;
    R.a = rec_args(p);
    R.prep_blocks = rec_prep_wgs(p->N, R.a.nblk);
    reps.push_back(R);
    bc.push_back((p->N + 63) / 64);
    be.push_back(p->N);
    bp.push_back(2 * R.prep_blocks);
    bf.push_back(p->N);
    b->ps.push_back(p);
    b->image.push_back(it.image);
  }
  b->nrep = (int)reps.size();
  const size_t sz = sizeof(ExRep);
  if (hipGetLastError() != hipSuccess || batch_table_make(b->chain, reps.data(), sz, bc) ||
      batch_table_make(b->enk, reps.data(), sz, be) || batch_table_make(b->prep, reps.data(), sz, bp) ||
      batch_table_make(b->force, reps.data(), sz, bf) || hipStreamSynchronize(rt().s0) != hipSuccess) {
    exch_batch_free(b);
    return nullptr;
  }
  b->reps = (ExRep *)b->chain.reps;
  return b;
}

void exch_batch_free(ExchBatch *b) {
  if (!b) return;
  batch_table_free(b->chain);
  batch_table_free(b->enk);
  batch_table_free(b->prep);
  batch_table_free(b->force);
  delete b;
}

// cycle energies (k_chain, k_scan_dd, k_enk_fill of every replica)
int exch_batch_enk(ExchBatch *b, hipStream_t st) {
  hipLaunchKernelGGL(kb_chain, dim3(b->chain.total), dim3(256), 0, st, (const ExRep *)b->chain.reps, b->chain.start,
                     b->nrep);
  hipLaunchKernelGGL(kb_scan_dd, dim3(b->nrep), dim3(256), 0, st, (const ExRep *)b->chain.reps);
  hipLaunchKernelGGL(kb_enk_fill, dim3(b->enk.total), dim3(256), 0, st, (const ExRep *)b->enk.reps, b->enk.start,
                     b->nrep);
  XMD_CHECK(hipGetLastError());
  for (size_t i = 0; i < b->ps.size(); ++i) {
    DevPimd *p = b->ps[i];
    p->enk_dev = true;
    p->last_image = b->image[i];
    p->enk_version++;
  }
  return 0;
}

// forward and backward recursions of every replica
int exch_batch_recursion(ExchBatch *b, hipStream_t st) {
  hipLaunchKernelGGL(kb_rec_prep, dim3(b->prep.total), dim3(256), 0, st, (const ExRep *)b->prep.reps, b->prep.start,
                     b->nrep);
  hipLaunchKernelGGL(kb_rec_solo, dim3(2 * b->nrep), dim3(256), 0, st, (const ExRep *)b->prep.reps);
  XMD_CHECK(hipGetLastError());
  for (DevPimd *p : b->ps) {
    p->vbn_dev = true;
    p->back_dev = true;
    p->back_version = p->enk_version;
  }
  return 0;
}

int exch_batch_force(ExchBatch *b, hipStream_t st) {
  hipLaunchKernelGGL(kb_exch_force, dim3(b->force.total), dim3(256), 0, st, (const ExRep *)b->force.reps,
                     b->force.start, b->nrep);
  XMD_CHECK(hipGetLastError());
  return 0;
}

int exch_batch_status(ExchBatch *b) {
  int rc = 0;
  for (DevPimd *p : b->ps)
    if (recursion_status(p)) rc = -1;
  return rc;
}

// ENk filled on the host (driver-provided) -> device, both layouts.
static int ensure_enk(DevPimd *p) {
  if (p->enk_dev) return 0;
  size_t tri = (size_t)p->N * (p->N + 1) / 2;
  hipStream_t st = rt().s0;
  XMD_CHECK(hipMemcpyAsync(p->ENk, p->host->ENk[0], tri * sizeof(double), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_enk_transpose, dim3(p->N), dim3(256), 0, st, p->ENk, p->ET, p->N);
  XMD_CHECK(hipGetLastError());
  XMD_CHECK(hipStreamSynchronize(st));
  p->enk_dev = true;
  p->enk_version++;
  return 0;
}

static int ensure_vbn(DevPimd *p) {
  if (p->vbn_dev) return 0;
  hipStream_t st = rt().s0;
  XMD_CHECK(hipMemcpyAsync(p->VBN, p->host->VBN, (p->N + 1) * sizeof(double), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_vbn_to_L, dim3((p->N + 256) / 256), dim3(256), 0, st, p->VBN, p->Lf, p->N, p->host->beta);
  XMD_CHECK(hipGetLastError());
  XMD_CHECK(hipStreamSynchronize(st));
  p->vbn_dev = true;
  return 0;
}

static size_t est_lds(int N) { return (size_t)(RB * (RB - 1) + 2 * RB + 2 * (N + 1)) * sizeof(double); }

static int run_estimator(DevPimd *p, const double *Bm, double *out) {
  if (est_lds(p->N) > 65536) {
    fprintf(stderr, "xmd: estimator recursion supports N <= 2015\n");
    return -1;
  }
  RecArgs a = rec_args(p);
  double *d_out = p->work + 4 * (p->N + 1);
  hipLaunchKernelGGL(k_estimator, dim3(1), dim3(RT), est_lds(p->N), rt().s0, a, Bm, d_out);
  XMD_CHECK(hipGetLastError());
  XMD_CHECK(hipMemcpyAsync(out, d_out, sizeof(double), hipMemcpyDeviceToHost, rt().s0));
  XMD_CHECK(hipStreamSynchronize(rt().s0));
  return 0;
}

static int fill_enk2(DevPimd *p, int image, int forward) {
  DevSim *d = p->ds;
  const int N = p->N, P = p->P;
  Geo g = geo_of(p, image);
  double om = p->host->omegaP;
  hipStream_t st = rt().s0;
  double *link = p->work + 2 * (N + 1);
  hipLaunchKernelGGL(k_chain2, dim3((N + 255) / 256), dim3(256), 0, st, d->X, d->M, N, P, om, g, forward, link);
  XMD_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_scan_dd, dim3(1), dim3(256), 0, st, link, N, p->C2pre);
  XMD_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_enk2_fill, dim3(N), dim3(256), 0, st, d->X, d->M, N, P, om, g, forward, p->C2pre, p->ENk2);
  XMD_CHECK(hipGetLastError());
  p->enk2_dev = true;
  return 0;
}

int connection_matrix(DevPimd *p, double *host_G) {
  size_t nn = (size_t)p->N * p->N;
  if (!p->Gm) XMD_CHECK(hipMalloc(&p->Gm, nn * sizeof(double)));
  RecArgs a = rec_args(p);
  hipLaunchKernelGGL(k_connection_matrix, dim3(p->N), dim3(256), 0, rt().s0, a, p->Gm);
  XMD_CHECK(hipGetLastError());
  XMD_CHECK(hipMemcpyAsync(host_G, p->Gm, nn * sizeof(double), hipMemcpyDeviceToHost, rt().s0));
  XMD_CHECK(hipStreamSynchronize(rt().s0));
  return 0;
}

int backward_potentials(DevPimd *p, double *host_Vt) {
  std::vector<double> R(p->N + 1);
  XMD_CHECK(hipMemcpy(R.data(), p->Rb, (p->N + 1) * sizeof(double), hipMemcpyDeviceToHost));
  for (int i = 0; i <= p->N; ++i) host_Vt[i] = -R[i] / p->host->beta;
  return 0;
}

// ------------------------------------------------ single-value helpers
//
// The reference's per-element helpers of the exchange path, evaluated on the
// device state: one cycle energy, one link-virial cycle term, the log-sum-exp
// shifts and one bead's spring gradient.  A single workgroup each; the sums
// are tree reductions (the reference sums sequentially).

__device__ double block_sum(double v, double *red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  return red[0];
}

__device__ double block_min(double v, double *red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  return red[0];
}

// E_{N2}^{(k)} by the explicit cycle walk (md_pimd_x.c:569-591): springs from
// bead (l, j) to its successor on the ring of particles N2-k+1..N2.
__global__ __launch_bounds__(256) void k_cycle_energy(const double *__restrict__ X, const double *__restrict__ M, int N,
                                                     int P, int N2, int k, double om, Geo g, double *__restrict__ out) {
  __shared__ double red[256];
  const int nb = N * P, l0 = N2 - k + 1;
  double e = 0;
  for (int t = threadIdx.x; t < k * P; t += blockDim.x) {
    const int l = l0 + t / P, j = t % P + 1;  // 1-based
    const int sa = (j - 1) * N + (l - 1);
    int sb;
    if (j < P)
      sb = j * N + (l - 1);
    else
      sb = (l == N2) ? (l0 - 1) : l;  // bead 1 of the next particle on the ring
    e += spring(X, nb, sa, sb, M[sa], om, g);
  }
  e = block_sum(e, red);
  if (threadIdx.x == 0) out[0] = e;
}

// Link-virial cycle term (md_pimd_x.c:3959-3987 with forward = 1, and
// md_pimd_x.c:4376-4420 for one explicit bead triple, pbc = 0):
// sum over the listed triples (a, b, c) of 1/2 m_a w^2 (x_a - x_b).(x_a - x_c).
__global__ __launch_bounds__(256) void k_cycle_virial(const double *__restrict__ X, const double *__restrict__ M, int N,
                                                     int P, int N2, int k, int sa1, int sb1, int sc1, double om, Geo g,
                                                     double *__restrict__ out) {
  __shared__ double red[256];
  const int nb = N * P;
  double e = 0;
  const int n = sa1 >= 0 ? 1 : k;
  for (int t = threadIdx.x; t < n; t += blockDim.x) {
    int sa, sb, sc;
    if (sa1 >= 0) {
      sa = sa1;
      sb = sb1;
      sc = sc1;
    } else {
      const int l = N2 - k + 1 + t;  // 1-based
      sa = (l == N2) ? (N2 - k) : l;  // successor of bead (l, P): bead 1 of the next particle
      sb = (P - 1) * N + (l - 1);
      sc = l - 1;
    }
    const double m = M[sa];
    const double dx = comp(X[sa] - X[sb], g.L0, g.image), ex = comp(X[sa] - X[sc], g.L0, g.image);
    const double dy = comp(X[nb + sa] - X[nb + sb], g.L1, g.image), ey = comp(X[nb + sa] - X[nb + sc], g.L1, g.image);
    const double dz = comp(X[2 * nb + sa] - X[2 * nb + sb], g.L2, g.image),
                 ez = comp(X[2 * nb + sa] - X[2 * nb + sc], g.L2, g.image);
    e += 0.5 * m * om * om * dx * ex;
    e += 0.5 * m * om * om * dy * ey;
    e += 0.5 * m * om * om * dz * ez;
  }
  e = block_sum(e, red);
  if (threadIdx.x == 0) out[0] = e;
}

// Log-sum-exp shifts of the recursions (md_pimd_x.c:688-808 forward,
// 2213-2335 backward, 4652-4663 at a second xi):
//  forward : min_k [-(k-1) ln xi + beta (E_n^(k) + V(n-k))],              k = 1..n
//  backward: min_l [-(l-u) ln xi + ln l + beta (E_l^(l-u+1) + V(l))],       l = u..N
// (xi = 0: the k = 1 / l = u term alone); the running minimum starts at 1e10
// as in the reference, so terms with a NaN logarithm never enter.
__global__ __launch_bounds__(256) void k_lse_shift(const double *__restrict__ ENk, const double *__restrict__ V, int N,
                                                  int n, int backward, double beta, double vi,
                                                  double *__restrict__ out) {
  __shared__ double red[256];
  double res = 1e10;
  if (vi == 0.0) {
    if (threadIdx.x == 0)
      res = backward ? log((double)n) + beta * (ENk[(size_t)(n - 1) * n / 2] + V[n])
                     : beta * (ENk[(size_t)(n - 1) * n / 2] + V[n - 1]);
  } else {
    const double lv = log(vi);
    const int cnt = backward ? N - n + 1 : n;
    for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
      double tmp;
      if (backward) {
        const int l = n + t, k = t + 1;
        tmp = -(k - 1) * lv + log((double)l) + beta * (ENk[(size_t)(l - 1) * l / 2 + k - 1] + V[l]);
      } else {
        const int k = t + 1;
        tmp = -(k - 1) * lv + beta * (ENk[(size_t)(n - 1) * n / 2 + k - 1] + V[n - k]);
      }
      if (tmp < res) res = tmp;
    }
  }
  res = block_min(res, red);
  if (threadIdx.x == 0) out[0] = res;
}

// Spring gradient of one bead (md_pimd_x.c:967-995, 2344-2367):
// m_a w^2 [(x_a - x_b) w_b + (x_a - x_c) w_c], w = 1 or 0 per link.
__global__ void k_spring_grad(const double *__restrict__ X, const double *__restrict__ M, int nb, int sa, int sb,
                              int sc, int wb, int wc, double om, Geo g, double *__restrict__ out) {
  const int d = threadIdx.x;
  if (d >= 3) return;
  const double L = d == 0 ? g.L0 : (d == 1 ? g.L1 : g.L2);
  const double m = M[sa];
  const double xa = X[d * nb + sa];
  double r = 0;
  if (wb) r += m * om * om * comp(xa - X[d * nb + sb], L, g.image);
  if (wc) r += m * om * om * comp(xa - X[d * nb + sc], L, g.image);
  out[d] = r;
}

}  // namespace xmd

using namespace xmd;

namespace {

// Reference bead index (l-1) P + (j-1) -> device slot (j-1) N + (l-1).
int ref_to_slot(int index, int N, int P) { return (index % P) * N + index / P; }

double fetch_scalar(DevPimd *p, int n, double *host) {
  double v[3] = {0, 0, 0};
  if (hipMemcpyAsync(v, p->work + 4 * (p->N + 1), n * sizeof(double), hipMemcpyDeviceToHost, rt().s0) !=
          hipSuccess ||
      hipStreamSynchronize(rt().s0) != hipSuccess)
    return NAN;
  if (host)
    for (int i = 0; i < n; ++i) host[i] = v[i];
  return v[0];
}

DevPimd *helper_pimd(md_pimd_t_x *pimd) {
  if (!pimd || ensure_pimd_device(pimd)) return nullptr;
  return pimd_state(pimd, false);
}

// V (N+1 host values) into device scratch after the helpers' output slot.
double *stage_host_vector(DevPimd *p, const double *V) {
  double *d = p->work + 4 * (p->N + 1) + 8;
  if (hipMemcpyAsync(d, V, (p->N + 1) * sizeof(double), hipMemcpyHostToDevice, rt().s0) != hipSuccess) return nullptr;
  return d;
}

}  // namespace

extern "C" {

int md_pimd_fast_fill_ENk_x(md_pimd_t_x *pimd, int image) {
  if (ensure_pimd_device(pimd)) return -1;
  return launch_enk(pimd_state(pimd, false), image, rt().s0);
}

// The cycle-walk algorithm (md_pimd_x.c:593-665) defines the same E_n^(k);
// the engine fills it with the closed form above.
int md_pimd_fill_ENk_x(md_pimd_t_x *pimd, int image) { return md_pimd_fast_fill_ENk_x(pimd, image); }

int md_pimd_fill_VB_noalloc_x(int N, md_num_t_x beta, md_num_t_x vi, const md_num_t_x *d_ENk, md_num_t_x *d_VBN,
                              md_num_t_x *d_Vt) {
  if (!rt().ready || !g_scratch[rt().device]) {
    fprintf(stderr, "xmd: md_pimd_fill_VB_noalloc_x needs the library set up first (md_init_setup_x)\n");
    return -1;
  }
  if (N < 1 || N > NA_NMAX) {
    fprintf(stderr, "xmd: md_pimd_fill_VB_noalloc_x supports 1 <= N <= %d (got %d)\n", NA_NMAX, N);
    return -1;
  }
  if (!(beta > 0) || !(vi >= 0) || !d_ENk || !d_VBN) {
    fprintf(stderr, "xmd: md_pimd_fill_VB_noalloc_x needs beta > 0, vi >= 0 and device buffers for ENk and VBN\n");
    return -1;
  }
  return noalloc_fill_vb(N, beta, vi, d_ENk, d_VBN, d_Vt);
}

int md_pimd_fill_VB_x(md_pimd_t_x *pimd) {
  if (ensure_pimd_device(pimd)) return -1;
  DevPimd *p = pimd_state(pimd, false);
  if (ensure_enk(p)) return -1;
  // forward and backward passes are independent: one launch, two workgroups
  return launch_recursion(p, true, true, rt().s0);
}

int md_pimd_fast_fill_force_VB_x(md_pimd_t_x *pimd, int image) {
  if (ensure_pimd_device(pimd)) return -1;
  DevPimd *p = pimd_state(pimd, false);
  if (ensure_enk(p) || ensure_vbn(p)) return -1;
  if (!p->back_dev || p->back_version != p->enk_version)
    if (launch_recursion(p, false, true, rt().s0)) return -1;
  md_simulation_t_x *sim = pimd->sim;
  double L[3] = {box_L(sim, 0), box_L(sim, 1), box_L(sim, 2)};
  if (launch_ring_springs(p, image, L, rt().s0)) return -1;
  return launch_exchange_force(p, image, rt().s0, false);
}

// The forward-mode gradient of md_pimd_x.c:997-1109 gives the same forces
// as the backward-pass formulation.
int md_pimd_fill_force_VB_x(md_pimd_t_x *pimd, int image) { return md_pimd_fast_fill_force_VB_x(pimd, image); }

int md_pimd_calc_VBN_energy_x(md_pimd_t_x *pimd, md_stats_t_x *stats) {
  if (ensure_pimd_device(pimd)) return -1;
  DevPimd *p = pimd_state(pimd, false);
  if (ensure_enk(p) || ensure_vbn(p)) return -1;
  double r = 0;
  if (run_estimator(p, p->ENk, &r)) return -1;
  stats->es[stats->N - 1] += r;
  stats->es[stats->N - 1] += MD_DIMENSION_X * pimd->N * pimd->P / (2 * pimd->beta);
  return 0;
}

int md_pimd_fast_fill_ENk2_x(md_pimd_t_x *pimd, int image) {
  if (ensure_pimd_device(pimd)) return -1;
  return fill_enk2(pimd_state(pimd, false), image, 0);
}

int md_pimd_fill_ENk2_x(md_pimd_t_x *pimd, int image) {
  if (ensure_pimd_device(pimd)) return -1;
  return fill_enk2(pimd_state(pimd, false), image, 1);
}

int md_pimd_calc_vi_virial_energy_x(md_pimd_t_x *pimd, md_stats_t_x *stats, int image) {
  if (ensure_pimd_device(pimd)) return -1;
  DevPimd *p = pimd_state(pimd, false);
  if (ensure_enk(p) || ensure_vbn(p)) return -1;
  if (!p->enk2_dev) {
    size_t tri = (size_t)p->N * (p->N + 1) / 2;
    XMD_CHECK(hipMemcpy(p->ENk2, pimd->ENk2[0], tri * sizeof(double), hipMemcpyHostToDevice));
    p->enk2_dev = true;
  }
  double r = 0;
  if (run_estimator(p, p->ENk2, &r)) return -1;
  DevSim *d = p->ds;
  int nblk = (int)std::min<long>((p->N * (long)(p->P - 1) + 255) / 256, 1024);
  if (nblk < 1) nblk = 1;
  hipLaunchKernelGGL(k_vir_term, dim3(nblk), dim3(256), 0, rt().s0, d->X, d->F, d->M, p->N, p->P, geo_of(p, image),
                     d->red);
  XMD_CHECK(hipGetLastError());
  double v = 0;
  if (device_sum(d->red, nblk, &v)) return -1;
  stats->es[stats->N - 1] += r;
  stats->es[stats->N - 1] += MD_DIMENSION_X * pimd->N / (2 * pimd->beta);
  stats->es[stats->N - 1] += v / 2.0;
  return 0;
}

// md_pimd_x.c:5452-5517 (pbc = 0): 2 res[N] / 3 + N / beta + (bead-1-relative
// virial) / 3 + (box-centred virial) / 3, with res the energy estimator's
// ENk2 recursion.
int md_pimd_calc_vi_virial_pressure_x(md_pimd_t_x *pimd, md_stats_t_x *stats, int image) {
  if (ensure_pimd_device(pimd)) return -1;
  DevPimd *p = pimd_state(pimd, false);
  if (!(pimd->sim->box && pimd->sim->box->type == MD_ND_RECT_BOX_X)) {
    fprintf(stderr, "xmd: the virial pressure needs a rectangular box\n");
    return -1;
  }
  if (ensure_enk(p) || ensure_vbn(p)) return -1;
  if (!p->enk2_dev) {
    size_t tri = (size_t)p->N * (p->N + 1) / 2;
    XMD_CHECK(hipMemcpy(p->ENk2, pimd->ENk2[0], tri * sizeof(double), hipMemcpyHostToDevice));
    p->enk2_dev = true;
  }
  double r = 0;
  if (run_estimator(p, p->ENk2, &r)) return -1;
  DevSim *d = p->ds;
  const Geo g = geo_of(p, image);
  int nblk = (int)std::min<long>((p->N * (long)(p->P - 1) + 255) / 256, 1024);
  if (nblk < 1) nblk = 1;
  hipLaunchKernelGGL(k_vir_term, dim3(nblk), dim3(256), 0, rt().s0, d->X, d->F, d->M, p->N, p->P, g, d->red);
  XMD_CHECK(hipGetLastError());
  double v = 0;
  if (device_sum(d->red, nblk, &v)) return -1;
  const int nb = p->N * p->P;
  int nblk2 = std::min((nb + 255) / 256, 1024);
  hipLaunchKernelGGL(k_box_virial, dim3(nblk2), dim3(256), 0, rt().s0, d->X, d->F, d->M, nb, g, d->red);
  XMD_CHECK(hipGetLastError());
  double w = 0;
  if (device_sum(d->red, nblk2, &w)) return -1;
  stats->es[stats->N - 1] += 2.0 * r / MD_DIMENSION_X;
  stats->es[stats->N - 1] += pimd->N / pimd->beta;
  stats->es[stats->N - 1] += v / MD_DIMENSION_X;
  stats->es[stats->N - 1] += w / MD_DIMENSION_X;
  return 0;
}

// Sign-reweighting factors for n values of xi' in one launch (one workgroup
// each): out[i] = sign_{xi'_i}(N) exp(-beta V_{|xi'_i|}(N) + beta V_B(N)).
static int vi_sign_grid(DevPimd *p, int n, const double *vi2, double *out) {
  if (n <= 0) return 0;
  if (2 * (p->N + 1) * sizeof(double) > 65536) {
    fprintf(stderr, "xmd: sign reweighting supports N <= 4094\n");
    return -1;
  }
  if (n > 512) {
    fprintf(stderr, "xmd: at most 512 xi' values per call\n");
    return -1;
  }
  RecArgs a = rec_args(p);
  double *d_vi = p->work + 4 * (p->N + 1) + 8, *d_out = d_vi + 512;
  hipStream_t st = rt().s0;
  XMD_CHECK(hipMemcpyAsync(d_vi, vi2, n * sizeof(double), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_vi_sign, dim3(n), dim3(256), 2 * (p->N + 1) * sizeof(double), st, a, d_vi, d_out);
  XMD_CHECK(hipGetLastError());
  XMD_CHECK(hipMemcpyAsync(out, d_out, n * sizeof(double), hipMemcpyDeviceToHost, st));
  XMD_CHECK(hipStreamSynchronize(st));
  return 0;
}

int md_pimd_calc_vi_sign_grid_x(md_pimd_t_x *pimd, int n, const md_num_t_x *vi2, md_num_t_x *factor) {
  if (!vi2 || !factor || ensure_pimd_device(pimd)) return -1;
  DevPimd *p = pimd_state(pimd, false);
  if (ensure_enk(p) || ensure_vbn(p)) return -1;
  return vi_sign_grid(p, n, vi2, factor);
}

int md_pimd_calc_vi_sign_x(md_pimd_t_x *pimd, md_stats_t_x *e_s, md_stats_t_x *s_s, md_num_t_x vi2) {
  if (ensure_pimd_device(pimd)) return -1;
  DevPimd *p = pimd_state(pimd, false);
  if (ensure_enk(p) || ensure_vbn(p)) return -1;
  double s = 0;
  if (vi_sign_grid(p, 1, &vi2, &s)) return -1;
  e_s->es[0] *= s;
  s_s->es[0] *= s;
  return 0;
}

// Single-value helpers of the exchange path (md_pimd_x.h:121,125,128,137,139,156,161,171),
// evaluated on the device state.  Particle indices are the reference's
// (l - 1) P + (j - 1).  NaN is returned when the device call fails.
md_num_t_x md_pimd_ENk_x(md_pimd_t_x *pimd, int N2, int k, int image) {
  DevPimd *p = helper_pimd(pimd);
  if (!p || N2 < 1 || N2 > p->N || k < 1 || k > N2) return NAN;
  hipLaunchKernelGGL(k_cycle_energy, dim3(1), dim3(256), 0, rt().s0, p->ds->X, p->ds->M, p->N, p->P, N2, k,
                     pimd->omegaP, geo_of(p, image), p->work + 4 * (p->N + 1));
  if (hipGetLastError() != hipSuccess) return NAN;
  return fetch_scalar(p, 1, nullptr);
}

md_num_t_x md_pimd_ENk2_x(md_pimd_t_x *pimd, int N2, int k, int image) {
  DevPimd *p = helper_pimd(pimd);
  if (!p || N2 < 1 || N2 > p->N || k < 1 || k > N2) return NAN;
  hipLaunchKernelGGL(k_cycle_virial, dim3(1), dim3(256), 0, rt().s0, p->ds->X, p->ds->M, p->N, p->P, N2, k, -1, -1,
                     -1, pimd->omegaP, geo_of(p, image), p->work + 4 * (p->N + 1));
  if (hipGetLastError() != hipSuccess) return NAN;
  return fetch_scalar(p, 1, nullptr);
}

md_num_t_x md_pimd_fast_ENk2_x(md_pimd_t_x *pimd, int index, int index2, int index3, int image) {
  DevPimd *p = helper_pimd(pimd);
  const int nb = p ? p->N * p->P : 0;
  if (!p || index < 0 || index >= nb || index2 < 0 || index2 >= nb || index3 < 0 || index3 >= nb) return NAN;
  if (pimd->pbc) {
    fprintf(stderr, "xmd: winding-number (pbc) cycle terms are not part of this engine\n");
    return NAN;
  }
  hipLaunchKernelGGL(k_cycle_virial, dim3(1), dim3(64), 0, rt().s0, p->ds->X, p->ds->M, p->N, p->P, 1, 1,
                     ref_to_slot(index, p->N, p->P), ref_to_slot(index2, p->N, p->P),
                     ref_to_slot(index3, p->N, p->P), pimd->omegaP, geo_of(p, image), p->work + 4 * (p->N + 1));
  if (hipGetLastError() != hipSuccess) return NAN;
  return fetch_scalar(p, 1, nullptr);
}

md_num_t_x md_pimd_xminE_x(md_pimd_t_x *pimd, int N2) {
  DevPimd *p = helper_pimd(pimd);
  if (!p || N2 < 1 || N2 > p->N || ensure_enk(p) || ensure_vbn(p)) return NAN;
  hipLaunchKernelGGL(k_lse_shift, dim3(1), dim3(256), 0, rt().s0, p->ENk, p->VBN, p->N, N2, 0, pimd->beta, pimd->vi,
                     p->work + 4 * (p->N + 1));
  if (hipGetLastError() != hipSuccess) return NAN;
  return fetch_scalar(p, 1, nullptr);
}

md_num_t_x md_pimd_xminE2_x(md_pimd_t_x *pimd, int N2, md_num_t_x *VBN2, md_num_t_x vi2) {
  DevPimd *p = helper_pimd(pimd);
  if (!p || !VBN2 || N2 < 1 || N2 > p->N || ensure_enk(p)) return NAN;
  double *V = stage_host_vector(p, VBN2);
  if (!V) return NAN;
  hipLaunchKernelGGL(k_lse_shift, dim3(1), dim3(256), 0, rt().s0, p->ENk, V, p->N, N2, 0, pimd->beta, vi2,
                     p->work + 4 * (p->N + 1));
  if (hipGetLastError() != hipSuccess) return NAN;
  return fetch_scalar(p, 1, nullptr);
}

md_num_t_x md_pimd_fast_xminE_x(md_pimd_t_x *pimd, int u, md_num_t_x *V) {
  DevPimd *p = helper_pimd(pimd);
  if (!p || !V || u < 1 || u > p->N || ensure_enk(p)) return NAN;
  double *Vd = stage_host_vector(p, V);
  if (!Vd) return NAN;
  hipLaunchKernelGGL(k_lse_shift, dim3(1), dim3(256), 0, rt().s0, p->ENk, Vd, p->N, u, 1, pimd->beta, pimd->vi,
                     p->work + 4 * (p->N + 1));
  if (hipGetLastError() != hipSuccess) return NAN;
  return fetch_scalar(p, 1, nullptr);
}

void md_pimd_dENk_x(md_pimd_t_x *pimd, int N2, int k, int l, int j, int image, md_num_t_x *dENk) {
  for (int i = 0; i < MD_DIMENSION_X; ++i) dENk[i] = 0;
  DevPimd *p = helper_pimd(pimd);
  if (!p || !(l >= N2 - k + 1 && l <= N2) || j < 1 || j > p->P || N2 > p->N) return;
  const int P = p->P;
  const int a = MD_PIMD_INDEX_X(l, j, P), b = md_pimd_next_index_x(l, j, N2, k, P),
            c = md_pimd_prev_index_x(l, j, N2, k, P);
  hipLaunchKernelGGL(k_spring_grad, dim3(1), dim3(64), 0, rt().s0, p->ds->X, p->ds->M, p->ds->nb,
                     ref_to_slot(a, p->N, P), ref_to_slot(b, p->N, P), ref_to_slot(c, p->N, P), 1, 1, pimd->omegaP,
                     geo_of(p, image), p->work + 4 * (p->N + 1));
  if (hipGetLastError() != hipSuccess) return;
  fetch_scalar(p, 3, dENk);
}

void md_pimd_fast_dENk_x(md_pimd_t_x *pimd, int index, int index2, int index3, int image, md_num_t_x *dENk) {
  for (int i = 0; i < MD_DIMENSION_X; ++i) dENk[i] = 0;
  DevPimd *p = helper_pimd(pimd);
  const int nb = p ? p->N * p->P : 0;
  if (!p || index < 0 || index >= nb || index2 < 0 || index2 >= nb || index3 < 0 || index3 >= nb) return;
  const int P = p->P;
  // a link that is cut at slice cut_Pj is dropped (md_pimd_x.c:2354-2357)
  const int wb = (index + 1 != index2 || (index % P) + 1 != pimd->cut_Pj) ? 1 : 0;
  const int wc = (index - 1 != index3 || (index % P) + 1 != pimd->cut_Pj + 1) ? 1 : 0;
  hipLaunchKernelGGL(k_spring_grad, dim3(1), dim3(64), 0, rt().s0, p->ds->X, p->ds->M, p->ds->nb,
                     ref_to_slot(index, p->N, P), ref_to_slot(index2, p->N, P), ref_to_slot(index3, p->N, P), wb, wc,
                     pimd->omegaP, geo_of(p, image), p->work + 4 * (p->N + 1));
  if (hipGetLastError() != hipSuccess) return;
  fetch_scalar(p, 3, dENk);
}

}  // extern "C"
