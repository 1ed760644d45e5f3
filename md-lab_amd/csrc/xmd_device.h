// Device helpers shared by the engine's kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace xmd {

// md_simulation_x.c:858-865: image count with the reference's truncation.
__device__ __forceinline__ int dv_img_count(double x, double L) {
  int c = (int)(fabs(x) / L);
  if (x > L) return -c;
  if (x < 0) return c + 1;
  return 0;
}

// md_simulation_x.c:70-81
__device__ __forceinline__ double dv_min_image(double a, double L) {
  double s = a + L / 2;
  s += L * dv_img_count(s, L);
  return s - L / 2;
}

// Batched launches cover many independent systems ("replicas") with one
// grid: replica r owns the blocks start[r] <= b < start[r+1] (start holds
// nrep + 1 prefix sums).  Returns r for block b < start[nrep].
__device__ __forceinline__ int batch_rep(const int *__restrict__ start, int nrep, int b) {
  int lo = 0, hi = nrep - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= b)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Broadcast lane `lane` (wave-uniform) of a double through two scalar reads.
__device__ __forceinline__ double readlane_d(double v, int lane) {
  int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Double-double arithmetic (Knuth TwoSum / Dekker) for prefix sums whose
// differences must keep full double precision.
struct dd {
  double hi, lo;
};

__device__ __forceinline__ dd dd_add(dd a, dd b) {
  double s = a.hi + b.hi;
  double bb = s - a.hi;
  double e = (a.hi - (s - bb)) + (b.hi - bb);
  e += a.lo + b.lo;
  double h = s + e;
  return dd{h, e - (h - s)};
}

__device__ __forceinline__ double dd_sub_to_double(dd a, dd b) {
  double s = a.hi - b.hi;
  double bb = s - a.hi;
  double e = (a.hi - (s - bb)) + (-b.hi - bb);
  return s + (e + (a.lo - b.lo));
}

}  // namespace xmd
