// Wave-level (64-lane) reductions and uniform-value helpers for gfx950.
#pragma once
#include <hip/hip_runtime.h>

namespace kkt {

// ------------------------------------------------------------------------
// wave helpers (64-wide)
// ------------------------------------------------------------------------
__device__ inline double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}
template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
// Lane 0's value in every lane.  Applied to values that are equal across the
// wave (LDS broadcasts, wave max) so the compiler sees uniform branch and loop
// conditions and emits scalar control flow around barriers.
__device__ inline int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline double uni(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)b);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ inline double readlane_d(double v, int lane) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = __builtin_amdgcn_readlane((unsigned)b, lane);
  const unsigned hi = __builtin_amdgcn_readlane((unsigned)(b >> 32), lane);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
// Row-scan reductions with DPP (GFX9 controls: row_shr:n = 0x110+n, row_bcast:15
// = 0x142 on rows 1 and 3, row_bcast:31 = 0x143 on rows 2 and 3).  A lane with no
// source keeps its own value; after the six steps lane 63 holds the result.
template <int CTRL, int RMASK>
__device__ inline double dpp_d(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const int lo = (int)(unsigned)b, hi = (int)(unsigned)(b >> 32);
  const unsigned plo = (unsigned)__builtin_amdgcn_update_dpp(lo, lo, CTRL, RMASK, 0xF, false);
  const unsigned phi = (unsigned)__builtin_amdgcn_update_dpp(hi, hi, CTRL, RMASK, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)phi << 32) | plo));
}
__device__ inline double wave_max_u(double v) {
  v = fmax(v, dpp_d<0x111, 0xF>(v));
  v = fmax(v, dpp_d<0x112, 0xF>(v));
  v = fmax(v, dpp_d<0x114, 0xF>(v));
  v = fmax(v, dpp_d<0x118, 0xF>(v));
  v = fmax(v, dpp_d<0x142, 0xA>(v));
  v = fmax(v, dpp_d<0x143, 0xC>(v));
  return readlane_d(v, 63);
}
// largest value and the lowest lane holding it; uniform result
__device__ inline void wave_argmax_u(double& v, int& i) {
  const double mx = wave_max_u(v);
  const unsigned long long hit = __ballot(v == mx);
  v = mx;
  i = __ffsll((unsigned long long)hit) - 1;
}


}  // namespace kkt
