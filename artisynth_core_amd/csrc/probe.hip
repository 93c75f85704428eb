// HBM streaming probe (measurement only; not part of the solver's C ABI):
// the achievable read rate of one launch on this GPU, the ceiling the solve's
// bandwidth-bound kernels are compared with.  Each lane keeps G 16-byte loads
// in flight (all issued before the first use), consecutive lanes read
// consecutive 16-byte chunks, and the grid covers the buffer once (no
// grid-stride loop), so a launch is one pass of streaming reads.  The lane's
// sum goes to one 8-byte slot per workgroup so the loads are not eliminated.
#include <hip/hip_runtime.h>

template <int G>
__global__ __launch_bounds__(256) void k_stream_read(const float4* __restrict__ src, long long nchunks,
                                                     float* __restrict__ sink) {
  const long long base = ((long long)blockIdx.x * 256 * G) + threadIdx.x;
  float4 v[G];
#pragma unroll
  for (int q = 0; q < G; q++) {
    const long long c = base + (long long)q * 256;
    v[q] = src[c < nchunks ? c : 0];
  }
  float acc = 0.f;
#pragma unroll
  for (int q = 0; q < G; q++) acc += (v[q].x + v[q].y) + (v[q].z + v[q].w);
  __shared__ float sm[256];
  sm[threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.f;
    for (int i = 0; i < 256; i++) t += sm[i];
    sink[blockIdx.x] = t;
  }
}

extern "C" {
// Reads `bytes` (a multiple of 16) from src once with `inflight` (1, 2, 4, 8 or
// 16) loads in flight per lane; sink must hold bytes / (4096 * inflight) + 1
// floats.  Returns 0, or -1 on a bad argument or launch failure.
int kkt_probe_stream_read(const void* src, long long bytes, int inflight, float* sink, hipStream_t st) {
  if (!src || !sink || bytes <= 0 || (bytes & 15)) return -1;
  const long long nchunks = bytes / 16;
  const long long per = 256LL * inflight;
  const unsigned grid = (unsigned)((nchunks + per - 1) / per);
  const float4* s = (const float4*)src;
  switch (inflight) {
    case 1: hipLaunchKernelGGL(k_stream_read<1>, dim3(grid), dim3(256), 0, st, s, nchunks, sink); break;
    case 2: hipLaunchKernelGGL(k_stream_read<2>, dim3(grid), dim3(256), 0, st, s, nchunks, sink); break;
    case 4: hipLaunchKernelGGL(k_stream_read<4>, dim3(grid), dim3(256), 0, st, s, nchunks, sink); break;
    case 8: hipLaunchKernelGGL(k_stream_read<8>, dim3(grid), dim3(256), 0, st, s, nchunks, sink); break;
    case 16: hipLaunchKernelGGL(k_stream_read<16>, dim3(grid), dim3(256), 0, st, s, nchunks, sink); break;
    default: return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
}
