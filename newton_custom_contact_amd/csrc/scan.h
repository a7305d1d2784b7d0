// One-workgroup exclusive scan over per-segment counts, shared by the collide and contact-match kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace nmc {

// Exclusive scans of N per-segment count arrays by one 1024-lane workgroup; writes the totals.  A null
// input array scans as zeros, a null output array is skipped.  Each lane scans 4 consecutive entries, so
// a pass covers 4096 segments.
constexpr int SCAN_BLOCK = 1024;
constexpr int SCAN_ITEMS = 4;
template <int N>
__device__ void segment_scan(const int* const (&in)[N], int* const (&out)[N], int n, int (&total)[N]) {
  __shared__ int s_w[N][SCAN_BLOCK / 64];
  __shared__ int s_carry[N];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (threadIdx.x < N) s_carry[threadIdx.x] = 0;
  __syncthreads();
  for (int base = 0; base < n; base += SCAN_BLOCK * SCAN_ITEMS) {
    const int i0 = base + threadIdx.x * SCAN_ITEMS;
    int v[N][SCAN_ITEMS], sum[N], inc[N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
      sum[k] = 0;
#pragma unroll
      for (int t = 0; t < SCAN_ITEMS; ++t) {
        v[k][t] = (in[k] && i0 + t < n) ? in[k][i0 + t] : 0;
        sum[k] += v[k][t];
      }
      inc[k] = sum[k];
    }
    for (int off = 1; off < 64; off <<= 1) {
#pragma unroll
      for (int k = 0; k < N; ++k) {
        int t = __shfl_up(inc[k], off, 64);
        if (lane >= off) inc[k] += t;
      }
    }
    if (lane == 63) {
#pragma unroll
      for (int k = 0; k < N; ++k) s_w[k][w] = inc[k];
    }
    __syncthreads();
    int c[N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
      c[k] = s_carry[k];
      for (int q = 0; q < w; ++q) c[k] += s_w[k][q];
      int ex = c[k] + inc[k] - sum[k];
      if (out[k]) {
#pragma unroll
        for (int t = 0; t < SCAN_ITEMS; ++t) {
          if (i0 + t < n) out[k][i0 + t] = ex;
          ex += v[k][t];
        }
      }
    }
    __syncthreads();
    if (threadIdx.x == SCAN_BLOCK - 1) {
#pragma unroll
      for (int k = 0; k < N; ++k) s_carry[k] = c[k] + inc[k];
    }
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < N; ++k) total[k] = s_carry[k];
}

}  // namespace nmc
