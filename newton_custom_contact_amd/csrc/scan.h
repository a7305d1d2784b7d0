// One-workgroup exclusive scan over per-segment counts, shared by the collide, broad-phase and
// contact-match kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace nmc {

// Exclusive scans of N per-segment count arrays by one 1024-lane workgroup.  Each lane scans 8
// consecutive entries, so one pass covers 8192 segments (larger counts take several passes).
//
// segment_scan_fn: load(i, v) fills v[0..N) with the counts of segment i (i < n); emit(i, ex, v) receives
// the exclusive prefixes ex[0..N) of segment i with its counts; the totals come back in total[].  Doing the
// per-segment follow-up work in emit saves the second pass over global memory a separate loop would take.
constexpr int SCAN_BLOCK = 1024;
constexpr int SCAN_ITEMS = 8;
template <int N, class Load, class Emit>
__device__ void segment_scan_fn(int n, Load load, Emit emit, int (&total)[N]) {
  __shared__ int s_w[N][SCAN_BLOCK / 64];
  __shared__ int s_carry[N];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (threadIdx.x < N) s_carry[threadIdx.x] = 0;
  __syncthreads();
  for (int base = 0; base < n; base += SCAN_BLOCK * SCAN_ITEMS) {
    const int i0 = base + threadIdx.x * SCAN_ITEMS;
    int v[SCAN_ITEMS][N], sum[N], inc[N];
#pragma unroll
    for (int k = 0; k < N; ++k) sum[k] = 0;
#pragma unroll
    for (int t = 0; t < SCAN_ITEMS; ++t) {
#pragma unroll
      for (int k = 0; k < N; ++k) v[t][k] = 0;
      if (i0 + t < n) load(i0 + t, v[t]);
#pragma unroll
      for (int k = 0; k < N; ++k) sum[k] += v[t][k];
    }
#pragma unroll
    for (int k = 0; k < N; ++k) inc[k] = sum[k];
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
    int c[N], ex[N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
      c[k] = s_carry[k];
      for (int q = 0; q < w; ++q) c[k] += s_w[k][q];
      ex[k] = c[k] + inc[k] - sum[k];
    }
#pragma unroll
    for (int t = 0; t < SCAN_ITEMS; ++t) {
      if (i0 + t < n) emit(i0 + t, ex, v[t]);
#pragma unroll
      for (int k = 0; k < N; ++k) ex[k] += v[t][k];
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

// Plain form: N input arrays (a null one scans as zeros) into N output arrays (a null one is skipped).
template <int N>
__device__ void segment_scan(const int* const (&in)[N], int* const (&out)[N], int n, int (&total)[N]) {
  segment_scan_fn<N>(
      n,
      [&](int i, int (&v)[N]) {
#pragma unroll
        for (int k = 0; k < N; ++k) v[k] = in[k] ? in[k][i] : 0;
      },
      [&](int i, const int (&ex)[N], const int (&)[N]) {
#pragma unroll
        for (int k = 0; k < N; ++k)
          if (out[k]) out[k][i] = ex[k];
      },
      total);
}

}  // namespace nmc
