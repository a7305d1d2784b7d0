// Broad-phase building blocks shared by the collide pipeline and the standalone broad phases: the
// sortable float key, the collision-group and filter-pair tests of broad_phase_common.py, lane groups and
// an in-LDS bitonic sort.
#pragma once
#include <hip/hip_runtime.h>

namespace nmc {

constexpr int CBLOCK = 256;
constexpr int BP_EXPLICIT_ITEMS = 4;  // explicit broad phase: input pairs per lane per block

__device__ __forceinline__ unsigned sortable_float(float f) {
  if (f == 0.0f) f = 0.0f;  // -0 and +0 compare equal: give them one key
  unsigned u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ bool group_pair(int a, int b) {
  if (a == 0 || b == 0) return false;
  if (a > 0) return a == b || b < 0;
  return a != b;
}

// First filter-pair entry whose first shape is >= a (entries are sorted (a, b) pairs with a < b).
__device__ __forceinline__ int filter_lower_bound(const int* __restrict__ fp, int n, int a) {
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (fp[2 * mid] < a) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// is_pair_excluded (broad_phase_common.py:82-113) over the entries [e0, e1) whose first shape is a.
__device__ __forceinline__ bool excluded(int b, const int* __restrict__ fp, int e0, int e1) {
  for (int e = e0; e < e1; ++e) {
    int mb = fp[2 * e + 1];
    if (mb == b) return true;
    if (mb > b) break;
  }
  return false;
}

// is_pair_excluded (broad_phase_common.py:82-113): binary search of the canonical pair (a, b) in a list of
// n lexicographically sorted (min, max) pairs.
__device__ __forceinline__ bool pair_listed(const int* __restrict__ fp, int n, int a, int b) {
  int lo = 0, hi = n - 1;
  while (lo <= hi) {
    const int mid = (lo + hi) >> 1;
    const int ma = fp[2 * mid], mb = fp[2 * mid + 1];
    if (a == ma && b == mb) return true;
    if (a < ma || (a == ma && b < mb)) hi = mid - 1; else lo = mid + 1;
  }
  return false;
}

// Lane groups that run one segment: the whole workgroup (barrier) or one wave (wave-local ordering of
// its LDS accesses; a wave's LDS operations complete in issue order).
struct BlockGroup {
  int lane;
  static constexpr int size = CBLOCK;
  __device__ __forceinline__ void sync() const { __syncthreads(); }
};
struct WaveGroup {
  int lane;
  static constexpr int size = 64;
  __device__ __forceinline__ void sync() const {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
};

// In-LDS bitonic sort of n2 (power of two) keys, ascending.
template <class K, class G = BlockGroup>
__device__ void bitonic_sort(K* keys, int n2, G g = G{static_cast<int>(threadIdx.x)}) {
  for (int k = 2; k <= n2; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = g.lane; i < n2; i += G::size) {
        int ixj = i ^ j;
        if (ixj > i) {
          K a = keys[i], b = keys[ixj];
          bool up = (i & k) == 0;
          if ((a > b) == up) {
            keys[i] = b;
            keys[ixj] = a;
          }
        }
      }
      g.sync();
    }
  }
}

}  // namespace nmc
