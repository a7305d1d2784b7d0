// CPU oracle -- TEST INFRASTRUCTURE ONLY.  Standalone broad phases over given AABBs, restated from the
// reference's kernels as one sequential pass each:
//   all pairs  newton/_src/geometry/broad_phase_nxn.py:68-214 (lower-triangular pair index per world,
//              world / group tests, check_aabb_overlap with per-shape gaps, filter list)
//   SAP        newton/_src/geometry/broad_phase_sap.py:43-394, 510-677 (projection on the fixed axis with the
//              gap added to the half extents, stable sort of the lower bounds per segment, range = first
//              sorted lower bound >= upper (binary search), sweep over the range with the same tests)
//   explicit   broad_phase_nxn.py:28-65 (AABB overlap of each given pair)
//   common     broad_phase_common.py:18-180
// Pairs are appended in the order a single thread visits them.
#pragma once
#include <algorithm>
#include <cstdint>
#include <vector>

#include "oracle_collide.h"

namespace oracle {

struct BroadInput {
  const float* lower;  // S x 3
  const float* upper;
  const float* gap;    // S or null
  const int* group;
  const int* world;
  const int* filter;   // F x 2 sorted, or null
  int filter_count;
};

inline bool broad_overlap(const BroadInput& in, int a, int b) {
  const float ga = in.gap ? in.gap[a] : 0.0f, gb = in.gap ? in.gap[b] : 0.0f;
  const float c = ga + gb;
  const float* la = in.lower + 3 * a;
  const float* ua = in.upper + 3 * a;
  const float* lb = in.lower + 3 * b;
  const float* ub = in.upper + 3 * b;
  return la[0] <= ub[0] + c && ua[0] >= lb[0] - c && la[1] <= ub[1] + c && ua[1] >= lb[1] - c &&
         la[2] <= ub[2] + c && ua[2] >= lb[2] - c;
}

inline void broad_push(std::vector<int>& out, int a, int b) {
  out.push_back(a);
  out.push_back(b);
}

// _nxn_broadphase_kernel for every thread id in order.
inline void broad_all_pairs(const BroadInput& in, const int* index_map, const int* slice_ends, int segments,
                            std::vector<int>& out) {
  const int regular = segments - 1;
  for (int w = 0; w < segments; ++w) {
    const int start = w > 0 ? slice_ends[w - 1] : 0;
    const int n = slice_ends[w] - start;
    for (int r = 0; r < n; ++r) {
      for (int c = r + 1; c < n; ++c) {
        const int t1 = index_map[start + r], t2 = index_map[start + c];
        const int a = std::min(t1, t2), b = std::max(t1, t2);
        const int w1 = in.world[a], w2 = in.world[b];
        if (w1 == -1 && w2 == -1 && !(w >= regular)) continue;
        if (!test_world_and_group_pair(w1, w2, in.group[a], in.group[b])) continue;
        if (!broad_overlap(in, a, b)) continue;
        if (in.filter_count > 0 && is_pair_excluded(a, b, in.filter, in.filter_count)) continue;
        broad_push(out, a, b);
      }
    }
  }
}

inline void broad_sap(const BroadInput& in, const int* index_map, const int* slice_ends, int segments,
                      std::vector<int>& out) {
  const V3<float> dir = normalize(V3<float>{0.5935f, 0.7790f, 0.1235f});
  const int regular = segments - 1;
  for (int w = 0; w < segments; ++w) {
    const int start = w > 0 ? slice_ends[w - 1] : 0;
    const int n = slice_ends[w] - start;
    std::vector<float> plo(n), phi(n);
    std::vector<int> order(n);
    for (int l = 0; l < n; ++l) {
      const int s = index_map[start + l];
      const V3<float> lo{in.lower[3 * s], in.lower[3 * s + 1], in.lower[3 * s + 2]};
      const V3<float> up{in.upper[3 * s], in.upper[3 * s + 1], in.upper[3 * s + 2]};
      const float g = in.gap ? in.gap[s] : 0.0f;
      V3<float> half = 0.5f * (up - lo);
      half = V3<float>{half.x + g, half.y + g, half.z + g};
      const float radius = dot(dir, half);
      const float center = dot(dir, 0.5f * (lo + up));
      plo[l] = center - radius;
      phi[l] = center + radius;
      order[l] = l;
    }
    std::vector<uint64_t> okey(n);
    for (int l = 0; l < n; ++l) okey[l] = sap_order_key(plo[l]);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return okey[a] < okey[b]; });
    std::vector<float> sorted(n);
    for (int i = 0; i < n; ++i) sorted[i] = plo[order[i]];
    for (int i = 0; i < n; ++i) {
      const float upper = phi[order[i]];
      // binary_search_segment over [i + 1, n): first sorted lower bound that is not below upper
      int lo = i + 1, hi = n;
      while (lo < hi) {
        const int mid = (lo + hi) / 2;
        if (sorted[mid] < upper) lo = mid + 1; else hi = mid;
      }
      for (int j = i + 1; j < lo; ++j) {
        const int t1 = index_map[start + order[i]], t2 = index_map[start + order[j]];
        if (t1 == t2) continue;
        const int a = std::min(t1, t2), b = std::max(t1, t2);
        const int w1 = in.world[a], w2 = in.world[b];
        if (w1 == -1 && w2 == -1 && !(w >= regular)) continue;
        if (!test_world_and_group_pair(w1, w2, in.group[a], in.group[b])) continue;
        if (in.filter_count > 0 && is_pair_excluded(a, b, in.filter, in.filter_count)) continue;
        if (!broad_overlap(in, a, b)) continue;
        broad_push(out, a, b);
      }
    }
  }
}

inline void broad_explicit(const BroadInput& in, const int* pairs, int pair_count, std::vector<int>& out) {
  for (int e = 0; e < pair_count; ++e) {
    const int a = pairs[2 * e], b = pairs[2 * e + 1];
    if (broad_overlap(in, a, b)) broad_push(out, a, b);
  }
}

}  // namespace oracle
