// CPU oracle -- TEST INFRASTRUCTURE ONLY.  Frame-to-frame contact matching restatement.
//
// Follows newton/_src/geometry/contact_match.py step by step:
//   _lower_bound_int64 / pair range          :130-150, :280-292
//   closest midpoint within thresholds       :238-324 (world midpoint 0.5 (world(p0) + world(p1)),
//                                             `dist_sq <= best and ndot >= threshold`, ascending prev index)
//   packed claim + atomic_min                :160-214, :314-320 (float_flip(dist_sq) << 32 | key low 32 bits)
//   resolve (winner keeps, losers MATCH_BROKEN) :341-370
//   sticky replay of matched rows            :497-520
//   report (new: match < 0; broken: prev not won) :536-566
//   save sorted state (keys, midpoints, normals, sticky columns) :432-480
// The previous frame is kept in the reference's layout: sorted by the 64-bit key over the whole buffer.
#pragma once
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

#include "oracle_collide.h"

namespace oracle {

struct MatcherState {
  float pos_threshold_sq = 0.0f, normal_dot_threshold = 0.0f;
  bool sticky = false;
  std::vector<int64_t> keys;  // previous frame, key-sorted
  std::vector<V3<float>> pos, normal, p0, p1, o0, o1;
};

inline uint32_t float_flip(float f) {
  uint32_t i;
  std::memcpy(&i, &f, 4);
  uint32_t mask = (uint32_t)(-(int32_t)(i >> 31)) | 0x80000000u;
  return i ^ mask;
}

inline int64_t pack_claim(float dist_sq, int64_t key) {
  return (int64_t(float_flip(dist_sq)) << 32) | (key & int64_t(0xFFFFFFFF));
}

inline int lower_bound_keys(const std::vector<int64_t>& keys, int lo, int hi, int64_t target) {
  while (lo < hi) {
    int mid = lo + (hi - lo) / 2;
    if (keys[mid] < target) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// One frame.  Contacts are given key-sorted (the deterministic pipeline's order); arrays are updated in
// place by the sticky replay.  Outputs: match_index (into the previous key-sorted buffer, or -1 / -2),
// matched_key (the previous contact's key, or -1), broken_keys (previous contacts no one won, ascending).
inline void contact_match(MatcherState& S, const OracleModel* m, const float* body_q, int n, const int64_t* keys,
                          const int* shape0, const int* shape1, float* point0, float* point1, float* offset0,
                          float* offset1, float* normal, int* match_index, int64_t* matched_key,
                          std::vector<int64_t>& broken_keys) {
  auto wpoint = [&](int shape, const float* p) {
    V3<float> v{p[0], p[1], p[2]};
    int b = m->shape_body[shape];
    if (b == -1) return v;
    Xf<float> X = load_xf_t<float, float>(body_q + 7 * b);
    return xf_point(X, v);
  };
  const int n_old = int(S.keys.size());
  std::vector<int64_t> claim(n_old, INT64_MAX);
  std::vector<float> best_d2(n);
  for (int t = 0; t < n; ++t) {
    match_index[t] = -1;
    if (n_old == 0) continue;
    const int64_t key = keys[t];
    const int64_t prefix = key & ~int64_t(0x7FFFFF);
    const int lo = lower_bound_keys(S.keys, 0, n_old, prefix);
    const int hi = lower_bound_keys(S.keys, lo, n_old, prefix + int64_t(0x800000));
    if (lo >= hi) continue;
    const V3<float> pw = float(0.5) * (wpoint(shape0[t], point0 + 3 * t) + wpoint(shape1[t], point1 + 3 * t));
    const V3<float> nn{normal[3 * t], normal[3 * t + 1], normal[3 * t + 2]};
    int best = -1;
    float bd = S.pos_threshold_sq;
    for (int j = lo; j < hi; ++j) {
      V3<float> diff = pw - S.pos[j];
      float d2 = dot(diff, diff);
      float nd = dot(nn, S.normal[j]);
      if (d2 <= bd && nd >= S.normal_dot_threshold) {
        bd = d2;
        best = j;
      }
    }
    if (best >= 0) {
      match_index[t] = best;
      best_d2[t] = bd;
      claim[best] = std::min(claim[best], pack_claim(bd, key));
    } else {
      match_index[t] = -2;
    }
  }
  std::vector<char> won(n_old, 0);
  for (int t = 0; t < n; ++t) {
    const int c = match_index[t];
    matched_key[t] = -1;
    if (c < 0) continue;
    if ((claim[c] & int64_t(0xFFFFFFFF)) == (keys[t] & int64_t(0xFFFFFFFF))) {
      won[c] = 1;
      matched_key[t] = S.keys[c];
    } else {
      match_index[t] = -2;
    }
  }
  broken_keys.clear();
  for (int j = 0; j < n_old; ++j)
    if (!won[j]) broken_keys.push_back(S.keys[j]);
  if (S.sticky) {
    for (int t = 0; t < n; ++t) {
      const int c = match_index[t];
      if (c < 0) continue;
      const V3<float>* src[5] = {&S.p0[c], &S.p1[c], &S.o0[c], &S.o1[c], &S.normal[c]};
      float* dst[5] = {point0, point1, offset0, offset1, normal};
      for (int k = 0; k < 5; ++k) {
        dst[k][3 * t] = src[k]->x;
        dst[k][3 * t + 1] = src[k]->y;
        dst[k][3 * t + 2] = src[k]->z;
      }
    }
  }
  // save this frame (already key-sorted)
  S.keys.assign(keys, keys + n);
  S.pos.resize(n);
  S.normal.resize(n);
  if (S.sticky) {
    S.p0.resize(n);
    S.p1.resize(n);
    S.o0.resize(n);
    S.o1.resize(n);
  }
  for (int t = 0; t < n; ++t) {
    S.pos[t] = float(0.5) * (wpoint(shape0[t], point0 + 3 * t) + wpoint(shape1[t], point1 + 3 * t));
    S.normal[t] = V3<float>{normal[3 * t], normal[3 * t + 1], normal[3 * t + 2]};
    if (S.sticky) {
      S.p0[t] = V3<float>{point0[3 * t], point0[3 * t + 1], point0[3 * t + 2]};
      S.p1[t] = V3<float>{point1[3 * t], point1[3 * t + 1], point1[3 * t + 2]};
      S.o0[t] = V3<float>{offset0[3 * t], offset0[3 * t + 1], offset0[3 * t + 2]};
      S.o1[t] = V3<float>{offset1[3 * t], offset1[3 * t + 1], offset1[3 * t + 2]};
    }
  }
}

}  // namespace oracle
