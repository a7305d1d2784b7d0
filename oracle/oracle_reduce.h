// CPU oracle -- TEST INFRASTRUCTURE ONLY.  Restatement of the reference's mesh contact reduction.
//
// Follows:
//   constants, face normals, get_slot, project_point_to_plane, get_spatial_direction_2d,
//   compute_voxel_index, float_flip             newton/_src/geometry/contact_reduction.py:60-407
//   key / value packing (fast, deterministic)   newton/_src/geometry/contact_reduction_global.py:232-432
//   encode_oct / decode_oct                     newton/_src/geometry/contact_reduction_global.py:435-480
//   reduce_contact_in_hashtable (mesh-plane)    newton/_src/geometry/contact_reduction_global.py:968-1059
//   export_and_reduce_contact_centered (mesh-mesh)
//                                               newton/_src/geometry/contact_reduction_global.py:1092-1227
//   export_reduced_contacts_kernel              newton/_src/geometry/contact_reduction_global.py:1327-1440
//   voxel resolution of a shape                 newton/_src/sim/builder.py:9936-9961
//
// The reference keeps one hash table over all shape pairs: key (shape_a, shape_b, bin), bins 0-19 the
// icosahedron normal bins (6 spatial-extreme slots + 1 max-depth slot each), bins 20-34 groups of 7 voxel
// slots; every slot keeps the largest packed (score, fingerprint, contact id) by atomic max, and the export
// writes each distinct winning contact once.  Here one pair's table is a map from key to its 7 slots, filled
// in candidate order with the same "replace when larger" rule, then exported.  Candidate order cannot change
// a slot's maximum; the contact id only decides between equal packed scores, which the deterministic packing
// never has (fingerprints are unique in a pair) and the fast packing breaks by the buffer slot an atomic
// counter hands out, i.e. by scheduling.  This restatement uses the contact's sub key as that id in the fast
// packing (one of the outcomes the reference can produce), the buffer position in the deterministic packing.
// The reference's table can fill (max_triangle_pairs / 4 entries; failed inserts skip the reduction); the
// restatement has no such limit.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <vector>

#include "oracle_math.h"

namespace oracle {

constexpr int RED_BINS = 20;      // NUM_NORMAL_BINS (icosahedron)
constexpr int RED_DIRS = 6;       // NUM_SPATIAL_DIRECTIONS
constexpr int RED_VOXELS = 100;   // NUM_VOXEL_DEPTH_SLOTS
constexpr int RED_VPK = RED_DIRS + 1;  // VALUES_PER_KEY
constexpr int RED_SLOTS = RED_BINS * RED_VPK + RED_VOXELS;  // 240 (<= MAX_CONTACTS_PER_PAIR 255)
constexpr float RED_BETA = 0.0001f;   // BETA_THRESHOLD

// _ICOSAHEDRON_NORMALS (contact_reduction.py:149-177): top cap 0-4, equatorial belt 5-14, bottom cap 15-19
inline const float* red_face_normals() {
  static const float F[RED_BINS * 3] = {
      0.49112338f,  0.79465455f,  0.35682216f,  -0.18759243f, 0.79465450f,  0.57735026f,
      -0.60706190f, 0.79465450f,  0.0f,         -0.18759237f, 0.79465450f,  -0.57735026f,
      0.49112340f,  0.79465455f,  -0.35682210f, 0.98224690f,  -0.18759257f, 0.0f,
      0.79465440f,  0.18759239f,  -0.57735030f, 0.30353096f,  -0.18759252f, 0.93417233f,
      0.79465440f,  0.18759243f,  0.57735030f,  -0.79465450f, -0.18759249f, 0.57735030f,
      -0.30353105f, 0.18759243f,  0.93417240f,  -0.79465440f, -0.18759240f, -0.57735030f,
      -0.98224690f, 0.18759254f,  0.0f,         0.30353096f,  -0.18759250f, -0.93417233f,
      -0.30353084f, 0.18759246f,  -0.93417240f, 0.18759249f,  -0.79465440f, 0.57735026f,
      -0.49112338f, -0.79465450f, 0.35682213f,  -0.49112338f, -0.79465455f, -0.35682213f,
      0.18759243f,  -0.79465440f, -0.57735026f, 0.60706200f,  -0.79465440f, 0.0f};
  return F;
}

template <class T> V3<T> red_face(int i) {
  const float* F = red_face_normals();
  return V3<T>{T(F[3 * i]), T(F[3 * i + 1]), T(F[3 * i + 2])};
}

// get_slot, icosahedron branch (contact_reduction.py:272-295): search the cap(s) the normal's y selects,
// first strictly larger dot product wins.
template <class T> int red_get_slot(V3<T> n) {
  const T up = n.y;
  int s, e;
  if (up > T(0.65f)) {
    s = 0; e = 5;
  } else if (up < T(-0.65f)) {
    s = 15; e = 20;
  } else if (up >= T(0)) {
    s = 0; e = 15;
  } else {
    s = 5; e = 20;
  }
  int best = s;
  T md = dot(n, red_face<T>(s));
  for (int i = s + 1; i < e; ++i) {
    const T d = dot(n, red_face<T>(i));
    if (d > md) {
      md = d;
      best = i;
    }
  }
  return best;
}

// project_point_to_plane (contact_reduction.py:308-331)
template <class T> V2<T> red_project(int bin, V3<T> p) {
  const V3<T> fn = red_face<T>(bin);
  const V3<T> ref = tabs(fn.y) < T(0.9f) ? V3<T>{T(0), T(1), T(0)} : V3<T>{T(1), T(0), T(0)};
  const V3<T> u = normalize(ref - dot(ref, fn) * fn);
  const V3<T> v = cross(fn, u);
  return V2<T>{dot(p, u), dot(p, v)};
}

// get_spatial_direction_2d (contact_reduction.py:334-345): angle = i * (2 pi / 6) in float32; its cosine and
// sine correctly rounded to float32 (the reference evaluates the device's sinf/cosf of the same angle).
inline V2<float> red_direction(int i) {
  const float pi = 3.14159265358979323846f;
  const float step = (2.0f * pi) / 6.0f;
  const float a = float(i) * step;
  return V2<float>{float(std::cos(double(a))), float(std::sin(double(a)))};
}

// float_flip (contact_reduction.py:104-110): order-preserving float -> uint32 (-0 below +0)
inline uint32_t red_float_flip(float f) {
  uint32_t i;
  std::memcpy(&i, &f, 4);
  const uint32_t mask = uint32_t(-int32_t(i >> 31)) | 0x80000000u;
  return i ^ mask;
}

// make_contact_value (contact_reduction_global.py:296-432): fast (full score, 32-bit id) or deterministic
// (22-bit score, 22-bit fingerprint, 20-bit id)
inline uint64_t red_value(float score, int fingerprint, uint32_t id, bool det) {
  const uint32_t f = red_float_flip(score);
  if (det)
    return (uint64_t(f >> 10) << 42) | ((uint64_t(uint32_t(fingerprint)) & 0x3FFFFFull) << 20) | (uint64_t(id) & 0xFFFFFull);
  return (uint64_t(f) << 32) | uint64_t(id);
}

// compute_voxel_index (contact_reduction.py:363-407); float -> int truncation saturated so that NaN and
// out-of-range values land where the clamp puts them
template <class T> int red_voxel_index(V3<T> pl, V3<T> lo, V3<T> hi, const int res[3]) {
  const V3<T> size = hi - lo;
  T r[3] = {T(0), T(0), T(0)};
  if (size.x > T(1e-6f)) r[0] = (pl.x - lo.x) / size.x;
  if (size.y > T(1e-6f)) r[1] = (pl.y - lo.y) / size.y;
  if (size.z > T(1e-6f)) r[2] = (pl.z - lo.z) / size.z;
  int v[3];
  for (int k = 0; k < 3; ++k) {
    const T t = r[k] * T(float(res[k]));
    v[k] = t >= T(0) ? (t < T(float(res[k])) ? int(t) : res[k] - 1) : 0;
    v[k] = v[k] < 0 ? 0 : (v[k] > res[k] - 1 ? res[k] - 1 : v[k]);
  }
  const int idx = v[0] + v[1] * res[0] + v[2] * res[0] * res[1];
  return idx < 0 ? 0 : (idx > RED_VOXELS - 1 ? RED_VOXELS - 1 : idx);
}

// encode_oct / decode_oct (contact_reduction_global.py:435-480)
template <class T> V2<T> red_encode_oct(V3<T> n) {
  const T l1 = tabs(n.x) + tabs(n.y) + tabs(n.z);
  if (l1 < T(1.0e-20f)) return V2<T>{T(0), T(0)};
  const T inv = T(1) / l1;
  T ox = n.x * inv, oy = n.y * inv;
  const T oz = n.z * inv;
  if (oz < T(0)) {
    const T sx = ox < T(0) ? T(-1) : T(1);
    const T sy = oy < T(0) ? T(-1) : T(1);
    const T nx = (T(1) - tabs(oy)) * sx;
    const T ny = (T(1) - tabs(ox)) * sy;
    ox = nx;
    oy = ny;
  }
  return V2<T>{ox, oy};
}

template <class T> V3<T> red_decode_oct(V2<T> e) {
  const T nz = T(1) - tabs(e.x) - tabs(e.y);
  T nx = e.x, ny = e.y;
  if (nz < T(0)) {
    const T sx = nx < T(0) ? T(-1) : T(1);
    const T sy = ny < T(0) ? T(-1) : T(1);
    const T ax = (T(1) - tabs(ny)) * sx;
    const T ay = (T(1) - tabs(nx)) * sy;
    nx = ax;
    ny = ay;
  }
  return normalize(V3<T>{nx, ny, nz});
}

// One candidate contact of a mesh pair, as the reducer's buffer holds it.
template <class T> struct RedCand {
  V3<T> pos;      // world point
  V3<T> normal;   // a -> b
  T depth;        // signed distance
  int sub;        // fingerprint = sort sub key
  // where the candidate competes (filled by the caller's path): normal used for the bin, the point the bin
  // projects, the beta test, the voxel
  V3<T> bin_normal;
  V3<T> proj_point;
  bool use_beta;
  int voxel;
};

// The reduction of one pair's candidates: slot maxima (one table entry per (bin) key), then the distinct
// winners in entry order.  Returns the winners' positions in `cand`.
template <class T> std::vector<int> red_reduce(const std::vector<RedCand<T>>& cand, bool det) {
  // key: bin (0..19) or 20 + voxel group; value: 7 slots
  std::map<int, std::vector<uint64_t>> table;
  auto update = [&](int key, int slot, uint64_t v) {
    std::vector<uint64_t>& e = table[key];
    if (e.empty()) e.assign(RED_VPK, 0ull);
    if (e[slot] < v) e[slot] = v;
  };
  for (size_t c = 0; c < cand.size(); ++c) {
    const RedCand<T>& k = cand[c];
    const uint32_t id = det ? uint32_t(c) : uint32_t(k.sub);
    const int bin = red_get_slot(k.bin_normal);
    const V2<T> p2 = red_project(bin, k.proj_point);
    for (int d = 0; d < RED_DIRS; ++d) {
      if (!k.use_beta) continue;
      const V2<float> dir = red_direction(d);
      const T score = p2.x * T(dir.x) + p2.y * T(dir.y);
      update(bin, d, red_value(float(score), k.sub, id, det));
    }
    update(bin, RED_DIRS, red_value(float(-k.depth), k.sub, id, det));
    const int group = k.voxel / RED_VPK, local = k.voxel % RED_VPK;
    update(RED_BINS + group, local, red_value(float(-k.depth), k.sub, id, det));
  }
  // export: per entry, its slots' distinct contacts; a contact exported once over all entries
  std::vector<int> out;
  std::vector<char> done(cand.size(), 0);
  for (const auto& kv : table) {
    for (int s = 0; s < RED_VPK; ++s) {
      const uint64_t v = kv.second[s];
      if (v == 0ull) continue;
      int c = -1;
      if (det) {
        c = int(v & 0xFFFFFull);
      } else {
        const int sub = int(uint32_t(v & 0xFFFFFFFFull));
        for (size_t q = 0; q < cand.size(); ++q)
          if (cand[q].sub == sub) { c = int(q); break; }
      }
      if (c < 0 || done[c]) continue;
      done[c] = 1;
      out.push_back(c);
    }
  }
  return out;
}

// compute_voxel_resolution_from_aabb (builder.py:9936-9961): voxels of about equal size, at most `budget`
inline void red_voxel_resolution(const double lo[3], const double hi[3], int budget, int out[3]) {
  double size[3];
  for (int k = 0; k < 3; ++k) size[k] = std::max(hi[k] - lo[k], 1e-6);
  const double volume = size[0] * size[1] * size[2];
  double v = std::pow(volume / budget, 1.0 / 3.0);
  v = std::max(v, 1e-6);
  int n[3];
  for (int k = 0; k < 3; ++k) {
    // Python round(): half to even
    const double x = size[k] / v;
    double r = std::nearbyint(x);
    n[k] = std::max(1, int(r));
  }
  while ((long long)n[0] * n[1] * n[2] > budget) {
    if (n[0] >= n[1] && n[0] >= n[2] && n[0] > 1) --n[0];
    else if (n[1] >= n[2] && n[1] > 1) --n[1];
    else if (n[2] > 1) --n[2];
    else break;
  }
  for (int k = 0; k < 3; ++k) out[k] = n[k];
}

}  // namespace oracle
