// Mesh contact reduction (CollisionPipeline reduce_contacts=True) on gfx950, for one mesh pair held by one
// workgroup.
//
// The reference reduces mesh contacts through one global hash table (newton/_src/geometry/
// contact_reduction_global.py:671-1580; constants and helpers contact_reduction.py:60-407): for a shape pair,
// every candidate contact competes for
//   * its normal bin (the best of 20 icosahedron face normals): 6 scan-direction slots (largest projection of
//     the contact point on the face plane onto directions k pi / 3; only contacts shallower than a beta
//     threshold compete) and 1 max-depth slot;
//   * its voxel (the pair's first-shape grid of at most 100 voxels): the deepest contact;
// each slot keeping the largest packed (score, fingerprint, id) value, and every distinct winner is written
// once.  Here the pair's whole table is 240 64-bit slots in LDS (bin b scan k at 7 b + k, max depth at 7 b + 6,
// voxel v at 140 + v), updated with LDS 64-bit atomic max; no global hash table, no insert failures.  The
// value's low field carries the candidate's sub key, so a winner is re-evaluated from it.
#pragma once
#include <hip/hip_runtime.h>

#include "nmc_internal.h"
#include "nmc_math.h"

namespace nmc {

constexpr int RED_BINS = 20;                 // NUM_NORMAL_BINS (icosahedron)
constexpr int RED_DIRS = 6;                  // NUM_SPATIAL_DIRECTIONS
constexpr int RED_VOXELS = 100;              // NUM_VOXEL_DEPTH_SLOTS
constexpr int RED_VPK = RED_DIRS + 1;        // slots per normal bin
constexpr int RED_SLOTS = RED_BINS * RED_VPK + RED_VOXELS;  // 240 <= MAX_CONTACTS_PER_PAIR (255)
constexpr float RED_BETA = 0.0001f;          // BETA_THRESHOLD [m]

// Icosahedron face normals (contact_reduction.py:149-177): top cap 0-4, belt 5-14, bottom cap 15-19.
static __constant__ float c_red_face[RED_BINS * 3] = {
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

__device__ __forceinline__ f3 red_face(int i) { return f3{c_red_face[3 * i], c_red_face[3 * i + 1], c_red_face[3 * i + 2]}; }

// get_slot (contact_reduction.py:272-295): the candidate range by the normal's y, first strictly larger dot.
__device__ __forceinline__ int red_bin(f3 n) {
  const float up = n.y;
  int s, e;
  if (up > 0.65f) {
    s = 0; e = 5;
  } else if (up < -0.65f) {
    s = 15; e = 20;
  } else if (up >= 0.0f) {
    s = 0; e = 15;
  } else {
    s = 5; e = 20;
  }
  int best = s;
  float md = dot(n, red_face(s));
  for (int i = s + 1; i < e; ++i) {
    const float d = dot(n, red_face(i));
    if (d > md) {
      md = d;
      best = i;
    }
  }
  return best;
}

// project_point_to_plane (contact_reduction.py:308-331)
__device__ __forceinline__ void red_project(int bin, f3 p, float& px, float& py) {
  const f3 fn = red_face(bin);
  const f3 ref = fabsf(fn.y) < 0.9f ? f3{0.0f, 1.0f, 0.0f} : f3{1.0f, 0.0f, 0.0f};
  const f3 u = normalize(ref - dot(ref, fn) * fn);
  const f3 v = cross(fn, u);
  px = dot(p, u);
  py = dot(p, v);
}

// float_flip (contact_reduction.py:104-110): order-preserving float -> uint32, -0 below +0
__device__ __forceinline__ unsigned red_flip(float f) {
  const unsigned i = __float_as_uint(f);
  const unsigned mask = static_cast<unsigned>(-static_cast<int>(i >> 31)) | 0x80000000u;
  return i ^ mask;
}

// make_contact_value (contact_reduction_global.py:296-432) with the sub key as fingerprint and id
__device__ __forceinline__ unsigned long long red_value(float score, int sub, int mode) {
  const unsigned f = red_flip(score);
  const unsigned long long s = static_cast<unsigned>(sub);
  if (mode == 2) return (static_cast<unsigned long long>(f >> 10) << 42) | ((s & 0x3FFFFFull) << 20) | (s & 0xFFFFFull);
  return (static_cast<unsigned long long>(f) << 32) | s;
}

// the sub key of a slot's winner
__device__ __forceinline__ int red_winner(unsigned long long v, int mode) {
  return mode == 2 ? static_cast<int>((v >> 20) & 0x3FFFFFull) : static_cast<int>(v & 0xFFFFFFFFull);
}

// compute_voxel_index (contact_reduction.py:363-407), clamped to the 100 slots; the float -> int truncation
// saturates (NaN and negatives land in cell 0, overflow in the last cell, where the clamp puts them)
__device__ __forceinline__ int red_voxel(f3 pl, f3 lo, f3 hi, int nx, int ny, int nz) {
  const f3 size = hi - lo;
  float r[3] = {0.0f, 0.0f, 0.0f};
  if (size.x > 1e-6f) r[0] = (pl.x - lo.x) / size.x;
  if (size.y > 1e-6f) r[1] = (pl.y - lo.y) / size.y;
  if (size.z > 1e-6f) r[2] = (pl.z - lo.z) / size.z;
  const int n[3] = {nx, ny, nz};
  int v[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float t = r[k] * static_cast<float>(n[k]);
    v[k] = t >= 0.0f ? (t < static_cast<float>(n[k]) ? static_cast<int>(t) : n[k] - 1) : 0;
    v[k] = v[k] < 0 ? 0 : (v[k] > n[k] - 1 ? n[k] - 1 : v[k]);
  }
  const int idx = v[0] + v[1] * nx + v[2] * nx * ny;
  return idx < 0 ? 0 : (idx > RED_VOXELS - 1 ? RED_VOXELS - 1 : idx);
}

// encode_oct / decode_oct (contact_reduction_global.py:435-480): the reducer stores normals octahedrally
__device__ __forceinline__ f3 red_oct_roundtrip(f3 n) {
  const float l1 = fabsf(n.x) + fabsf(n.y) + fabsf(n.z);
  float ex = 0.0f, ey = 0.0f;
  if (!(l1 < 1.0e-20f)) {
    const float inv = 1.0f / l1;
    ex = n.x * inv;
    ey = n.y * inv;
    const float oz = n.z * inv;
    if (oz < 0.0f) {
      const float sx = ex < 0.0f ? -1.0f : 1.0f;
      const float sy = ey < 0.0f ? -1.0f : 1.0f;
      const float ax = (1.0f - fabsf(ey)) * sx;
      const float ay = (1.0f - fabsf(ex)) * sy;
      ex = ax;
      ey = ay;
    }
  }
  const float nz = 1.0f - fabsf(ex) - fabsf(ey);
  float nx = ex, ny = ey;
  if (nz < 0.0f) {
    const float sx = nx < 0.0f ? -1.0f : 1.0f;
    const float sy = ny < 0.0f ? -1.0f : 1.0f;
    const float ax = (1.0f - fabsf(ny)) * sx;
    const float ay = (1.0f - fabsf(nx)) * sy;
    nx = ax;
    ny = ay;
  }
  return normalize(f3{nx, ny, nz});
}

__device__ __forceinline__ void red_update(unsigned long long* slots, int k, unsigned long long v) {
  if (slots[k] < v) atomicMax(&slots[k], v);
}

// One candidate into the pair's LDS table: bin normal, projected point, beta test, voxel (reduce_contact_in_
// hashtable / export_and_reduce_contact_centered, contact_reduction_global.py:968-1227).
__device__ __forceinline__ void red_insert(unsigned long long* slots, const RedParams& rp, f3 bin_normal, f3 proj,
                                           float depth, bool use_beta, int voxel, int sub) {
  const int bin = red_bin(bin_normal);
  if (use_beta) {
    float px, py;
    red_project(bin, proj, px, py);
#pragma unroll
    for (int k = 0; k < RED_DIRS; ++k) {
      const float score = px * rp.dirs.c[k] + py * rp.dirs.s[k];
      red_update(slots, RED_VPK * bin + k, red_value(score, sub, rp.mode));
    }
  }
  const unsigned long long vd = red_value(-depth, sub, rp.mode);
  red_update(slots, RED_VPK * bin + RED_DIRS, vd);
  red_update(slots, RED_BINS * RED_VPK + voxel, vd);
}

}  // namespace nmc
