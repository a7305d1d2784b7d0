// Device functions of the SDF narrow phase: dense-grid sampling and gradient, the edge search, and
// the grid cook.  Reference: newton/_src/geometry/sdf_texture.py:776-955 (texture_sample_sdf /
// texture_sample_sdf_grad: clamp to the grid box, trilinear blend of the 8 corner values, plus the
// distance to the box outside it), sdf_contact.py:51-82 (scale_sdf_result_to_world), :448-460 (edge
// bounding sphere), :483-681 (do_edge_sdf_collision: Brent's method, 5 iterations, endpoint check).
#pragma once
#include "nmc_internal.h"
#include "nmc_math.h"

namespace nmc {

// One mesh's dense grid: (cells + 1) vertices per axis from `lower`, spacing `voxel`, x fastest.
struct SdfGrid {
  const float* v;
  f3 lower, upper;
  float inv_dx;
  int cx, cy, cz;  // cells per axis
};

__device__ __forceinline__ SdfGrid sdf_grid(const NmcModelDesc& d, int mesh) {
  SdfGrid g;
  g.v = d.mesh_sdf + d.mesh_sdf_start[mesh];
  g.lower = load3(d.mesh_sdf_lower + 3 * mesh);
  const float vx = d.mesh_sdf_voxel[mesh];
  g.cx = d.mesh_sdf_cells[3 * mesh];
  g.cy = d.mesh_sdf_cells[3 * mesh + 1];
  g.cz = d.mesh_sdf_cells[3 * mesh + 2];
  g.upper = f3{g.lower.x + vx * (float)g.cx, g.lower.y + vx * (float)g.cy, g.lower.z + vx * (float)g.cz};
  g.inv_dx = 1.0f / vx;
  return g;
}

struct SdfCell {
  float c[8];  // v000 v100 v010 v110 v001 v101 v011 v111
  float tx, ty, tz;
  float diff_mag;
  f3 diff;
};

__device__ __forceinline__ SdfCell sdf_cell(const SdfGrid& g, f3 p) {
  SdfCell s;
  f3 cl{clampf(p.x, g.lower.x, g.upper.x), clampf(p.y, g.lower.y, g.upper.y), clampf(p.z, g.lower.z, g.upper.z)};
  s.diff = p - cl;
  s.diff_mag = len(s.diff);
  f3 f = (cl - g.lower) * g.inv_dx;
  float fx = clampf(f.x, 0.0f, (float)g.cx);
  float fy = clampf(f.y, 0.0f, (float)g.cy);
  float fz = clampf(f.z, 0.0f, (float)g.cz);
  int ix = max(0, min((int)floorf(fx), g.cx - 1));
  int iy = max(0, min((int)floorf(fy), g.cy - 1));
  int iz = max(0, min((int)floorf(fz), g.cz - 1));
  s.tx = fx - (float)ix;
  s.ty = fy - (float)iy;
  s.tz = fz - (float)iz;
  const int nx = g.cx + 1, ny = g.cy + 1;
  const float* b = g.v + ((long long)iz * ny + iy) * nx + ix;
  const long long sy = nx, sz = (long long)nx * ny;
  s.c[0] = b[0];
  s.c[1] = b[1];
  s.c[2] = b[sy];
  s.c[3] = b[sy + 1];
  s.c[4] = b[sz];
  s.c[5] = b[sz + 1];
  s.c[6] = b[sz + sy];
  s.c[7] = b[sz + sy + 1];
  return s;
}

__device__ __forceinline__ float sdf_trilinear(const SdfCell& s) {
  float c00 = s.c[0] + (s.c[1] - s.c[0]) * s.tx;
  float c10 = s.c[2] + (s.c[3] - s.c[2]) * s.tx;
  float c01 = s.c[4] + (s.c[5] - s.c[4]) * s.tx;
  float c11 = s.c[6] + (s.c[7] - s.c[6]) * s.tx;
  float c0 = c00 + (c10 - c00) * s.ty;
  float c1 = c01 + (c11 - c01) * s.ty;
  return c0 + (c1 - c0) * s.tz;
}

__device__ __forceinline__ float sdf_sample(const SdfGrid& g, f3 p) {
  SdfCell s = sdf_cell(g, p);
  return sdf_trilinear(s) + s.diff_mag;
}

__device__ __forceinline__ float sdf_sample_grad(const SdfGrid& g, f3 p, f3& grad) {
  SdfCell s = sdf_cell(g, p);
  float val = sdf_trilinear(s);
  float omtx = 1.0f - s.tx, omty = 1.0f - s.ty, omtz = 1.0f - s.tz;
  const float* c = s.c;
  float gx = omty * omtz * (c[1] - c[0]) + s.ty * omtz * (c[3] - c[2]) + omty * s.tz * (c[5] - c[4]) +
             s.ty * s.tz * (c[7] - c[6]);
  float gy = omtx * omtz * (c[2] - c[0]) + s.tx * omtz * (c[3] - c[1]) + omtx * s.tz * (c[6] - c[4]) +
             s.tx * s.tz * (c[7] - c[5]);
  float gz = omtx * omty * (c[4] - c[0]) + s.tx * omty * (c[5] - c[1]) + omtx * s.ty * (c[6] - c[2]) +
             s.tx * s.ty * (c[7] - c[3]);
  grad = f3{gx * g.inv_dx, gy * g.inv_dx, gz * g.inv_dx};
  if (s.diff_mag > 0.0f) {
    val = val + s.diff_mag;
    grad = s.diff / s.diff_mag;
  }
  return val;
}

// Deepest point of the edge v0 -> v1 in the SDF (Brent's method on t in [0, 1], 5 iterations, started at
// the midpoint whose value the culling pass computed), then the nearer endpoint when the minimum sits
// near one.  Returns the SDF value and writes the point.
__device__ float edge_sdf_search(const SdfGrid& g, f3 v0, f3 v1, float midpoint_sdf, f3& point) {
  const float golden = 0.3819660112501051f;
  f3 e = v1 - v0;
  float a = 0.0f, b = 1.0f, x = 0.5f, w = 0.5f, vb = 0.5f;
  float fx = midpoint_sdf, fw = fx, fv = fx;
  float dstep = 0.0f, estep = 0.0f;
#pragma unroll 1
  for (int it = 0; it < 5; ++it) {
    float m = 0.5f * (a + b);
    float tol = 1.0e-2f * fabsf(x) + 1.0e-8f;
    float tol2 = 2.0f * tol;
    if (fabsf(x - m) <= tol2 - 0.5f * (b - a)) break;
    bool parabolic = false;
    float pn = 0.0f, qd = 0.0f;
    if (fabsf(estep) > tol) {
      float r = (x - w) * (fx - fv);
      qd = (x - vb) * (fx - fw);
      pn = (x - vb) * qd - (x - w) * r;
      qd = 2.0f * (qd - r);
      if (qd > 0.0f) pn = -pn; else qd = -qd;
      if (fabsf(pn) < 0.5f * fabsf(qd * estep)) {
        float trial = pn / qd;
        float ut = x + trial;
        if (ut - a >= tol2 && b - ut >= tol2) parabolic = true;
      }
    }
    if (parabolic) {
      estep = dstep;
      dstep = pn / qd;
    } else {
      estep = x >= m ? a - x : b - x;
      dstep = golden * estep;
    }
    float u;
    if (fabsf(dstep) >= tol) u = x + dstep;
    else u = dstep > 0.0f ? x + tol : x - tol;
    float fu = sdf_sample(g, v0 + e * u);
    if (fu <= fx) {
      if (u < x) b = x; else a = x;
      vb = w; fv = fw;
      w = x; fw = fx;
      x = u; fx = fu;
    } else {
      if (u < x) a = u; else b = u;
      if (fu <= fw || w == x) {
        vb = w; fv = fw;
        w = u; fw = fu;
      } else if (fu <= fv || vb == x || vb == w) {
        vb = u; fv = fu;
      }
    }
  }
  float bt = x, bf = fx;
  if (x < 0.2f || x > 0.8f) {
    float ct = x < 0.5f ? 0.0f : 1.0f;
    float fe = sdf_sample(g, v0 + e * ct);
    if (fe < bf) { bt = ct; bf = fe; }
  }
  point = v0 + e * bt;
  return bf;
}

// ------------------------------------------------------------------ grid cook
// Closest point on triangle abc to p (region tests on the barycentric coordinates).
__device__ __forceinline__ f3 closest_on_triangle(f3 p, f3 a, f3 b, f3 c) {
  f3 ab = b - a, ac = c - a, ap = p - a;
  float d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.0f && d2 <= 0.0f) return a;
  f3 bp = p - b;
  float d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0.0f && d4 <= d3) return b;
  float vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
    float v = d1 / (d1 - d3);
    return a + ab * v;
  }
  f3 cp = p - c;
  float d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0.0f && d5 <= d6) return c;
  float vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
    float w = d2 / (d2 - d6);
    return a + ac * w;
  }
  float va = d3 * d6 - d5 * d4;
  if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
    float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    return b + (c - b) * w;
  }
  float denom = 1.0f / (va + vb + vc);
  float v = vb * denom, w = vc * denom;
  return a + ab * v + ac * w;
}

// Solid angle of triangle abc seen from the origin (Van Oosterom & Strackee, 1983).
__device__ __forceinline__ float solid_angle(f3 a, f3 b, f3 c) {
  float la = len(a), lb = len(b), lc = len(c);
  float num = dot(a, cross(b, c));
  float den = la * lb * lc + dot(a, b) * lc + dot(a, c) * lb + dot(b, c) * la;
  return 2.0f * atan2f(num, den);
}

}  // namespace nmc
