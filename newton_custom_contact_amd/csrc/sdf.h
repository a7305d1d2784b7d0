// Device functions of the SDF narrow phase: sparse-texture sampling and gradient, the edge search, and
// the cook's distance query.  Reference: newton/_src/geometry/sdf_texture.py:597-955 (_read_cell_corners,
// texture_sample_sdf, texture_sample_sdf_grad: clamp to the SDF box, trilinear blend of the 8 corner values
// of the fine cell or of the coarse subgrid cell, plus the distance to the box outside it),
// sdf_contact.py:51-82 (scale_sdf_result_to_world), :448-460 (edge bounding sphere), :483-681
// (do_edge_sdf_collision: Brent's method, 5 iterations, endpoint check).
#pragma once
#include "nmc_internal.h"
#include "nmc_math.h"

namespace nmc {

// One mesh's sparse SDF (TextureSDFData, sdf_texture.py:48-87): a coarse grid of the field at the corners of
// 8^3-cell subgrids, an indirection slot per subgrid, and a texture of the narrow-band subgrids' 9^3 samples.
// coarse: (d+1)(h+1)(w+1) values, x fastest; slots: w x h x d, z fastest (the reference's (w, h, d) array);
// texels: ts^3 samples, x fastest, float32 / unorm16 / unorm8 (fmt = bytes per texel).  A slot at or above
// SLOT_LINEAR (demoted or far-field subgrid) is sampled from the coarse grid; any other holds the texture block
// coordinates (x | y << 10 | z << 20) of the subgrid's samples.
constexpr unsigned SDF_SLOT_EMPTY = 0xFFFFFFFFu;
constexpr unsigned SDF_SLOT_LINEAR = 0xFFFFFFFEu;
constexpr int SDF_SUBGRID = 8;

struct SdfGrid {
  const float* coarse;
  const unsigned* slots;
  const unsigned char* texels;
  f3 lower, upper, inv_dx;
  float qmin, qrange;  // subgrids_min_sdf_value, subgrids_sdf_value_range
  int w, h, d, ts, fmt;
};

__device__ __forceinline__ SdfGrid sdf_grid(const NmcModelDesc& d, int mesh) {
  SdfGrid g;
  const int64_t* st = d.mesh_sdf_start + 3 * mesh;
  g.coarse = d.mesh_sdf_coarse + st[0];
  g.slots = d.mesh_sdf_slots + st[1];
  g.texels = d.mesh_sdf_texels + st[2];
  const float* b = d.mesh_sdf_box + 9 * mesh;
  g.lower = f3{b[0], b[1], b[2]};
  g.upper = f3{b[3], b[4], b[5]};
  g.inv_dx = f3{b[6], b[7], b[8]};
  g.qmin = d.mesh_sdf_quant[2 * mesh];
  g.qrange = d.mesh_sdf_quant[2 * mesh + 1];
  const int* dm = d.mesh_sdf_dims + 5 * mesh;
  g.w = dm[0];
  g.h = dm[1];
  g.d = dm[2];
  g.ts = dm[3];
  g.fmt = dm[4];
  return g;
}

// A texel as the reference's point-sampled texture read returns it: float32 as stored, unorm16 / unorm8
// normalised to [0, 1].
__device__ __forceinline__ float sdf_texel(const SdfGrid& g, long long i) {
  if (g.fmt == 4) return reinterpret_cast<const float*>(g.texels)[i];
  if (g.fmt == 2) return (float)reinterpret_cast<const unsigned short*>(g.texels)[i] / 65535.0f;
  return (float)g.texels[i] / 255.0f;
}

struct SdfCell {
  float c[8];  // v000 v100 v010 v110 v001 v101 v011 v111
  float tx, ty, tz;
  bool fine;   // read from the subgrid texture (values still in texture units)
  float diff_mag;
  f3 diff;
};

// _read_cell_corners (sdf_texture.py:597-701) on the clamped point: the fine cell of the point, its subgrid's
// slot, then the 8 texels of the cell, or the 8 coarse values of the subgrid with coarse weights.
__device__ __forceinline__ SdfCell sdf_cell(const SdfGrid& g, f3 p) {
  SdfCell s;
  const f3 cl{clampf(p.x, g.lower.x, g.upper.x), clampf(p.y, g.lower.y, g.upper.y), clampf(p.z, g.lower.z, g.upper.z)};
  s.diff = p - cl;
  s.diff_mag = len(s.diff);
  const f3 q = cl - g.lower;
  const f3 f{q.x * g.inv_dx.x, q.y * g.inv_dx.y, q.z * g.inv_dx.z};
  const float vx = (float)g.w * 8.0f, vy = (float)g.h * 8.0f, vz = (float)g.d * 8.0f;
  const float fx = clampf(f.x, 0.0f, vx), fy = clampf(f.y, 0.0f, vy), fz = clampf(f.z, 0.0f, vz);
  const int ix = max(0, min((int)floorf(fx), (int)vx - 1));
  const int iy = max(0, min((int)floorf(fy), (int)vy - 1));
  const int iz = max(0, min((int)floorf(fz), (int)vz - 1));
  s.tx = fx - (float)ix;
  s.ty = fy - (float)iy;
  s.tz = fz - (float)iz;
  const int xb = max(0, min((int)((float)ix * 0.125f), g.w - 1));
  const int yb = max(0, min((int)((float)iy * 0.125f), g.h - 1));
  const int zb = max(0, min((int)((float)iz * 0.125f), g.d - 1));
  const unsigned slot = g.slots[((long long)xb * g.h + yb) * g.d + zb];
  if (slot >= SDF_SLOT_LINEAR) {
    s.fine = false;
    s.tx = fx * 0.125f - (float)xb;
    s.ty = fy * 0.125f - (float)yb;
    s.tz = fz * 0.125f - (float)zb;
    const long long sx = g.w + 1, sy = (long long)(g.w + 1) * (g.h + 1);
    const float* b = g.coarse + zb * sy + yb * sx + xb;
    s.c[0] = b[0];
    s.c[1] = b[1];
    s.c[2] = b[sx];
    s.c[3] = b[sx + 1];
    s.c[4] = b[sy];
    s.c[5] = b[sy + 1];
    s.c[6] = b[sy + sx];
    s.c[7] = b[sy + sx + 1];
  } else {
    s.fine = true;
    const long long ox = (long long)(slot & 0x3FFu) * 9 + (ix - xb * 8);
    const long long oy = (long long)((slot >> 10) & 0x3FFu) * 9 + (iy - yb * 8);
    const long long oz = (long long)((slot >> 20) & 0x3FFu) * 9 + (iz - zb * 8);
    const long long ts = g.ts, i0 = (oz * ts + oy) * ts + ox;
    s.c[0] = sdf_texel(g, i0);
    s.c[1] = sdf_texel(g, i0 + 1);
    s.c[2] = sdf_texel(g, i0 + ts);
    s.c[3] = sdf_texel(g, i0 + ts + 1);
    s.c[4] = sdf_texel(g, i0 + ts * ts);
    s.c[5] = sdf_texel(g, i0 + ts * ts + 1);
    s.c[6] = sdf_texel(g, i0 + ts * ts + ts);
    s.c[7] = sdf_texel(g, i0 + ts * ts + ts + 1);
  }
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

// texture_sample_sdf (sdf_texture.py:776-894): the blend in texture units, de-quantised once, plus the
// distance to the SDF box outside it.
__device__ __forceinline__ float sdf_sample(const SdfGrid& g, f3 p) {
  SdfCell s = sdf_cell(g, p);
  float v = sdf_trilinear(s);
  if (s.fine) v = v * g.qrange + g.qmin;
  return v + s.diff_mag;
}

// texture_sample_sdf_grad (sdf_texture.py:897-955): every corner de-quantised first, the analytic trilinear
// gradient scaled by the fine cell's inverse size (coarse cells too), replaced by the direction to the box
// outside it.
__device__ __forceinline__ float sdf_sample_grad(const SdfGrid& g, f3 p, f3& grad) {
  SdfCell s = sdf_cell(g, p);
  if (s.fine) {
#pragma unroll
    for (int k = 0; k < 8; ++k) s.c[k] = s.c[k] * g.qrange + g.qmin;
  }
  float val = sdf_trilinear(s);
  float omtx = 1.0f - s.tx, omty = 1.0f - s.ty, omtz = 1.0f - s.tz;
  const float* c = s.c;
  float gx = omty * omtz * (c[1] - c[0]) + s.ty * omtz * (c[3] - c[2]) + omty * s.tz * (c[5] - c[4]) +
             s.ty * s.tz * (c[7] - c[6]);
  float gy = omtx * omtz * (c[2] - c[0]) + s.tx * omtz * (c[3] - c[1]) + omtx * s.tz * (c[6] - c[4]) +
             s.tx * s.tz * (c[7] - c[5]);
  float gz = omtx * omty * (c[4] - c[0]) + s.tx * omty * (c[5] - c[1]) + omtx * s.ty * (c[6] - c[2]) +
             s.tx * s.ty * (c[7] - c[3]);
  grad = f3{gx * g.inv_dx.x, gy * g.inv_dx.y, gz * g.inv_dx.z};
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
