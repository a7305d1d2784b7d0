// CPU oracle -- TEST INFRASTRUCTURE ONLY.  Triangle-mesh contact restatement.
//
// Follows:
//   texture_sample_sdf / _grad, _read_cell_corners    newton/_src/geometry/sdf_texture.py:597-955
//   scale_sdf_result_to_world                 newton/_src/geometry/sdf_contact.py:51-82
//   get_edge_bounding_sphere, edge culling    newton/_src/geometry/sdf_contact.py:447-460, 1010-1045
//   do_edge_sdf_collision (Brent + endpoints) newton/_src/geometry/sdf_contact.py:526-681
//   mesh-mesh contact per edge                newton/_src/geometry/sdf_contact.py:870-1150
//   mesh-plane contact per vertex             newton/_src/geometry/narrow_phase.py:1121-1190
//   SDF cook (distance to the nearest triangle, winding-number sign; coarse grid, narrow-band occupancy,
//   linearity demotion, texture addresses, quantised packing)
//                                             newton/_src/geometry/sdf_texture.py:119-566, 959-1308
//   reduce_contacts=False: every candidate the writer keeps   narrow_phase.py:1090-1190, sdf_contact.py:830-1065
//   reduce_contacts=True: the normal-bin / voxel reduction    oracle_reduce.h (contact_reduction_global.py)
// A pair's contacts are written in sub-key order.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "oracle_collide.h"
#include "oracle_reduce.h"

namespace oracle {

// The reference's texture SDF (TextureSDFData, sdf_texture.py:48-87): coarse grid at the subgrid corners
// ((d+1)(h+1)(w+1), x fastest), one slot per subgrid (w x h x d, z fastest), texture of 9^3-sample blocks
// (ts^3, x fastest; fmt = bytes per texel: float32, unorm16, unorm8).
constexpr uint32_t SLOT_EMPTY = 0xFFFFFFFFu;
constexpr uint32_t SLOT_LINEAR = 0xFFFFFFFEu;

template <class T> struct SdfG {
  const float* coarse;
  const uint32_t* slots;
  const unsigned char* texels;
  V3<T> lower, upper, inv_dx;
  T qmin, qrange;
  int w, h, d, ts, fmt;
};

template <class T> SdfG<T> sdf_grid(const OracleModel* m, int mesh) {
  SdfG<T> g;
  const int64_t* st = m->mesh_sdf_start + 3 * mesh;
  g.coarse = m->mesh_sdf_coarse + st[0];
  g.slots = m->mesh_sdf_slots + st[1];
  g.texels = m->mesh_sdf_texels + st[2];
  const float* b = m->mesh_sdf_box + 9 * mesh;
  g.lower = V3<T>{T(b[0]), T(b[1]), T(b[2])};
  g.upper = V3<T>{T(b[3]), T(b[4]), T(b[5])};
  g.inv_dx = V3<T>{T(b[6]), T(b[7]), T(b[8])};
  g.qmin = T(m->mesh_sdf_quant[2 * mesh]);
  g.qrange = T(m->mesh_sdf_quant[2 * mesh + 1]);
  const int* dm = m->mesh_sdf_dims + 5 * mesh;
  g.w = dm[0];
  g.h = dm[1];
  g.d = dm[2];
  g.ts = dm[3];
  g.fmt = dm[4];
  return g;
}

// wp.texture_sample with closest filtering at a texel centre: the stored float, or a unorm texel normalised
// to [0, 1].
template <class T> T sdf_texel(const SdfG<T>& g, long long i) {
  if (g.fmt == 4) {
    float v;
    std::memcpy(&v, g.texels + 4 * i, 4);
    return T(v);
  }
  if (g.fmt == 2) {
    uint16_t v;
    std::memcpy(&v, g.texels + 2 * i, 2);
    return T(v) / T(65535);
  }
  return T(g.texels[i]) / T(255);
}

template <class T> struct SdfCellO {
  T c[8];
  T tx, ty, tz;
  bool fine;
  T diff_mag;
  V3<T> diff;
};

// texture_sample_sdf / _read_cell_corners (sdf_texture.py:597-701, 776-894): clamp to the SDF box, fine cell
// of the point, slot of its subgrid; a texture block gives the cell's 8 texels, a linear or empty slot the
// subgrid's 8 coarse values with the weights of the coarse cell.
template <class T> SdfCellO<T> sdf_cell(const SdfG<T>& g, V3<T> p) {
  SdfCellO<T> s;
  V3<T> cl{tclamp(p.x, g.lower.x, g.upper.x), tclamp(p.y, g.lower.y, g.upper.y), tclamp(p.z, g.lower.z, g.upper.z)};
  s.diff = p - cl;
  s.diff_mag = length(s.diff);
  V3<T> q = cl - g.lower;
  V3<T> f{q.x * g.inv_dx.x, q.y * g.inv_dx.y, q.z * g.inv_dx.z};
  const T fvx = T(g.w) * T(8), fvy = T(g.h) * T(8), fvz = T(g.d) * T(8);
  T fx = tclamp(f.x, T(0), fvx), fy = tclamp(f.y, T(0), fvy), fz = tclamp(f.z, T(0), fvz);
  int ix = std::max(0, std::min(int(std::floor(fx)), int(fvx) - 1));
  int iy = std::max(0, std::min(int(std::floor(fy)), int(fvy) - 1));
  int iz = std::max(0, std::min(int(std::floor(fz)), int(fvz) - 1));
  s.tx = fx - T(ix);
  s.ty = fy - T(iy);
  s.tz = fz - T(iz);
  const T f2c = T(1) / T(8);
  int xb = std::max(0, std::min(int(T(ix) * f2c), g.w - 1));
  int yb = std::max(0, std::min(int(T(iy) * f2c), g.h - 1));
  int zb = std::max(0, std::min(int(T(iz) * f2c), g.d - 1));
  uint32_t slot = g.slots[((long long)xb * g.h + yb) * g.d + zb];
  if (slot >= SLOT_LINEAR) {
    s.fine = false;
    V3<T> cf{fx * f2c, fy * f2c, fz * f2c};
    s.tx = cf.x - T(xb);
    s.ty = cf.y - T(yb);
    s.tz = cf.z - T(zb);
    const long long sx = g.w + 1, sy = (long long)(g.w + 1) * (g.h + 1);
    const long long o[8] = {0, 1, sx, sx + 1, sy, sy + 1, sy + sx, sy + sx + 1};
    const float* b = g.coarse + zb * sy + yb * sx + xb;
    for (int k = 0; k < 8; ++k) s.c[k] = T(b[o[k]]);
  } else {
    s.fine = true;
    // tex_o + l + 0.5 is a texel centre: point sampling reads texel block * 9 + l
    const long long ox = (long long)(slot & 0x3FFu) * 9 + (ix - xb * 8);
    const long long oy = (long long)((slot >> 10) & 0x3FFu) * 9 + (iy - yb * 8);
    const long long oz = (long long)((slot >> 20) & 0x3FFu) * 9 + (iz - zb * 8);
    const long long ts = g.ts;
    const long long o[8] = {0, 1, ts, ts + 1, ts * ts, ts * ts + 1, ts * ts + ts, ts * ts + ts + 1};
    const long long i0 = (oz * ts + oy) * ts + ox;
    for (int k = 0; k < 8; ++k) s.c[k] = sdf_texel(g, i0 + o[k]);
  }
  return s;
}

template <class T> T sdf_trilinear(const SdfCellO<T>& s) {
  T c00 = s.c[0] + (s.c[1] - s.c[0]) * s.tx;
  T c10 = s.c[2] + (s.c[3] - s.c[2]) * s.tx;
  T c01 = s.c[4] + (s.c[5] - s.c[4]) * s.tx;
  T c11 = s.c[6] + (s.c[7] - s.c[6]) * s.tx;
  T c0 = c00 + (c10 - c00) * s.ty;
  T c1 = c01 + (c11 - c01) * s.ty;
  return c0 + (c1 - c0) * s.tz;
}

// texture_sample_sdf (sdf_texture.py:776-894): a texture cell's blend is de-quantised after blending.
template <class T> T sdf_sample(const SdfG<T>& g, V3<T> p) {
  SdfCellO<T> s = sdf_cell(g, p);
  T v = sdf_trilinear(s);
  if (s.fine) v = v * g.qrange + g.qmin;
  return v + s.diff_mag;
}

// texture_sample_sdf_grad (sdf_texture.py:897-955): corners de-quantised before blending
// (apply_subgrid_sdf_scale in _read_cell_corners), gradient scaled by the fine inverse cell size.
template <class T> T sdf_sample_grad(const SdfG<T>& g, V3<T> p, V3<T>& grad) {
  SdfCellO<T> s = sdf_cell(g, p);
  if (s.fine)
    for (int k = 0; k < 8; ++k) s.c[k] = s.c[k] * g.qrange + g.qmin;
  T val = sdf_trilinear(s);
  T omtx = T(1) - s.tx, omty = T(1) - s.ty, omtz = T(1) - s.tz;
  const T* c = s.c;
  T gx = omty * omtz * (c[1] - c[0]) + s.ty * omtz * (c[3] - c[2]) + omty * s.tz * (c[5] - c[4]) +
         s.ty * s.tz * (c[7] - c[6]);
  T gy = omtx * omtz * (c[2] - c[0]) + s.tx * omtz * (c[3] - c[1]) + omtx * s.tz * (c[6] - c[4]) +
         s.tx * s.tz * (c[7] - c[5]);
  T gz = omtx * omty * (c[4] - c[0]) + s.tx * omty * (c[5] - c[1]) + omtx * s.ty * (c[6] - c[2]) +
         s.tx * s.ty * (c[7] - c[3]);
  grad = V3<T>{gx * g.inv_dx.x, gy * g.inv_dx.y, gz * g.inv_dx.z};
  if (s.diff_mag > T(0)) {
    val = val + s.diff_mag;
    grad = s.diff / s.diff_mag;
  }
  return val;
}

// do_edge_sdf_collision (sdf_contact.py:526-681): Brent on t in [0, 1] from the midpoint, 5 iterations,
// then the nearer endpoint when the minimum lies in the outer fifths.
template <class T> T edge_sdf_search(const SdfG<T>& g, V3<T> v0, V3<T> v1, T midpoint_sdf, V3<T>& point) {
  const T golden = T(0.3819660112501051);
  V3<T> e = v1 - v0;
  T a = T(0), b = T(1), x = T(0.5), w = T(0.5), vb = T(0.5);
  T fx = midpoint_sdf, fw = fx, fv = fx;
  T dstep = T(0), estep = T(0);
  for (int it = 0; it < 5; ++it) {
    T mid = T(0.5) * (a + b);
    T tol = T(1.0e-2) * tabs(x) + T(1.0e-8);
    T tol2 = T(2) * tol;
    if (tabs(x - mid) <= tol2 - T(0.5) * (b - a)) break;
    bool parabolic = false;
    T pn = T(0), qd = T(0);
    if (tabs(estep) > tol) {
      T r = (x - w) * (fx - fv);
      qd = (x - vb) * (fx - fw);
      pn = (x - vb) * qd - (x - w) * r;
      qd = T(2) * (qd - r);
      if (qd > T(0)) pn = -pn; else qd = -qd;
      if (tabs(pn) < T(0.5) * tabs(qd * estep)) {
        T trial = pn / qd;
        T ut = x + trial;
        if (ut - a >= tol2 && b - ut >= tol2) parabolic = true;
      }
    }
    if (parabolic) {
      estep = dstep;
      dstep = pn / qd;
    } else {
      estep = x >= mid ? a - x : b - x;
      dstep = golden * estep;
    }
    T u;
    if (tabs(dstep) >= tol) u = x + dstep;
    else u = dstep > T(0) ? x + tol : x - tol;
    T fu = sdf_sample(g, v0 + e * u);
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
  T bt = x, bf = fx;
  if (x < T(0.2) || x > T(0.8)) {
    T ct = x < T(0.5) ? T(0) : T(1);
    T fe = sdf_sample(g, v0 + e * ct);
    if (fe < bf) { bt = ct; bf = fe; }
  }
  point = v0 + e * bt;
  return bf;
}

// ---------------------------------------------------------------- SDF cook
// Closest point on triangle abc (Voronoi regions of the barycentric coordinates, Ericson 2005, 5.1.5).
template <class T> V3<T> closest_on_triangle(V3<T> p, V3<T> a, V3<T> b, V3<T> c) {
  V3<T> ab = b - a, ac = c - a, ap = p - a;
  T d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= T(0) && d2 <= T(0)) return a;
  V3<T> bp = p - b;
  T d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= T(0) && d4 <= d3) return b;
  T vc = d1 * d4 - d3 * d2;
  if (vc <= T(0) && d1 >= T(0) && d3 <= T(0)) {
    T v = d1 / (d1 - d3);
    return a + ab * v;
  }
  V3<T> cp = p - c;
  T d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= T(0) && d5 <= d6) return c;
  T vb = d5 * d2 - d1 * d6;
  if (vb <= T(0) && d2 >= T(0) && d6 <= T(0)) {
    T w = d2 / (d2 - d6);
    return a + ac * w;
  }
  T va = d3 * d6 - d5 * d4;
  if (va <= T(0) && (d4 - d3) >= T(0) && (d5 - d6) >= T(0)) {
    T w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    return b + (c - b) * w;
  }
  T denom = T(1) / (va + vb + vc);
  T v = vb * denom, w = vc * denom;
  return a + ab * v + ac * w;
}

// Solid angle of triangle abc seen from the origin (Van Oosterom & Strackee 1983).
template <class T> T solid_angle(V3<T> a, V3<T> b, V3<T> c) {
  T la = length(a), lb = length(b), lc = length(c);
  T num = dot(a, cross(b, c));
  T den = la * lb * lc + dot(a, b) * lc + dot(a, c) * lb + dot(b, c) * la;
  return T(2) * std::atan2(num, den);
}

// get_distance_to_mesh (sdf_utils.py:640-651) at every vertex of a fine grid: vertex (ix, iy, iz) at
// lower + (ix cell[0], iy cell[1], iz cell[2]) (the position arithmetic of the cook kernels,
// sdf_texture.py:398-402), nearest-triangle distance, sign of `w > threshold` for the generalized winding
// number w, flipped back for a negative threshold.  n[0..2] vertices per axis, x fastest.
template <class T>
void sdf_cook(const float* vtx, int nv, const int* tri, int nt, const float lower[3], const float cell[3], const int n[3],
              float threshold, T* out) {
  const long long nx = n[0], ny = n[1], nz = n[2];
  const T pi = T(3.14159265358979323846);
  for (long long i = 0; i < nx * ny * nz; ++i) {
    int ix = int(i % nx), iy = int((i / nx) % ny), iz = int(i / (nx * ny));
    V3<T> p{T(lower[0]) + T(ix) * T(cell[0]), T(lower[1]) + T(iy) * T(cell[1]), T(lower[2]) + T(iz) * T(cell[2])};
    T best = T(3.0e38f), omega = T(0);
    for (int t = 0; t < nt; ++t) {
      int i0 = tri[3 * t], i1 = tri[3 * t + 1], i2 = tri[3 * t + 2];
      if (i0 < 0 || i0 >= nv || i1 < 0 || i1 >= nv || i2 < 0 || i2 >= nv) continue;
      V3<T> a{T(vtx[3 * i0]), T(vtx[3 * i0 + 1]), T(vtx[3 * i0 + 2])};
      V3<T> b{T(vtx[3 * i1]), T(vtx[3 * i1 + 1]), T(vtx[3 * i1 + 2])};
      V3<T> c{T(vtx[3 * i2]), T(vtx[3 * i2 + 1]), T(vtx[3 * i2 + 2])};
      V3<T> q = closest_on_triangle(p, a, b, c);
      T dd = length(p - q);
      best = dd < best ? dd : best;
      omega += solid_angle(a - p, b - p, c - p);
    }
    T wn = omega / (T(4) * pi);
    const bool above = wn > T(threshold);
    out[i] = (threshold >= 0.0f ? above : !above) ? -best : best;
  }
}

// build_sparse_sdf_from_mesh after its mesh queries (sdf_texture.py:1017-1160), kernel by kernel, on the fine
// grid of (w, h, d) subgrids (8 w + 1 vertices along x, x fastest).  Every query point of the reference is a
// fine vertex: subgrid corners (_build_coarse_sdf_from_mesh_kernel), centres (_check_subgrid_occupied_kernel,
// (8 b + 4) cells) and samples.  Returns the number of required subgrids.
inline int sdf_subgrids(const float* fine, int w, int h, int d, float band_lo, float band_hi, float lin_thr,
                        float* coarse, int* occupied, int* required, int* linear, int* address) {
  const long long nx = 8LL * w + 1, ny = 8LL * h + 1;
  auto F = [&](long long x, long long y, long long z) { return fine[(z * ny + y) * nx + x]; };
  const int bgx = w + 1, bgy = h + 1, bgz = d + 1;
  // _build_coarse_sdf_from_mesh_kernel (:329-360)
  for (int t = 0; t < bgx * bgy * bgz; ++t) {
    int z = t / (bgx * bgy), r = t - z * bgx * bgy, y = r / bgx, x = r - y * bgx;
    coarse[t] = F(8LL * x, 8LL * y, 8LL * z);
  }
  const int total = w * h * d;
  // _check_subgrid_occupied_kernel (:198-224) with _is_in_narrow_band (:164-169): wp.sign(0) is +1
  for (int t = 0; t < total; ++t) {
    int z = t / (w * h), r = t - z * w * h, y = r / w, x = r - y * w;
    float c = F(8LL * x + 4, 8LL * y + 4, 8LL * z + 4);
    bool in = (c < 0.0f ? -1.0f : 1.0f) > 0.0f ? c < band_hi : c > band_lo;
    required[t] = in ? 1 : 0;
    occupied[t] = required[t];
    linear[t] = 0;
  }
  if (lin_thr > 0.0f) {
    // _accumulate_subgrid_linearity_error_kernel (:227-304) with _interp_coarse_sdf (:119-161)
    std::vector<float> err(total, 0.0f);
    const float inv_cpsg = 1.0f / float(8);
    for (int t = 0; t < total; ++t) {
      if (!required[t]) continue;
      int bz = t / (w * h), r = t - bz * w * h, by = r / w, bx = r - by * w;
      for (int k = 0; k < 729; ++k) {
        int lz = k / 81, rr = k - lz * 81, ly = rr / 9, lx = rr - ly * 9;
        float mesh_val = F(8LL * bx + lx, 8LL * by + ly, 8LL * bz + lz);
        float cfx = float(bx) + float(lx) * inv_cpsg, cfy = float(by) + float(ly) * inv_cpsg;
        float cfz = float(bz) + float(lz) * inv_cpsg;
        int x0 = std::max(0, std::min(int(std::floor(cfx)), bgx - 2));
        int y0 = std::max(0, std::min(int(std::floor(cfy)), bgy - 2));
        int z0 = std::max(0, std::min(int(std::floor(cfz)), bgz - 2));
        float tx = tclamp(cfx - float(x0), 0.0f, 1.0f), ty = tclamp(cfy - float(y0), 0.0f, 1.0f);
        float tz = tclamp(cfz - float(z0), 0.0f, 1.0f);
        auto B = [&](int x, int y, int z) { return coarse[(z * bgy + y) * bgx + x]; };
        float v000 = B(x0, y0, z0), v100 = B(x0 + 1, y0, z0), v010 = B(x0, y0 + 1, z0), v110 = B(x0 + 1, y0 + 1, z0);
        float v001 = B(x0, y0, z0 + 1), v101 = B(x0 + 1, y0, z0 + 1), v011 = B(x0, y0 + 1, z0 + 1);
        float v111 = B(x0 + 1, y0 + 1, z0 + 1);
        float c00 = v000 * (1.0f - tx) + v100 * tx;
        float c10 = v010 * (1.0f - tx) + v110 * tx;
        float c01 = v001 * (1.0f - tx) + v101 * tx;
        float c11 = v011 * (1.0f - tx) + v111 * tx;
        float c0 = c00 * (1.0f - ty) + c10 * ty;
        float c1 = c01 * (1.0f - ty) + c11 * ty;
        float coarse_val = c0 * (1.0f - tz) + c1 * tz;
        err[t] = std::max(err[t], std::fabs(mesh_val - coarse_val));
      }
    }
    // _apply_subgrid_linearity_kernel (:307-326)
    for (int t = 0; t < total; ++t) {
      if (!required[t]) continue;
      if (err[t] < lin_thr) {
        linear[t] = 1;
        required[t] = 0;
      }
    }
  }
  // exclusive prefix sum of the required flags (:1163-1166)
  int run = 0;
  for (int t = 0; t < total; ++t) {
    address[t] = run;
    run += required[t];
  }
  return run;
}

// _populate_subgrid_texture_{float32,uint16,uint8}_kernel (:363-566) with _write_subgrid_slot (:172-195), then
// the SLOT_LINEAR tags of demoted subgrids (:1275-1288).  slots: (w, h, d), z fastest; texels: ts^3, x fastest,
// ts = 9 blocks.
inline void sdf_pack(const float* fine, int w, int h, int d, const int* required, const int* linear,
                     const int* address, int blocks, int fmt, float qmin, float qinv, uint32_t* slots,
                     unsigned char* texels) {
  const long long nx = 8LL * w + 1, ny = 8LL * h + 1;
  const long long ts = 9LL * blocks;
  for (long long t = 0; t < (long long)w * h * d; ++t) slots[t] = SLOT_EMPTY;
  for (int t = 0; t < w * h * d; ++t) {
    int bz = t / (w * h), r = t - bz * w * h, by = r / w, bx = r - by * w;
    long long si = ((long long)bx * h + by) * d + bz;
    if (linear[t]) slots[si] = SLOT_LINEAR;
    if (!required[t]) continue;
    int a = address[t];
    int az = a / (blocks * blocks), ar = a - az * blocks * blocks, ay = ar / blocks, ax = ar - ay * blocks;
    slots[si] = uint32_t(ax) | (uint32_t(ay) << 10) | (uint32_t(az) << 20);
    for (int k = 0; k < 729; ++k) {
      int lz = k / 81, rr = k - lz * 81, ly = rr / 9, lx = rr - ly * 9;
      float v = fine[((8LL * bz + lz) * ny + 8LL * by + ly) * nx + 8LL * bx + lx];
      long long ti = ((long long)(az * 9 + lz) * ts + (ay * 9 + ly)) * ts + (ax * 9 + lx);
      if (fmt == 4) {
        std::memcpy(texels + 4 * ti, &v, 4);
      } else {
        float vn = tclamp((v - qmin) * qinv, 0.0f, 1.0f);
        if (fmt == 2) {
          uint16_t q = uint16_t(vn * 65535.0f);
          std::memcpy(texels + 2 * ti, &q, 2);
        } else {
          texels[ti] = (unsigned char)(vn * 255.0f);
        }
      }
    }
  }
}

// ---------------------------------------------------------------- mesh pairs
template <class T> struct MeshCandO {
  V3<T> center, normal;
  T dist;
  int sub;
};

// One mesh pair (sa = mesh, sb = mesh or infinite plane).  Candidates are the contacts the reference's mesh
// kernels produce (vertex below gap + margins; edge whose search distance is below gap + margins).  Unreduced
// (mesh_reduce 0): the candidates the writer keeps (collide.py:62-148).  Reduced (1 fast, 2 deterministic
// packing): the reduction's winners, exported with the octahedral round trip of their normal, that the writer
// keeps (contact_reduction_global.py:1327-1440).  Mesh-plane candidates compete as reduce_contact_in_hashtable
// places them (bin of the stored normal, world point, beta scaled by the mesh's AABB diagonal, voxel of the
// mesh); mesh-mesh candidates as export_and_reduce_contact_centered does (bin of the normal, point relative to
// the midpoint of the two shape origins, beta 1e-4 m, voxel of the shape whose edge made the contact).
template <class T>
void mesh_pair(const OracleModel* m, const std::vector<ShapeWorld<T>>& sw, int sa, int sb, Writer<T>& w) {
  const ShapeWorld<T>& A = sw[sa];
  const ShapeWorld<T>& B = sw[sb];
  const T gap_sum = T(m->shape_gap[sa]) + T(m->shape_gap[sb]);
  const int reduce = m->mesh_reduce;
  ContactData<T> c;
  c.r_a = T(0);
  c.r_b = T(0);
  c.m_a = A.margin;
  c.m_b = B.margin;
  c.shape_a = sa;
  c.shape_b = sb;
  c.gap_sum = gap_sum;
  c.sub_key = 0;
  std::vector<MeshCandO<T>> kept;   // unreduced
  std::vector<RedCand<T>> cand;     // reduced
  auto aabb = [&](int s, V3<T>& lo, V3<T>& hi) {
    lo = V3<T>{T(m->shape_aabb_lower[3 * s]), T(m->shape_aabb_lower[3 * s + 1]), T(m->shape_aabb_lower[3 * s + 2])};
    hi = V3<T>{T(m->shape_aabb_upper[3 * s]), T(m->shape_aabb_upper[3 * s + 1]), T(m->shape_aabb_upper[3 * s + 2])};
  };
  auto keep_plain = [&](V3<T> center, V3<T> normal, T dist, int sub) {
    c.center = center;
    c.normal = normal;
    c.dist = dist;
    if (passes_gap(c)) kept.push_back(MeshCandO<T>{center, normal, dist, sub});
  };
  if (m->shape_type[sb] == PLANE) {
    const int mesh = m->shape_mesh[sa];
    const int v0 = m->mesh_vertex_start[mesh], nv = m->mesh_vertex_start[mesh + 1] - v0;
    const Xf<T> Xp_inv = xf_inverse(B.X);
    const V3<T> pn = xf_vector(B.X, V3<T>{T(0), T(0), T(1)});
    const T total_margin = A.margin + B.margin;
    V3<T> lo, hi;
    aabb(sa, lo, hi);
    const T beta = T(RED_BETA) * length(hi - lo);
    const Xf<T> X_inv = xf_inverse(A.X);
    for (int v = 0; v < nv; ++v) {
      const float* pv = m->mesh_vertex + 3 * (v0 + v);
      V3<T> vl{T(pv[0]) * A.gscale.x, T(pv[1]) * A.gscale.y, T(pv[2]) * A.gscale.z};
      V3<T> vw = xf_point(A.X, vl);
      V3<T> vp = xf_point(Xp_inv, vw);
      V3<T> pp = xf_point(B.X, V3<T>{vp.x, vp.y, T(0)});
      T dist = dot(vw - pp, pn);
      if (!(dist < gap_sum + total_margin)) continue;
      const V3<T> center = (vw + pp) * T(0.5);
      if (!reduce) {
        keep_plain(center, -pn, dist, v);
        continue;
      }
      RedCand<T> k;
      k.pos = center;
      k.normal = -pn;
      k.depth = dist;
      k.sub = v;
      k.bin_normal = red_decode_oct(red_encode_oct(k.normal));
      k.proj_point = center;
      k.use_beta = dist < beta;
      k.voxel = red_voxel_index(xf_point(X_inv, center), lo, hi, m->shape_voxel_res + 3 * sa);
      cand.push_back(k);
    }
  } else {
    for (int mode = 0; mode < 2; ++mode) {
      const ShapeWorld<T>& Tr = mode == 0 ? A : B;
      const ShapeWorld<T>& S = mode == 0 ? B : A;
      const int tri_shape = mode == 0 ? sa : sb;
      const int tri_mesh = m->shape_mesh[tri_shape];
      const int sdf_mesh = m->shape_mesh[mode == 0 ? sb : sa];
      const SdfG<T> G = sdf_grid<T>(m, sdf_mesh);
      // a grid cooked from scaled vertices is sampled with unit scale (sdf_contact.py:918-930)
      const V3<T> sc = (m->mesh_sdf_flags[sdf_mesh] & 1) ? V3<T>{T(1), T(1), T(1)} : S.gscale;
      const V3<T> ss{tmax(sc.x, T(1e-10f)), tmax(sc.y, T(1e-10f)), tmax(sc.z, T(1e-10f))};
      const V3<T> inv_s{T(1) / ss.x, T(1) / ss.y, T(1) / ss.z};
      const T min_s = tmin(tmin(ss.x, ss.y), ss.z);
      const Xf<T> X_ts = xf_mul(xf_inverse(S.X), Tr.X);
      const T threshold = gap_sum + Tr.margin + S.margin;
      const T threshold_u = threshold / min_s;
      const int e0 = m->mesh_edge_start[tri_mesh], ne = m->mesh_edge_start[tri_mesh + 1] - e0;
      const int vb = m->mesh_vertex_start[tri_mesh];
      // reduction frame of this side (sdf_contact.py:1286-1324)
      const V3<T> midpoint = (Tr.X.p + S.X.p) * T(0.5);
      const Xf<T> X_ws_tri = xf_inverse(Tr.X);
      V3<T> tlo, thi;
      aabb(tri_shape, tlo, thi);
      for (int e = 0; e < ne; ++e) {
        const int ia = m->mesh_edge[2 * (e0 + e)], ib = m->mesh_edge[2 * (e0 + e) + 1];
        const float* qa = m->mesh_vertex + 3 * (vb + ia);
        const float* qb = m->mesh_vertex + 3 * (vb + ib);
        V3<T> pa{T(qa[0]) * Tr.gscale.x, T(qa[1]) * Tr.gscale.y, T(qa[2]) * Tr.gscale.z};
        V3<T> pb{T(qb[0]) * Tr.gscale.x, T(qb[1]) * Tr.gscale.y, T(qb[2]) * Tr.gscale.z};
        V3<T> ea = xf_point(X_ts, pa), eb = xf_point(X_ts, pb);
        ea = V3<T>{ea.x * inv_s.x, ea.y * inv_s.y, ea.z * inv_s.z};
        eb = V3<T>{eb.x * inv_s.x, eb.y * inv_s.y, eb.z * inv_s.z};
        V3<T> mid = (ea + eb) * T(0.5);
        T rad = length(eb - ea) * T(0.5);
        T cull = rad + threshold_u;
        V3<T> cl{tclamp(mid.x, G.lower.x, G.upper.x), tclamp(mid.y, G.lower.y, G.upper.y),
                 tclamp(mid.z, G.lower.z, G.upper.z)};
        if (length_sq(mid - cl) > cull * cull) continue;
        T msdf = sdf_sample(G, mid);
        if (!(msdf <= cull)) continue;
        V3<T> pu;
        T du0 = edge_sdf_search(G, ea, eb, msdf, pu);
        if (!(du0 * min_s < threshold)) continue;
        V3<T> gu;
        T du = sdf_sample_grad(G, pu, gu);
        T dist = du * min_s;
        V3<T> gsc{gu.x * inv_s.x, gu.y * inv_s.y, gu.z * inv_s.z};
        T gl = length(gsc);
        V3<T> dir = gl > T(0) ? gsc / gl : gu;
        V3<T> pw = xf_point(S.X, V3<T>{pu.x * ss.x, pu.y * ss.y, pu.z * ss.z});
        V3<T> dw = xf_vector(S.X, dir);
        T dl = length(dw);
        if (dl > T(0)) {
          dw = dw / dl;
        } else {
          V3<T> fb = pw - S.X.p;
          T fl = length(fb);
          dw = fl > T(0) ? fb / fl : V3<T>{T(0), T(1), T(0)};
        }
        const V3<T> normal = mode == 0 ? -dw : dw;
        const int sub = (e << 2) | (mode << 1);
        if (!reduce) {
          keep_plain(pw, normal, dist, sub);
          continue;
        }
        RedCand<T> k;
        k.pos = pw;
        k.normal = normal;
        k.depth = dist;
        k.sub = sub;
        k.bin_normal = normal;
        k.proj_point = pw - midpoint;
        k.use_beta = dist < T(RED_BETA);
        k.voxel = red_voxel_index(xf_point(X_ws_tri, pw), tlo, thi, m->shape_voxel_res + 3 * tri_shape);
        cand.push_back(k);
      }
    }
  }
  if (reduce) {
    for (int q : red_reduce(cand, reduce == 2)) {
      const RedCand<T>& k = cand[q];
      kept.push_back(MeshCandO<T>{k.pos, red_decode_oct(red_encode_oct(k.normal)), k.depth, k.sub});
    }
  }
  std::sort(kept.begin(), kept.end(), [](const MeshCandO<T>& x, const MeshCandO<T>& y) { return x.sub < y.sub; });
  for (const MeshCandO<T>& k : kept) {
    c.center = k.center;
    c.normal = k.normal;
    c.dist = k.dist;
    c.sub_key = k.sub;
    write_contact(c, w, reduce != 0);
  }
}

}  // namespace oracle
