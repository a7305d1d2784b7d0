// Narrow-phase device functions: analytic primitive colliders, MPR, GJK and the multi-contact
// manifold, plus the ContactData -> Contacts record conversion.
//
// Behaviour follows (file:line in the reference):
//   narrow_phase_primitive_kernel routing      newton/_src/geometry/narrow_phase.py:136-605
//   analytic colliders                         newton/_src/geometry/collision_primitive.py:48-439, 1161-1215
//   infinite-plane bsphere cull / box proxy    newton/_src/geometry/collision_core.py:493-567, 568-725
//   compute_gjk_mpr_contacts (Minkowski cores) newton/_src/geometry/collision_core.py:278-393
//   solve_convex_multi_contact (MPR inflate)   newton/_src/geometry/collision_convex.py:26-135
//   solve_mpr_core                             newton/_src/geometry/mpr.py:289-505
//   solve_closest_distance_core                newton/_src/geometry/simplex_solver.py:60-492
//   build_manifold + polygon clipping          newton/_src/geometry/multicontact.py:28-956
//   post_process_minkowski_only                newton/_src/geometry/collision_core.py:144-169
//   write_contact                              newton/_src/sim/collide.py:62-148
// One call processes one broad-phase pair on one lane; every contact it keeps lands in one of five
// slots reserved for that pair, with slot k carrying sub_key order, so the compaction that follows
// reproduces the reference's sort-key order without any sort.
#pragma once
#include "nmc_internal.h"
#include "nmc_math.h"

namespace nmc {

struct GShape {
  int type;
  f3 scale;
};

// Per-shape world data computed by compute_shape_aabbs (collide.py:150-275).
struct ShapeW {
  xform X;
  f3 gscale;
  float margin;
  f3 lo, hi;
};

struct CData {  // ContactData (contact_data.py:17-56)
  f3 center, normal;
  float dist, r_a, r_b, m_a, m_b, gap_sum;
  int shape_a, shape_b, sub_key;
};

// Output of one pair: up to 5 records in sub-key order.
// Destination of one pair's contacts: five consecutive staging slots starting at slot0.
struct WriteCtx {
  const float* __restrict__ shape_gap;
  StageBuffers st;
  long long slot0;
  int n;  // records written so far
};

// ALL: the ellipsoid, cylinder and cone maps are compiled in (kernels of models without those shapes leave them
// out, as the reference's lean support does, collide.py:743-759: fewer registers in the MPR/GJK loops)
template <bool ALL = true>
__device__ __forceinline__ f3 support(const GShape& g, f3 d) {
  if (g.type == GEO_BOX) {
    float thr = 1.0e-10f * len(d);
    float sx = d.x >= -thr ? 1.0f : -1.0f;
    float sy = d.y >= -thr ? 1.0f : -1.0f;
    float sz = d.z >= -thr ? 1.0f : -1.0f;
    return {sx * g.scale.x, sy * g.scale.y, sz * g.scale.z};
  }
  if (g.type == GEO_SPHERE) {
    f3 n = len_sq(d) > 1.0e-12f ? normalize(d) : f3{1.0f, 0.0f, 0.0f};
    return n * g.scale.x;
  }
  if (g.type == GEO_CAPSULE) {
    f3 n = len_sq(d) > 1.0e-12f ? normalize(d) : f3{1.0f, 0.0f, 0.0f};
    f3 r = n * g.scale.x;
    if (d.z >= 0.0f) return r + f3{0.0f, 0.0f, g.scale.y};
    return r + f3{0.0f, 0.0f, -g.scale.y};
  }
  if (g.type == GEO_PLANE) {
    float sx = d.x >= 0.0f ? 1.0f : -1.0f;
    float sy = d.y >= 0.0f ? 1.0f : -1.0f;
    return {sx * g.scale.x, sy * g.scale.y, 0.0f};
  }
  if (ALL && g.type == GEO_ELLIPSOID) {
    // (a^2 dx, b^2 dy, c^2 dz) / |(a dx, b dy, c dz)| (support_function.py:218-239)
    const float a = g.scale.x, b = g.scale.y, c = g.scale.z;
    if (len_sq(d) > 1.0e-12f) {
      const float adx = a * d.x, bdy = b * d.y, cdz = c * d.z;
      const float den_sq = adx * adx + bdy * bdy + cdz * cdz;
      if (den_sq > 1.0e-12f) {
        const float den = sqrtf(den_sq);
        return {(a * a) * d.x / den, (b * b) * d.y / den, (c * c) * d.z / den};
      }
    }
    return {a, 0.0f, 0.0f};
  }
  if (ALL && g.type == GEO_CYLINDER) {
    // rim point in the direction's XY projection, on the cap the direction's z picks (support_function.py:240-260)
    const float r = g.scale.x, hh = g.scale.y;
    const f3 dxy{d.x, d.y, 0.0f};
    f3 lat{r, 0.0f, 0.0f};
    if (len_sq(dxy) > 1.0e-12f) {
      const f3 nxy = normalize(dxy);
      lat = f3{nxy.x * r, nxy.y * r, 0.0f};
    }
    if (d.z > 0.0f) return {lat.x, lat.y, hh};
    if (d.z < 0.0f) return {lat.x, lat.y, -hh};
    return lat;
  }
  if (ALL && g.type == GEO_CONE) {
    // apex (0, 0, hh) when dz >= r / (2 hh) |d_xy|, else the base rim at z = -hh (support_function.py:262-285)
    const float r = g.scale.x, hh = g.scale.y;
    const f3 dxy{d.x, d.y, 0.0f};
    const float lxy = len(dxy);
    const float k = hh > 1.0e-12f ? r / (2.0f * hh) : 0.0f;
    if (lxy <= 1.0e-12f) return d.z >= 0.0f ? f3{0.0f, 0.0f, hh} : f3{r, 0.0f, -hh};
    if (d.z >= k * lxy) return {0.0f, 0.0f, hh};
    const f3 nxy = dxy / lxy;
    return {nxy.x * r, nxy.y * r, -hh};
  }
  return {0.0f, 0.0f, 0.0f};
}

// compute_tight_aabb_from_support (collision_core.py:396-491) for the analytic shape types.
__device__ __forceinline__ void tight_aabb(const GShape& g, qf q, f3 pos, f3& lo, f3& hi) {
  m33 R = qmat(q);
  f3 lx{R.m[0][0], R.m[0][1], R.m[0][2]};
  f3 ly{R.m[1][0], R.m[1][1], R.m[1][2]};
  f3 lz{R.m[2][0], R.m[2][1], R.m[2][2]};
  float max_x = dot(lx, support(g, lx));
  float max_y = dot(ly, support(g, ly));
  float max_z = dot(lz, support(g, lz));
  float min_x = dot(lx, support(g, -lx));
  float min_y = dot(ly, support(g, -ly));
  float min_z = dot(lz, support(g, -lz));
  lo = f3{min_x, min_y, min_z} + pos;
  hi = f3{max_x, max_y, max_z} + pos;
}

__device__ __forceinline__ ShapeW shape_world(int s, const float* __restrict__ body_q, const NmcModelDesc& d) {
  ShapeW r;
  int b = d.shape_body[s];
  int type = d.shape_type[s];
  xform Xs = load_xform(d.shape_transform + 7 * s);
  r.X = b == -1 ? Xs : xmul(load_xform(body_q + 7 * b), Xs);
  float margin = d.shape_margin[s];
  float eg = margin + d.shape_gap[s];
  f3 mvec{eg, eg, eg};
  f3 scale = load3(d.shape_scale + 3 * s);
  bool inf_plane = type == GEO_PLANE && scale.x == 0.0f && scale.y == 0.0f;
  f3 gscale = scale;
  if (inf_plane) {
    float rad = d.shape_radius[s];
    f3 he{rad, rad, rad};
    r.lo = r.X.p - he - mvec;
    r.hi = r.X.p + he + mvec;
  } else if (type == GEO_MESH) {
    // local AABB (scale baked in) rotated into the world frame (collide.py:215-245)
    f3 llo = load3(d.shape_aabb_lower + 3 * s), lhi = load3(d.shape_aabb_upper + 3 * s);
    f3 c = (llo + lhi) * 0.5f, h = (lhi - llo) * 0.5f;
    f3 wc = rot(r.X.q, c) + r.X.p;
    f3 r0 = rot(r.X.q, f3{1.0f, 0.0f, 0.0f}), r1 = rot(r.X.q, f3{0.0f, 1.0f, 0.0f});
    f3 r2 = rot(r.X.q, f3{0.0f, 0.0f, 1.0f});
    f3 wh{fabsf(r0.x) * h.x + fabsf(r1.x) * h.y + fabsf(r2.x) * h.z,
          fabsf(r0.y) * h.x + fabsf(r1.y) * h.y + fabsf(r2.y) * h.z,
          fabsf(r0.z) * h.x + fabsf(r1.z) * h.y + fabsf(r2.z) * h.z};
    r.lo = wc - wh - mvec;
    r.hi = wc + wh + mvec;
  } else {
    GShape g;
    g.type = type;
    if (type == GEO_PLANE) gscale = f3{scale.x * 0.5f, scale.y * 0.5f, 0.0f};
    g.scale = gscale;
    f3 lo, hi;
    tight_aabb(g, r.X.q, r.X.p, lo, hi);
    r.lo = lo - mvec;
    r.hi = hi + mvec;
  }
  r.gscale = gscale;
  r.margin = margin;
  return r;
}

// Shape world data as kept in StageBuffers.shape_wc (20 floats per shape).
__device__ __forceinline__ void store_wc(float* __restrict__ o, const ShapeW& w) {
  store_xform(o, w.X);
  store3(o + 7, w.gscale);
  o[10] = w.margin;
  store3(o + 11, w.lo);
  store3(o + 14, w.hi);
}

__device__ __forceinline__ ShapeW load_wc(const float* __restrict__ o) {
  ShapeW s;
  s.X = load_xform(o);
  s.gscale = load3(o + 7);
  s.margin = o[10];
  s.lo = load3(o + 11);
  s.hi = load3(o + 14);
  return s;
}

// ------------------------------------------------------------------------- contact records
__device__ __forceinline__ float contact_d(const CData& c, f3& n, f3& aw, f3& bw) {
  float total = c.r_a + c.r_b + c.m_a + c.m_b;
  n = normalize(c.normal);
  aw = c.center - n * (0.5f * c.dist + c.r_a);
  bw = c.center + n * (0.5f * c.dist + c.r_b);
  f3 diff = bw - aw;
  float distance = dot(diff, n);
  return distance - total;
}

__device__ __forceinline__ bool passes_gap(const CData& c) {
  f3 n, aw, bw;
  float d = contact_d(c, n, aw, bw);
  return d <= c.gap_sum;
}

// write_contact (collide.py:62-148), first half: the staging record of contact c at slot s holds the
// world-frame points, the unit normal, both offsets and the sub key; the compaction applies the inverse
// body transforms (the second half) as it writes the contact out.
__device__ __forceinline__ void write_record(const CData& c, f3 n, f3 aw, f3 bw, const WriteCtx& w, long long s) {
  float off_a = c.r_a + c.m_a, off_b = c.r_b + c.m_b;
  float4* r = w.st.rec + REC_F4 * s;
  r[0] = make_float4(aw.x, aw.y, aw.z, bw.x);
  r[1] = make_float4(bw.y, bw.z, n.x, n.y);
  r[2] = make_float4(n.z, off_a, off_b, __int_as_float(c.sub_key));
}

// write_contact (collide.py:62-148) into the pair's next slot.
__device__ __forceinline__ void emit(const CData& c, WriteCtx& w, bool check_gap) {
  f3 n, aw, bw;
  float d = contact_d(c, n, aw, bw);
  float gap = w.shape_gap[c.shape_a] + w.shape_gap[c.shape_b];
  if (check_gap && d > gap) return;
  if (w.n >= 5) return;
  write_record(c, n, aw, bw, w, w.slot0 + w.n);
  w.n += 1;
}

__device__ __forceinline__ CData post_process(CData c, int type_a, int type_b) {
  f3 n = c.normal;
  if (type_a == GEO_SPHERE || type_a == GEO_CAPSULE) {
    c.center = c.center + n * (c.r_a * 0.5f);
    c.dist = c.dist - c.r_a;
  }
  if (type_b == GEO_SPHERE || type_b == GEO_CAPSULE) {
    c.center = c.center - n * (c.r_b * 0.5f);
    c.dist = c.dist - c.r_b;
  }
  return c;
}

// ------------------------------------------------------------------------- analytic colliders
__device__ __forceinline__ f3 closest_segment_point(f3 a, f3 b, f3 pt) {
  f3 ab = b - a;
  float t = dot(pt - a, ab) / (dot(ab, ab) + 1e-6f);
  return a + clampf(t, 0.0f, 1.0f) * ab;
}

__device__ __forceinline__ void plane_sphere(f3 n, f3 pp, f3 sp, float r, float& dist, f3& pos) {
  dist = dot(sp - pp, n) - r;
  pos = sp - n * (r + 0.5f * dist);
}

__device__ __forceinline__ void sphere_sphere(f3 p1, float r1, f3 p2, float r2, float& dist, f3& pos, f3& n) {
  f3 dir = p2 - p1;
  float d = len(dir);
  n = d == 0.0f ? f3{1.0f, 0.0f, 0.0f} : dir / d;
  d = d - (r1 + r2);
  pos = p1 + n * (r1 + 0.5f * d);
  dist = d;
}

__device__ __forceinline__ void capsule_capsule(f3 p1, f3 ax1, float r1, float hl1, f3 p2, f3 ax2, float r2, float hl2,
                                float dist[2], f3 pos[2], f3& normal) {
  const float MINVAL = 1e-15f;
  dist[0] = dist[1] = MAXVAL;
  pos[0] = pos[1] = f3{0.0f, 0.0f, 0.0f};
  normal = f3{0.0f, 0.0f, 0.0f};
  f3 axis1 = ax1 * hl1, axis2 = ax2 * hl2;
  f3 dif = p1 - p2;
  float ma = dot(axis1, axis1);
  float mb = -dot(axis1, axis2);
  float mc = dot(axis2, axis2);
  float u = -dot(axis1, dif);
  float v = dot(axis2, dif);
  float det = ma * mc - mb * mb;
  if (fabsf(det) >= MINVAL) {
    float inv_det = 1.0f / det;
    float x1 = (mc * u - mb * v) * inv_det;
    float x2 = (ma * v - mb * u) * inv_det;
    if (x1 > 1.0f) { x1 = 1.0f; x2 = (v - mb) / mc; }
    else if (x1 < -1.0f) { x1 = -1.0f; x2 = (v + mb) / mc; }
    if (x2 > 1.0f) { x2 = 1.0f; x1 = clampf((u - mb) / ma, -1.0f, 1.0f); }
    else if (x2 < -1.0f) { x2 = -1.0f; x1 = clampf((u + mb) / ma, -1.0f, 1.0f); }
    f3 vec1 = p1 + axis1 * x1, vec2 = p2 + axis2 * x2;
    sphere_sphere(vec1, r1, vec2, r2, dist[0], pos[0], normal);
  } else {
    f3 vec1 = p1 + axis1;
    float x2 = clampf((v - mb) / mc, -1.0f, 1.0f);
    f3 vec2 = p2 + axis2 * x2;
    sphere_sphere(vec1, r1, vec2, r2, dist[0], pos[0], normal);
    vec1 = p1 - axis1;
    x2 = clampf((v + mb) / mc, -1.0f, 1.0f);
    vec2 = p2 + axis2 * x2;
    f3 n2;
    sphere_sphere(vec1, r1, vec2, r2, dist[1], pos[1], n2);
  }
}

// The four kept corners live in named registers (k0..k3) and are addressed by compile-time index only: an
// array indexed by the running slot number would be placed in scratch memory.  Same selection as the
// reference: the first four corners within the margin, then each further one replaces the current worst
// (largest distance) when it is deeper.
__device__ __forceinline__ void plane_box(f3 n, f3 pp, f3 bp, const m33& R, f3 size, float margin, float dist[4], f3 pos[4]) {
  float center_dist = dot(bp - pp, n);
  float d0 = MAXVAL, d1 = MAXVAL, d2 = MAXVAL, d3 = MAXVAL;
  const f3 zero{0.0f, 0.0f, 0.0f};
  f3 q0 = zero, q1 = zero, q2 = zero, q3 = zero;
  int ncontact = 0, worst = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    f3 corner{(i & 1) ? size.x : -size.x, (i & 2) ? size.y : -size.y, (i & 4) ? size.z : -size.z};
    corner = mv(R, corner);
    float ldist = dot(n, corner);
    float cdist = center_dist + ldist;
    if (cdist > margin) continue;
    f3 cpos = corner + bp - (0.5f * n) * cdist;
    const float dw = worst == 0 ? d0 : (worst == 1 ? d1 : (worst == 2 ? d2 : d3));
    int slot;
    if (ncontact < 4) {
      slot = ncontact;
    } else if (cdist < dw) {
      slot = worst;
    } else {
      continue;
    }
    if (slot == 0) { d0 = cdist; q0 = cpos; }
    else if (slot == 1) { d1 = cdist; q1 = cpos; }
    else if (slot == 2) { d2 = cdist; q2 = cpos; }
    else { d3 = cdist; q3 = cpos; }
    if (ncontact < 4) {
      // dw is an earlier corner's distance (worst < ncontact): unchanged by the store above
      if (ncontact == 0 || cdist > dw) worst = ncontact;
      ncontact += 1;
    } else {
      worst = 0;
      if (d1 > d0) worst = 1;
      if (d2 > (worst == 0 ? d0 : d1)) worst = 2;
      if (d3 > (worst == 0 ? d0 : (worst == 1 ? d1 : d2))) worst = 3;
    }
  }
  dist[0] = d0; dist[1] = d1; dist[2] = d2; dist[3] = d3;
  pos[0] = q0; pos[1] = q1; pos[2] = q2; pos[3] = q3;
}

// collide_plane_ellipsoid (collision_primitive.py:334-363): the ellipsoid's support point against the plane's
// normal, through the unit sphere it is the image of.
__device__ __forceinline__ void plane_ellipsoid(f3 n, f3 pp, f3 ep, const m33& R, f3 size, float& dist, f3& pos) {
  const f3 t = mtv(R, n);
  const f3 ss = -normalize(f3{t.x * size.x, t.y * size.y, t.z * size.z});
  const f3 p = ep + mv(R, f3{ss.x * size.x, ss.y * size.y, ss.z * size.z});
  dist = dot(n, p - pp);
  pos = p - (n * dist) * 0.5f;
}

// collide_sphere_cylinder (collision_primitive.py:443-513): side, cap or rim, by where the centre projects.
__device__ __forceinline__ void sphere_cylinder(f3 sp, float sr, f3 cp, f3 axis, float cr, float chh, float& dist, f3& pos, f3& n) {
  const f3 vec = sp - cp;
  const float x = dot(vec, axis);
  const f3 a_proj = axis * x;
  const f3 p_proj = vec - a_proj;
  const float pps = dot(p_proj, p_proj);
  bool side = fabsf(x) < chh;
  bool cap = pps < cr * cr;
  if (side && cap) {
    const float dcap = chh - fabsf(x);
    const float drad = cr - sqrtf(pps);
    if (dcap < drad) side = false;
    else cap = false;
  }
  if (side) {
    sphere_sphere(sp, sr, cp + a_proj, cr, dist, pos, n);
  } else if (cap) {
    const f3 pc = x > 0.0f ? cp + axis * chh : cp - axis * chh;
    const f3 pn = x > 0.0f ? axis : -axis;
    plane_sphere(pn, pc, sp, sr, dist, pos);
    n = -pn;
  } else {
    const float l = sqrtf(pps);
    const float inv = 1.0f / (l != 0.0f ? l : 1.0e-15f);  // safe_div (newton/_src/math/__init__.py:282)
    const f3 rim = p_proj * (cr * inv);
    const f3 cap_off = axis * ((x < 0.0f ? -1.0f : 1.0f) * chh);
    sphere_sphere(sp, sr, (cp + cap_off) + rim, 0.0f, dist, pos, n);
  }
}

// collide_plane_cylinder (collision_primitive.py:516-669): the deepest rim point, then (axis within 22.5 degrees
// of the plane normal) a tripod on the near cap in a world-anchored orientation, or (rolling) the two side
// generators through the deepest radial direction and the near-cap rim point facing the plane; extra points
// closer than 1 % of max(radius, half height) to the deepest one are dropped.  Kept points live in named
// registers (no runtime-indexed array).
__device__ __forceinline__ void plane_cylinder(f3 n, f3 pp, f3 cp, f3 axis_in, float cr, float chh, float dist[4], f3 pos[4]) {
  f3 axis = axis_in;
  float dna = dot(n, axis);
  if (dna > 0.0f) {
    axis = -axis;
    dna = -dna;
  }
  const f3 capc = cp + axis * chh;
  f3 pal = -n + axis * dna;
  const float pal2 = dot(pal, pal);
  const bool has_align = pal2 > 1.0e-10f;
  if (has_align) pal = pal * (1.0f / sqrtf(pal2));
  f3 ref{1.0f, 0.0f, 0.0f};
  if (fabsf(dot(axis, ref)) > 0.9f) ref = f3{0.0f, 1.0f, 0.0f};
  const f3 pfix = normalize(ref - axis * dot(axis, ref));
  const float abs_dot = -dna;
  const bool flat = abs_dot >= (float)0.92387953251128674;  // cos(22.5 degrees)
  const f3 dperp = has_align ? pal : pfix;
  const f3 dpt = capc + dperp * cr;
  const float dd = dot(dpt - pp, n);
  const f3 dpos = dpt - n * (dd * 0.5f);
  const f3 zero{0.0f, 0.0f, 0.0f};
  float d1 = MAXVAL, d2 = MAXVAL, d3 = MAXVAL;
  f3 q1 = zero, q2 = zero, q3 = zero;
  int nc = 1;
  const float mt = 0.01f * fmax2(cr, chh);
  const float mt2 = mt * mt;
  auto add = [&](f3 pt) {
    const float d = dot(pt - pp, n);
    const f3 q = pt - n * (d * 0.5f);
    if (nc < 4 && len_sq(q - dpos) > mt2) {
      if (nc == 1) { d1 = d; q1 = q; }
      else if (nc == 2) { d2 = d; q2 = q; }
      else { d3 = d; q3 = q; }
      nc += 1;
    }
  };
  if (flat) {
    const f3 u = pfix * cr, v = cross(axis, pfix) * cr;
    const float c120 = -0.5f, s120 = 0.8660254f;
    add(capc + u);
    add((capc + c120 * u) + s120 * v);
    add((capc + c120 * u) - s120 * v);
  } else {
    const f3 pr = has_align ? pal : pfix;
    const f3 u = pr * cr, v = cross(axis, pr) * cr;
    add((cp + axis * chh) + u);
    add((cp - axis * chh) + u);
    const f3 pv = capc + v, nv = capc - v;
    const float dpv = dot(pv - pp, n), dnv = dot(nv - pp, n);
    add(dpv <= dnv ? pv : nv);
  }
  dist[0] = dd; dist[1] = d1; dist[2] = d2; dist[3] = d3;
  pos[0] = dpos; pos[1] = q1; pos[2] = q2; pos[3] = q3;
}

__device__ void sphere_box(f3 sp, float r, f3 bp, const m33& R, f3 size, float& dist, f3& pos_out, f3& normal) {
  f3 center = mtv(R, sp - bp);
  f3 clamped = vmax(-size, vmin(size, center));
  f3 cdir = clamped - center;
  float cn = len(cdir);
  if (cn != 0.0f) cdir = cdir / cn;
  f3 pos;
  if (cn <= 1e-6f) {
    float closest = 2.0f * (size.x + size.y + size.z);
    int k = 0;
    for (int i = 0; i < 6; ++i) {
      float fd = fabsf(((i % 2) ? 1.0f : -1.0f) * get(size, i / 2) - get(center, i / 2));
      if (closest > fd) { closest = fd; k = i; }
    }
    f3 nearest{0.0f, 0.0f, 0.0f};
    put(nearest, k / 2, (k % 2) ? -1.0f : 1.0f);
    pos = center + nearest * (r - closest) / 2.0f;
    normal = mv(R, nearest);
    dist = -closest - r;
  } else {
    f3 deepest = center + cdir * r;
    pos = 0.5f * (clamped + deepest);
    normal = mv(R, cdir);
    dist = cn - r;
  }
  pos_out = bp + mv(R, pos);
}

// ------------------------------------------------------------------------- MPR
struct Vert { f3 B, A; };  // A = BtoA

template <bool ALL>
__device__ __forceinline__ Vert mink(const GShape& ga, const GShape& gb, f3 dir, qf ob, f3 pb, float extend) {
  Vert v;
  f3 pa = support<ALL>(ga, dir);
  f3 nd = -dir;
  f3 tmp = rot_inv(ob, nd);
  f3 sb = rot(ob, support<ALL>(gb, tmp));
  v.B = sb + pb;
  if (extend != 0.0f) {
    f3 d = normalize(dir) * extend * 0.5f;
    pa = pa + d;
    v.B = v.B - d;
  }
  v.A = pa - v.B;
  return v;
}

__device__ __forceinline__ Vert geo_center(qf ob, f3 pb) {
  f3 zero{0.0f, 0.0f, 0.0f};
  Vert c;
  f3 cbw = pb + rot(ob, zero);
  c.B = cbw;
  c.A = zero - cbw;
  return c;
}

struct CvxRes {
  bool flag;
  f3 pa, pb, n;
  float value;
};

template <bool ALL>
__device__ CvxRes mpr_core(const GShape& ga, const GShape& gb, qf ob, f3 pb, float extend) {
  const int MAX_ITER = 30;
  const float CE = 1e-5f;
  const float NE = 1e-16f;
  float pen = 0.0f;
  f3 zero{0.0f, 0.0f, 0.0f};
  f3 point_a = zero, point_b = zero, normal;
  Vert v0 = geo_center(ob, pb);
  normal = v0.A;
  if (len_sq(normal) < NE) {
    float best_dot = -1.0e30f;
    f3 best_dir{1.0f, 0.0f, 0.0f};
    for (int ax = 0; ax < 3; ++ax) {
      f3 probe = zero;
      put(probe, ax, 1.0f);
      Vert sv = mink<ALL>(ga, gb, probe, ob, pb, extend);
      float d = dot(sv.A, probe);
      if (d > best_dot) { best_dot = d; best_dir = probe; }
    }
    v0.A = best_dir * 1e-05f;
  }
  normal = -v0.A;
  Vert v1 = mink<ALL>(ga, gb, normal, ob, pb, extend);
  point_a = v1.B + v1.A;
  point_b = v1.B;
  if (dot(v1.A, normal) <= 0.0f) return {false, point_a, point_b, normal, pen};
  normal = cross(v1.A, v0.A);
  if (len_sq(normal) < NE * NE) {
    normal = v1.A - v0.A;
    normal = normalize(normal);
    pen = dot(v1.A, normal);
    return {true, point_a, point_b, normal, pen};
  }
  Vert v2 = mink<ALL>(ga, gb, normal, ob, pb, extend);
  if (dot(v2.A, normal) <= 0.0f) return {false, point_a, point_b, normal, pen};
  f3 t1 = v1.A - v0.A;
  f3 t2 = v2.A - v0.A;
  normal = cross(t1, t2);
  float dist = dot(normal, v0.A);
  if (dist > 0.0f) {
    Vert tv = v1;
    v1 = v2;
    v2 = tv;
    normal = -normal;
  }
  int phase1 = 0, phase2 = 0;
  bool hit = false;
  Vert v3;
  while (true) {
    if (phase1 > MAX_ITER) return {false, point_a, point_b, normal, pen};
    phase1 += 1;
    v3 = mink<ALL>(ga, gb, normal, ob, pb, extend);
    if (dot(v3.A, normal) <= 0.0f) return {false, point_a, point_b, normal, pen};
    t1 = cross(v1.A, v3.A);
    if (dot(t1, v0.A) < 0.0f) {
      v2 = v3;
      t1 = v1.A - v0.A;
      t2 = v3.A - v0.A;
      normal = cross(t1, t2);
      continue;
    }
    t1 = cross(v3.A, v2.A);
    if (dot(t1, v0.A) < 0.0f) {
      v1 = v3;
      t1 = v3.A - v0.A;
      t2 = v2.A - v0.A;
      normal = cross(t1, t2);
      continue;
    }
    break;
  }
  while (true) {
    phase2 += 1;
    t1 = v2.A - v1.A;
    t2 = v3.A - v1.A;
    normal = cross(t1, t2);
    float normal_sq = len_sq(normal);
    if (normal_sq < NE * NE) return {false, point_a, point_b, normal, pen};
    if (!hit) {
      float d = dot(normal, v1.A);
      hit = d >= 0.0f;
    }
    Vert v4 = mink<ALL>(ga, gb, normal, ob, pb, extend);
    f3 t3 = v4.A - v3.A;
    float delta = dot(t3, normal);
    pen = dot(v4.A, normal);
    if (delta * delta <= CE * CE * normal_sq || pen <= 0.0f || phase2 > MAX_ITER) {
      if (hit) {
        float inv_normal = 1.0f / sqrtf(normal_sq);
        pen *= inv_normal;
        normal = normal * inv_normal;
        t3 = cross(v1.A, t1);
        float gamma = dot(t3, normal) * inv_normal;
        t3 = cross(t2, v1.A);
        float beta = dot(t3, normal) * inv_normal;
        float alpha = 1.0f - gamma - beta;
        point_a = alpha * (v1.B + v1.A) + beta * (v2.B + v2.A) + gamma * (v3.B + v3.A);
        point_b = alpha * v1.B + beta * v2.B + gamma * v3.B;
      }
      return {hit, point_a, point_b, normal, pen};
    }
    t1 = cross(v4.A, v0.A);
    float dd = dot(t1, v1.A);
    if (dd >= 0.0f) {
      dd = dot(t1, v2.A);
      if (dd >= 0.0f) v1 = v4; else v3 = v4;
    } else {
      dd = dot(t1, v3.A);
      if (dd >= 0.0f) v2 = v4; else v1 = v4;
    }
  }
}

// ------------------------------------------------------------------------- GJK
// Lanes per workgroup of the kernels that run the narrow phase (the per-lane LDS buffers are sized by it).
constexpr int kNarrowBlock = 256;

// The simplex's vertices live in LDS, one lane's float f at s_smp[f * kNarrowBlock + t] (slot numbers are
// run-time values, and a register array indexed by them would be placed in scratch memory): A of slot i at
// floats 3i..3i+2, B at 12 + 3i..; barycentric weights and the slot mask stay in registers, written
// through the compile-time-indexed selects of sset.
struct SimplexVerts {
  float* base;  // s_smp + lane
  __device__ __forceinline__ f3 get(int f0) const {
    return {base[f0 * kNarrowBlock], base[(f0 + 1) * kNarrowBlock], base[(f0 + 2) * kNarrowBlock]};
  }
  __device__ __forceinline__ void put(int f0, f3 v) const {
    base[f0 * kNarrowBlock] = v.x;
    base[(f0 + 1) * kNarrowBlock] = v.y;
    base[(f0 + 2) * kNarrowBlock] = v.z;
  }
  __device__ __forceinline__ f3 A(int i) const { return get(3 * i); }
  __device__ __forceinline__ f3 B(int i) const { return get(12 + 3 * i); }
  __device__ __forceinline__ void setA(int i, f3 v) const { put(3 * i, v); }
  __device__ __forceinline__ void setB(int i, f3 v) const { put(12 + 3 * i, v); }
};

struct Simplex {
  SimplexVerts v;
  float bc[4];
  unsigned mask;
};

__device__ __forceinline__ void sset(float (&a)[4], int i, float v) {
  a[0] = i == 0 ? v : a[0];
  a[1] = i == 1 ? v : a[1];
  a[2] = i == 2 ? v : a[2];
  a[3] = i == 3 ? v : a[3];
}

__device__ __forceinline__ f3 smp_segment(const Simplex& s, int i0, int i1, float (&bc)[4], unsigned& mask) {
  const float EPS = 1e-8f;
  f3 a = s.v.A(i0), b = s.v.A(i1);
  f3 edge = b - a;
  float vsq = len_sq(edge);
  bool degenerate = vsq < EPS;
  float denom = degenerate ? EPS : vsq;
  float t = -dot(a, edge) / denom;
  float l0 = 1.0f - t, l1 = t;
  mask = (1u << i0) | (1u << i1);
  bc[0] = bc[1] = bc[2] = bc[3] = 0.0f;
  if (l0 < 0.0f || degenerate) { mask = 1u << i1; l0 = 0.0f; l1 = 1.0f; }
  else if (l1 < 0.0f) { mask = 1u << i0; l0 = 1.0f; l1 = 0.0f; }
  sset(bc, i0, l0);
  sset(bc, i1, l1);
  return l0 * a + l1 * b;
}

__device__ __forceinline__ f3 smp_triangle(const Simplex& s, int i0, int i1, int i2, float (&bc)[4], unsigned& mask) {
  const float EPS = 1e-8f;
  f3 a = s.v.A(i0), b = s.v.A(i1), c = s.v.A(i2);
  f3 u = a - b, w = a - c;
  f3 normal = cross(u, w);
  float t = len_sq(normal);
  bool degenerate = t < EPS;
  float denom = degenerate ? EPS : t;
  float it = 1.0f / denom;
  f3 c1 = cross(u, a), c2 = cross(a, w);
  float l2 = dot(c1, normal) * it;
  float l1 = dot(c2, normal) * it;
  float l0 = 1.0f - l2 - l1;
  float best = 1e30f;
  f3 closest_pt{0.0f, 0.0f, 0.0f};
  float bct[4];
  bc[0] = bc[1] = bc[2] = bc[3] = 0.0f;
  mask = 0;
  unsigned m;
  if (l0 < 0.0f || degenerate) {
    f3 cl = smp_segment(s, i1, i2, bct, m);
    float dist = len_sq(cl);
    if (dist < best) {
#pragma unroll
      for (int k = 0; k < 4; ++k) bc[k] = bct[k];
      mask = m; best = dist; closest_pt = cl;
    }
  }
  if (l1 < 0.0f || degenerate) {
    f3 cl = smp_segment(s, i0, i2, bct, m);
    float dist = len_sq(cl);
    if (dist < best) {
#pragma unroll
      for (int k = 0; k < 4; ++k) bc[k] = bct[k];
      mask = m; best = dist; closest_pt = cl;
    }
  }
  if (l2 < 0.0f || degenerate) {
    f3 cl = smp_segment(s, i0, i1, bct, m);
    float dist = len_sq(cl);
    if (dist < best) {
#pragma unroll
      for (int k = 0; k < 4; ++k) bc[k] = bct[k];
      mask = m; closest_pt = cl;
    }
  }
  if (mask != 0) return closest_pt;
  sset(bc, i0, l0);
  sset(bc, i1, l1);
  sset(bc, i2, l2);
  mask = (1u << i0) | (1u << i1) | (1u << i2);
  return l0 * a + l1 * b + l2 * c;
}

__device__ __forceinline__ float det4(f3 a, f3 b, f3 c, f3 d) { return dot(b - a, cross(c - a, d - a)); }

__device__ __forceinline__ f3 smp_tetra(const Simplex& s, float (&bc)[4], unsigned& mask) {
  const float EPS = 1e-8f;
  f3 v0 = s.v.A(0), v1 = s.v.A(1), v2 = s.v.A(2), v3 = s.v.A(3);
  float det_t = det4(v0, v1, v2, v3);
  bool degenerate = fabsf(det_t) < EPS;
  float denom = degenerate ? EPS : det_t;
  float inv = 1.0f / denom;
  f3 zero{0.0f, 0.0f, 0.0f};
  float lam[4];
  lam[0] = det4(zero, v1, v2, v3) * inv;
  lam[1] = det4(v0, zero, v2, v3) * inv;
  lam[2] = det4(v0, v1, zero, v3) * inv;
  lam[3] = 1.0f - lam[0] - lam[1] - lam[2];
  float best = 1e30f;
  f3 closest_pt = zero;
  float bct[4];
  bc[0] = bc[1] = bc[2] = bc[3] = 0.0f;
  mask = 0;
  unsigned m;
  const int F[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    if (lam[f] < 0.0f || degenerate) {
      f3 cl = smp_triangle(s, F[f][0], F[f][1], F[f][2], bct, m);
      float dist = len_sq(cl);
      if (dist < best) {
#pragma unroll
        for (int k = 0; k < 4; ++k) bc[k] = bct[k];
        mask = m;
        if (f < 3) best = dist;
        closest_pt = cl;
      }
    }
  }
  if (mask != 0) return closest_pt;
  bc[0] = lam[0]; bc[1] = lam[1]; bc[2] = lam[2]; bc[3] = lam[3];
  mask = 15u;
  return zero;
}

__device__ __forceinline__ void smp_closest(const Simplex& s, f3& pa, f3& pb) {
  pa = f3{0.0f, 0.0f, 0.0f};
  pb = pa;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if ((s.mask & (1u << i)) == 0) continue;
    float b = s.bc[i];
    const f3 sb = s.v.B(i);
    pa = pa + b * (sb + s.v.A(i));
    pb = pb + b * sb;
  }
}

template <bool ALL>
__device__ CvxRes gjk_core(const GShape& ga, const GShape& gb, qf ob, f3 pb, float extend) {
  const int MAX_ITER = 30;
  const float CE = 1e-4f;
  const float EPS = 1e-8f;
  f3 zero{0.0f, 0.0f, 0.0f};
  f3 point_a = zero, point_b = zero, normal = zero;
  __shared__ float s_smp[24 * kNarrowBlock];
  Simplex s;
  s.v = SimplexVerts{s_smp + threadIdx.x};
#pragma unroll
  for (int i = 0; i < 4; ++i) { s.v.setB(i, zero); s.v.setA(i, zero); s.bc[i] = 0.0f; }
  s.mask = 0;
  int iter = MAX_ITER;
  Vert center = geo_center(ob, pb);
  f3 v = center.A;
  float dist_sq = len_sq(v);
  f3 last_search{1.0f, 0.0f, 0.0f};
  while (iter > 0) {
    iter -= 1;
    if (dist_sq < CE * CE) {
      smp_closest(s, point_a, point_b);
      return {false, point_a, point_b, zero, 0.0f};
    }
    f3 search = -v;
    last_search = search;
    Vert w = mink<ALL>(ga, gb, search, ob, pb, extend);
    f3 wv = w.A;
    float delta_dist = dot(v, v - wv);
    if (delta_dist < CE * sqrtf(dist_sq)) break;
    bool dup = false;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (!dup && (s.mask & (1u << i))) {
        if (len_sq(s.v.A(i) - wv) < CE * CE) dup = true;
      }
    }
    if (dup) break;
    // used slots in ascending order, then the last free slot (idx0..idx3)
    int use_count = 0, free_slot = 0;
    int idx0 = 0, idx1 = 0, idx2 = 0, idx3 = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (s.mask & (1u << i)) {
        idx0 = use_count == 0 ? i : idx0;
        idx1 = use_count == 1 ? i : idx1;
        idx2 = use_count == 2 ? i : idx2;
        idx3 = use_count == 3 ? i : idx3;
        use_count += 1;
      } else {
        free_slot = i;
      }
    }
    idx0 = use_count == 0 ? free_slot : idx0;
    idx1 = use_count == 1 ? free_slot : idx1;
    idx2 = use_count == 2 ? free_slot : idx2;
    idx3 = use_count == 3 ? free_slot : idx3;
    use_count += 1;
    s.v.setB(free_slot, w.B);
    s.v.setA(free_slot, w.A);
    f3 closest = zero;
    bool success = true;
    float bc[4];
    unsigned mask;
    if (use_count == 1) {
      closest = s.v.A(idx0);
      s.mask = 1u << idx0;
      sset(s.bc, idx0, 1.0f);
    } else if (use_count == 2) {
      closest = smp_segment(s, idx0, idx1, bc, mask);
#pragma unroll
      for (int k = 0; k < 4; ++k) s.bc[k] = bc[k];
      s.mask = mask;
    } else if (use_count == 3) {
      closest = smp_triangle(s, idx0, idx1, idx2, bc, mask);
#pragma unroll
      for (int k = 0; k < 4; ++k) s.bc[k] = bc[k];
      s.mask = mask;
    } else if (use_count == 4) {
      closest = smp_tetra(s, bc, mask);
#pragma unroll
      for (int k = 0; k < 4; ++k) s.bc[k] = bc[k];
      s.mask = mask;
      success = !(mask == 15u);
    } else {
      success = false;
    }
    (void)idx3;
    if (!success) {
      smp_closest(s, point_a, point_b);
      return {false, point_a, point_b, zero, 0.0f};
    }
    v = closest;
    dist_sq = len_sq(v);
  }
  float distance = sqrtf(dist_sq);
  smp_closest(s, point_a, point_b);
  f3 delta = point_b - point_a;
  float dls = len_sq(delta);
  if (dls > EPS * EPS) {
    normal = delta * (1.0f / sqrtf(dls));
  } else if (distance > CE) {
    normal = v * (-1.0f / distance);
  } else {
    float nsq = len_sq(last_search);
    normal = nsq > 0.0f ? last_search * (1.0f / sqrtf(nsq)) : f3{1.0f, 0.0f, 0.0f};
  }
  return {true, point_a, point_b, normal, distance};
}

// ------------------------------------------------------------------------- manifold
__device__ __forceinline__ float sarea(f2 a, f2 b, f2 q) { return (b.x - a.x) * (q.y - a.y) - (b.y - a.y) * (q.x - a.x); }

__device__ __forceinline__ f2 isect(f2 s, f2 e, f2 a, f2 b) {
  float sa = sarea(s, e, a);
  float sb = sarea(s, e, b);
  float t = fabsf(sa) / fabsf(sa - sb);
  return (1.0f - t) * a + t * b;
}

// Clip buffer entries.  Two convex polygons of at most five vertices intersect in at most ten, so the
// guards below never fire on well-formed input; they keep every index inside the buffer regardless.
constexpr int kClipMax = 12;

// The clip buffer of one lane, kept in LDS rather than in scratch (the polygon loops index it by value):
// entry i of lane t sits at s_poly[i * kNarrowBlock + t], so a wave's accesses to one entry hit
// consecutive 8-byte words.
struct LanePoly {
  f2* base;
  __device__ __forceinline__ f2& operator[](int i) const { return base[i * kNarrowBlock]; }
  __device__ __forceinline__ LanePoly operator+(int o) const { return LanePoly{base + o * kNarrowBlock}; }
};

// Sutherland-Hodgman clip of `loop` (in place, in the shared 12-entry buffer) by the line s->e.
template <class Buf>
__device__ int trim_in_place(f2 s, f2 e, const Buf& buf, int lo, int loop_count) {
  if (loop_count < 3) return loop_count;
  f2 ia{0.0f, 0.0f}, ib{0.0f, 0.0f};
  int change_a = -1, change_b = -1;
  bool keep = false;
  bool prev_outside = sarea(s, e, buf[lo]) <= 0.0f;
  for (int i = 0; i < loop_count; ++i) {
    int nx = (i + 1) % loop_count;
    bool outside = sarea(s, e, buf[lo + nx]) <= 0.0f;
    if (outside != prev_outside) {
      f2 inter = isect(s, e, buf[lo + i], buf[lo + nx]);
      if (change_a < 0) { change_a = i; keep = !prev_outside; ia = inter; }
      else { change_b = i; ib = inter; }
    }
    prev_outside = outside;
  }
  int nlc;
  if (change_a >= 0 && change_b >= 0) {
    int li = -1;
    nlc = loop_count;
    int i = 0;
    while (i < loop_count) {
      if (keep) { li += 1; if (li < kClipMax && i < kClipMax) buf[lo + li] = buf[lo + i]; }
      if (i == change_a || i == change_b) {
        f2 pt = i == change_a ? ia : ib;
        if (li == i && !keep) {
          li += 1;
          int j = nlc < kClipMax - 1 ? nlc : kClipMax - 1;
          while (j > li) { buf[lo + j] = buf[lo + j - 1]; j -= 1; }
          if (li < kClipMax) buf[lo + li] = pt;
          nlc += 1;
          i += 1;
          change_b += 1;
          loop_count += 1;
        } else {
          li += 1;
          if (li < kClipMax) buf[lo + li] = pt;
        }
        keep = !keep;
      }
      i += 1;
    }
    nlc = li + 1 < kClipMax ? li + 1 : kClipMax;
  } else if (prev_outside) {
    nlc = 0;
  } else {
    nlc = loop_count;
  }
  return nlc;
}

template <class Buf>
__device__ __forceinline__ bool seg_rect(const Buf& buf, int o) {
  const float move = 1e-5f;
  f2 p0 = buf[o], p1 = buf[o + 1];
  float dx = p1.x - p0.x, dy = p1.y - p0.y;
  float dl = sqrtf(dx * dx + dy * dy);
  if (!(dl > 1e-10f)) return false;
  float px = -dy / dl, py = dx / dl;
  float ox = px * move, oy = py * move;
  buf[o] = f2{p0.x - ox, p0.y - oy};
  buf[o + 1] = f2{p1.x - ox, p1.y - oy};
  buf[o + 2] = f2{p1.x + ox, p1.y + oy};
  buf[o + 3] = f2{p0.x + ox, p0.y + oy};
  return true;
}

// trim polygon lives at buf[5..], loop at buf[0..] (the reference aliases them the same way)
template <class Buf>
__device__ int trim_all(const Buf& buf, int tcount, int lcount) {
  const int T0 = 5;
  if (tcount <= 1) return lcount < 1 ? lcount : 1;
  if (tcount == 2) {
    if (!seg_rect(buf, T0)) return lcount < 1 ? lcount : 1;
    tcount = 4;
  }
  if (lcount == 2) {
    if (!seg_rect(buf, 0)) return lcount < 1 ? lcount : 1;
    lcount = 4;
  }
  int cur = lcount;
  f2 t0 = buf[T0];
  for (int i = 0; i < tcount; ++i) {
    f2 s = buf[T0 + i];
    f2 e = i == tcount - 1 ? t0 : buf[T0 + i + 1];
    cur = trim_in_place(s, e, buf, 0, cur);
  }
  return cur;
}

template <class Buf>
__device__ __forceinline__ void calipers(const Buf& hull, int n, int out[4]) {
  int p1 = 0, p3 = 1;
  f2 d0 = hull[p1] - hull[p3];
  float max_dist_sq = d0.x * d0.x + d0.y * d0.y;
  const float tie = 1.0e-3f;
  int j = 1;
  for (int i = 0; i < n; ++i) {
    f2 hi = hull[i], hi1 = hull[(i + 1) % n];
    while (true) {
      f2 hj = hull[j], hj1 = hull[(j + 1) % n];
      float a1 = sarea(hi, hi1, hj1);
      float a0 = sarea(hi, hi1, hj);
      if (a1 > a0) j = (j + 1) % n; else break;
    }
    f2 hj = hull[j];
    f2 d1{hi.x - hj.x, hi.y - hj.y};
    float ds1 = d1.x * d1.x + d1.y * d1.y;
    if (ds1 > max_dist_sq * (1.0f + tie)) { max_dist_sq = ds1; p1 = i; p3 = j; }
    f2 d2{hi1.x - hj.x, hi1.y - hj.y};
    float ds2 = d2.x * d2.x + d2.y * d2.y;
    if (ds2 > max_dist_sq * (1.0f + tie)) { max_dist_sq = ds2; p1 = (i + 1) % n; p3 = j; }
  }
  int p2 = 0, p4 = 0;
  float ma1 = 0.0f, ma2 = 0.0f;
  for (int i = 0; i < n; ++i) {
    float area = sarea(hull[p1], hull[p3], hull[i]);
    if (area > ma1 * (1.0f + tie)) { ma1 = area; p2 = i; }
    else if (-area > ma2 * (1.0f + tie)) { ma2 = -area; p4 = i; }
  }
  out[0] = p1; out[1] = p2; out[2] = p3; out[3] = p4;
}

template <class Buf>
__device__ int remove_zero_edges(const Buf& loop, int count, float eps) {
  if (count < 2) return 0;
  int w = 0;
  for (int r = 1; r < count; ++r) {
    if (len_sq(loop[r] - loop[w]) > eps) { w += 1; loop[w] = loop[r]; }
  }
  int nc;
  if (w > 0) nc = len_sq(loop[w] - loop[0]) < eps ? w : w + 1;
  else nc = w + 1;
  if (nc < 2) nc = 0;
  return nc;
}

template <class Buf>
__device__ __forceinline__ int add_nodup(const Buf& arr, int count, f2 v, float eps, bool& added) {
  added = false;
  if (count > 0 && len_sq(arr[0] - v) < eps) return count;
  if (count > 1 && len_sq(arr[count - 1] - v) < eps) return count;
  arr[count] = v;
  added = true;
  return count + 1;
}

struct Tracker { f3 ref, prev, normal; float area; };
__device__ __forceinline__ void track(Tracker& t, f3 p, int id) {
  if (id == 0) { t.ref = p; t.area = 0.0f; }
  else if (id == 1) { t.prev = p; }
  else {
    f3 e1 = t.prev - t.ref, e2 = p - t.ref;
    f3 c = cross(e1, e2);
    float a = dot(c, c);
    if (a > t.area) { t.area = a; t.normal = c; }
    t.prev = p;
  }
}
struct Proj { float d; f3 n; };

__device__ __forceinline__ f3 seg_proj_normal(f3 seg, f3 ref) {
  f3 right = cross(seg, ref);
  f3 n = cross(right, seg);
  float l = len(n);
  return l > 1e-12f ? n / l : ref;
}

__device__ void make_proj(const Tracker& ta, f3 anchor_a, const Tracker& tb, f3 anchor_b, f3 cn, Proj& pa, Proj& pb) {
  f3 zero{0.0f, 0.0f, 0.0f};
  pa = Proj{0.0f, zero};
  pb = Proj{0.0f, zero};
  if (ta.area == 0.0f && tb.area == 0.0f) {
    f3 da = ta.prev - ta.ref, db = tb.prev - tb.ref;
    f3 ppa = 0.5f * (ta.ref + ta.prev);
    pa.n = seg_proj_normal(da, cn);
    pa.d = -dot(ppa, pa.n);
    f3 ppb = 0.5f * (tb.ref + tb.prev);
    pb.n = seg_proj_normal(db, cn);
    pb.d = -dot(ppb, pb.n);
    return;
  }
  if (ta.area > 0.0f) {
    float l = sqrtf(fmax2(1.0e-12f, ta.area));
    pa.n = ta.normal / l;
    pa.d = -dot(anchor_a, pa.n);
  }
  if (tb.area > 0.0f) {
    float l = sqrtf(fmax2(1.0e-12f, tb.area));
    pb.n = tb.normal / l;
    pb.d = -dot(anchor_b, pb.n);
  }
  if (ta.area == 0.0f) {
    f3 dir = ta.prev - ta.ref;
    f3 pp = 0.5f * (ta.ref + ta.prev);
    pa.n = seg_proj_normal(dir, pb.n);
    pa.d = -dot(pp, pa.n);
  }
  if (tb.area == 0.0f) {
    f3 dir = tb.prev - tb.ref;
    f3 pp = 0.5f * (tb.ref + tb.prev);
    pb.n = seg_proj_normal(dir, pa.n);
    pb.d = -dot(pp, pb.n);
  }
}

__device__ __forceinline__ f3 ray_plane(f3 o, f3 dir, float pd, f3 pn) {
  float denom = dot(dir, pn);
  if (fabsf(denom) < 1.0e-12f) return o;
  float t = -(dot(o, pn) + pd) / denom;
  return o + dir * t;
}

__device__ __forceinline__ void onb(f3 n, f3& b1, f3& b2) {
  if (n.z < 0.0f) {
    float a = 1.0f / (1.0f - n.z);
    float b = n.x * n.y * a;
    b1 = f3{1.0f - n.x * n.x * a, -b, n.x};
    b2 = f3{b, n.y * n.y * a - 1.0f, -n.y};
  } else {
    float a = 1.0f / (1.0f + n.z);
    float b = -n.x * n.y * a;
    b1 = f3{1.0f - n.x * n.x * a, b, -n.x};
    b2 = f3{b, 1.0f - n.y * n.y * a, -n.y};
  }
}

struct PairCtx {
  GShape ga, gb;
  qf oa, ob_world;
  f3 pa_world, pb_world;
  CData tmpl;
};

// post_process_axial_on_discrete_contact (collision_core.py:172-275): the Minkowski adjustment, then, for a
// cylinder or cone against a flat-faced shape (box, plane or the infinite plane's box proxy) in a rolling pose,
// the point projected onto the plane through the axial shape's centre spanned by its axis and the normal.
// Rolling: a cylinder's axis within 2 degrees of perpendicular to the normal; a cone's axis at its half angle
// atan2(r, 2 hh) from the normal plane to within 2 degrees, tested through sin(half angle -/+ 2 degrees)
// = sin a cos 2 -/+ cos a sin 2 with sin a = r / L, cos a = 2 hh / L, L = |(r, 2 hh)| (the same thresholds,
// without a trigonometric call).
template <bool ALL>
__device__ __forceinline__ CData post_process(CData c, const PairCtx& P) {
  const f3 normal = c.normal;
  c = post_process(c, P.ga.type, P.gb.type);
  if (!ALL) return c;
  const int ta = P.ga.type, tb = P.gb.type;
  const bool da = ta == GEO_BOX || ta == GEO_PLANE, db = tb == GEO_BOX || tb == GEO_PLANE;
  const bool xa = ta == GEO_CYLINDER || ta == GEO_CONE, xb = tb == GEO_CYLINDER || tb == GEO_CONE;
  if (!((da && xb) || (db && xa))) return c;
  const bool b_axial = da && xb;
  const f3 axis = rot(b_axial ? P.ob_world : P.oa, f3{0.0f, 0.0f, 1.0f});
  const GShape& g = b_axial ? P.gb : P.ga;
  const f3 spos = b_axial ? P.pb_world : P.pa_world;
  const f3 an = b_axial ? normal : -normal;
  const float and_ = fabsf(dot(axis, an));
  const float SIN2 = (float)0.034899496702500969, COS2 = (float)0.99939082701909576;
  bool rolling;
  if (g.type == GEO_CONE) {
    const float hh2 = 2.0f * g.scale.y;
    const float l = sqrtf(g.scale.x * g.scale.x + hh2 * hh2);
    const float sa = g.scale.x / l, ca = hh2 / l;
    rolling = and_ >= sa * COS2 - ca * SIN2 && and_ <= sa * COS2 + ca * SIN2;
  } else {
    rolling = and_ <= SIN2;
  }
  if (rolling) {
    const f3 ppn = normalize(cross(axis, an));
    c.center = c.center - ppn * dot(c.center - spos, ppn);  // project_point_onto_plane (collision_core.py:51-65)
  }
  return c;
}

// Directions of the five tilted support probes (cos and sin of k 2 pi / 5), in constant memory: the probe
// loop index is wave-uniform, so the loads are scalar.
static __constant__ float c_probe_cos[5] = {1.0f, (float)0.30901699437494745, (float)-0.8090169943749473,
                                            (float)-0.8090169943749476, (float)0.30901699437494723};
static __constant__ float c_probe_sin[5] = {0.0f, (float)0.9510565162951535, (float)0.5877852522924732,
                                            (float)-0.587785252292473, (float)-0.9510565162951536};

template <bool ALL>
__device__ void manifold(const PairCtx& P, qf rel_ob, f3 rel_pb, f3 p_a, f3 p_b, f3 normal, WriteCtx& w) {
  const float EPS = 0.00001f;
  const float SIN_TILT = (float)0.03489949670250097;
  const float COS_TILT = (float)0.9993908270190958;
  const float COS_DEEP = (float)0.9999984769132877;
  f3 ta, tb;
  onb(normal, ta, tb);
  f3 zero{0.0f, 0.0f, 0.0f};
  Tracker tra{zero, zero, zero, 0.0f}, trb{zero, zero, zero, 0.0f};
  f3 center = 0.5f * (p_a + p_b);
  __shared__ f2 s_poly[kClipMax * kNarrowBlock];
  const LanePoly buf{s_poly + threadIdx.x};
  for (int i = 0; i < kClipMax; ++i) buf[i] = f2{0.0f, 0.0f};
  int a_count = 0, b_count = 0;
  f3 lnb = rot_inv(rel_ob, -normal);
  f3 ltab = rot_inv(rel_ob, -ta);
  f3 ltbb = rot_inv(rel_ob, -tb);
  for (int e = 0; e < 5; ++e) {
    float c = c_probe_cos[e], s = c_probe_sin[e];
    float ct = COS_TILT;
    float cs = c * SIN_TILT, ss = s * SIN_TILT;
    f3 dir_a = normal * ct + cs * ta + ss * tb;
    f3 pt_a = support<ALL>(P.ga, dir_a);
    f3 pr_a = pt_a - center;
    f2 a2{dot(ta, pr_a), dot(tb, pr_a)};
    bool added;
    a_count = add_nodup(buf + 5, a_count, a2, EPS, added);
    if (added) track(tra, pt_a, a_count - 1);
    f3 ldb = lnb * ct + cs * ltab + ss * ltbb;
    f3 pbl = support<ALL>(P.gb, ldb);
    f3 pt_b = rot(rel_ob, pbl) + rel_pb;
    f3 pr_b = pt_b - center;
    f2 b2{dot(ta, pr_b), dot(tb, pr_b)};
    b_count = add_nodup(buf, b_count, b2, EPS, added);
    if (added) track(trb, pt_b, b_count - 1);
  }
  f3 normal_world = rot(P.oa, normal);
  int count_out;
  float normal_dot;
  if (a_count < 2 || b_count < 2) {
    count_out = 0;
    normal_dot = 0.0f;
  } else {
    Proj pa, pb;
    make_proj(tra, p_a, trb, p_b, normal, pa, pb);
    if (fabsf(dot(normal, pa.n)) < COS_TILT || fabsf(dot(normal, pb.n)) < COS_TILT) {
      count_out = 0;
      normal_dot = 0.0f;
    } else {
      normal_dot = fabsf(dot(pa.n, pb.n));
      int loop_count = trim_all(buf, a_count, b_count);
      loop_count = remove_zero_edges(buf, loop_count, EPS);
      if (loop_count > 1) {
        int res[4] = {0, 1, 2, 3};
        if (loop_count > 4) {
          calipers(buf, loop_count, res);
          loop_count = 4;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          if (i >= loop_count) break;
          int ia = res[i];
          f3 pl = buf[ia].x * ta + buf[ia].y * tb + center;
          f3 a = ray_plane(pl, normal, pa.d, pa.n);
          f3 b = ray_plane(pl, normal, pb.d, pb.n);
          f3 cpl = 0.5f * (a + b);
          float sd = dot(b - a, normal);
          f3 cpw = rot(P.oa, cpl) + P.pa_world;
          CData cd = P.tmpl;
          cd.center = cpw;
          cd.normal = normal_world;
          cd.dist = sd;
          cd.sub_key = (P.tmpl.sub_key << 3) | i;
          cd = post_process<ALL>(cd, P);
          emit(cd, w, true);
        }
      } else {
        normal_dot = 0.0f;
        loop_count = 0;
      }
      count_out = loop_count < 4 ? loop_count : 4;
    }
  }
  if (normal_dot < COS_DEEP || count_out == 0) {
    f3 dcl = 0.5f * (p_a + p_b);
    float dsd = dot(p_b - p_a, normal);
    f3 dcw = rot(P.oa, dcl) + P.pa_world;
    CData cd = P.tmpl;
    cd.center = dcw;
    cd.normal = normal_world;
    cd.dist = dsd;
    cd.sub_key = (P.tmpl.sub_key << 3) | count_out;
    cd = post_process<ALL>(cd, P);
    emit(cd, w, true);
  }
}

template <bool ALL>
__device__ void solve_convex(const PairCtx& P, float threshold, bool skip_multi, WriteCtx& w) {
  qf rel_ob = qmul(qinv(P.oa), P.ob_world);
  f3 rel_pb = rot_inv(P.oa, P.pb_world - P.pa_world);
  float margin_sum = P.tmpl.m_a + P.tmpl.m_b;
  const float eps = 1.0e-4f;
  float enlarge = margin_sum <= 0.0f ? eps : (margin_sum < eps ? 2.0f * eps : 0.0f);
  CvxRes r = mpr_core<ALL>(P.ga, P.gb, rel_ob, rel_pb, enlarge);
  f3 point_a, point_b, normal;
  float sd;
  if (r.flag) {
    sd = -r.value + enlarge;
    float half = enlarge * 0.5f;
    point_a = r.pa - r.n * half;
    point_b = r.pb + r.n * half;
    normal = r.n;
  } else {
    CvxRes g = gjk_core<ALL>(P.ga, P.gb, rel_ob, rel_pb, 0.0f);
    point_a = g.pa;
    point_b = g.pb;
    normal = g.n;
    sd = g.value;
  }
  if (skip_multi || sd > threshold) {
    f3 point = 0.5f * (point_a + point_b);
    point = rot(P.oa, point) + P.pa_world;
    f3 nws = rot(P.oa, normal);
    CData cd = P.tmpl;
    cd.center = point;
    cd.normal = nws;
    cd.dist = sd;
    cd.sub_key = P.tmpl.sub_key << 3;
    cd = post_process<ALL>(cd, P);
    emit(cd, w, true);
    return;
  }
  manifold<ALL>(P, rel_ob, rel_pb, point_a, point_b, normal, w);
}

// Type pairs with an analytic collider (narrow_phase.py:327-513); every other primitive pair goes through MPR/GJK
// and the manifold (narrow_phase.py:595-600), as does an analytic pair whose collider reports no point.
__device__ __forceinline__ bool analytic_pair(int ta, int tb) {
  if (ta == GEO_PLANE)
    return tb == GEO_SPHERE || tb == GEO_ELLIPSOID || tb == GEO_BOX || tb == GEO_CAPSULE || tb == GEO_CYLINDER;
  if (ta == GEO_SPHERE) return tb == GEO_SPHERE || tb == GEO_CAPSULE || tb == GEO_CYLINDER || tb == GEO_BOX;
  return ta == GEO_CAPSULE && tb == GEO_CAPSULE;
}
__device__ __forceinline__ bool convex_route(int ta, int tb) { return !analytic_pair(ta, tb); }

// Route class of a pair: 0 analytic, 1 MPR/GJK, 2 mesh (mesh-mesh or mesh-infinite plane,
// narrow_phase.py:290-320; other mesh pairs are refused when the model is created).
__device__ __forceinline__ int route_class(int ta, int tb) {
  if (ta == GEO_MESH || tb == GEO_MESH) return 2;
  return convex_route(ta, tb) ? 1 : 0;
}

// One broad-phase pair (sa, sb already ordered by GeoType) -> up to 5 records at w.slot0.
template <bool ALL>
__device__ void narrow_pair(int sa, int sb, int ta, int tb, const ShapeW& A, const ShapeW& B, float gap_sum,
                            WriteCtx& w) {
  w.n = 0;
  f3 scale_a = A.gscale, scale_b = B.gscale;
  f3 pos_a = A.X.p, pos_b = B.X.p;
  qf qa = A.X.q, qb = B.X.q;
  bool plane_a = ta == GEO_PLANE;
  float r_a = (ta == GEO_SPHERE || ta == GEO_CAPSULE) ? scale_a.x : 0.0f;
  float r_b = (tb == GEO_SPHERE || tb == GEO_CAPSULE) ? scale_b.x : 0.0f;
  float cd[4] = {MAXVAL, MAXVAL, MAXVAL, MAXVAL};
  f3 zero{0.0f, 0.0f, 0.0f};
  f3 cp[4] = {zero, zero, zero, zero};
  f3 cn = zero;
  const f3 ez{0.0f, 0.0f, 1.0f};
  if (plane_a && tb == GEO_SPHERE) {
    f3 pn = rot(qa, ez);
    plane_sphere(pn, pos_a, pos_b, scale_b.x, cd[0], cp[0]);
    cn = pn;
  } else if (ALL && plane_a && tb == GEO_ELLIPSOID) {
    f3 pn = rot(qa, ez);
    plane_ellipsoid(pn, pos_a, pos_b, qmat(qb), scale_b, cd[0], cp[0]);
    cn = pn;
  } else if (plane_a && tb == GEO_BOX) {
    f3 pn = rot(qa, ez);
    m33 R = qmat(qb);
    plane_box(pn, pos_a, pos_b, R, scale_b, gap_sum, cd, cp);
    cn = pn;
  } else if (ta == GEO_SPHERE && tb == GEO_SPHERE) {
    sphere_sphere(pos_a, scale_a.x, pos_b, scale_b.x, cd[0], cp[0], cn);
  } else if (plane_a && tb == GEO_CAPSULE) {
    f3 pn = rot(qa, ez);
    f3 axis = rot(qb, ez);
    f3 seg = axis * scale_b.y;
    plane_sphere(pn, pos_a, pos_b + seg, scale_b.x, cd[0], cp[0]);
    plane_sphere(pn, pos_a, pos_b - seg, scale_b.x, cd[1], cp[1]);
    cn = pn;
  } else if (ALL && plane_a && tb == GEO_CYLINDER) {
    f3 pn = rot(qa, ez);
    plane_cylinder(pn, pos_a, pos_b, rot(qb, ez), scale_b.x, scale_b.y, cd, cp);
    cn = pn;
  } else if (ta == GEO_SPHERE && tb == GEO_CAPSULE) {
    f3 axis = rot(qb, ez);
    f3 seg = axis * scale_b.y;
    f3 pt = closest_segment_point(pos_b - seg, pos_b + seg, pos_a);
    sphere_sphere(pos_a, scale_a.x, pt, scale_b.x, cd[0], cp[0], cn);
  } else if (ta == GEO_CAPSULE && tb == GEO_CAPSULE) {
    f3 ax1 = rot(qa, ez), ax2 = rot(qb, ez);
    float d2[2];
    f3 p2[2];
    capsule_capsule(pos_a, ax1, scale_a.x, scale_a.y, pos_b, ax2, scale_b.x, scale_b.y, d2, p2, cn);
    cd[0] = d2[0]; cp[0] = p2[0]; cd[1] = d2[1]; cp[1] = p2[1];
  } else if (ALL && ta == GEO_SPHERE && tb == GEO_CYLINDER) {
    sphere_cylinder(pos_a, scale_a.x, pos_b, rot(qb, ez), scale_b.x, scale_b.y, cd[0], cp[0], cn);
  } else if (ta == GEO_SPHERE && tb == GEO_BOX) {
    m33 R = qmat(qb);
    sphere_box(pos_a, scale_a.x, pos_b, R, scale_b, cd[0], cp[0], cn);
  }
  int num = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) num += cd[k] < MAXVAL ? 1 : 0;
  if (num > 0) {
    CData c;
    c.normal = cn;
    c.r_a = r_a;
    c.r_b = r_b;
    c.m_a = A.margin;
    c.m_b = B.margin;
    c.shape_a = sa;
    c.shape_b = sb;
    c.gap_sum = gap_sum;
    bool valid[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      valid[k] = false;
      if (cd[k] < MAXVAL) {
        c.center = cp[k];
        c.dist = cd[k];
        valid[k] = passes_gap(c);
      }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (!valid[k]) continue;
      c.center = cp[k];
      c.dist = cd[k];
      c.sub_key = k;
      emit(c, w, false);
    }
    return;
  }
  bool inf_a = ta == GEO_PLANE && scale_a.x == 0.0f && scale_a.y == 0.0f;
  bool inf_b = tb == GEO_PLANE && scale_b.x == 0.0f && scale_b.y == 0.0f;
  if (inf_a && inf_b) return;
  float bs_ra = 0.0f, bs_rb = 0.0f;
  if (inf_a || inf_b) {
    f3 ca = 0.5f * (A.lo + A.hi), cb = 0.5f * (B.lo + B.hi);
    bs_ra = len(0.5f * (A.hi - A.lo));
    bs_rb = len(0.5f * (B.hi - B.lo));
    f3 ppos = inf_a ? pos_a : pos_b;
    qf pq = inf_a ? qa : qb;
    f3 oc = inf_a ? cb : ca;
    float orad = inf_a ? bs_rb : bs_ra;
    f3 pn = rot(pq, ez);
    float center_dist = dot(oc - ppos, pn);
    if (!(center_dist <= orad)) return;
  }
  PairCtx P;
  P.ga = GShape{ta, scale_a};
  P.gb = GShape{tb, scale_b};
  P.oa = qa;
  P.ob_world = qb;
  P.pa_world = pos_a;
  P.pb_world = pos_b;
  if (inf_a) {
    float other = bs_rb + gap_sum;
    float lateral = other * 10.0f, depth = other * 10.0f;
    P.ga = GShape{GEO_BOX, f3{lateral, lateral, depth}};
    f3 pn = rot(qa, ez);
    float dn = dot(pos_b - pos_a, pn);
    f3 surf = pos_b - pn * dn;
    P.pa_world = surf - pn * depth;
  }
  if (inf_b) {
    float other = bs_ra + gap_sum;
    float lateral = other * 10.0f, depth = other * 10.0f;
    P.gb = GShape{GEO_BOX, f3{lateral, lateral, depth}};
    f3 pn = rot(qb, ez);
    float dn = dot(pos_a - pos_b, pn);
    f3 surf = pos_a - pn * dn;
    P.pb_world = surf - pn * depth;
  }
  float rea = 0.0f, reb = 0.0f;
  const float small_radius = 0.0001f;
  if (P.ga.type == GEO_SPHERE || P.ga.type == GEO_CAPSULE) { rea = P.ga.scale.x; P.ga.scale.x = small_radius; }
  if (P.gb.type == GEO_SPHERE || P.gb.type == GEO_CAPSULE) { reb = P.gb.scale.x; P.gb.scale.x = small_radius; }
  CData& t = P.tmpl;
  t.center = zero;
  t.normal = zero;
  t.dist = 0.0f;
  t.r_a = rea;
  t.r_b = reb;
  t.m_a = A.margin;
  t.m_b = B.margin;
  t.shape_a = sa;
  t.shape_b = sb;
  t.gap_sum = gap_sum;
  t.sub_key = 0;
  float threshold = gap_sum + rea + reb + A.margin + B.margin;
  bool skip_multi = P.ga.type == GEO_SPHERE || P.gb.type == GEO_SPHERE || P.ga.type == GEO_ELLIPSOID ||
                    P.gb.type == GEO_ELLIPSOID;
  solve_convex<ALL>(P, threshold, skip_multi, w);
}

}  // namespace nmc
