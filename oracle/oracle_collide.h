// CPU oracle -- TEST INFRASTRUCTURE ONLY.  Collision pipeline restatement.
//
// Follows, function by function:
//   compute_shape_aabbs                      newton/_src/sim/collide.py:150-275
//   compute_tight_aabb_from_support          newton/_src/geometry/collision_core.py:396-491
//   support_map (box/sphere/capsule/plane)   newton/_src/geometry/support_function.py:106-304
//   SAP project / sort / range / sweep       newton/_src/geometry/broad_phase_sap.py:43-394, 510-677
//   check_aabb_overlap, group/world tests    newton/_src/geometry/broad_phase_common.py:18-180
//   narrow_phase_primitive_kernel            newton/_src/geometry/narrow_phase.py:136-605
//   analytic colliders                       newton/_src/geometry/collision_primitive.py:48-439, 1161-1215
//   narrow_phase_kernel_gjk_mpr              newton/_src/geometry/narrow_phase.py:608-778
//   find_contacts / plane->box proxy         newton/_src/geometry/collision_core.py:508-725
//   compute_gjk_mpr_contacts                 newton/_src/geometry/collision_core.py:278-393
//   solve_convex_multi_contact               newton/_src/geometry/collision_convex.py:26-135
//   solve_mpr_core                           newton/_src/geometry/mpr.py:289-505
//   solve_closest_distance_core              newton/_src/geometry/simplex_solver.py:60-492
//   build_manifold and clipping              newton/_src/geometry/multicontact.py:28-956
//   post_process (Minkowski sphere/capsule)  newton/_src/geometry/collision_core.py:144-275
//   write_contact                            newton/_src/sim/collide.py:62-148
//   make_contact_sort_key                    newton/_src/geometry/contact_data.py:59-83
#pragma once
#include <algorithm>
#include <vector>

#include "oracle.h"
#include "oracle_math.h"

namespace oracle {

enum { PLANE = 1, SPHERE = 3, CAPSULE = 4, ELLIPSOID = 5, CYLINDER = 6, BOX = 7, MESH = 8, CONE = 9 };
enum { COLLIDE_SHAPES = 2 };
static const double MAXVAL = 1.0e10;

template <class T> struct Shape {  // GenericShapeData
  int type;
  V3<T> scale;
};

template <class T> struct ContactData {
  V3<T> center;
  V3<T> normal;
  T dist;
  T r_a, r_b;
  T m_a, m_b;
  int shape_a, shape_b;
  T gap_sum;
  int sub_key;
};

template <class T> struct Contact {
  int shape0, shape1;
  V3<T> point0, point1, offset0, offset1, normal;
  T margin0, margin1;
  int64_t key;
};

inline int64_t make_sort_key(int a, int b, int sub) {
  return ((int64_t(a) & 0xFFFFF) << 43) | ((int64_t(b) & 0xFFFFF) << 23) | (int64_t(sub) & 0x7FFFFF);
}

// ---------------------------------------------------------------- support map (support_function.py:106)
template <class T> V3<T> support(const Shape<T>& g, V3<T> d) {
  const T eps = T(1.0e-12);
  if (g.type == BOX) {
    T thr = T(1.0e-10) * length(d);
    T sx = d.x >= -thr ? T(1) : T(-1);
    T sy = d.y >= -thr ? T(1) : T(-1);
    T sz = d.z >= -thr ? T(1) : T(-1);
    return {sx * g.scale.x, sy * g.scale.y, sz * g.scale.z};
  }
  if (g.type == SPHERE) {
    V3<T> n = length_sq(d) > eps ? normalize(d) : V3<T>{T(1), T(0), T(0)};
    return n * g.scale.x;
  }
  if (g.type == CAPSULE) {
    V3<T> n = length_sq(d) > eps ? normalize(d) : V3<T>{T(1), T(0), T(0)};
    V3<T> r = n * g.scale.x;
    if (d.z >= T(0)) return r + V3<T>{T(0), T(0), g.scale.y};
    return r + V3<T>{T(0), T(0), -g.scale.y};
  }
  if (g.type == PLANE) {
    T sx = d.x >= T(0) ? T(1) : T(-1);
    T sy = d.y >= T(0) ? T(1) : T(-1);
    return {sx * g.scale.x, sy * g.scale.y, T(0)};
  }
  if (g.type == ELLIPSOID) {  // support_function.py:218-239
    const T a = g.scale.x, b = g.scale.y, c = g.scale.z;
    if (length_sq(d) > eps) {
      const T adx = a * d.x, bdy = b * d.y, cdz = c * d.z;
      const T den_sq = adx * adx + bdy * bdy + cdz * cdz;
      if (den_sq > eps) {
        const T den = tsqrt(den_sq);
        return {(a * a) * d.x / den, (b * b) * d.y / den, (c * c) * d.z / den};
      }
    }
    return {a, T(0), T(0)};
  }
  if (g.type == CYLINDER) {  // support_function.py:240-260
    const T r = g.scale.x, hh = g.scale.y;
    const V3<T> dxy{d.x, d.y, T(0)};
    V3<T> lat{r, T(0), T(0)};
    if (length_sq(dxy) > eps) {
      const V3<T> nxy = normalize(dxy);
      lat = V3<T>{nxy.x * r, nxy.y * r, T(0)};
    }
    if (d.z > T(0)) return {lat.x, lat.y, hh};
    if (d.z < T(0)) return {lat.x, lat.y, -hh};
    return lat;
  }
  if (g.type == CONE) {  // support_function.py:262-285
    const T r = g.scale.x, hh = g.scale.y;
    const V3<T> dxy{d.x, d.y, T(0)};
    const T lxy = length(dxy);
    const T k = hh > eps ? r / (T(2) * hh) : T(0);
    if (lxy <= eps) return d.z >= T(0) ? V3<T>{T(0), T(0), hh} : V3<T>{r, T(0), -hh};
    if (d.z >= k * lxy) return {T(0), T(0), hh};
    const V3<T> nxy = dxy / lxy;
    return {nxy.x * r, nxy.y * r, -hh};
  }
  return {T(0), T(0), T(0)};
}

// ---------------------------------------------------------------- per-shape world data (collide.py:150)
template <class T> struct ShapeWorld {
  Xf<T> X;       // geom_xform
  V3<T> gscale;  // geom_data.xyz
  T margin;      // geom_data.w
  V3<T> lo, hi;  // expanded AABB
};

template <class T> Xf<T> load_xf(const float* a) {
  return Xf<T>{{T(a[0]), T(a[1]), T(a[2])}, {T(a[3]), T(a[4]), T(a[5]), T(a[6])}};
}
template <class T, class S> Xf<T> load_xf_t(const S* a) {
  return Xf<T>{{T(a[0]), T(a[1]), T(a[2])}, {T(a[3]), T(a[4]), T(a[5]), T(a[6])}};
}

template <class T> void tight_aabb(const Shape<T>& g, Q<T> q, V3<T> pos, V3<T>& lo, V3<T>& hi) {
  M3<T> R = quat_to_matrix(q);
  // rows of R are the world axes in local coordinates (transpose(R) columns)
  V3<T> lx{R.m[0][0], R.m[0][1], R.m[0][2]};
  V3<T> ly{R.m[1][0], R.m[1][1], R.m[1][2]};
  V3<T> lz{R.m[2][0], R.m[2][1], R.m[2][2]};
  T max_x = dot(lx, support(g, lx));
  T max_y = dot(ly, support(g, ly));
  T max_z = dot(lz, support(g, lz));
  T min_x = dot(lx, support(g, -lx));
  T min_y = dot(ly, support(g, -ly));
  T min_z = dot(lz, support(g, -lz));
  lo = V3<T>{min_x, min_y, min_z} + pos;
  hi = V3<T>{max_x, max_y, max_z} + pos;
}

template <class T, class S>
ShapeWorld<T> compute_shape_world(const OracleModel* m, const S* body_q, int s) {
  ShapeWorld<T> r;
  int b = m->shape_body[s];
  int type = m->shape_type[s];
  Xf<T> Xs = load_xf<T>(m->shape_transform + 7 * s);
  r.X = b == -1 ? Xs : xf_mul(load_xf_t<T, S>(body_q + 7 * b), Xs);
  V3<T> pos = r.X.p;
  Q<T> q = r.X.q;
  T margin = T(m->shape_margin[s]);
  T eg = margin + T(m->shape_gap[s]);
  V3<T> mvec{eg, eg, eg};
  V3<T> scale{T(m->shape_scale[3 * s]), T(m->shape_scale[3 * s + 1]), T(m->shape_scale[3 * s + 2])};
  bool inf_plane = type == PLANE && scale.x == T(0) && scale.y == T(0);
  V3<T> gscale = scale;
  if (inf_plane) {
    T rad = T(m->shape_radius[s]);
    V3<T> he{rad, rad, rad};
    r.lo = pos - he - mvec;
    r.hi = pos + he + mvec;
  } else if (type == MESH) {
    // pre-computed local AABB rotated into the world frame (collide.py:215-245)
    const float* l = m->shape_aabb_lower + 3 * s;
    const float* u = m->shape_aabb_upper + 3 * s;
    V3<T> llo{T(l[0]), T(l[1]), T(l[2])}, lhi{T(u[0]), T(u[1]), T(u[2])};
    V3<T> c = (llo + lhi) * T(0.5), h = (lhi - llo) * T(0.5);
    V3<T> wc = qrot(q, c) + pos;
    V3<T> r0 = qrot(q, V3<T>{T(1), T(0), T(0)}), r1 = qrot(q, V3<T>{T(0), T(1), T(0)});
    V3<T> r2 = qrot(q, V3<T>{T(0), T(0), T(1)});
    V3<T> wh{tabs(r0.x) * h.x + tabs(r1.x) * h.y + tabs(r2.x) * h.z,
             tabs(r0.y) * h.x + tabs(r1.y) * h.y + tabs(r2.y) * h.z,
             tabs(r0.z) * h.x + tabs(r1.z) * h.y + tabs(r2.z) * h.z};
    r.lo = wc - wh - mvec;
    r.hi = wc + wh + mvec;
  } else {
    Shape<T> g;
    g.type = type;
    if (type == PLANE) gscale = V3<T>{scale.x * T(0.5), scale.y * T(0.5), T(0)};
    g.scale = gscale;
    V3<T> lo, hi;
    tight_aabb(g, q, pos, lo, hi);
    r.lo = lo - mvec;
    r.hi = hi + mvec;
  }
  r.gscale = gscale;
  r.margin = margin;
  return r;
}

// ---------------------------------------------------------------- broad phase (SAP)
inline bool test_group_pair(int a, int b) {
  if (a == 0 || b == 0) return false;
  if (a > 0) return a == b || b < 0;
  return a != b;
}
inline bool test_world_and_group_pair(int wa, int wb, int ga, int gb) {
  if (wa != -1 && wb != -1 && wa != wb) return false;
  return test_group_pair(ga, gb);
}
inline bool is_pair_excluded(int a, int b, const int* fp, int n) {
  if (n <= 0) return false;
  int lo = 0, hi = n - 1;
  while (lo <= hi) {
    int mid = (lo + hi) >> 1;
    int ma = fp[2 * mid], mb = fp[2 * mid + 1];
    if (a == ma && b == mb) return true;
    bool less = a < ma || (a == ma && b < mb);
    if (less) hi = mid - 1; else lo = mid + 1;
  }
  return false;
}
template <class T> bool aabb_overlap(const ShapeWorld<T>& a, const ShapeWorld<T>& b) {
  const T c = T(0) + T(0);
  return a.lo.x <= b.hi.x + c && a.hi.x >= b.lo.x - c && a.lo.y <= b.hi.y + c && a.hi.y >= b.lo.y - c &&
         a.lo.z <= b.hi.z + c && a.hi.z >= b.lo.z - c;
}

template <class T>
void sap_pairs(const OracleModel* m, const std::vector<ShapeWorld<T>>& sw, std::vector<std::pair<int, int>>& out) {
  V3<T> dir = normalize(V3<T>{T(0.5935f), T(0.7790f), T(0.1235f)});
  int regular = m->segment_count - 1;
  for (int seg = 0; seg < m->segment_count; ++seg) {
    int s0 = m->segment_start[seg], n = m->segment_start[seg + 1] - s0;
    std::vector<T> plo(n), phi(n);
    std::vector<int> order(n);
    for (int l = 0; l < n; ++l) {
      const ShapeWorld<T>& a = sw[m->segment_shape[s0 + l]];
      V3<T> half = T(0.5) * (a.hi - a.lo);
      half = V3<T>{half.x + T(0), half.y + T(0), half.z + T(0)};
      T radius = dot(dir, half);
      T center = dot(dir, T(0.5) * (a.lo + a.hi));
      plo[l] = center - radius;
      phi[l] = center + radius;
      order[l] = l;
    }
    std::vector<uint64_t> okey(n);
    for (int l = 0; l < n; ++l) okey[l] = sap_order_key(plo[l]);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return okey[a] < okey[b]; });
    // "sap" sweeps while lower_j < upper_i; "nxn" (broad_phase_nxn.py:129-214) and "explicit"
    // (:28-65, the pairs of a list, AABB test only) visit every pair of the segment
    const bool sweep = m->broad_phase_mode == 0, listed = m->broad_phase_mode == 2;
    for (int i = 0; i < n; ++i) {
      T upper = phi[order[i]];
      for (int j = i + 1; j < n && (!sweep || plo[order[j]] < upper); ++j) {
        int t1 = m->segment_shape[s0 + order[i]], t2 = m->segment_shape[s0 + order[j]];
        if (t1 == t2) continue;
        int a = std::min(t1, t2), b = std::max(t1, t2);
        int w1 = m->shape_world[a], w2 = m->shape_world[b];
        if (w1 == -1 && w2 == -1 && seg < regular) continue;
        if (listed) {
          if (!is_pair_excluded(a, b, m->explicit_pairs, m->explicit_pair_count)) continue;  // membership
        } else {
          if (!test_world_and_group_pair(w1, w2, m->shape_group[a], m->shape_group[b])) continue;
          if (is_pair_excluded(a, b, m->filter_pairs, m->filter_count)) continue;
        }
        if (!aabb_overlap(sw[a], sw[b])) continue;
        out.push_back({a, b});
      }
    }
  }
}

// ---------------------------------------------------------------- analytic colliders
template <class T> struct NormLen { V3<T> v; T n; };
template <class T> NormLen<T> normalize_with_norm(V3<T> x) {
  T n = length(x);
  if (n == T(0)) return {x, T(0)};
  return {x / n, n};
}

template <class T> V3<T> closest_segment_point(V3<T> a, V3<T> b, V3<T> pt) {
  V3<T> ab = b - a;
  T t = dot(pt - a, ab) / (dot(ab, ab) + T(1e-6));
  return a + tclamp(t, T(0), T(1)) * ab;
}

template <class T> void plane_sphere(V3<T> n, V3<T> pp, V3<T> sp, T r, T& dist, V3<T>& pos) {
  dist = dot(sp - pp, n) - r;
  pos = sp - n * (r + T(0.5) * dist);
}

template <class T> void sphere_sphere(V3<T> p1, T r1, V3<T> p2, T r2, T& dist, V3<T>& pos, V3<T>& n) {
  V3<T> dir = p2 - p1;
  T d = length(dir);
  n = d == T(0) ? V3<T>{T(1), T(0), T(0)} : dir / d;
  d = d - (r1 + r2);
  pos = p1 + n * (r1 + T(0.5) * d);
  dist = d;
}

template <class T>
void capsule_capsule(V3<T> p1, V3<T> ax1, T r1, T hl1, V3<T> p2, V3<T> ax2, T r2, T hl2, T dist[2], V3<T> pos[2],
                     V3<T>& normal) {
  const T MINVAL = T(1e-15);
  dist[0] = dist[1] = T(MAXVAL);
  pos[0] = pos[1] = V3<T>{T(0), T(0), T(0)};
  normal = V3<T>{T(0), T(0), T(0)};
  V3<T> axis1 = ax1 * hl1, axis2 = ax2 * hl2;
  V3<T> dif = p1 - p2;
  T ma = dot(axis1, axis1);
  T mb = -dot(axis1, axis2);
  T mc = dot(axis2, axis2);
  T u = -dot(axis1, dif);
  T v = dot(axis2, dif);
  T det = ma * mc - mb * mb;
  if (tabs(det) >= MINVAL) {
    T inv_det = T(1) / det;
    T x1 = (mc * u - mb * v) * inv_det;
    T x2 = (ma * v - mb * u) * inv_det;
    if (x1 > T(1)) { x1 = T(1); x2 = (v - mb) / mc; }
    else if (x1 < T(-1)) { x1 = T(-1); x2 = (v + mb) / mc; }
    if (x2 > T(1)) { x2 = T(1); x1 = tclamp((u - mb) / ma, T(-1), T(1)); }
    else if (x2 < T(-1)) { x2 = T(-1); x1 = tclamp((u + mb) / ma, T(-1), T(1)); }
    V3<T> vec1 = p1 + axis1 * x1, vec2 = p2 + axis2 * x2;
    sphere_sphere(vec1, r1, vec2, r2, dist[0], pos[0], normal);
  } else {
    V3<T> vec1 = p1 + axis1;
    T x2 = tclamp((v - mb) / mc, T(-1), T(1));
    V3<T> vec2 = p2 + axis2 * x2;
    sphere_sphere(vec1, r1, vec2, r2, dist[0], pos[0], normal);
    vec1 = p1 - axis1;
    x2 = tclamp((v + mb) / mc, T(-1), T(1));
    vec2 = p2 + axis2 * x2;
    V3<T> n2;
    sphere_sphere(vec1, r1, vec2, r2, dist[1], pos[1], n2);
  }
}

template <class T>
void plane_box(V3<T> n, V3<T> pp, V3<T> bp, const M3<T>& R, V3<T> size, T margin, T dist[4], V3<T> pos[4]) {
  T center_dist = dot(bp - pp, n);
  for (int i = 0; i < 4; ++i) { dist[i] = T(MAXVAL); pos[i] = V3<T>{T(0), T(0), T(0)}; }
  int ncontact = 0, worst = 0;
  for (int i = 0; i < 8; ++i) {
    V3<T> corner{(i & 1) ? size.x : -size.x, (i & 2) ? size.y : -size.y, (i & 4) ? size.z : -size.z};
    corner = mv(R, corner);
    T ldist = dot(n, corner);
    T cdist = center_dist + ldist;
    if (cdist > margin) continue;
    V3<T> cpos = corner + bp - (T(0.5) * n) * cdist;
    if (ncontact < 4) {
      dist[ncontact] = cdist;
      pos[ncontact] = cpos;
      if (ncontact == 0 || cdist > dist[worst]) worst = ncontact;
      ncontact += 1;
    } else if (cdist < dist[worst]) {
      dist[worst] = cdist;
      pos[worst] = cpos;
      worst = 0;
      if (dist[1] > dist[worst]) worst = 1;
      if (dist[2] > dist[worst]) worst = 2;
      if (dist[3] > dist[worst]) worst = 3;
    }
  }
}

// collide_plane_ellipsoid (collision_primitive.py:334-363)
template <class T>
void plane_ellipsoid(V3<T> n, V3<T> pp, V3<T> ep, const M3<T>& R, V3<T> size, T& dist, V3<T>& pos) {
  const V3<T> t = mtv(R, n);
  const V3<T> ss = -normalize(V3<T>{t.x * size.x, t.y * size.y, t.z * size.z});
  const V3<T> p = ep + mv(R, V3<T>{ss.x * size.x, ss.y * size.y, ss.z * size.z});
  dist = dot(n, p - pp);
  pos = p - (n * dist) * T(0.5);
}

// collide_sphere_cylinder (collision_primitive.py:443-513)
template <class T>
void sphere_cylinder(V3<T> sp, T sr, V3<T> cp, V3<T> axis, T cr, T chh, T& dist, V3<T>& pos, V3<T>& n) {
  const V3<T> vec = sp - cp;
  const T x = dot(vec, axis);
  const V3<T> a_proj = axis * x;
  const V3<T> p_proj = vec - a_proj;
  const T pps = dot(p_proj, p_proj);
  bool side = tabs(x) < chh;
  bool cap = pps < cr * cr;
  if (side && cap) {
    const T dcap = chh - tabs(x);
    const T drad = cr - tsqrt(pps);
    if (dcap < drad) side = false;
    else cap = false;
  }
  if (side) {
    sphere_sphere(sp, sr, cp + a_proj, cr, dist, pos, n);
  } else if (cap) {
    const V3<T> pc = x > T(0) ? cp + axis * chh : cp - axis * chh;
    const V3<T> pn = x > T(0) ? axis : -axis;
    plane_sphere(pn, pc, sp, sr, dist, pos);
    n = -pn;
  } else {
    const T l = tsqrt(pps);
    const T inv = T(1) / (l != T(0) ? l : T(1.0e-15));  // safe_div (newton/_src/math/__init__.py:282)
    const V3<T> rim = p_proj * (cr * inv);
    const V3<T> cap_off = axis * ((x < T(0) ? T(-1) : T(1)) * chh);
    sphere_sphere(sp, sr, (cp + cap_off) + rim, T(0), dist, pos, n);
  }
}

// collide_plane_cylinder (collision_primitive.py:516-669): deepest rim point, then the flat-mode tripod or the
// rolling-mode generators and facing cap-rim point, each kept when farther than 1 % of max(r, hh) from the deepest.
template <class T>
void plane_cylinder(V3<T> n, V3<T> pp, V3<T> cp, V3<T> axis_in, T cr, T chh, T dist[4], V3<T> pos[4]) {
  V3<T> axis = axis_in;
  T dna = dot(n, axis);
  if (dna > T(0)) {
    axis = -axis;
    dna = -dna;
  }
  const V3<T> capc = cp + axis * chh;
  V3<T> pal = -n + axis * dna;
  const T pal2 = dot(pal, pal);
  const bool has_align = pal2 > T(1.0e-10);
  if (has_align) pal = pal * (T(1) / tsqrt(pal2));
  V3<T> ref{T(1), T(0), T(0)};
  if (tabs(dot(axis, ref)) > T(0.9)) ref = V3<T>{T(0), T(1), T(0)};
  const V3<T> pfix = normalize(ref - axis * dot(axis, ref));
  const T abs_dot = -dna;
  const bool flat = abs_dot >= T(0.92387953251128674);  // cos(22.5 degrees)
  const V3<T> dperp = has_align ? pal : pfix;
  const V3<T> dpt = capc + dperp * cr;
  const T dd = dot(dpt - pp, n);
  const V3<T> dpos = dpt - n * (dd * T(0.5));
  const V3<T> zero{T(0), T(0), T(0)};
  for (int k = 0; k < 4; ++k) {
    dist[k] = T(MAXVAL);
    pos[k] = zero;
  }
  dist[0] = dd;
  pos[0] = dpos;
  int nc = 1;
  const T mt = T(0.01) * tmax(cr, chh);
  const T mt2 = mt * mt;
  auto add = [&](V3<T> pt) {
    const T d = dot(pt - pp, n);
    const V3<T> q = pt - n * (d * T(0.5));
    if (nc < 4 && length_sq(q - dpos) > mt2) {
      dist[nc] = d;
      pos[nc] = q;
      nc += 1;
    }
  };
  if (flat) {
    const V3<T> u = pfix * cr, v = cross(axis, pfix) * cr;
    const T c120 = T(-0.5), s120 = T(0.8660254f);
    add(capc + u);
    add((capc + c120 * u) + s120 * v);
    add((capc + c120 * u) - s120 * v);
  } else {
    const V3<T> pr = has_align ? pal : pfix;
    const V3<T> u = pr * cr, v = cross(axis, pr) * cr;
    add((cp + axis * chh) + u);
    add((cp - axis * chh) + u);
    const V3<T> pv = capc + v, nv = capc - v;
    const T dpv = dot(pv - pp, n), dnv = dot(nv - pp, n);
    add(dpv <= dnv ? pv : nv);
  }
}

template <class T>
void sphere_box(V3<T> sp, T r, V3<T> bp, const M3<T>& R, V3<T> size, T& dist, V3<T>& pos_out, V3<T>& normal) {
  V3<T> center = mtv(R, sp - bp);
  V3<T> clamped = vmax(-size, vmin(size, center));
  NormLen<T> cd = normalize_with_norm(clamped - center);
  V3<T> pos;
  if (cd.n <= T(1e-6)) {
    T closest = T(2) * (size.x + size.y + size.z);
    int k = 0;
    for (int i = 0; i < 6; ++i) {
      T fd = tabs(((i % 2) ? T(1) : T(-1)) * comp(size, i / 2) - comp(center, i / 2));
      if (closest > fd) { closest = fd; k = i; }
    }
    V3<T> nearest{T(0), T(0), T(0)};
    set_comp(nearest, k / 2, (k % 2) ? T(-1) : T(1));
    pos = center + nearest * (r - closest) / T(2);
    normal = mv(R, nearest);
    dist = -closest - r;
  } else {
    V3<T> deepest = center + cd.v * r;
    pos = T(0.5) * (clamped + deepest);
    normal = mv(R, cd.v);
    dist = cd.n - r;
  }
  pos_out = bp + mv(R, pos);
}

// ---------------------------------------------------------------- MPR / GJK (mpr.py, simplex_solver.py)
template <class T> struct Vert { V3<T> B, BtoA; };

template <class T> V3<T> support_b(const Shape<T>& gb, V3<T> dir, Q<T> ob, V3<T> pb) {
  V3<T> tmp = qrot_inv(ob, dir);
  V3<T> r = support(gb, tmp);
  r = qrot(ob, r);
  return r + pb;
}

template <class T> Vert<T> minkowski_support(const Shape<T>& ga, const Shape<T>& gb, V3<T> dir, Q<T> ob, V3<T> pb,
                                             T extend) {
  Vert<T> v;
  V3<T> pa = support(ga, dir);
  v.B = support_b(gb, -dir, ob, pb);
  if (extend != T(0)) {
    V3<T> d = normalize(dir) * extend * T(0.5);
    pa = pa + d;
    v.B = v.B - d;
  }
  v.BtoA = pa - v.B;
  return v;
}

template <class T> Vert<T> geometric_center(Q<T> ob, V3<T> pb) {
  Vert<T> c;
  V3<T> zero{T(0), T(0), T(0)};
  V3<T> cbw = pb + qrot(ob, zero);
  c.B = cbw;
  c.BtoA = zero - cbw;
  return c;
}

template <class T> struct ConvexResult {
  bool flag;
  V3<T> pa, pb, n;
  T value;
};

template <class T>
ConvexResult<T> solve_mpr_core(const Shape<T>& ga, const Shape<T>& gb, Q<T> ob, V3<T> pb, T extend) {
  const int MAX_ITER = 30;
  const T CE = T(1e-5);
  const T NE = T(1e-16);
  T pen = T(0);
  V3<T> zero{T(0), T(0), T(0)};
  V3<T> point_a = zero, point_b = zero, normal = zero;
  Vert<T> v0 = geometric_center(ob, pb);
  normal = v0.BtoA;
  if (length_sq(normal) < NE) {
    T best_dot = T(-1.0e30);
    V3<T> best_dir{T(1), T(0), T(0)};
    for (int ax = 0; ax < 3; ++ax) {
      V3<T> probe = zero;
      set_comp(probe, ax, T(1));
      Vert<T> sv = minkowski_support(ga, gb, probe, ob, pb, extend);
      T d = dot(sv.BtoA, probe);
      if (d > best_dot) { best_dot = d; best_dir = probe; }
    }
    v0.BtoA = best_dir * T(1e-05);
  }
  normal = -v0.BtoA;
  Vert<T> v1 = minkowski_support(ga, gb, normal, ob, pb, extend);
  point_a = v1.B + v1.BtoA;
  point_b = v1.B;
  if (dot(v1.BtoA, normal) <= T(0)) return {false, point_a, point_b, normal, pen};
  normal = cross(v1.BtoA, v0.BtoA);
  if (length_sq(normal) < NE * NE) {
    normal = v1.BtoA - v0.BtoA;
    normal = normalize(normal);
    pen = dot(v1.BtoA, normal);
    return {true, point_a, point_b, normal, pen};
  }
  Vert<T> v2 = minkowski_support(ga, gb, normal, ob, pb, extend);
  if (dot(v2.BtoA, normal) <= T(0)) return {false, point_a, point_b, normal, pen};
  V3<T> temp1 = v1.BtoA - v0.BtoA;
  V3<T> temp2 = v2.BtoA - v0.BtoA;
  normal = cross(temp1, temp2);
  T dist = dot(normal, v0.BtoA);
  if (dist > T(0)) {
    std::swap(v1, v2);
    normal = -normal;
  }
  int phase1 = 0, phase2 = 0;
  bool hit = false;
  Vert<T> v3;
  while (true) {
    if (phase1 > MAX_ITER) return {false, point_a, point_b, normal, pen};
    phase1 += 1;
    v3 = minkowski_support(ga, gb, normal, ob, pb, extend);
    if (dot(v3.BtoA, normal) <= T(0)) return {false, point_a, point_b, normal, pen};
    temp1 = cross(v1.BtoA, v3.BtoA);
    if (dot(temp1, v0.BtoA) < T(0)) {
      v2 = v3;
      temp1 = v1.BtoA - v0.BtoA;
      temp2 = v3.BtoA - v0.BtoA;
      normal = cross(temp1, temp2);
      continue;
    }
    temp1 = cross(v3.BtoA, v2.BtoA);
    if (dot(temp1, v0.BtoA) < T(0)) {
      v1 = v3;
      temp1 = v3.BtoA - v0.BtoA;
      temp2 = v2.BtoA - v0.BtoA;
      normal = cross(temp1, temp2);
      continue;
    }
    break;
  }
  Vert<T> v4;
  while (true) {
    phase2 += 1;
    temp1 = v2.BtoA - v1.BtoA;
    temp2 = v3.BtoA - v1.BtoA;
    normal = cross(temp1, temp2);
    T normal_sq = length_sq(normal);
    if (normal_sq < NE * NE) return {false, point_a, point_b, normal, pen};
    if (!hit) {
      T d = dot(normal, v1.BtoA);
      hit = d >= T(0);
    }
    v4 = minkowski_support(ga, gb, normal, ob, pb, extend);
    V3<T> temp3 = v4.BtoA - v3.BtoA;
    T delta = dot(temp3, normal);
    pen = dot(v4.BtoA, normal);
    if (delta * delta <= CE * CE * normal_sq || pen <= T(0) || phase2 > MAX_ITER) {
      if (hit) {
        T inv_normal = T(1) / tsqrt(normal_sq);
        pen *= inv_normal;
        normal = normal * inv_normal;
        temp3 = cross(v1.BtoA, temp1);
        T gamma = dot(temp3, normal) * inv_normal;
        temp3 = cross(temp2, v1.BtoA);
        T beta = dot(temp3, normal) * inv_normal;
        T alpha = T(1) - gamma - beta;
        point_a = alpha * (v1.B + v1.BtoA) + beta * (v2.B + v2.BtoA) + gamma * (v3.B + v3.BtoA);
        point_b = alpha * v1.B + beta * v2.B + gamma * v3.B;
      }
      return {hit, point_a, point_b, normal, pen};
    }
    temp1 = cross(v4.BtoA, v0.BtoA);
    T dd = dot(temp1, v1.BtoA);
    if (dd >= T(0)) {
      dd = dot(temp1, v2.BtoA);
      if (dd >= T(0)) v1 = v4; else v3 = v4;
    } else {
      dd = dot(temp1, v3.BtoA);
      if (dd >= T(0)) v2 = v4; else v1 = v4;
    }
  }
}

template <class T> struct Simplex {
  V3<T> B[4], A[4];  // A holds BtoA
  T bc[4];
  unsigned mask;
};

template <class T> V3<T> closest_segment(const Simplex<T>& s, int i0, int i1, T bc[4], unsigned& mask) {
  const T EPS = T(1e-8);
  V3<T> a = s.A[i0], b = s.A[i1];
  V3<T> edge = b - a;
  T vsq = length_sq(edge);
  bool degenerate = vsq < EPS;
  T denom = degenerate ? EPS : vsq;
  T t = -dot(a, edge) / denom;
  T l0 = T(1) - t, l1 = t;
  mask = (1u << i0) | (1u << i1);
  for (int k = 0; k < 4; ++k) bc[k] = T(0);
  if (l0 < T(0) || degenerate) { mask = 1u << i1; l0 = T(0); l1 = T(1); }
  else if (l1 < T(0)) { mask = 1u << i0; l0 = T(1); l1 = T(0); }
  bc[i0] = l0;
  bc[i1] = l1;
  return l0 * a + l1 * b;
}

template <class T> V3<T> closest_triangle(const Simplex<T>& s, int i0, int i1, int i2, T bc[4], unsigned& mask) {
  const T EPS = T(1e-8);
  V3<T> a = s.A[i0], b = s.A[i1], c = s.A[i2];
  V3<T> u = a - b, w = a - c;
  V3<T> normal = cross(u, w);
  T t = length_sq(normal);
  bool degenerate = t < EPS;
  T denom = degenerate ? EPS : t;
  T it = T(1) / denom;
  V3<T> c1 = cross(u, a), c2 = cross(a, w);
  T l2 = dot(c1, normal) * it;
  T l1 = dot(c2, normal) * it;
  T l0 = T(1) - l2 - l1;
  T best = T(1e30);
  V3<T> closest_pt{T(0), T(0), T(0)};
  T bct[4];
  for (int k = 0; k < 4; ++k) bc[k] = T(0);
  mask = 0;
  unsigned m;
  if (l0 < T(0) || degenerate) {
    V3<T> cl = closest_segment(s, i1, i2, bct, m);
    T dist = length_sq(cl);
    if (dist < best) { for (int k = 0; k < 4; ++k) bc[k] = bct[k]; mask = m; best = dist; closest_pt = cl; }
  }
  if (l1 < T(0) || degenerate) {
    V3<T> cl = closest_segment(s, i0, i2, bct, m);
    T dist = length_sq(cl);
    if (dist < best) { for (int k = 0; k < 4; ++k) bc[k] = bct[k]; mask = m; best = dist; closest_pt = cl; }
  }
  if (l2 < T(0) || degenerate) {
    V3<T> cl = closest_segment(s, i0, i1, bct, m);
    T dist = length_sq(cl);
    if (dist < best) { for (int k = 0; k < 4; ++k) bc[k] = bct[k]; mask = m; closest_pt = cl; }
  }
  if (mask != 0) return closest_pt;
  bc[i0] = l0; bc[i1] = l1; bc[i2] = l2;
  mask = (1u << i0) | (1u << i1) | (1u << i2);
  return l0 * a + l1 * b + l2 * c;
}

template <class T> T determinant(V3<T> a, V3<T> b, V3<T> c, V3<T> d) { return dot(b - a, cross(c - a, d - a)); }

template <class T> V3<T> closest_tetrahedron(const Simplex<T>& s, T bc[4], unsigned& mask) {
  const T EPS = T(1e-8);
  V3<T> v0 = s.A[0], v1 = s.A[1], v2 = s.A[2], v3 = s.A[3];
  T det_t = determinant(v0, v1, v2, v3);
  bool degenerate = tabs(det_t) < EPS;
  T denom = degenerate ? EPS : det_t;
  T inv = T(1) / denom;
  V3<T> zero{T(0), T(0), T(0)};
  T l0 = determinant(zero, v1, v2, v3) * inv;
  T l1 = determinant(v0, zero, v2, v3) * inv;
  T l2 = determinant(v0, v1, zero, v3) * inv;
  T l3 = T(1) - l0 - l1 - l2;
  T best = T(1e30);
  V3<T> closest_pt = zero;
  T bct[4];
  for (int k = 0; k < 4; ++k) bc[k] = T(0);
  mask = 0;
  unsigned m;
  const int faces[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
  const T lam[4] = {l0, l1, l2, l3};
  for (int f = 0; f < 4; ++f) {
    if (lam[f] < T(0) || degenerate) {
      V3<T> cl = closest_triangle(s, faces[f][0], faces[f][1], faces[f][2], bct, m);
      T dist = length_sq(cl);
      if (dist < best) {
        for (int k = 0; k < 4; ++k) bc[k] = bct[k];
        mask = m;
        if (f < 3) best = dist;
        closest_pt = cl;
      }
    }
  }
  if (mask != 0) return closest_pt;
  bc[0] = l0; bc[1] = l1; bc[2] = l2; bc[3] = l3;
  mask = 15u;
  return zero;
}

template <class T> void simplex_get_closest(const Simplex<T>& s, V3<T>& pa, V3<T>& pb) {
  pa = V3<T>{T(0), T(0), T(0)};
  pb = pa;
  for (int i = 0; i < 4; ++i) {
    if ((s.mask & (1u << i)) == 0) continue;
    T b = s.bc[i];
    pa = pa + b * (s.B[i] + s.A[i]);
    pb = pb + b * s.B[i];
  }
}

template <class T>
ConvexResult<T> solve_gjk_core(const Shape<T>& ga, const Shape<T>& gb, Q<T> ob, V3<T> pb, T extend) {
  const int MAX_ITER = 30;
  const T CE = T(1e-4);
  const T EPS = T(1e-8);
  V3<T> zero{T(0), T(0), T(0)};
  T distance = T(0);
  V3<T> point_a = zero, point_b = zero, normal = zero;
  Simplex<T> s;
  for (int i = 0; i < 4; ++i) { s.B[i] = zero; s.A[i] = zero; s.bc[i] = T(0); }
  s.mask = 0;
  int iter = MAX_ITER;
  Vert<T> center = geometric_center(ob, pb);
  V3<T> v = center.BtoA;
  T dist_sq = length_sq(v);
  V3<T> last_search = {T(1), T(0), T(0)};
  while (iter > 0) {
    iter -= 1;
    if (dist_sq < CE * CE) {
      simplex_get_closest(s, point_a, point_b);
      return {false, point_a, point_b, zero, T(0)};
    }
    V3<T> search = -v;
    last_search = search;
    Vert<T> w = minkowski_support(ga, gb, search, ob, pb, extend);
    V3<T> wv = w.BtoA;
    T delta_dist = dot(v, v - wv);
    if (delta_dist < CE * tsqrt(dist_sq)) break;
    bool dup = false;
    for (int i = 0; i < 4; ++i) {
      if (s.mask & (1u << i)) {
        if (length_sq(s.A[i] - wv) < CE * CE) { dup = true; break; }
      }
    }
    if (dup) break;
    int use_count = 0, free_slot = 0;
    int idx[4] = {0, 0, 0, 0};
    for (int i = 0; i < 4; ++i) {
      if (s.mask & (1u << i)) idx[use_count++] = i;
      else free_slot = i;
    }
    idx[use_count] = free_slot;
    use_count += 1;
    s.B[free_slot] = w.B;
    s.A[free_slot] = w.BtoA;
    V3<T> closest = zero;
    bool success = true;
    T bc[4];
    unsigned mask;
    if (use_count == 1) {
      closest = s.A[idx[0]];
      s.mask = 1u << idx[0];
      s.bc[idx[0]] = T(1);
    } else if (use_count == 2) {
      closest = closest_segment(s, idx[0], idx[1], bc, mask);
      for (int k = 0; k < 4; ++k) s.bc[k] = bc[k];
      s.mask = mask;
    } else if (use_count == 3) {
      closest = closest_triangle(s, idx[0], idx[1], idx[2], bc, mask);
      for (int k = 0; k < 4; ++k) s.bc[k] = bc[k];
      s.mask = mask;
    } else if (use_count == 4) {
      closest = closest_tetrahedron(s, bc, mask);
      for (int k = 0; k < 4; ++k) s.bc[k] = bc[k];
      s.mask = mask;
      success = !(mask == 15u);
    } else {
      success = false;
    }
    if (!success) {
      simplex_get_closest(s, point_a, point_b);
      return {false, point_a, point_b, zero, T(0)};
    }
    v = closest;
    dist_sq = length_sq(v);
  }
  distance = tsqrt(dist_sq);
  simplex_get_closest(s, point_a, point_b);
  V3<T> delta = point_b - point_a;
  T dls = length_sq(delta);
  if (dls > EPS * EPS) {
    normal = delta * (T(1) / tsqrt(dls));
  } else if (distance > CE) {
    normal = v * (T(-1) / distance);
  } else {
    T nsq = length_sq(last_search);
    normal = nsq > T(0) ? last_search * (T(1) / tsqrt(nsq)) : V3<T>{T(1), T(0), T(0)};
  }
  return {true, point_a, point_b, normal, distance};
}

// ---------------------------------------------------------------- contact writing
template <class T> struct Writer {
  const OracleModel* m;
  std::vector<Xf<T>> body_q;
  std::vector<Contact<T>>* out;
};

template <class T> T contact_distance_d(const ContactData<T>& c, V3<T>& n, V3<T>& aw, V3<T>& bw) {
  T total = c.r_a + c.r_b + c.m_a + c.m_b;
  n = normalize(c.normal);
  aw = c.center - n * (T(0.5) * c.dist + c.r_a);
  bw = c.center + n * (T(0.5) * c.dist + c.r_b);
  V3<T> diff = bw - aw;
  T distance = dot(diff, n);
  return distance - total;
}

template <class T> bool passes_gap(const ContactData<T>& c) {
  V3<T> n, aw, bw;
  T d = contact_distance_d(c, n, aw, bw);
  return d <= c.gap_sum;
}

// write_contact (collide.py:62-148); check_gap selects the output_index < 0 path
template <class T> void write_contact(const ContactData<T>& c, Writer<T>& w, bool check_gap) {
  V3<T> n, aw, bw;
  T d = contact_distance_d(c, n, aw, bw);
  T gap = T(w.m->shape_gap[c.shape_a]) + T(w.m->shape_gap[c.shape_b]);
  if (check_gap && d > gap) return;
  T off_a = c.r_a + c.m_a, off_b = c.r_b + c.m_b;
  Contact<T> o;
  o.shape0 = c.shape_a;
  o.shape1 = c.shape_b;
  int b0 = w.m->shape_body[c.shape_a], b1 = w.m->shape_body[c.shape_b];
  Xf<T> Xa = b0 == -1 ? xf_identity<T>() : xf_inverse(w.body_q[b0]);
  Xf<T> Xb = b1 == -1 ? xf_identity<T>() : xf_inverse(w.body_q[b1]);
  o.point0 = xf_point(Xa, aw);
  o.point1 = xf_point(Xb, bw);
  o.offset0 = xf_vector(Xa, off_a * n);
  o.offset1 = xf_vector(Xb, (-off_b) * n);
  o.normal = n;
  o.margin0 = off_a;
  o.margin1 = off_b;
  o.key = make_sort_key(c.shape_a, c.shape_b, c.sub_key);
  w.out->push_back(o);
}

template <class T> ContactData<T> post_process(ContactData<T> c, int type_a, int type_b) {
  V3<T> n = c.normal;
  if (type_a == SPHERE || type_a == CAPSULE) {
    c.center = c.center + n * (c.r_a * T(0.5));
    c.dist = c.dist - c.r_a;
  }
  if (type_b == SPHERE || type_b == CAPSULE) {
    c.center = c.center - n * (c.r_b * T(0.5));
    c.dist = c.dist - c.r_b;
  }
  return c;
}

// ---------------------------------------------------------------- manifold (multicontact.py)
template <class T> T signed_area(V2<T> a, V2<T> b, V2<T> q) {
  return (b.x - a.x) * (q.y - a.y) - (b.y - a.y) * (q.x - a.x);
}

template <class T> V2<T> intersection_point(V2<T> s, V2<T> e, V2<T> a, V2<T> b) {
  T sa = signed_area(s, e, a);
  T sb = signed_area(s, e, b);
  T interp = tabs(sa) / tabs(sa - sb);
  return (T(1) - interp) * a + interp * b;
}

// The clip buffer holds 12 entries (two convex polygons of <= 5 vertices meet in <= 10); the index
// guards never fire on well-formed input.
static const int kClipMax = 12;

template <class T> void insert_v2(V2<T>* arr, int count, int index, V2<T> el) {
  int i = count;
  while (i > index) { arr[i] = arr[i - 1]; i -= 1; }
  if (index < kClipMax) arr[index] = el;
}

template <class T> int trim_in_place(V2<T> s, V2<T> e, V2<T>* loop, int loop_count) {
  if (loop_count < 3) return loop_count;
  V2<T> ia{T(0), T(0)}, ib{T(0), T(0)};
  int change_a = -1, change_b = -1;
  bool keep = false;
  bool prev_outside = signed_area(s, e, loop[0]) <= T(0);
  for (int i = 0; i < loop_count; ++i) {
    int nx = (i + 1) % loop_count;
    bool outside = signed_area(s, e, loop[nx]) <= T(0);
    if (outside != prev_outside) {
      V2<T> inter = intersection_point(s, e, loop[i], loop[nx]);
      if (change_a < 0) { change_a = i; keep = !prev_outside; ia = inter; }
      else { change_b = i; ib = inter; }
    }
    prev_outside = outside;
  }
  int new_loop_count;
  if (change_a >= 0 && change_b >= 0) {
    int li = -1;
    new_loop_count = loop_count;
    int i = 0;
    while (i < loop_count) {
      if (keep) { li += 1; if (li < kClipMax && i < kClipMax) loop[li] = loop[i]; }
      if (i == change_a || i == change_b) {
        V2<T> pt = i == change_a ? ia : ib;
        if (li == i && !keep) {
          li += 1;
          insert_v2(loop, new_loop_count < kClipMax - 1 ? new_loop_count : kClipMax - 1, li, pt);
          new_loop_count += 1;
          i += 1;
          change_b += 1;
          loop_count += 1;
        } else {
          li += 1;
          if (li < kClipMax) loop[li] = pt;
        }
        keep = !keep;
      }
      i += 1;
    }
    new_loop_count = li + 1 < kClipMax ? li + 1 : kClipMax;
  } else if (prev_outside) {
    new_loop_count = 0;
  } else {
    new_loop_count = loop_count;
  }
  return new_loop_count;
}

template <class T> bool segment_to_rect(V2<T>* p) {
  const T move = T(1e-5);
  V2<T> p0 = p[0], p1 = p[1];
  T dx = p1.x - p0.x, dy = p1.y - p0.y;
  T dl = tsqrt(dx * dx + dy * dy);
  if (!(dl > T(1e-10))) return false;
  T px = -dy / dl, py = dx / dl;
  T ox = px * move, oy = py * move;
  p[0] = V2<T>{p0.x - ox, p0.y - oy};
  p[1] = V2<T>{p1.x - ox, p1.y - oy};
  p[2] = V2<T>{p1.x + ox, p1.y + oy};
  p[3] = V2<T>{p0.x + ox, p0.y + oy};
  return true;
}

template <class T> int trim_all_in_place(V2<T>* trim, int tcount, V2<T>* loop, int lcount) {
  if (tcount <= 1) return std::min(1, lcount);
  if (tcount == 2) {
    if (!segment_to_rect(trim)) return std::min(1, lcount);
    tcount = 4;
  }
  if (lcount == 2) {
    if (!segment_to_rect(loop)) return std::min(1, lcount);
    lcount = 4;
  }
  int cur = lcount;
  V2<T> t0 = trim[0];
  for (int i = 0; i < tcount; ++i) {
    V2<T> s = trim[i];
    V2<T> e = i == tcount - 1 ? t0 : trim[i + 1];
    cur = trim_in_place(s, e, loop, cur);
  }
  return cur;
}

template <class T> void calipers(const V2<T>* hull, int n, int out[4]) {
  int p1 = 0, p3 = 1;
  V2<T> d0 = hull[p1] - hull[p3];
  T max_dist_sq = d0.x * d0.x + d0.y * d0.y;
  const T tie = T(1.0e-3);
  int j = 1;
  for (int i = 0; i < n; ++i) {
    V2<T> hi = hull[i], hi1 = hull[(i + 1) % n];
    while (true) {
      V2<T> hj = hull[j], hj1 = hull[(j + 1) % n];
      T a1 = signed_area(hi, hi1, hj1);
      T a0 = signed_area(hi, hi1, hj);
      if (a1 > a0) j = (j + 1) % n; else break;
    }
    V2<T> hj = hull[j];
    V2<T> d1{hi.x - hj.x, hi.y - hj.y};
    T ds1 = d1.x * d1.x + d1.y * d1.y;
    if (ds1 > max_dist_sq * (T(1) + tie)) { max_dist_sq = ds1; p1 = i; p3 = j; }
    V2<T> d2{hi1.x - hj.x, hi1.y - hj.y};
    T ds2 = d2.x * d2.x + d2.y * d2.y;
    if (ds2 > max_dist_sq * (T(1) + tie)) { max_dist_sq = ds2; p1 = (i + 1) % n; p3 = j; }
  }
  int p2 = 0, p4 = 0;
  T ma1 = T(0), ma2 = T(0);
  for (int i = 0; i < n; ++i) {
    T area = signed_area(hull[p1], hull[p3], hull[i]);
    if (area > ma1 * (T(1) + tie)) { ma1 = area; p2 = i; }
    else if (-area > ma2 * (T(1) + tie)) { ma2 = -area; p4 = i; }
  }
  out[0] = p1; out[1] = p2; out[2] = p3; out[3] = p4;
}

template <class T> int remove_zero_length_edges(V2<T>* loop, int count, T eps) {
  if (count < 2) return 0;
  int w = 0;
  for (int r = 1; r < count; ++r) {
    if (length_sq(loop[r] - loop[w]) > eps) { w += 1; loop[w] = loop[r]; }
  }
  int nc;
  if (w > 0) nc = length_sq(loop[w] - loop[0]) < eps ? w : w + 1;
  else nc = w + 1;
  if (nc < 2) nc = 0;
  return nc;
}

template <class T> int add_avoid_dup(V2<T>* arr, int count, V2<T> v, T eps, bool& added) {
  added = false;
  if (count > 0 && length_sq(arr[0] - v) < eps) return count;
  if (count > 1 && length_sq(arr[count - 1] - v) < eps) return count;
  arr[count] = v;
  added = true;
  return count + 1;
}

template <class T> struct Tracker { V3<T> ref, prev, normal; T area; };
template <class T> void tracker_update(Tracker<T>& t, V3<T> p, int id) {
  if (id == 0) { t.ref = p; t.area = T(0); }
  else if (id == 1) { t.prev = p; }
  else {
    V3<T> e1 = t.prev - t.ref, e2 = p - t.ref;
    V3<T> c = cross(e1, e2);
    T a = dot(c, c);
    if (a > t.area) { t.area = a; t.normal = c; }
    t.prev = p;
  }
}
template <class T> struct Projector { T d; V3<T> n; };

template <class T> V3<T> seg_projector_normal(V3<T> seg, V3<T> ref) {
  V3<T> right = cross(seg, ref);
  V3<T> n = cross(right, seg);
  T l = length(n);
  return l > T(1e-12) ? n / l : ref;
}

template <class T> void make_projectors(const Tracker<T>& ta, V3<T> anchor_a, const Tracker<T>& tb, V3<T> anchor_b,
                                        V3<T> cn, Projector<T>& pa, Projector<T>& pb) {
  V3<T> zero{T(0), T(0), T(0)};
  pa = Projector<T>{T(0), zero};
  pb = Projector<T>{T(0), zero};
  if (ta.area == T(0) && tb.area == T(0)) {
    V3<T> da = ta.prev - ta.ref, db = tb.prev - tb.ref;
    V3<T> ppa = T(0.5) * (ta.ref + ta.prev);
    pa.n = seg_projector_normal(da, cn);
    pa.d = -dot(ppa, pa.n);
    V3<T> ppb = T(0.5) * (tb.ref + tb.prev);
    pb.n = seg_projector_normal(db, cn);
    pb.d = -dot(ppb, pb.n);
    return;
  }
  if (ta.area > T(0)) {
    T l = tsqrt(tmax(T(1.0e-12), ta.area));
    pa.n = ta.normal / l;
    pa.d = -dot(anchor_a, pa.n);
  }
  if (tb.area > T(0)) {
    T l = tsqrt(tmax(T(1.0e-12), tb.area));
    pb.n = tb.normal / l;
    pb.d = -dot(anchor_b, pb.n);
  }
  if (ta.area == T(0)) {
    V3<T> dir = ta.prev - ta.ref;
    V3<T> pp = T(0.5) * (ta.ref + ta.prev);
    pa.n = seg_projector_normal(dir, pb.n);
    pa.d = -dot(pp, pa.n);
  }
  if (tb.area == T(0)) {
    V3<T> dir = tb.prev - tb.ref;
    V3<T> pp = T(0.5) * (tb.ref + tb.prev);
    pb.n = seg_projector_normal(dir, pa.n);
    pb.d = -dot(pp, pb.n);
  }
}

template <class T> V3<T> ray_plane(V3<T> o, V3<T> dir, T pd, V3<T> pn) {
  T denom = dot(dir, pn);
  if (tabs(denom) < T(1.0e-12)) return o;
  T t = -(dot(o, pn) + pd) / denom;
  return o + dir * t;
}

template <class T> struct PairCtx {
  Shape<T> ga, gb;
  Q<T> oa, ob_world;
  V3<T> pa_world, pb_world;
  ContactData<T> tmpl;
};

// post_process_axial_on_discrete_contact (collision_core.py:172-275): Minkowski adjustment, then the rolling
// projection for a cylinder or cone against a box / plane.  The cone's thresholds sin(atan2(r, 2 hh) -/+ 2 deg)
// are evaluated as sin a cos 2 -/+ cos a sin 2 with sin a = r / L, cos a = 2 hh / L, L = |(r, 2 hh)|.
template <class T> ContactData<T> post_process(ContactData<T> c, const PairCtx<T>& P) {
  const V3<T> normal = c.normal;
  c = post_process(c, P.ga.type, P.gb.type);
  const int ta = P.ga.type, tb = P.gb.type;
  const bool da = ta == BOX || ta == PLANE, db = tb == BOX || tb == PLANE;
  const bool xa = ta == CYLINDER || ta == CONE, xb = tb == CYLINDER || tb == CONE;
  if (!((da && xb) || (db && xa))) return c;
  const bool b_axial = da && xb;
  const V3<T> axis = qrot(b_axial ? P.ob_world : P.oa, V3<T>{T(0), T(0), T(1)});
  const Shape<T>& g = b_axial ? P.gb : P.ga;
  const V3<T> spos = b_axial ? P.pb_world : P.pa_world;
  const V3<T> an = b_axial ? normal : -normal;
  const T and_ = tabs(dot(axis, an));
  const T SIN2 = T(0.034899496702500969), COS2 = T(0.99939082701909576);
  bool rolling;
  if (g.type == CONE) {
    const T hh2 = T(2) * g.scale.y;
    const T l = tsqrt(g.scale.x * g.scale.x + hh2 * hh2);
    const T sa = g.scale.x / l, ca = hh2 / l;
    rolling = and_ >= sa * COS2 - ca * SIN2 && and_ <= sa * COS2 + ca * SIN2;
  } else {
    rolling = and_ <= SIN2;
  }
  if (rolling) {
    const V3<T> ppn = normalize(cross(axis, an));
    c.center = c.center - ppn * dot(c.center - spos, ppn);  // project_point_onto_plane (collision_core.py:51-65)
  }
  return c;
}

template <class T> void orthonormal_basis(V3<T> n, V3<T>& b1, V3<T>& b2) {
  if (n.z < T(0)) {
    T a = T(1) / (T(1) - n.z);
    T b = n.x * n.y * a;
    b1 = V3<T>{T(1) - n.x * n.x * a, -b, n.x};
    b2 = V3<T>{b, n.y * n.y * a - T(1), -n.y};
  } else {
    T a = T(1) / (T(1) + n.z);
    T b = -n.x * n.y * a;
    b1 = V3<T>{T(1) - n.x * n.x * a, b, -n.x};
    b2 = V3<T>{b, T(1) - n.y * n.y * a, -n.y};
  }
}

template <class T> int build_manifold(const PairCtx<T>& P, Q<T> rel_ob, V3<T> rel_pb, V3<T> p_a, V3<T> p_b,
                                      V3<T> normal, Writer<T>& w) {
  const T EPS = T(0.00001);
  const T SIN_TILT = T(0.03489949670250097);
  const T COS_TILT = T(0.9993908270190958);
  const T COS_DEEP = T(0.9999984769132877);
  const T PC[5] = {T(1.0), T(0.30901699437494745), T(-0.8090169943749473), T(-0.8090169943749476),
                   T(0.30901699437494723)};
  const T PS[5] = {T(0.0), T(0.9510565162951535), T(0.5877852522924732), T(-0.587785252292473),
                   T(-0.9510565162951536)};
  V3<T> ta, tb;
  orthonormal_basis(normal, ta, tb);
  V3<T> zero{T(0), T(0), T(0)};
  Tracker<T> tra{zero, zero, zero, T(0)}, trb{zero, zero, zero, T(0)};
  V3<T> center = T(0.5) * (p_a + p_b);
  V2<T> buf[12];
  for (int i = 0; i < 12; ++i) buf[i] = V2<T>{T(0), T(0)};
  V2<T>* bbuf = buf;
  V2<T>* abuf = buf + 5;
  int a_count = 0, b_count = 0;
  V3<T> lnb = qrot_inv(rel_ob, -normal);
  V3<T> ltab = qrot_inv(rel_ob, -ta);
  V3<T> ltbb = qrot_inv(rel_ob, -tb);
  for (int e = 0; e < 5; ++e) {
    T c = PC[e], s = PS[e];
    T ct = COS_TILT;
    T cs = c * SIN_TILT, ss = s * SIN_TILT;
    V3<T> dir_a = normal * ct + cs * ta + ss * tb;
    V3<T> pt_a = support(P.ga, dir_a);
    V3<T> pr_a = pt_a - center;
    V2<T> a2{dot(ta, pr_a), dot(tb, pr_a)};
    bool added;
    a_count = add_avoid_dup(abuf, a_count, a2, EPS, added);
    if (added) tracker_update(tra, pt_a, a_count - 1);
    V3<T> ldb = lnb * ct + cs * ltab + ss * ltbb;
    V3<T> pbl = support(P.gb, ldb);
    V3<T> pt_b = qrot(rel_ob, pbl) + rel_pb;
    V3<T> pr_b = pt_b - center;
    V2<T> b2{dot(ta, pr_b), dot(tb, pr_b)};
    b_count = add_avoid_dup(bbuf, b_count, b2, EPS, added);
    if (added) tracker_update(trb, pt_b, b_count - 1);
  }
  V3<T> normal_world = qrot(P.oa, normal);
  int count_out;
  T normal_dot;
  if (a_count < 2 || b_count < 2) {
    count_out = 0;
    normal_dot = T(0);
  } else {
    Projector<T> pa, pb;
    make_projectors(tra, p_a, trb, p_b, normal, pa, pb);
    if (tabs(dot(normal, pa.n)) < COS_TILT || tabs(dot(normal, pb.n)) < COS_TILT) {
      count_out = 0;
      normal_dot = T(0);
    } else {
      normal_dot = tabs(dot(pa.n, pb.n));
      int loop_count = trim_all_in_place(abuf, a_count, bbuf, b_count);
      loop_count = remove_zero_length_edges(bbuf, loop_count, EPS);
      if (loop_count > 1) {
        int res[4] = {0, 1, 2, 3};
        if (loop_count > 4) {
          calipers(bbuf, loop_count, res);
          loop_count = 4;
        }
        for (int i = 0; i < loop_count; ++i) {
          int ia = res[i];
          V3<T> pl = bbuf[ia].x * ta + bbuf[ia].y * tb + center;
          V3<T> a = ray_plane(pl, normal, pa.d, pa.n);
          V3<T> b = ray_plane(pl, normal, pb.d, pb.n);
          V3<T> cpl = T(0.5) * (a + b);
          T sd = dot(b - a, normal);
          V3<T> cpw = qrot(P.oa, cpl) + P.pa_world;
          ContactData<T> cd = P.tmpl;
          cd.center = cpw;
          cd.normal = normal_world;
          cd.dist = sd;
          cd.sub_key = (P.tmpl.sub_key << 3) | i;
          cd = post_process(cd, P);
          write_contact(cd, w, true);
        }
      } else {
        normal_dot = T(0);
        loop_count = 0;
      }
      count_out = std::min(loop_count, 4);
    }
  }
  if (normal_dot < COS_DEEP || count_out == 0) {
    V3<T> dcl = T(0.5) * (p_a + p_b);
    T dsd = dot(p_b - p_a, normal);
    V3<T> dcw = qrot(P.oa, dcl) + P.pa_world;
    ContactData<T> cd = P.tmpl;
    cd.center = dcw;
    cd.normal = normal_world;
    cd.dist = dsd;
    cd.sub_key = (P.tmpl.sub_key << 3) | count_out;
    cd = post_process(cd, P);
    write_contact(cd, w, true);
    count_out += 1;
  }
  return count_out;
}

// solve_convex_multi_contact (collision_convex.py:26-135)
template <class T> void solve_convex(const PairCtx<T>& P, T threshold, bool skip_multi, Writer<T>& w) {
  Q<T> rel_ob = qmul(qinv(P.oa), P.ob_world);
  V3<T> rel_pb = qrot_inv(P.oa, P.pb_world - P.pa_world);
  T margin_sum = P.tmpl.m_a + P.tmpl.m_b;
  const T eps = T(1.0e-4);
  T enlarge = margin_sum <= T(0) ? eps : (margin_sum < eps ? T(2) * eps : T(0));
  ConvexResult<T> r = solve_mpr_core(P.ga, P.gb, rel_ob, rel_pb, enlarge);
  V3<T> point_a, point_b, normal;
  T sd;
  if (r.flag) {
    sd = -r.value + enlarge;
    T half = enlarge * T(0.5);
    point_a = r.pa - r.n * half;
    point_b = r.pb + r.n * half;
    normal = r.n;
  } else {
    ConvexResult<T> g = solve_gjk_core(P.ga, P.gb, rel_ob, rel_pb, T(0));
    point_a = g.pa;
    point_b = g.pb;
    normal = g.n;
    sd = g.value;
  }
  if (skip_multi || sd > threshold) {
    V3<T> point = T(0.5) * (point_a + point_b);
    point = qrot(P.oa, point) + P.pa_world;
    V3<T> nws = qrot(P.oa, normal);
    ContactData<T> cd = P.tmpl;
    cd.center = point;
    cd.normal = nws;
    cd.dist = sd;
    cd.sub_key = P.tmpl.sub_key << 3;
    cd = post_process(cd, P);
    write_contact(cd, w, true);
    return;
  }
  build_manifold(P, rel_ob, rel_pb, point_a, point_b, normal, w);
}

// ---------------------------------------------------------------- narrow phase per pair
template <class T>
void mesh_pair(const OracleModel* m, const std::vector<ShapeWorld<T>>& sw, int sa, int sb, Writer<T>& w);

template <class T> void narrow_pair(const OracleModel* m, const std::vector<ShapeWorld<T>>& sw, int sa, int sb,
                                    Writer<T>& w) {
  if (sa == sb || sa < 0 || sb < 0) return;
  int ta = m->shape_type[sa], tb = m->shape_type[sb];
  if (ta > tb) { std::swap(sa, sb); std::swap(ta, tb); }
  // mesh pairs (narrow_phase.py:290-320): mesh-mesh, and mesh-infinite plane kept as (mesh, plane)
  if (ta == MESH || tb == MESH) {
    if (ta == PLANE) std::swap(sa, sb);
    mesh_pair(m, sw, sa, sb, w);
    return;
  }
  const ShapeWorld<T>& A = sw[sa];
  const ShapeWorld<T>& B = sw[sb];
  V3<T> scale_a = A.gscale, scale_b = B.gscale;
  V3<T> pos_a = A.X.p, pos_b = B.X.p;
  Q<T> qa = A.X.q, qb = B.X.q;
  T gap_sum = T(m->shape_gap[sa]) + T(m->shape_gap[sb]);
  bool plane_a = ta == PLANE;
  T r_a = (ta == SPHERE || ta == CAPSULE) ? scale_a.x : T(0);
  T r_b = (tb == SPHERE || tb == CAPSULE) ? scale_b.x : T(0);
  T cd[4] = {T(MAXVAL), T(MAXVAL), T(MAXVAL), T(MAXVAL)};
  V3<T> zero{T(0), T(0), T(0)};
  V3<T> cp[4] = {zero, zero, zero, zero};
  V3<T> cn = zero;
  V3<T> ez{T(0), T(0), T(1)};
  if (plane_a && tb == SPHERE) {
    V3<T> pn = qrot(qa, ez);
    plane_sphere(pn, pos_a, pos_b, scale_b.x, cd[0], cp[0]);
    cn = pn;
  } else if (plane_a && tb == ELLIPSOID) {
    V3<T> pn = qrot(qa, ez);
    plane_ellipsoid(pn, pos_a, pos_b, quat_to_matrix(qb), scale_b, cd[0], cp[0]);
    cn = pn;
  } else if (plane_a && tb == BOX) {
    V3<T> pn = qrot(qa, ez);
    M3<T> R = quat_to_matrix(qb);
    plane_box(pn, pos_a, pos_b, R, scale_b, gap_sum, cd, cp);
    cn = pn;
  } else if (ta == SPHERE && tb == SPHERE) {
    sphere_sphere(pos_a, scale_a.x, pos_b, scale_b.x, cd[0], cp[0], cn);
  } else if (plane_a && tb == CAPSULE) {
    V3<T> pn = qrot(qa, ez);
    V3<T> axis = qrot(qb, ez);
    V3<T> seg = axis * scale_b.y;
    plane_sphere(pn, pos_a, pos_b + seg, scale_b.x, cd[0], cp[0]);
    plane_sphere(pn, pos_a, pos_b - seg, scale_b.x, cd[1], cp[1]);
    cn = pn;
  } else if (plane_a && tb == CYLINDER) {
    V3<T> pn = qrot(qa, ez);
    plane_cylinder(pn, pos_a, pos_b, qrot(qb, ez), scale_b.x, scale_b.y, cd, cp);
    cn = pn;
  } else if (ta == SPHERE && tb == CAPSULE) {
    V3<T> axis = qrot(qb, ez);
    V3<T> seg = axis * scale_b.y;
    V3<T> pt = closest_segment_point(pos_b - seg, pos_b + seg, pos_a);
    sphere_sphere(pos_a, scale_a.x, pt, scale_b.x, cd[0], cp[0], cn);
  } else if (ta == CAPSULE && tb == CAPSULE) {
    V3<T> ax1 = qrot(qa, ez), ax2 = qrot(qb, ez);
    T d2[2];
    V3<T> p2[2];
    capsule_capsule(pos_a, ax1, scale_a.x, scale_a.y, pos_b, ax2, scale_b.x, scale_b.y, d2, p2, cn);
    cd[0] = d2[0]; cp[0] = p2[0]; cd[1] = d2[1]; cp[1] = p2[1];
  } else if (ta == SPHERE && tb == CYLINDER) {
    sphere_cylinder(pos_a, scale_a.x, pos_b, qrot(qb, ez), scale_b.x, scale_b.y, cd[0], cp[0], cn);
  } else if (ta == SPHERE && tb == BOX) {
    M3<T> R = quat_to_matrix(qb);
    sphere_box(pos_a, scale_a.x, pos_b, R, scale_b, cd[0], cp[0], cn);
  }
  int num = 0;
  for (int k = 0; k < 4; ++k) num += cd[k] < T(MAXVAL) ? 1 : 0;
  if (num > 0) {
    ContactData<T> c;
    c.normal = cn;
    c.r_a = r_a;
    c.r_b = r_b;
    c.m_a = A.margin;
    c.m_b = B.margin;
    c.shape_a = sa;
    c.shape_b = sb;
    c.gap_sum = gap_sum;
    bool valid[4];
    for (int k = 0; k < 4; ++k) {
      valid[k] = false;
      if (cd[k] < T(MAXVAL)) {
        c.center = cp[k];
        c.dist = cd[k];
        valid[k] = passes_gap(c);
      }
    }
    for (int k = 0; k < 4; ++k) {
      if (!valid[k]) continue;
      c.center = cp[k];
      c.dist = cd[k];
      c.sub_key = k;
      write_contact(c, w, false);
    }
    return;
  }
  // ---- GJK / MPR route (narrow_phase.py:608-778)
  bool inf_a = ta == PLANE && scale_a.x == T(0) && scale_a.y == T(0);
  bool inf_b = tb == PLANE && scale_b.x == T(0) && scale_b.y == T(0);
  if (inf_a && inf_b) return;
  T bs_ra = T(0), bs_rb = T(0);
  if (inf_a || inf_b) {
    V3<T> ca = T(0.5) * (A.lo + A.hi), cb = T(0.5) * (B.lo + B.hi);
    bs_ra = length(T(0.5) * (A.hi - A.lo));
    bs_rb = length(T(0.5) * (B.hi - B.lo));
    V3<T> ppos = inf_a ? pos_a : pos_b;
    Q<T> pq = inf_a ? qa : qb;
    V3<T> oc = inf_a ? cb : ca;
    T orad = inf_a ? bs_rb : bs_ra;
    V3<T> pn = qrot(pq, ez);
    T center_dist = dot(oc - ppos, pn);
    if (!(center_dist <= orad)) return;
  }
  PairCtx<T> P;
  P.ga = Shape<T>{ta, scale_a};
  P.gb = Shape<T>{tb, scale_b};
  P.oa = qa;
  P.ob_world = qb;
  P.pa_world = pos_a;
  P.pb_world = pos_b;
  if (inf_a) {
    T other = bs_rb + gap_sum;
    T lateral = other * T(10), depth = other * T(10);
    P.ga = Shape<T>{BOX, V3<T>{lateral, lateral, depth}};
    V3<T> pn = qrot(qa, ez);
    T dn = dot(pos_b - pos_a, pn);
    V3<T> surf = pos_b - pn * dn;
    P.pa_world = surf - pn * depth;
  }
  if (inf_b) {
    T other = bs_ra + gap_sum;
    T lateral = other * T(10), depth = other * T(10);
    P.gb = Shape<T>{BOX, V3<T>{lateral, lateral, depth}};
    V3<T> pn = qrot(qb, ez);
    T dn = dot(pos_a - pos_b, pn);
    V3<T> surf = pos_a - pn * dn;
    P.pb_world = surf - pn * depth;
  }
  T rea = T(0), reb = T(0);
  const T small_radius = T(0.0001);
  if (P.ga.type == SPHERE || P.ga.type == CAPSULE) { rea = P.ga.scale.x; P.ga.scale.x = small_radius; }
  if (P.gb.type == SPHERE || P.gb.type == CAPSULE) { reb = P.gb.scale.x; P.gb.scale.x = small_radius; }
  ContactData<T>& t = P.tmpl;
  t.center = zero;
  t.normal = zero;
  t.dist = T(0);
  t.r_a = rea;
  t.r_b = reb;
  t.m_a = A.margin;
  t.m_b = B.margin;
  t.shape_a = sa;
  t.shape_b = sb;
  t.gap_sum = gap_sum;
  t.sub_key = 0;
  T threshold = gap_sum + rea + reb + A.margin + B.margin;
  bool skip_multi = P.ga.type == SPHERE || P.gb.type == SPHERE || P.ga.type == ELLIPSOID || P.gb.type == ELLIPSOID;
  solve_convex(P, threshold, skip_multi, w);
}

template <class T, class S>
int collide(const OracleModel* m, const S* body_q, std::vector<Contact<T>>& out) {
  std::vector<ShapeWorld<T>> sw(m->shape_count);
  for (int s = 0; s < m->shape_count; ++s) sw[s] = compute_shape_world<T, S>(m, body_q, s);
  std::vector<std::pair<int, int>> pairs;
  sap_pairs(m, sw, pairs);
  Writer<T> w;
  w.m = m;
  w.body_q.resize(m->body_count);
  for (int b = 0; b < m->body_count; ++b) w.body_q[b] = load_xf_t<T, S>(body_q + 7 * b);
  w.out = &out;
  for (auto& p : pairs) narrow_pair(m, sw, p.first, p.second, w);
  std::stable_sort(out.begin(), out.end(), [](const Contact<T>& a, const Contact<T>& b) { return a.key < b.key; });
  return int(out.size());
}

}  // namespace oracle
