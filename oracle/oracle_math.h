// CPU oracle -- TEST INFRASTRUCTURE ONLY.
//
// Scalar vector/quaternion/transform math for the CPU restatement of Newton's rigid-contact step.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load the oracle; the product
// path (newton_custom_contact_amd/csrc) never includes or links anything under oracle/.
//
// Every function is templated on the scalar type T (float for parity, double for the FP64 rerun).
// The arithmetic order of each operation is fixed here and is the one the HIP kernels use:
//   dot(a,b)          = (a.x*b.x + a.y*b.y) + a.z*b.z                         (Warp vec dot)
//   cross(a,b)        = (a.y*b.z - a.z*b.y, a.z*b.x - a.x*b.z, a.x*b.y - a.y*b.x)
//   normalize(a)      = l = sqrt(dot(a,a)); l > 0 ? a / l : 0                  (Warp normalize, kEps = 0)
//   quat_rotate(q,v)  = v*c + u*d + cross(u,v)*w*2, c = 2*w*w - 1, d = 2*dot(u,v)
//   quat_rotate_inv   = v*c + u*d - cross(u,v)*w*2
//   quat_mul          = Hamilton product, (x, y, z, w) layout
//   quat_to_matrix(q) = columns quat_rotate(q, e_x), quat_rotate(q, e_y), quat_rotate(q, e_z)
//   mat*vec           = (m[i][0]*v.x + m[i][1]*v.y) + m[i][2]*v.z
//   min(a,b) = a < b ? a : b, max(a,b) = a > b ? a : b, clamp(x,lo,hi) = min(max(lo,x),hi)
// These are the semantics Newton's kernels rely on from Warp's builtins (warp-lang 1.14, not vendored
// in the reference); they are restated from the published Warp builtin definitions.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

namespace oracle {

// Order of the SAP sort (broad_phase_sap.py:600-640 sorts with wp.utils.segmented_sort_pairs, a radix sort
// on the float keys' sign-flipped bit patterns): -NaN < -inf < ... < -0 = +0 < ... < +inf < +NaN, ties in
// index order under a stable sort.  For finite keys this is the order of operator<; unlike operator< it is a
// strict weak order when a key is NaN (a body whose state has diverged), so the sweep then stays the
// reference's.
inline uint64_t sap_order_key(float f) {
  if (f == 0.0f) f = 0.0f;
  uint32_t u;
  std::memcpy(&u, &f, sizeof u);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
inline uint64_t sap_order_key(double f) {
  if (f == 0.0) f = 0.0;
  uint64_t u;
  std::memcpy(&u, &f, sizeof u);
  return (u >> 63) ? ~u : (u | (uint64_t(1) << 63));
}

template <class T> struct V2 { T x, y; };
template <class T> struct V3 { T x, y, z; };
template <class T> struct Q { T x, y, z, w; };
template <class T> struct M3 { T m[3][3]; };
template <class T> struct Xf { V3<T> p; Q<T> q; };

template <class T> inline T tmin(T a, T b) { return a < b ? a : b; }
template <class T> inline T tmax(T a, T b) { return a > b ? a : b; }
template <class T> inline T tclamp(T x, T lo, T hi) { return tmin(tmax(lo, x), hi); }
template <class T> inline T tabs(T a) { return std::fabs(a); }
template <class T> inline T tsqrt(T a) { return std::sqrt(a); }

// asin / acos.  FP64: libm.  FP32: the Cephes single-precision method (S. L. Moshier, Cephes Math
// Library, asinf.c / acosf.c), evaluated operation by operation so that it is reproducible on any
// IEEE FP32 unit; within 2 ulp of the exact value (pinned against float64 arcsin/arccos in the tests).
inline double tasin(double x) { return std::asin(x); }
inline double tacos(double x) { return std::acos(x); }
inline float tasin(float x) {
  float a = std::fabs(x);
  bool big = a > 0.5f;
  float z = big ? 0.5f * (1.0f - a) : a * a;
  float r = big ? std::sqrt(z) : a;
  float poly = 4.2163199048e-2f;
  poly = poly * z + 2.4181311049e-2f;
  poly = poly * z + 4.5470025998e-2f;
  poly = poly * z + 7.4953002686e-2f;
  poly = poly * z + 1.6666752422e-1f;
  float y = poly * z * r + r;
  if (big) y = static_cast<float>(1.5707963267948966) - (y + y);
  return x < 0.0f ? -y : y;
}
inline float tacos(float x) {
  const float pi = static_cast<float>(3.14159265358979323846), half_pi = static_cast<float>(1.5707963267948966);
  if (x < -0.5f) return pi - 2.0f * tasin(std::sqrt(0.5f * (1.0f + x)));
  if (x > 0.5f) return 2.0f * tasin(std::sqrt(0.5f * (1.0f - x)));
  return half_pi - tasin(x);
}

template <class T> inline V3<T> v3(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> inline V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline V3<T> cross(V3<T> a, V3<T> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> inline T length_sq(V3<T> a) { return dot(a, a); }
template <class T> inline T length(V3<T> a) { return tsqrt(dot(a, a)); }
template <class T> inline V3<T> normalize(V3<T> a) {
  T l = length(a);
  if (l > T(0)) return a / l;
  return {T(0), T(0), T(0)};
}
template <class T> inline V3<T> vmin(V3<T> a, V3<T> b) { return {tmin(a.x, b.x), tmin(a.y, b.y), tmin(a.z, b.z)}; }
template <class T> inline V3<T> vmax(V3<T> a, V3<T> b) { return {tmax(a.x, b.x), tmax(a.y, b.y), tmax(a.z, b.z)}; }
template <class T> inline T comp(V3<T> a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
template <class T> inline void set_comp(V3<T>& a, int i, T v) {
  if (i == 0) a.x = v; else if (i == 1) a.y = v; else a.z = v;
}

template <class T> inline V2<T> operator+(V2<T> a, V2<T> b) { return {a.x + b.x, a.y + b.y}; }
template <class T> inline V2<T> operator-(V2<T> a, V2<T> b) { return {a.x - b.x, a.y - b.y}; }
template <class T> inline V2<T> operator*(T s, V2<T> a) { return {s * a.x, s * a.y}; }
template <class T> inline T length_sq(V2<T> a) { return a.x * a.x + a.y * a.y; }

template <class T> inline Q<T> qmul(Q<T> a, Q<T> b) {
  return {a.w * b.x + b.w * a.x + a.y * b.z - b.y * a.z, a.w * b.y + b.w * a.y + a.z * b.x - b.z * a.x,
          a.w * b.z + b.w * a.z + a.x * b.y - b.x * a.y, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
template <class T> inline Q<T> qadd(Q<T> a, Q<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
template <class T> inline Q<T> qscale(Q<T> a, T s) { return {a.x * s, a.y * s, a.z * s, a.w * s}; }
template <class T> inline Q<T> qscale_l(T s, Q<T> a) { return {s * a.x, s * a.y, s * a.z, s * a.w}; }
template <class T> inline Q<T> qinv(Q<T> q) { return {-q.x, -q.y, -q.z, q.w}; }
template <class T> inline Q<T> qnormalize(Q<T> q) {
  T l = tsqrt(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
  if (l > T(0)) return {q.x / l, q.y / l, q.z / l, q.w / l};
  return {T(0), T(0), T(0), T(0)};
}
template <class T> inline V3<T> qrot(Q<T> q, V3<T> x) {
  T c = T(2) * q.w * q.w - T(1);
  T d = T(2) * (q.x * x.x + q.y * x.y + q.z * x.z);
  return {x.x * c + q.x * d + (q.y * x.z - q.z * x.y) * q.w * T(2),
          x.y * c + q.y * d + (q.z * x.x - q.x * x.z) * q.w * T(2),
          x.z * c + q.z * d + (q.x * x.y - q.y * x.x) * q.w * T(2)};
}
template <class T> inline V3<T> qrot_inv(Q<T> q, V3<T> x) {
  T c = T(2) * q.w * q.w - T(1);
  T d = T(2) * (q.x * x.x + q.y * x.y + q.z * x.z);
  return {x.x * c + q.x * d - (q.y * x.z - q.z * x.y) * q.w * T(2),
          x.y * c + q.y * d - (q.z * x.x - q.x * x.z) * q.w * T(2),
          x.z * c + q.z * d - (q.x * x.y - q.y * x.x) * q.w * T(2)};
}
template <class T> inline M3<T> quat_to_matrix(Q<T> q) {
  V3<T> c0 = qrot(q, V3<T>{T(1), T(0), T(0)});
  V3<T> c1 = qrot(q, V3<T>{T(0), T(1), T(0)});
  V3<T> c2 = qrot(q, V3<T>{T(0), T(0), T(1)});
  M3<T> r;
  r.m[0][0] = c0.x; r.m[1][0] = c0.y; r.m[2][0] = c0.z;
  r.m[0][1] = c1.x; r.m[1][1] = c1.y; r.m[2][1] = c1.z;
  r.m[0][2] = c2.x; r.m[1][2] = c2.y; r.m[2][2] = c2.z;
  return r;
}
template <class T> inline V3<T> mv(const M3<T>& a, V3<T> v) {
  return {a.m[0][0] * v.x + a.m[0][1] * v.y + a.m[0][2] * v.z, a.m[1][0] * v.x + a.m[1][1] * v.y + a.m[1][2] * v.z,
          a.m[2][0] * v.x + a.m[2][1] * v.y + a.m[2][2] * v.z};
}
template <class T> inline V3<T> mtv(const M3<T>& a, V3<T> v) {  // transpose(a) @ v
  return {a.m[0][0] * v.x + a.m[1][0] * v.y + a.m[2][0] * v.z, a.m[0][1] * v.x + a.m[1][1] * v.y + a.m[2][1] * v.z,
          a.m[0][2] * v.x + a.m[1][2] * v.y + a.m[2][2] * v.z};
}
template <class T> inline M3<T> mscale_l(T s, const M3<T>& a) {
  M3<T> r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.m[i][j] = s * a.m[i][j];
  return r;
}

template <class T> inline Xf<T> xf_identity() { return Xf<T>{{T(0), T(0), T(0)}, {T(0), T(0), T(0), T(1)}}; }
template <class T> inline V3<T> xf_point(const Xf<T>& t, V3<T> p) { return qrot(t.q, p) + t.p; }
template <class T> inline V3<T> xf_vector(const Xf<T>& t, V3<T> v) { return qrot(t.q, v); }
template <class T> inline Xf<T> xf_mul(const Xf<T>& a, const Xf<T>& b) {
  return Xf<T>{qrot(a.q, b.p) + a.p, qmul(a.q, b.q)};
}
template <class T> inline Xf<T> xf_inverse(const Xf<T>& t) {
  Q<T> qi = qinv(t.q);
  return Xf<T>{-qrot(qi, t.p), qi};
}

}  // namespace oracle
