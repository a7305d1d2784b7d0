// Device math for the gfx950 rigid-contact kernels.
//
// Fixed arithmetic conventions (every kernel and the CPU oracle agree on them, which is what makes
// contact sets bit-identical; everything is compiled with -ffp-contract=off so no operation fuses):
//   dot(a,b)         = (a.x*b.x + a.y*b.y) + a.z*b.z
//   cross(a,b)       = (a.y*b.z - a.z*b.y, a.z*b.x - a.x*b.z, a.x*b.y - a.y*b.x)
//   normalize(a)     = l = sqrt(dot(a,a)); l > 0 ? a / l : 0
//   rot(q, v)        = v*c + u*d + cross(u,v)*w*2 with c = 2*w*w - 1, d = 2*dot(u,v), q = (u, w)
//   rot_inv(q, v)    = v*c + u*d - cross(u,v)*w*2
//   qmul             = Hamilton product on (x, y, z, w)
//   mat(q)           = columns rot(q, e_x), rot(q, e_y), rot(q, e_z)
//   M*v              = (m[i][0]*v.x + m[i][1]*v.y) + m[i][2]*v.z
//   fmin2(a,b) = a < b ? a : b, fmax2(a,b) = a > b ? a : b, clampf(x,lo,hi) = fmin2(fmax2(lo,x),hi)
// These are the semantics of the Warp builtins that Newton's kernels are written against.
// Division and sqrt are IEEE correctly rounded (hipcc default for f32 on gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nmc {

struct f3 { float x, y, z; };
struct f2 { float x, y; };
struct qf { float x, y, z, w; };
struct m33 { float m[3][3]; };
struct xform { f3 p; qf q; };

__device__ __forceinline__ float fmin2(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float fmax2(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return fmin2(fmax2(lo, x), hi); }

__device__ __forceinline__ f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 operator*(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ f3 operator/(f3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float len_sq(f3 a) { return dot(a, a); }
__device__ __forceinline__ float len(f3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ f3 normalize(f3 a) {
  float l = len(a);
  if (l > 0.0f) return a / l;
  return {0.0f, 0.0f, 0.0f};
}
__device__ __forceinline__ f3 vmin(f3 a, f3 b) { return {fmin2(a.x, b.x), fmin2(a.y, b.y), fmin2(a.z, b.z)}; }
__device__ __forceinline__ f3 vmax(f3 a, f3 b) { return {fmax2(a.x, b.x), fmax2(a.y, b.y), fmax2(a.z, b.z)}; }
__device__ __forceinline__ float get(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
__device__ __forceinline__ void put(f3& a, int i, float v) {
  if (i == 0) a.x = v; else if (i == 1) a.y = v; else a.z = v;
}

__device__ __forceinline__ f2 operator+(f2 a, f2 b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ f2 operator-(f2 a, f2 b) { return {a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ f2 operator*(float s, f2 a) { return {s * a.x, s * a.y}; }
__device__ __forceinline__ float len_sq(f2 a) { return a.x * a.x + a.y * a.y; }

__device__ __forceinline__ qf qmul(qf a, qf b) {
  return {a.w * b.x + b.w * a.x + a.y * b.z - b.y * a.z, a.w * b.y + b.w * a.y + a.z * b.x - b.z * a.x,
          a.w * b.z + b.w * a.z + a.x * b.y - b.x * a.y, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
__device__ __forceinline__ qf qadd(qf a, qf b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
__device__ __forceinline__ qf qscale(qf a, float s) { return {a.x * s, a.y * s, a.z * s, a.w * s}; }
__device__ __forceinline__ qf qinv(qf q) { return {-q.x, -q.y, -q.z, q.w}; }
__device__ __forceinline__ qf qnormalize(qf q) {
  float l = sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
  if (l > 0.0f) return {q.x / l, q.y / l, q.z / l, q.w / l};
  return {0.0f, 0.0f, 0.0f, 0.0f};
}
// asin / acos by the Cephes single-precision method (S. L. Moshier, Cephes Math Library, asinf.c and
// acosf.c): odd polynomial in x^2 on |x| <= 0.5, half-angle reduction through sqrt above; about 2 ulp.
// Written out operation by operation (no contraction) so every IEEE-conforming FP32 unit computes the
// same bits.
__device__ __forceinline__ float asin_c(float x) {
  float a = fabsf(x);
  float z, r;
  bool big = a > 0.5f;
  if (big) {
    z = 0.5f * (1.0f - a);
    r = sqrtf(z);
  } else {
    z = a * a;
    r = a;
  }
  float poly = (((4.2163199048e-2f * z + 2.4181311049e-2f) * z + 4.5470025998e-2f) * z + 7.4953002686e-2f) * z +
               1.6666752422e-1f;
  float y = poly * z * r + r;
  if (big) y = 1.5707963267948966f - (y + y);
  return x < 0.0f ? -y : y;
}
__device__ __forceinline__ float acos_c(float x) {
  if (x < -0.5f) return 3.14159265358979323846f - 2.0f * asin_c(sqrtf(0.5f * (1.0f + x)));
  if (x > 0.5f) return 2.0f * asin_c(sqrtf(0.5f * (1.0f - x)));
  return 1.5707963267948966f - asin_c(x);
}
__device__ __forceinline__ f3 rot(qf q, f3 x) {
  float c = 2.0f * q.w * q.w - 1.0f;
  float d = 2.0f * (q.x * x.x + q.y * x.y + q.z * x.z);
  return {x.x * c + q.x * d + (q.y * x.z - q.z * x.y) * q.w * 2.0f,
          x.y * c + q.y * d + (q.z * x.x - q.x * x.z) * q.w * 2.0f,
          x.z * c + q.z * d + (q.x * x.y - q.y * x.x) * q.w * 2.0f};
}
__device__ __forceinline__ f3 rot_inv(qf q, f3 x) {
  float c = 2.0f * q.w * q.w - 1.0f;
  float d = 2.0f * (q.x * x.x + q.y * x.y + q.z * x.z);
  return {x.x * c + q.x * d - (q.y * x.z - q.z * x.y) * q.w * 2.0f,
          x.y * c + q.y * d - (q.z * x.x - q.x * x.z) * q.w * 2.0f,
          x.z * c + q.z * d - (q.x * x.y - q.y * x.x) * q.w * 2.0f};
}
__device__ __forceinline__ m33 qmat(qf q) {
  f3 c0 = rot(q, f3{1.0f, 0.0f, 0.0f});
  f3 c1 = rot(q, f3{0.0f, 1.0f, 0.0f});
  f3 c2 = rot(q, f3{0.0f, 0.0f, 1.0f});
  m33 r;
  r.m[0][0] = c0.x; r.m[1][0] = c0.y; r.m[2][0] = c0.z;
  r.m[0][1] = c1.x; r.m[1][1] = c1.y; r.m[2][1] = c1.z;
  r.m[0][2] = c2.x; r.m[1][2] = c2.y; r.m[2][2] = c2.z;
  return r;
}
__device__ __forceinline__ f3 mv(const m33& a, f3 v) {
  return {a.m[0][0] * v.x + a.m[0][1] * v.y + a.m[0][2] * v.z, a.m[1][0] * v.x + a.m[1][1] * v.y + a.m[1][2] * v.z,
          a.m[2][0] * v.x + a.m[2][1] * v.y + a.m[2][2] * v.z};
}
__device__ __forceinline__ f3 mtv(const m33& a, f3 v) {
  return {a.m[0][0] * v.x + a.m[1][0] * v.y + a.m[2][0] * v.z, a.m[0][1] * v.x + a.m[1][1] * v.y + a.m[2][1] * v.z,
          a.m[0][2] * v.x + a.m[1][2] * v.y + a.m[2][2] * v.z};
}
__device__ __forceinline__ m33 mscale(float s, const m33& a) {
  m33 r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) r.m[i][j] = s * a.m[i][j];
  return r;
}

__device__ __forceinline__ xform xident() { return xform{{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f, 1.0f}}; }
__device__ __forceinline__ f3 xpoint(const xform& t, f3 p) { return rot(t.q, p) + t.p; }
__device__ __forceinline__ f3 xvector(const xform& t, f3 v) { return rot(t.q, v); }
__device__ __forceinline__ xform xmul(const xform& a, const xform& b) { return xform{rot(a.q, b.p) + a.p, qmul(a.q, b.q)}; }
__device__ __forceinline__ xform xinverse(const xform& t) {
  qf qi = qinv(t.q);
  return xform{-rot(qi, t.p), qi};
}

__device__ __forceinline__ xform load_xform(const float* __restrict__ a) {
  return xform{{a[0], a[1], a[2]}, {a[3], a[4], a[5], a[6]}};
}
__device__ __forceinline__ void store_xform(float* __restrict__ a, const xform& t) {
  a[0] = t.p.x; a[1] = t.p.y; a[2] = t.p.z;
  a[3] = t.q.x; a[4] = t.q.y; a[5] = t.q.z; a[6] = t.q.w;
}
__device__ __forceinline__ f3 load3(const float* __restrict__ a) { return {a[0], a[1], a[2]}; }
__device__ __forceinline__ void store3(float* __restrict__ a, f3 v) { a[0] = v.x; a[1] = v.y; a[2] = v.z; }
__device__ __forceinline__ m33 load_m33(const float* __restrict__ a) {
  m33 r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) r.m[i][j] = a[3 * i + j];
  return r;
}

}  // namespace nmc
