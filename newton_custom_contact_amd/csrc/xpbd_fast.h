// World-frame arithmetic for the XPBD contact iterations on gfx950.
//
// The same method as the reference-order solve (solve_body_contact_positions, kernels.py:2062-2299, and
// apply_body_deltas, kernels.py:809-879: the same rows, clamps, relaxation and Jacobi accumulate-then-apply
// order), evaluated with fewer VALU instructions:
//   * each body's pose is held once per iteration as a 3x3 rotation matrix (the linear map of Warp's
//     quat_rotate, (2w^2-1) I + 2w [u]x + 2 u u^T, for any q) beside its origin, world centre of mass and
//     world-frame inverse inertia R I^-1 R^T, so a contact row costs one quadratic form per body where the
//     reference rotates the lever arm into the body frame and multiplies by I^-1;
//   * static shapes use a frame slot of their own (identity, zero inverse mass and inertia): no branches on
//     the body index;
//   * products and sums fuse into FMAs; 1/x and 1/sqrt(x) are the hardware v_rcp_f32 / v_rsq_f32 (1 ulp).
// Results agree with the reference-order arithmetic to rounding (the solver's parity tolerance is 1e-5
// relative per step), not bit for bit.
#pragma once
#include <hip/hip_runtime.h>

#include "nmc_math.h"

namespace nmc {
namespace fastm {

__device__ __forceinline__ float dot3(f3 a, f3 b) { return fmaf(a.x, b.x, fmaf(a.y, b.y, a.z * b.z)); }
__device__ __forceinline__ f3 cross3(f3 a, f3 b) {
  return {fmaf(a.y, b.z, -(a.z * b.y)), fmaf(a.z, b.x, -(a.x * b.z)), fmaf(a.x, b.y, -(a.y * b.x))};
}
// s * a + y
__device__ __forceinline__ f3 axpy(float s, f3 a, f3 y) { return {fmaf(s, a.x, y.x), fmaf(s, a.y, y.y), fmaf(s, a.z, y.z)}; }
__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ f3 xyz(float4 v) { return {v.x, v.y, v.z}; }

// Squared lengths below this are treated as zero where the reference normalises a vector (a row whose
// error is below 1e-15 m changes nothing at FP32 resolution); keeps v_rsq_f32 away from denormals.
constexpr float TINY_SQ = 1.0e-30f;

// Per-body frame in LDS: 6 float4.
//   r0, r1, r2  rows of the rotation matrix, body origin in .w
//   cm          world centre of mass, effective inverse mass in .w
//   w0, w1      world-frame effective inverse inertia (xx, xy, xz, yy), (yz, zz, 0, 0)
constexpr int FRAME_F4 = 6;

__device__ __forceinline__ void rot_rows(qf q, float4& r0, float4& r1, float4& r2) {
  const float c = fmaf(2.0f * q.w, q.w, -1.0f);
  const float x2 = 2.0f * q.x, y2 = 2.0f * q.y, z2 = 2.0f * q.z, w2 = 2.0f * q.w;
  r0.x = fmaf(x2, q.x, c);
  r0.y = fmaf(x2, q.y, -(w2 * q.z));
  r0.z = fmaf(x2, q.z, w2 * q.y);
  r1.x = fmaf(y2, q.x, w2 * q.z);
  r1.y = fmaf(y2, q.y, c);
  r1.z = fmaf(y2, q.z, -(w2 * q.x));
  r2.x = fmaf(z2, q.x, -(w2 * q.y));
  r2.y = fmaf(z2, q.y, w2 * q.x);
  r2.z = fmaf(z2, q.z, c);
}

// R p + origin
__device__ __forceinline__ f3 fpoint(float4 r0, float4 r1, float4 r2, f3 p) {
  return {fmaf(r0.x, p.x, fmaf(r0.y, p.y, fmaf(r0.z, p.z, r0.w))),
          fmaf(r1.x, p.x, fmaf(r1.y, p.y, fmaf(r1.z, p.z, r1.w))),
          fmaf(r2.x, p.x, fmaf(r2.y, p.y, fmaf(r2.z, p.z, r2.w)))};
}
// R v
__device__ __forceinline__ f3 fvec(float4 r0, float4 r1, float4 r2, f3 v) {
  return {fmaf(r0.x, v.x, fmaf(r0.y, v.y, r0.z * v.z)), fmaf(r1.x, v.x, fmaf(r1.y, v.y, r1.z * v.z)),
          fmaf(r2.x, v.x, fmaf(r2.y, v.y, r2.z * v.z))};
}
// R^T v
__device__ __forceinline__ f3 fvec_t(float4 r0, float4 r1, float4 r2, f3 v) {
  return {fmaf(r0.x, v.x, fmaf(r1.x, v.y, r2.x * v.z)), fmaf(r0.y, v.x, fmaf(r1.y, v.y, r2.y * v.z)),
          fmaf(r0.z, v.x, fmaf(r1.z, v.y, r2.z * v.z))};
}
__device__ __forceinline__ f3 fmv(const m33& a, f3 v) {
  return {fmaf(a.m[0][0], v.x, fmaf(a.m[0][1], v.y, a.m[0][2] * v.z)),
          fmaf(a.m[1][0], v.x, fmaf(a.m[1][1], v.y, a.m[1][2] * v.z)),
          fmaf(a.m[2][0], v.x, fmaf(a.m[2][1], v.y, a.m[2][2] * v.z))};
}
// a^T W a for the symmetric W held in (w0, w1)
__device__ __forceinline__ float quad(float4 w0, float4 w1, f3 a) {
  const float t0 = fmaf(w0.y, a.y, w0.z * a.z);
  float s = a.x * fmaf(2.0f, t0, w0.x * a.x);
  s = fmaf(a.y, fmaf(2.0f, w1.x * a.z, w0.w * a.y), s);
  return fmaf(w1.y * a.z, a.z, s);
}

// Frame of a body from its rotation rows (origin in .w), world centre of mass c, body-frame inverse inertia J
// and inverse mass inv_m (both already zero for kinematic bodies).
__device__ __forceinline__ void frame_store(float4* F, float4 r0, float4 r1, float4 r2, f3 c, const m33& J,
                                            float inv_m) {
  // T = R J, W = T R^T (upper triangle)
  const f3 t0 = {fmaf(r0.x, J.m[0][0], fmaf(r0.y, J.m[1][0], r0.z * J.m[2][0])),
                 fmaf(r0.x, J.m[0][1], fmaf(r0.y, J.m[1][1], r0.z * J.m[2][1])),
                 fmaf(r0.x, J.m[0][2], fmaf(r0.y, J.m[1][2], r0.z * J.m[2][2]))};
  const f3 t1 = {fmaf(r1.x, J.m[0][0], fmaf(r1.y, J.m[1][0], r1.z * J.m[2][0])),
                 fmaf(r1.x, J.m[0][1], fmaf(r1.y, J.m[1][1], r1.z * J.m[2][1])),
                 fmaf(r1.x, J.m[0][2], fmaf(r1.y, J.m[1][2], r1.z * J.m[2][2]))};
  const f3 t2 = {fmaf(r2.x, J.m[0][0], fmaf(r2.y, J.m[1][0], r2.z * J.m[2][0])),
                 fmaf(r2.x, J.m[0][1], fmaf(r2.y, J.m[1][1], r2.z * J.m[2][1])),
                 fmaf(r2.x, J.m[0][2], fmaf(r2.y, J.m[1][2], r2.z * J.m[2][2]))};
  const f3 R0 = xyz(r0), R1 = xyz(r1), R2 = xyz(r2);
  F[0] = r0;
  F[1] = r1;
  F[2] = r2;
  F[3] = make_float4(c.x, c.y, c.z, inv_m);
  F[4] = make_float4(dot3(t0, R0), dot3(t0, R1), dot3(t0, R2), dot3(t1, R1));
  F[5] = make_float4(dot3(t1, R2), dot3(t2, R2), 0.0f, 0.0f);
}

// The frame of static shapes (body -1): identity, zero origin, zero inverse mass and inertia.
__device__ __forceinline__ void frame_static(float4* F) {
  F[0] = make_float4(1.0f, 0.0f, 0.0f, 0.0f);
  F[1] = make_float4(0.0f, 1.0f, 0.0f, 0.0f);
  F[2] = make_float4(0.0f, 0.0f, 1.0f, 0.0f);
  F[3] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  F[4] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  F[5] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// Effective-mass term of a row: -err / (dt w) * relax (compute_contact_constraint_delta, kernels.py:1945-1980)
__device__ __forceinline__ float row_delta(float err, float w, float relax, float dt) {
  const float dl = w > 0.0f ? -err * rcp(dt * w) : -err;
  return dl * relax;
}

}  // namespace fastm
}  // namespace nmc
