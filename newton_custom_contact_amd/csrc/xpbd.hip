// XPBD step kernel for gfx950: one workgroup advances one world through a whole
// SolverXPBD.step (newton/_src/solvers/xpbd/solver_xpbd.py:259-764):
//   joint forces -> integrate -> iterations x { contact solve + apply, joint solve + apply } -> store
// Body state stays in LDS for all iterations; HBM is touched once to load the step's inputs and once to
// store its outputs, plus the world's contact records (L2-resident across iterations).
//
// The reference accumulates per-body deltas with float atomics (kernels.py:2292-2297), which makes the
// sum order scheduling dependent.  Here every per-body sum runs in canonical order: contributions of
// contact c (both sides) land in LDS slots, each body owns the sorted list of its (contact, side) entries,
// and one lane per body adds them in ascending contact order (joints: ascending joint order, from the
// static body->joint table).  The result is deterministic and equals a sequential sweep over contacts.
#include <hip/hip_runtime.h>

#include "nmc_internal.h"
#include "nmc_math.h"
#include "xpbd_fast.h"

namespace nmc {

struct Spatial { f3 lin, ang; };
struct BodyState { xform q; f3 v, w; };

// integrate_rigid_body (newton/_src/solvers/solver.py:50-95)
__device__ __forceinline__ BodyState integrate_body(const BodyState& s, const Spatial& f, f3 com, const m33& I,
                                                   float inv_mass, const m33& inv_I, f3 g, float ang_damp, float dt) {
  f3 x0 = s.q.p;
  qf r0 = s.q.q;
  f3 x_com = x0 + rot(r0, com);
  float nz = inv_mass != 0.0f ? 1.0f : 0.0f;
  f3 v1 = s.v + (f.lin * inv_mass + g * nz) * dt;
  f3 x1 = x_com + v1 * dt;
  f3 wb = rot_inv(r0, s.w);
  f3 tb = rot_inv(r0, f.ang) - cross(wb, mv(I, wb));
  f3 w1 = rot(r0, wb + mv(inv_I, tb) * dt);
  qf wq{w1.x, w1.y, w1.z, 0.0f};
  qf r1 = qnormalize(qadd(r0, qscale(qscale(qmul(wq, r0), 0.5f), dt)));
  w1 = w1 * (1.0f - ang_damp * dt);
  BodyState o;
  o.q = xform{x1 - rot(r1, com), r1};
  o.v = v1;
  o.w = w1;
  return o;
}

// apply_body_deltas (newton/_src/solvers/xpbd/kernels.py:809-879)
__device__ __forceinline__ BodyState apply_deltas(const BodyState& s, f3 com, const m33& I, float inv_m,
                                                 const m33& inv_I, const Spatial& delta, bool has_weights,
                                                 float inv_weight, float dt) {
  if (inv_m == 0.0f) return s;
  qf q0 = s.q.q;
  float weight = 1.0f;
  if (has_weights && inv_weight > 0.0f) weight = 1.0f / inv_weight;
  f3 dp = delta.lin * (inv_m * weight);
  f3 dq = delta.ang * weight;
  f3 wb = rot_inv(q0, s.w);
  f3 dwb = mv(inv_I, rot_inv(q0, dq));
  f3 tb = cross(dwb, mv(I, wb + dwb)) + cross(wb, mv(I, dwb));
  f3 dw1 = rot(q0, dwb - mv(mscale(dt, inv_I), tb));
  f3 dwdt = dw1 * dt;
  qf dq4{dwdt.x, dwdt.y, dwdt.z, 0.0f};
  qf q1 = qadd(q0, qmul(qf{0.5f * dq4.x, 0.5f * dq4.y, 0.5f * dq4.z, 0.5f * dq4.w}, q0));
  q1 = qnormalize(q1);
  f3 x_com = s.q.p + rot(q0, com);
  f3 p1 = x_com + dp * dt;
  p1 = p1 - rot(q1, com);
  BodyState o;
  o.q = xform{p1, q1};
  f3 v1 = s.v + dp;
  f3 w1 = s.w + dw1;
  if (len(v1) < 1e-4f) v1 = f3{0.0f, 0.0f, 0.0f};
  if (len(w1) < 1e-4f) w1 = f3{0.0f, 0.0f, 0.0f};
  o.v = v1;
  o.w = w1;
  return o;
}

// compute_contact_constraint_delta (kernels.py:1945-1980)
__device__ __forceinline__ float contact_delta(float err, qf qa, qf qb, float mia, float mib, const m33& Ia,
                                               const m33& Ib, f3 la, f3 lb, f3 aa, f3 ab, float relax, float dt) {
  float denom = 0.0f;
  denom += len_sq(la) * mia;
  denom += len_sq(lb) * mib;
  f3 ra = rot_inv(qa, aa);
  f3 rb = rot_inv(qb, ab);
  denom += dot(ra, mv(Ia, ra));
  denom += dot(rb, mv(Ib, rb));
  float dl = -err;
  if (denom > 0.0f) dl /= dt * denom;
  return dl * relax;
}

__device__ __forceinline__ float positional_correction(float err, float derr, qf qa, qf qb, float mia, float mib,
                                                       const m33& Ia, const m33& Ib, f3 la, f3 lb, f3 aa, f3 ab,
                                                       float compliance, float damping, float dt) {
  float denom = 0.0f;
  denom += len_sq(la) * mia;
  denom += len_sq(lb) * mib;
  f3 ra = rot_inv(qa, aa);
  f3 rb = rot_inv(qb, ab);
  denom += dot(ra, mv(Ia, ra));
  denom += dot(rb, mv(Ib, rb));
  float alpha = compliance;
  float gamma = compliance * damping;
  float dl = -(err + alpha * 0.0f + gamma * derr);
  if (denom + alpha > 0.0f) dl /= (dt + gamma) * denom + alpha / dt;
  return dl;
}

__device__ __forceinline__ float angular_correction(float err, float derr, qf qa, qf qb, const m33& Ia, const m33& Ib,
                                                    f3 aa, f3 ab, float compliance, float damping, float dt) {
  float denom = 0.0f;
  f3 ra = rot_inv(qa, aa);
  f3 rb = rot_inv(qb, ab);
  denom += dot(ra, mv(Ia, ra));
  denom += dot(rb, mv(Ib, rb));
  float alpha = compliance;
  float gamma = compliance * damping;
  float dl = -(err + alpha * 0.0f + gamma * derr);
  if (denom + alpha > 0.0f) dl /= (dt + gamma) * denom + alpha / dt;
  return dl;
}

// Per-world views of body data (local body index).  The reference-order kernels copy the constant body data
// into LDS (inverse mass and inertia already effective); the world-frame kernels, whose LDS holds the per-body
// frames instead, read it from the model arrays (L2-resident) and zero the kinematic bodies' on the way.
struct WorldLds {
  float* q;      // 7
  float* qd;     // 6 (slot nb: zero, the static side of a contact)
  float* com;    // 3
  float* inv_m;  // 1 (effective when `eff`)
  float* inv_I;  // 9 (effective when `eff`)
  float* I;      // 9
  int* flags;    // (slot nb: 0)
  float4* F;     // world-frame arithmetic: fastm::FRAME_F4 per body, slot nb static
  bool eff;      // inv_m / inv_I hold effective values (LDS copies)
};

__device__ __forceinline__ xform lds_q(const WorldLds& L, int b) {
  const float* a = L.q + 7 * b;
  return xform{{a[0], a[1], a[2]}, {a[3], a[4], a[5], a[6]}};
}
__device__ __forceinline__ m33 lds_m33(const float* a) {
  m33 r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) r.m[i][j] = a[3 * i + j];
  return r;
}
// Effective inverse mass and inertia of local body b (solver.py:160-175: zero for kinematic bodies).
__device__ __forceinline__ float body_inv_m(const WorldLds& L, int b) {
  return (!L.eff && (L.flags[b] & BODY_KINEMATIC)) ? 0.0f : L.inv_m[b];
}
__device__ __forceinline__ m33 body_inv_I(const WorldLds& L, int b) {
  if (!L.eff && (L.flags[b] & BODY_KINEMATIC)) {
    m33 z;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) z.m[i][j] = 0.0f;
    return z;
  }
  return lds_m33(L.inv_I + 9 * b);
}

// solve_body_contact_positions (kernels.py:2062-2299) for one contact; bodies are local indices or -1.
__device__ bool contact_solve(const WorldLds& L, int la, int lb, f3 p0, f3 p1, f3 o0, f3 o1, f3 n, float thickness,
                              float mu, float mut, float mur, float relax, float dt, Spatial& da, Spatial& db,
                              float& lam_out) {
  const m33 Z = {{{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}}};
  xform Xa = la >= 0 ? lds_q(L, la) : xident();
  xform Xb = lb >= 0 ? lds_q(L, lb) : xident();
  f3 bx_a = xpoint(Xa, p0);
  f3 bx_b = xpoint(Xb, p1);
  float d = dot(n, bx_b - bx_a) - thickness;
  if (d >= 0.0f) return false;
  f3 zero{0.0f, 0.0f, 0.0f};
  float mia = 0.0f, mib = 0.0f;
  m33 Ia = Z, Ib = Z;
  f3 com_a = zero, com_b = zero, om_a = zero, om_b = zero;
  if (la >= 0) {
    com_a = load3(L.com + 3 * la); mia = body_inv_m(L, la); Ia = body_inv_I(L, la); om_a = load3(L.qd + 6 * la + 3);
  }
  if (lb >= 0) {
    com_b = load3(L.com + 3 * lb); mib = body_inv_m(L, lb); Ib = body_inv_I(L, lb); om_b = load3(L.qd + 6 * lb + 3);
  }
  // world centres of mass (the reference transforms them again for the friction rows; same values)
  const f3 wcom_a = xpoint(Xa, com_a), wcom_b = xpoint(Xb, com_b);
  f3 r_a = bx_a - wcom_a;
  f3 r_b = bx_b - wcom_b;
  f3 ang_a = -cross(r_a, n);
  f3 ang_b = cross(r_b, n);
  float lambda_n = contact_delta(d, Xa.q, Xb.q, mia, mib, Ia, Ib, -n, n, ang_a, ang_b, relax, dt);
  lam_out = lambda_n;
  f3 lda = (-n) * lambda_n, ldb = n * lambda_n;
  f3 ada = ang_a * lambda_n, adb = ang_b * lambda_n;
  if (mu > 0.0f) {
    bx_a = bx_a + xvector(Xa, o0);
    bx_b = bx_b + xvector(Xb, o1);
    f3 delta = bx_b - bx_a;
    f3 fd = delta - dot(n, delta) * n;
    r_a = bx_a - wcom_a;
    r_b = bx_b - wcom_b;
    f3 rel_v = zero;
    if (la >= 0 && (L.flags[la] & BODY_KINEMATIC)) {
      f3 va = load3(L.qd + 6 * la) + cross(load3(L.qd + 6 * la + 3), r_a);
      rel_v = rel_v - (va - dot(n, va) * n);
    }
    if (lb >= 0 && (L.flags[lb] & BODY_KINEMATIC)) {
      f3 vb = load3(L.qd + 6 * lb) + cross(load3(L.qd + 6 * lb + 3), r_b);
      rel_v = rel_v + (vb - dot(n, vb) * n);
    }
    fd = fd + rel_v * dt;
    f3 perp = normalize(fd);
    ang_a = -cross(r_a, perp);
    ang_b = cross(r_b, perp);
    float err = len(fd);
    if (err > 0.0f) {
      float lf = contact_delta(err, Xa.q, Xb.q, mia, mib, Ia, Ib, -perp, perp, ang_a, ang_b, relax, dt);
      lf = fmax2(lf, -lambda_n * mu);
      lda = lda - perp * lf;
      ldb = ldb + perp * lf;
      ada = ada + ang_a * lf;
      adb = adb + ang_b * lf;
    }
  }
  f3 dom = om_b - om_a;
  if (mut > 0.0f) {
    float err = dot(dom, n) * dt;
    if (fabsf(err) > 0.0f) {
      float lt = contact_delta(err, Xa.q, Xb.q, mia, mib, Ia, Ib, zero, zero, -n, n, relax, dt);
      lt = clampf(lt, -lambda_n * mut, lambda_n * mut);
      ada = ada - n * lt;
      adb = adb + n * lt;
    }
  }
  if (mur > 0.0f) {
    dom = dom - dot(n, dom) * n;
    float err = len(dom) * dt;
    if (err > 0.0f) {
      f3 rn = normalize(dom);
      float lr = contact_delta(err, Xa.q, Xb.q, mia, mib, Ia, Ib, zero, zero, -rn, rn, relax, dt);
      lr = fmax2(lr, -lambda_n * mur);
      ada = ada - rn * lr;
      adb = adb + rn * lr;
    }
  }
  da = Spatial{lda, ada};
  db = Spatial{ldb, adb};
  return true;
}

__device__ __forceinline__ void axis_limits_update(f3 axis, float lo_l, float up_l, f3& lo, f3& up) {
  f3 lt = axis * lo_l, ut = axis * up_l;
  f3 l = vmin(lt, ut), u = vmax(lt, ut);
  lo = vmin(lo, l);
  up = vmax(up, u);
}
__device__ __forceinline__ void weighted_target(f3 axis, float target, float weight, f3& targets, f3& weights) {
  f3 wa = axis * weight;
  targets = targets + wa * target;
  weights = weights + f3{fabsf(wa.x), fabsf(wa.y), fabsf(wa.z)};
}

struct JointAxes { f3 lo, up, tpos, kew, tvel, kdw; };

__device__ void joint_axes(const NmcModelDesc& d, int start, int count, const float* tposa, const float* tvela,
                           JointAxes& A) {
  f3 zero{0.0f, 0.0f, 0.0f};
  A = JointAxes{zero, zero, zero, zero, zero, zero};
  for (int k = 0; k < count; ++k) {
    int ai = start + k;
    f3 axis = load3(d.joint_axis + 3 * ai);
    if (k == 0) {
      f3 lt = axis * d.joint_limit_lower[ai], ut = axis * d.joint_limit_upper[ai];
      A.lo = vmin(lt, ut);
      A.up = vmax(lt, ut);
    } else {
      axis_limits_update(axis, d.joint_limit_lower[ai], d.joint_limit_upper[ai], A.lo, A.up);
    }
    float ke = d.joint_target_ke[ai], kd = d.joint_target_kd[ai];
    if (ke > 0.0f) weighted_target(axis, tposa[ai], ke, A.tpos, A.kew);
    if (kd > 0.0f) weighted_target(axis, tvela[ai], kd, A.tvel, A.kdw);
  }
  for (int i = 0; i < 3; ++i)
    if (get(A.kew, i) > 0.0f) put(A.tpos, i, get(A.tpos, i) / get(A.kew, i));
  for (int i = 0; i < 3; ++i)
    if (get(A.kdw, i) > 0.0f) put(A.tvel, i, get(A.tvel, i) / get(A.kdw, i));
}

// solve_body_joints (kernels.py:1441-1943) for one joint; parent/child given as local body indices
// (-1 = world).  Returns false when the joint writes nothing.
__device__ bool joint_solve(const NmcModelDesc& d, const WorldLds& L, int j, int lp, int lc, const float* tposa,
                            const float* tvela, float lin_compl, float ang_compl, float ang_relax, float lin_relax,
                            float dt, Spatial& dpo, Spatial& dco) {
  int type = d.joint_type[j];
  if (!d.joint_enabled[j]) return false;
  if (type == JT_FREE || type == JT_DISTANCE) return false;
  const m33 Z = {{{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}}};
  f3 zero{0.0f, 0.0f, 0.0f};
  xform X_pj = load_xform(d.joint_X_p + 7 * j);
  xform X_cj = load_xform(d.joint_X_c + 7 * j);
  xform X_wp = X_pj, pose_p = X_pj;
  float m_inv_p = 0.0f;
  m33 I_inv_p = Z;
  f3 com_p = zero, vel_p = zero, omega_p = zero;
  if (lp >= 0) {
    pose_p = lds_q(L, lp);
    X_wp = xmul(pose_p, X_wp);
    com_p = load3(L.com + 3 * lp);
    m_inv_p = body_inv_m(L, lp);
    I_inv_p = body_inv_I(L, lp);
    vel_p = load3(L.qd + 6 * lp);
    omega_p = load3(L.qd + 6 * lp + 3);
  }
  xform pose_c = lds_q(L, lc);
  xform X_wc = xmul(pose_c, X_cj);
  f3 com_c = load3(L.com + 3 * lc);
  float m_inv_c = body_inv_m(L, lc);
  m33 I_inv_c = body_inv_I(L, lc);
  f3 vel_c = load3(L.qd + 6 * lc), omega_c = load3(L.qd + 6 * lc + 3);
  if (m_inv_p == 0.0f && m_inv_c == 0.0f) return false;
  f3 ldp = zero, adp = zero, ldc = zero, adc = zero;
  xform rel_pose = xmul(xinverse(X_wp), X_wc);
  f3 rel_p = rel_pose.p;
  f3 x_c = X_wc.p;
  int axis_start = d.joint_qd_start[j];
  int lin_n = d.joint_dof_dim[2 * j], ang_n = d.joint_dof_dim[2 * j + 1];
  f3 world_com_p = xpoint(pose_p, com_p);
  {
    JointAxes A;
    joint_axes(d, axis_start, lin_n, tposa, tvela, A);
    m33 frame_p = qmat(X_wp.q);
    f3 r_p = x_c - world_com_p;
    f3 r_c = x_c - xpoint(pose_c, com_c);
    for (int dim = 0; dim < 3; ++dim) {
      float e = get(rel_p, dim);
      f3 lcv{frame_p.m[0][dim], frame_p.m[1][dim], frame_p.m[2][dim]};
      f3 lpv = -lcv;
      f3 ap = -cross(r_p, lcv);
      f3 ac = cross(r_c, lcv);
      float derr = dot(lpv, vel_p) + dot(lcv, vel_c) + dot(ap, omega_p) + dot(ac, omega_c);
      float err = 0.0f, compliance = lin_compl, damping = 0.0f;
      float derr_rel = derr - get(A.tvel, dim);
      float lower = get(A.lo, dim), upper = get(A.up, dim);
      if (e < lower) err = e - lower;
      else if (e > upper) err = e - upper;
      else {
        float target = clampf(get(A.tpos, dim), lower, upper);
        if (get(A.kew, dim) > 0.0f) { err = e - target; compliance = 1.0f / get(A.kew, dim); damping = get(A.kdw, dim); }
        else if (get(A.kdw, dim) > 0.0f) { compliance = 1.0f / get(A.kdw, dim); damping = get(A.kdw, dim); }
      }
      if (fabsf(err) > 1e-9f || fabsf(derr_rel) > 1e-9f) {
        float dl = positional_correction(err, derr_rel, pose_p.q, pose_c.q, m_inv_p, m_inv_c, I_inv_p, I_inv_c, lpv,
                                         lcv, ap, ac, compliance, damping, dt);
        ldp = ldp + lpv * (dl * lin_relax);
        adp = adp + ap * (dl * ang_relax);
        ldc = ldc + lcv * (dl * lin_relax);
        adc = adc + ac * (dl * ang_relax);
      }
    }
  }
  if (type == JT_FIXED || type == JT_PRISMATIC || type == JT_REVOLUTE || type == JT_D6) {
    qf q_p = X_wp.q, q_c = X_wc.q;
    if (q_p.x * q_c.x + q_p.y * q_c.y + q_p.z * q_c.z + q_p.w * q_c.w < 0.0f) q_c = qscale(q_c, -1.0f);
    qf rel_q = qmul(qinv(q_p), q_c);
    qf qtwist = qnormalize(qf{rel_q.x, 0.0f, 0.0f, rel_q.w});
    qf qswing = qmul(rel_q, qinv(qtwist));
    float s = sqrtf(rel_q.x * rel_q.x + rel_q.w * rel_q.w);
    float invs = 1.0f / s;
    float invscube = invs * invs * invs;
    float errs[3];
    errs[0] = 2.0f * asin_c(clampf(qtwist.x, -1.0f, 1.0f));
    errs[1] = qswing.y;
    errs[2] = qswing.z;
    qf g[3];
    g[0] = qf{invs - rel_q.x * rel_q.x * invscube, 0.0f, 0.0f, -(rel_q.w * rel_q.x) * invscube};
    g[1] = qf{-rel_q.w * (rel_q.w * rel_q.z + rel_q.x * rel_q.y) * invscube, rel_q.w * invs, -rel_q.x * invs,
              rel_q.x * (rel_q.w * rel_q.z + rel_q.x * rel_q.y) * invscube};
    g[2] = qf{rel_q.w * (rel_q.w * rel_q.y - rel_q.x * rel_q.z) * invscube, rel_q.x * invs, rel_q.w * invs,
              rel_q.x * (rel_q.z * rel_q.x - rel_q.w * rel_q.y) * invscube};
    g[0] = qscale(g[0], 2.0f / fabsf(qtwist.w));
    float swing_sq = qswing.w * qswing.w;
    if (swing_sq + 1.0e-4f < 1.0f) {
      float dd = sqrtf(1.0f - qswing.w * qswing.w);
      float theta = 2.0f * acos_c(clampf(qswing.w, -1.0f, 1.0f));
      float scale = theta / dd;
      errs[1] *= scale;
      errs[2] *= scale;
      g[1] = qscale(g[1], scale);
      g[2] = qscale(g[2], scale);
    }
    JointAxes A;
    joint_axes(d, axis_start + lin_n, ang_n, tposa, tvela, A);
    for (int dim = 0; dim < 3; ++dim) {
      float e = errs[dim];
      qf qc = qmul(qmul(qf{0.5f * q_p.x, 0.5f * q_p.y, 0.5f * q_p.z, 0.5f * q_p.w}, g[dim]), qinv(q_c));
      f3 ac{qc.x, qc.y, qc.z};
      f3 ap = -ac;
      float derr = dot(ap, omega_p) + dot(ac, omega_c);
      float err = 0.0f, compliance = ang_compl, damping = 0.0f;
      float ac_len = len(ac);
      float derr_rel = derr - get(A.tvel, dim) * ac_len;
      float lower = get(A.lo, dim), upper = get(A.up, dim);
      if (e < lower) err = e - lower;
      else if (e > upper) err = e - upper;
      else {
        float target = clampf(get(A.tpos, dim), lower, upper);
        if (get(A.kew, dim) > 0.0f) { err = e - target; compliance = 1.0f / get(A.kew, dim); damping = get(A.kdw, dim); }
        else if (get(A.kdw, dim) > 0.0f) { damping = get(A.kdw, dim); compliance = 1.0f / get(A.kdw, dim); }
      }
      float dl = angular_correction(err, derr_rel, pose_p.q, pose_c.q, I_inv_p, I_inv_c, ap, ac, compliance, damping,
                                    dt) * ang_relax;
      adp = adp + ap * dl;
      adc = adc + ac * dl;
    }
  }
  dpo = Spatial{ldp, adp};
  dco = Spatial{ldc, adc};
  return true;
}

// apply_body_deltas with an all-zero delta and no weights (the pass that follows the joint solve for
// bodies whose only joint is FREE, solver_xpbd.py:590-630).  With delta = 0 every correction term in
// apply_deltas is a signed zero, so the update reduces to renormalising q, re-deriving the origin from
// the unchanged COM, and the velocity clamp; all non-zero results are bit-identical to apply_deltas.
__device__ __forceinline__ BodyState apply_zero_delta(const BodyState& s, f3 com, float inv_m) {
  if (inv_m == 0.0f) return s;
  qf q1 = qnormalize(s.q.q);
  f3 x_com = s.q.p + rot(s.q.q, com);
  BodyState o;
  o.q = xform{x_com - rot(q1, com), q1};
  o.v = len(s.v) < 1e-4f ? f3{0.0f, 0.0f, 0.0f} : s.v;
  o.w = len(s.w) < 1e-4f ? f3{0.0f, 0.0f, 0.0f} : s.w;
  return o;
}


struct XpbdArgs {
  NmcModelDesc d;
  NmcXpbdParams p;
  const float* q_in;
  const float* qd_in;
  const float* f_in;
  const float* joint_f;
  const float* tpos;
  const float* tvel;
  NmcContacts c;
  float dt;
  float* q_out;
  float* qd_out;
  int key_cap;  // (contact, side) entries the LDS ordering buffer holds
  int joint_lds;  // joints per world with LDS delta slots (0 when no world has a non-FREE joint)
  int lds_bodies;  // bodies per world the launch's LDS layout holds
  const int* world_list;  // the launch's worlds (null: workgroup w runs world w)
  const int* bins;        // (first world, end world) per workgroup: a run of consecutive worlds solved as one
                          // (their bodies, joints and contacts lie back to back), or null
  unsigned* keys_global;  // ordering buffer of worlds whose entries exceed key_cap (2 per contact slot)
  long long keys_global_cap;
  int* counters;
  NmcLearnIO lrn;  // learned warm start I/O (all null / 0 for the plain step)
  NmcXpbdReport rep;  // XM_REPORT: contact impulses, joint impulses, body_parent_f
};

// Floats of the contribution region: per slot of a solve chunk 12 (da, db), or 10 in the split world-frame
// kernels ((da.lin, active) as one float4, then da.ang and db.ang), at least the 6 per body the body forces take
// there before the iterations.
__host__ __device__ inline int xpbd_contrib_floats(int chunk, int max_bodies, bool split = false) {
  const int a = (split ? 10 : 12) * chunk, b = 6 * (max_bodies > 0 ? max_bodies : 1);
  return a > b ? a : b;
}

// Solver modes (template parameter of the world kernel).  XM_REPORT is the plain step that also
// accumulates the reported contact and joint impulses.
enum { XM_PLAIN = 0, XM_LOG = 1, XM_PREP = 2, XM_WARM = 3, XM_REPORT = 4 };

// Per-contact data that is constant for the whole step (solve_body_contact_positions inputs).
// The body indices and the validity flag share one register: (la + 1) | (lb + 1) << 11 | valid << 22 (local
// indices below 2047; a world holds at most 512 bodies).
struct CReg {
  f3 p0, p1, o0, o1, n;
  float thick, mu, mut, mur;  // thick = margin0 + margin1
  int ids;
  __device__ __forceinline__ int la() const { return (ids & 0x7FF) - 1; }  // local body index, -1 = static
  __device__ __forceinline__ int lb() const { return ((ids >> 11) & 0x7FF) - 1; }
  __device__ __forceinline__ bool valid() const { return ((ids >> 22) & 1) != 0; }
};

// Contact at canonical position j (the world's range in key order); its buffer position is cpos(a.c, j).  A
// contact not stored (past the capacity in the global key layout) is inert.
__device__ __forceinline__ CReg load_contact(const XpbdArgs& a, int j, int b0, int nb) {
  const NmcModelDesc& d = a.d;
  CReg r;
  const int gc = cpos(a.c, j);
  if (gc < 0) {
    r.ids = 0;
    r.p0 = r.p1 = r.o0 = r.o1 = r.n = f3{0.0f, 0.0f, 0.0f};
    r.thick = r.mu = r.mut = r.mur = 0.0f;
    return r;
  }
  int s0 = a.c.shape0[gc], s1 = a.c.shape1[gc];
  int ba = s0 >= 0 ? d.shape_body[s0] : -1;
  int bb = s1 >= 0 ? d.shape_body[s1] : -1;
  // bodies outside this world never appear in its segment
  const int la = (ba >= b0 && ba - b0 < nb) ? ba - b0 : -1;
  const int lb = (bb >= b0 && bb - b0 < nb) ? bb - b0 : -1;
  r.ids = (la + 1) | ((lb + 1) << 11) | ((s0 != s1 && ba != bb) ? (1 << 22) : 0);
  int nz = 0;
  float mu = 0.0f, mut = 0.0f, mur = 0.0f;
  if (s0 >= 0) { nz += 1; mu += d.shape_mu[s0]; mut += d.shape_mu_torsional[s0]; mur += d.shape_mu_rolling[s0]; }
  if (s1 >= 0) { nz += 1; mu += d.shape_mu[s1]; mut += d.shape_mu_torsional[s1]; mur += d.shape_mu_rolling[s1]; }
  if (nz > 0) { mu /= (float)nz; mut /= (float)nz; mur /= (float)nz; }
  r.mu = mu; r.mut = mut; r.mur = mur;
  r.p0 = load3(a.c.point0 + 3 * gc);
  r.p1 = load3(a.c.point1 + 3 * gc);
  r.o0 = load3(a.c.offset0 + 3 * gc);
  r.o1 = load3(a.c.offset1 + 3 * gc);
  r.n = load3(a.c.normal + 3 * gc);
  r.thick = a.c.margin0[gc] + a.c.margin1[gc];
  return r;
}

__device__ __forceinline__ BodyState lds_state(const WorldLds& L, int lb) {
  return BodyState{lds_q(L, lb), load3(L.qd + 6 * lb), load3(L.qd + 6 * lb + 3)};
}
__device__ __forceinline__ void lds_store_state(const WorldLds& L, int lb, const BodyState& s) {
  store_xform(L.q + 7 * lb, s.q);
  store3(L.qd + 6 * lb, s.v);
  store3(L.qd + 6 * lb + 3, s.w);
}

// ---- world-frame arithmetic (FAST kernels)

// Frame of body lb from its LDS state.
__device__ __forceinline__ void frame_from_state(const WorldLds& L, int lb) {
  using namespace fastm;
  const xform X = lds_q(L, lb);
  float4 r0, r1, r2;
  rot_rows(X.q, r0, r1, r2);
  r0.w = X.p.x;
  r1.w = X.p.y;
  r2.w = X.p.z;
  const f3 c = fpoint(r0, r1, r2, load3(L.com + 3 * lb));
  frame_store(L.F + FRAME_F4 * lb, r0, r1, r2, c, body_inv_I(L, lb), body_inv_m(L, lb));
}

// solve_body_contact_positions (kernels.py:2062-2299) for one contact in world-frame arithmetic: the normal
// row, then friction (offset contact points, tangential drift plus kinematic surface velocity), torsional
// and rolling friction, each clamped by the normal multiplier as in the reference.  Body indices la/lb of
// -1 read the static frame at slot nb.  db.lin = -da.lin (exactly, as in the reference).
__device__ __forceinline__ bool contact_solve_fast(const WorldLds& L, const CReg& r, int nb, float relax, float dt,
                                                   Spatial& da, Spatial& db, float& lam_out) {
  using namespace fastm;
  const int ia = r.la() >= 0 ? r.la() : nb, ib = r.lb() >= 0 ? r.lb() : nb;
  const float4* A = L.F + FRAME_F4 * ia;
  const float4* B = L.F + FRAME_F4 * ib;
  const float4 a0 = A[0], a1 = A[1], a2 = A[2], b0 = B[0], b1 = B[1], b2 = B[2];
  f3 bx_a = fpoint(a0, a1, a2, r.p0);
  f3 bx_b = fpoint(b0, b1, b2, r.p1);
  const f3 n = r.n;
  const float d = dot3(n, bx_b - bx_a) - r.thick;
  if (d >= 0.0f) return false;
  const float4 ca = A[3], cb = B[3], wa0 = A[4], wa1 = A[5], wb0 = B[4], wb1 = B[5];
  const f3 ca3 = xyz(ca), cb3 = xyz(cb);
  const f3 ang_a = cross3(n, bx_a - ca3);  // -(r_a x n)
  const f3 ang_b = cross3(bx_b - cb3, n);
  const float nn = dot3(n, n);
  const float wn = fmaf(nn, ca.w, nn * cb.w) + quad(wa0, wa1, ang_a) + quad(wb0, wb1, ang_b);
  const float lam_n = row_delta(d, wn, relax, dt);
  lam_out = lam_n;
  f3 lin = n * (-lam_n);
  f3 anga = ang_a * lam_n, angb = ang_b * lam_n;
  if (r.mu > 0.0f) {
    bx_a = bx_a + fvec(a0, a1, a2, r.o0);
    bx_b = bx_b + fvec(b0, b1, b2, r.o1);
    const f3 delta = bx_b - bx_a;
    f3 fd = axpy(-dot3(n, delta), n, delta);
    const f3 ra = bx_a - ca3, rb = bx_b - cb3;
    const int fa_flags = L.flags[ia], fb_flags = L.flags[ib];
    if ((fa_flags | fb_flags) & BODY_KINEMATIC) {
      f3 rel_v{0.0f, 0.0f, 0.0f};
      if (fa_flags & BODY_KINEMATIC) {
        const f3 va = load3(L.qd + 6 * ia) + cross3(load3(L.qd + 6 * ia + 3), ra);
        rel_v = rel_v - axpy(-dot3(n, va), n, va);
      }
      if (fb_flags & BODY_KINEMATIC) {
        const f3 vb = load3(L.qd + 6 * ib) + cross3(load3(L.qd + 6 * ib + 3), rb);
        rel_v = rel_v + axpy(-dot3(n, vb), n, vb);
      }
      fd = axpy(dt, rel_v, fd);
    }
    const float s = dot3(fd, fd);
    if (s > TINY_SQ) {
      const float inv = rsq(s);
      const f3 perp = fd * inv;
      const f3 fa = cross3(perp, ra), fb = cross3(rb, perp);
      const float pp = dot3(perp, perp);
      const float wf = fmaf(pp, ca.w, pp * cb.w) + quad(wa0, wa1, fa) + quad(wb0, wb1, fb);
      float lf = row_delta(s * inv, wf, relax, dt);
      lf = fmax2(lf, -lam_n * r.mu);
      lin = axpy(-lf, perp, lin);
      anga = axpy(lf, fa, anga);
      angb = axpy(lf, fb, angb);
    }
  }
  if (r.mut > 0.0f || r.mur > 0.0f) {
    const f3 dom = load3(L.qd + 6 * ib + 3) - load3(L.qd + 6 * ia + 3);
    if (r.mut > 0.0f) {
      const float err = dot3(dom, n) * dt;
      if (fabsf(err) > 0.0f) {
        const float wt = quad(wa0, wa1, n) + quad(wb0, wb1, n);
        float lt = row_delta(err, wt, relax, dt);
        lt = clampf(lt, -lam_n * r.mut, lam_n * r.mut);
        anga = axpy(-lt, n, anga);
        angb = axpy(lt, n, angb);
      }
    }
    if (r.mur > 0.0f) {
      const f3 dr = axpy(-dot3(n, dom), n, dom);
      const float s = dot3(dr, dr);
      if (s > TINY_SQ) {
        const float inv = rsq(s);
        const f3 rn = dr * inv;
        const float wr = quad(wa0, wa1, rn) + quad(wb0, wb1, rn);
        float lr = row_delta(s * inv * dt, wr, relax, dt);
        lr = fmax2(lr, -lam_n * r.mur);
        anga = axpy(-lr, rn, anga);
        angb = axpy(lr, rn, angb);
      }
    }
  }
  da = Spatial{lin, anga};
  db = Spatial{-lin, angb};
  return true;
}

// apply_body_deltas (kernels.py:809-879) through the body's frame: the body-frame angular update with the
// gyroscopic term, the quaternion step and renormalisation, the origin re-derived from the moved centre of
// mass, the small-velocity clamps; then the new frame.
__device__ __forceinline__ void apply_fast(const WorldLds& L, int lb, const Spatial& delta, bool has_weights,
                                           float inv_weight, float dt) {
  using namespace fastm;
  float4* F = L.F + FRAME_F4 * lb;
  const float4 cm = F[3];
  const float inv_m = cm.w;
  if (inv_m == 0.0f) return;
  const float4 r0 = F[0], r1 = F[1], r2 = F[2];
  const m33 J = body_inv_I(L, lb), I = lds_m33(L.I + 9 * lb);
  const float weight = (has_weights && inv_weight > 0.0f) ? rcp(inv_weight) : 1.0f;
  const f3 dp = delta.lin * (inv_m * weight);
  const f3 dq = delta.ang * weight;
  const f3 w0 = load3(L.qd + 6 * lb + 3);
  const f3 wb = fvec_t(r0, r1, r2, w0);
  const f3 dwb = fmv(J, fvec_t(r0, r1, r2, dq));
  const f3 tb = cross3(dwb, fmv(I, wb + dwb)) + cross3(wb, fmv(I, dwb));
  const f3 dw1 = fvec(r0, r1, r2, axpy(-dt, fmv(J, tb), dwb));
  // q1 = normalize(q0 + (0.5 dt dw1, 0) * q0)
  const float* qs = L.q + 7 * lb;
  const f3 qv{qs[3], qs[4], qs[5]};
  const float qw = qs[6];
  const f3 h = dw1 * (0.5f * dt);
  const f3 hv = axpy(qw, h, cross3(h, qv));
  qf q1{qv.x + hv.x, qv.y + hv.y, qv.z + hv.z, qw - dot3(h, qv)};
  const float ql = fmaf(q1.x, q1.x, fmaf(q1.y, q1.y, fmaf(q1.z, q1.z, q1.w * q1.w)));
  if (ql > 0.0f) {
    const float inv = rsq(ql);
    q1 = qf{q1.x * inv, q1.y * inv, q1.z * inv, q1.w * inv};
  } else {
    q1 = qf{0.0f, 0.0f, 0.0f, 0.0f};
  }
  const f3 c1 = axpy(dt, dp, xyz(cm));
  float4 n0, n1, n2;
  rot_rows(q1, n0, n1, n2);
  const f3 p1 = c1 - fvec(n0, n1, n2, load3(L.com + 3 * lb));
  n0.w = p1.x;
  n1.w = p1.y;
  n2.w = p1.z;
  f3 v1 = load3(L.qd + 6 * lb) + dp;
  f3 w1 = w0 + dw1;
  if (dot3(v1, v1) < 1.0e-8f) v1 = f3{0.0f, 0.0f, 0.0f};
  if (dot3(w1, w1) < 1.0e-8f) w1 = f3{0.0f, 0.0f, 0.0f};
  store_xform(L.q + 7 * lb, xform{p1, q1});
  store3(L.qd + 6 * lb, v1);
  store3(L.qd + 6 * lb + 3, w1);
  frame_store(F, n0, n1, n2, c1, J, inv_m);
}

// Exclusive scan of cnt[0..n) into out[0..n] (out[n] = total) by the whole block.
__device__ void block_exclusive_scan(const int* cnt, int* out, int n, int* tmp) {
  __shared__ int s_carry;
  if (threadIdx.x == 0) s_carry = 0;
  __syncthreads();
  for (int base = 0; base < n; base += blockDim.x) {
    int i = base + threadIdx.x;
    int v = i < n ? cnt[i] : 0;
    tmp[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < (int)blockDim.x; off <<= 1) {
      int t = threadIdx.x >= (unsigned)off ? tmp[threadIdx.x - off] : 0;
      __syncthreads();
      tmp[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < n) out[i] = tmp[threadIdx.x] - v + s_carry;
    __syncthreads();
    if (threadIdx.x == blockDim.x - 1) s_carry += tmp[threadIdx.x];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[n] = s_carry;
  __syncthreads();
}

// Per-contact inputs of the warm-start network at the integrated state (index order of the restatement):
// d, normal effective inverse mass w, -d / (dt w) (0 unless d < 0), previous substep's multiplier of the
// matched contact, relative normal velocity, relative tangential speed, both inverse masses, both
// bodies' contact counts, normal z, friction, both static flags, both lever arms |r x n|.
__device__ void contact_features(const WorldLds& L, const CReg& r, int cnt_a, int cnt_b, float lam_prev, float dt,
                                 float* __restrict__ out) {
  const m33 Z = {{{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}}};
  float f[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) f[k] = 0.0f;
  if (r.valid()) {
    const f3 zero{0.0f, 0.0f, 0.0f};
    const int la = r.la(), lb = r.lb();
    xform Xa = la >= 0 ? lds_q(L, la) : xident();
    xform Xb = lb >= 0 ? lds_q(L, lb) : xident();
    f3 bx_a = xpoint(Xa, r.p0);
    f3 bx_b = xpoint(Xb, r.p1);
    float dd = dot(r.n, bx_b - bx_a) - r.thick;
    f3 com_a = la >= 0 ? load3(L.com + 3 * la) : zero;
    f3 com_b = lb >= 0 ? load3(L.com + 3 * lb) : zero;
    f3 r_a = bx_a - xpoint(Xa, com_a);
    f3 r_b = bx_b - xpoint(Xb, com_b);
    f3 ang_a = -cross(r_a, r.n);
    f3 ang_b = cross(r_b, r.n);
    float mia = la >= 0 ? body_inv_m(L, la) : 0.0f;
    float mib = lb >= 0 ? body_inv_m(L, lb) : 0.0f;
    m33 Ia = la >= 0 ? body_inv_I(L, la) : Z;
    m33 Ib = lb >= 0 ? body_inv_I(L, lb) : Z;
    float w = 0.0f;
    w += len_sq(-r.n) * mia;
    w += len_sq(r.n) * mib;
    f3 ra = rot_inv(Xa.q, ang_a);
    f3 rb = rot_inv(Xb.q, ang_b);
    w += dot(ra, mv(Ia, ra));
    w += dot(rb, mv(Ib, rb));
    f3 va = la >= 0 ? load3(L.qd + 6 * la) + cross(load3(L.qd + 6 * la + 3), r_a) : zero;
    f3 vb = lb >= 0 ? load3(L.qd + 6 * lb) + cross(load3(L.qd + 6 * lb + 3), r_b) : zero;
    f3 rv = vb - va;
    float vn = dot(r.n, rv);
    f3 vt = rv - r.n * vn;
    f[0] = dd;
    f[1] = w;
    f[2] = (dd < 0.0f && w > 0.0f) ? -dd / (dt * w) : 0.0f;
    f[3] = lam_prev;
    f[4] = vn;
    f[5] = len(vt);
    f[6] = mia;
    f[7] = mib;
    f[8] = la >= 0 ? (float)cnt_a : 0.0f;
    f[9] = lb >= 0 ? (float)cnt_b : 0.0f;
    f[10] = r.n.z;
    f[11] = r.mu;
    f[12] = mia == 0.0f ? 1.0f : 0.0f;
    f[13] = mib == 0.0f ? 1.0f : 0.0f;
    f[14] = len(ang_a);
    f[15] = len(ang_b);
  }
  float4* o = reinterpret_cast<float4*>(out);
  o[0] = make_float4(f[0], f[1], f[2], f[3]);
  o[1] = make_float4(f[4], f[5], f[6], f[7]);
  o[2] = make_float4(f[8], f[9], f[10], f[11]);
  o[3] = make_float4(f[12], f[13], f[14], f[15]);
}

// Penetration of one contact at the current state: max(0, -d) (0 for invalid contacts).
__device__ __forceinline__ float contact_penetration(const WorldLds& L, const CReg& r) {
  if (!r.valid()) return 0.0f;
  xform Xa = r.la() >= 0 ? lds_q(L, r.la()) : xident();
  xform Xb = r.lb() >= 0 ? lds_q(L, r.lb()) : xident();
  float dd = dot(r.n, xpoint(Xb, r.p1) - xpoint(Xa, r.p0)) - r.thick;
  return -dd > 0.0f ? -dd : 0.0f;
}

// Warm-start deltas of one contact: its predicted normal multiplier applied along the normal.
__device__ __forceinline__ bool warm_contrib(const WorldLds& L, const CReg& r, float lam, Spatial& da, Spatial& db) {
  if (!(r.valid() && lam > 0.0f)) return false;
  const f3 zero{0.0f, 0.0f, 0.0f};
  xform Xa = r.la() >= 0 ? lds_q(L, r.la()) : xident();
  xform Xb = r.lb() >= 0 ? lds_q(L, r.lb()) : xident();
  f3 bx_a = xpoint(Xa, r.p0);
  f3 bx_b = xpoint(Xb, r.p1);
  f3 com_a = r.la() >= 0 ? load3(L.com + 3 * r.la()) : zero;
  f3 com_b = r.lb() >= 0 ? load3(L.com + 3 * r.lb()) : zero;
  f3 ang_a = -cross(bx_a - xpoint(Xa, com_a), r.n);
  f3 ang_b = cross(bx_b - xpoint(Xb, com_b), r.n);
  da = Spatial{(-r.n) * lam, ang_a * lam};
  db = Spatial{r.n * lam, ang_b * lam};
  return true;
}

// BS lanes per world; each lane keeps CPL contacts in registers for the whole step and owns BPL bodies.
// MODE: XM_PLAIN the reference step; XM_LOG the same step, also writing features and per-contact normal
// multiplier sums; XM_PREP joint forces + integration + features, integrated state out; XM_WARM from the
// integrated state: warm-start pass, warm_iterations iterations, per-world residual check and cold
// fallback (integrated state, full iteration count).
// FAST: contact rows and body updates in world-frame arithmetic (xpbd_fast.h); otherwise the reference's
// operation order, bit-identical to a sequential CPU sweep compiled without contraction.  ART: the model has
// joints other than FREE / DISTANCE (without them the joint solve is not compiled in, which frees registers).
// The plain one-wave world-frame kernel without the joint solve is held to 128 VGPRs: with its LDS below 10 KiB a
// world (32 bodies), 16 worlds are then resident per CU (4 waves per SIMD) instead of 12, so 4096 worlds run in
// one round on 256 CUs.  The warm-started one (more state per lane) is held to 168 VGPRs, 12 worlds per CU
// instead of 8.
template <int BS, int CPL, int BPL, int MODE, bool FAST, bool ART>
__global__ void __launch_bounds__(BS)
__attribute__((amdgpu_waves_per_eu((BS == 64 && FAST && !ART && MODE == XM_PLAIN) ? 4 : ((BS == 64 && FAST && !ART && MODE == XM_WARM) ? 3 : 1))))
xpbd_world_kernel(XpbdArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int s_articulated;
  __shared__ float s_red[BS / 64];
  __shared__ int s_fallback;
  const NmcModelDesc& d = a.d;
  const NmcLearnIO& lrn = a.lrn;
  // world index; w == world_count owns the global bodies.  A bin of consecutive worlds w .. we - 1 is solved
  // as one: no body, joint or contact of one of its worlds refers to another, so the Jacobi iterations over
  // the bin are those of its worlds side by side.
  const int w = a.bins ? a.bins[2 * blockIdx.x] : (a.world_list ? a.world_list[blockIdx.x] : static_cast<int>(blockIdx.x));
  const int we = a.bins ? a.bins[2 * blockIdx.x + 1] : w + 1;
  const int b0 = d.world_body_first[w];
  const int nb = d.world_body_first[we - 1] + d.world_body_count[we - 1] - b0;
  const int j0 = d.world_joint_first[w];
  const int nj = d.world_joint_first[we - 1] + d.world_joint_count[we - 1] - j0;
  const int mb = a.lds_bodies > 0 ? a.lds_bodies : 1;
  const int mj = a.joint_lds;
  // CR contacts are held in registers for the whole step; each iteration solves CH of them at a time (one
  // register set per chunk in the world-frame kernels, whose contribution slots are then a quarter of the
  // LDS they would otherwise take; all of them at once in the reference-order kernels)
  constexpr bool SPLIT_SHAPE = FAST && BPL == 1;
  constexpr int CR = BS * CPL;
  constexpr int CH = SPLIT_SHAPE ? BS : CR;
  const float dt = a.dt;
  const int tid = threadIdx.x;

  WorldLds L;
  unsigned char* p = smem;
  L.F = reinterpret_cast<float4*>(p); p += sizeof(float4) * fastm::FRAME_F4 * (mb + 1);
  L.q = reinterpret_cast<float*>(p); p += sizeof(float) * 7 * mb;
  L.qd = reinterpret_cast<float*>(p); p += sizeof(float) * 6 * (mb + 1);
  if (FAST) {
    // constant body data from the model arrays (the per-body frames carry what the contact rows need); the
    // smaller LDS footprint keeps 16 one-wave worlds resident per CU
    L.com = const_cast<float*>(d.body_com) + 3ll * b0;
    L.inv_m = const_cast<float*>(d.body_inv_mass) + b0;
    L.inv_I = const_cast<float*>(d.body_inv_inertia) + 9ll * b0;
    L.I = const_cast<float*>(d.body_inertia) + 9ll * b0;
    L.eff = false;
  } else {
    L.com = reinterpret_cast<float*>(p); p += sizeof(float) * 3 * mb;
    L.inv_m = reinterpret_cast<float*>(p); p += sizeof(float) * mb;
    L.inv_I = reinterpret_cast<float*>(p); p += sizeof(float) * 9 * mb;
    L.I = reinterpret_cast<float*>(p); p += sizeof(float) * 9 * mb;
    L.eff = true;
  }
  L.flags = reinterpret_cast<int*>(p); p += sizeof(int) * (mb + 1);
  int* seg = reinterpret_cast<int*>(p); p += sizeof(int) * (mb + 1);
  int* cur = reinterpret_cast<int*>(p); p += sizeof(int) * (mb + 1);
  int* scan_tmp = reinterpret_cast<int*>(p); p += sizeof(int) * BS;
  p = smem + ((p - smem + 15) & ~static_cast<ptrdiff_t>(15));  // 16-byte aligned contribution slots
  // contribution slots; before the iterations the body forces use their last 6 mb floats (the report mode,
  // which uses them during the iterations, has them apart)
  const int cfl = xpbd_contrib_floats(CH, mb, SPLIT_SHAPE);
  float* contrib = reinterpret_cast<float*>(p); p += sizeof(float) * cfl;
  float* fl = contrib + (cfl - 6 * mb);
  if (MODE == XM_REPORT) { fl = reinterpret_cast<float*>(p); p += sizeof(float) * 6 * mb; }
  int* active = reinterpret_cast<int*>(p);
  if (!SPLIT_SHAPE) p += sizeof(int) * CH;
  float* jcontrib = reinterpret_cast<float*>(p); p += sizeof(float) * 12 * mj;
  int* jactive = reinterpret_cast<int*>(p); p += sizeof(int) * mj;
  unsigned* keys = reinterpret_cast<unsigned*>(p);

  const bool has_c = a.p.has_contacts != 0 && we - 1 < d.segment_count;
  const int cstart = has_c ? a.c.world_start[w] : 0;
  int nc = 0;
  if (has_c)
    for (int v = w; v < we; ++v) nc += a.c.world_count[v];  // the canonical layout puts worlds back to back
  // World-frame kernels, one body per lane: the per-body (contact, side) lists come from two bit masks per
  // body over the world's contacts (the body is shape 0's / shape 1's), built by LDS atomic OR (order-free)
  // in the contribution slots (free until the iterations), then written out in ascending order by the body's
  // lane: no atomic scatter, no sort.  Worlds whose masks do not fit there use the counting sort.
  const int mask_words = (nc + 31) >> 5;  // 32-bit words per body and side
  // (the masks share the contribution slots with the body forces at their end)
  const bool masks = SPLIT_SHAPE && 2 * nb * mask_words + 6 * mb <= cfl;
  if (2 * nc > a.key_cap) {
    // dense world: order its entries in the global buffer at its contact offset
    long long room = a.keys_global_cap - 2ll * cstart;
    if (2ll * nc > room) {
      if (tid == 0) atomicAdd(&a.counters[3], 1);
      nc = room > 0 ? static_cast<int>(room / 2) : 0;
    }
    keys = a.keys_global + 2ll * cstart;
  }
  if (tid == 0) {
    s_articulated = 0;
    s_fallback = 0;
  }
  if (MODE == XM_REPORT) {
    if (a.rep.contact_impulse)
      for (int k = tid; k < 6 * nc; k += BS) {
        const int pc = cpos(a.c, cstart + k / 6);
        if (pc >= 0) a.rep.contact_impulse[6ll * pc + k % 6] = 0.0f;
      }
    if (a.rep.body_parent_f)
      for (int k = tid; k < 6 * nj; k += BS) a.rep.joint_impulse[6ll * j0 + k] = 0.0f;
  }

  // ---- load (solver.py:160-175: kinematic bodies get zero effective inverse mass/inertia)
  const float* __restrict__ q_src = MODE == XM_WARM ? lrn.q_mid : a.q_in;
  const float* __restrict__ qd_src = MODE == XM_WARM ? lrn.qd_mid : a.qd_in;
  for (int lb = tid; lb < nb; lb += BS) {
    int b = b0 + lb;
    bool kin = (d.body_flags[b] & BODY_KINEMATIC) != 0;
    L.flags[lb] = d.body_flags[b];
    if (!FAST) {
      L.inv_m[lb] = kin ? 0.0f : d.body_inv_mass[b];
      for (int k = 0; k < 9; ++k) L.inv_I[9 * lb + k] = kin ? 0.0f : d.body_inv_inertia[9 * b + k];
      for (int k = 0; k < 3; ++k) L.com[3 * lb + k] = d.body_com[3 * b + k];
    }
    for (int k = 0; k < 7; ++k) L.q[7 * lb + k] = q_src[7 * b + k];
    for (int k = 0; k < 6; ++k) L.qd[6 * lb + k] = qd_src[6 * b + k];
    cur[lb] = 0;
  }
  if (tid == 0) {
    // slot nb: the static side of a contact
    for (int k = 0; k < 6; ++k) L.qd[6 * nb + k] = 0.0f;
    L.flags[nb] = 0;
    if (FAST) fastm::frame_static(L.F + fastm::FRAME_F4 * nb);
  }
  for (int jl = tid; jl < nj; jl += BS) {
    int j = j0 + jl;
    int t = d.joint_type[j];
    if (d.joint_enabled[j] && t != JT_FREE && t != JT_DISTANCE) s_articulated = 1;
  }
  // contacts held in registers for the whole step (chunk 0)
  CReg creg[CPL];
#pragma unroll
  for (int jj = 0; jj < CPL; ++jj) {
    int c = jj * BS + tid;
    creg[jj].ids = 0;
    if (c < nc) creg[jj] = load_contact(a, cstart + c, b0, nb);
  }
  __syncthreads();

  // ---- body_f + joint forces in joint order (kernels.py:890-1004)
  if (MODE != XM_WARM) {
    for (int lb = tid; lb < nb; lb += BS) {
      int b = b0 + lb;
      Spatial f{load3(a.f_in + 6 * b), load3(a.f_in + 6 * b + 3)};
      if (d.joint_count) {
        for (int e = d.body_joint_start[b]; e < d.body_joint_start[b + 1]; ++e) {
          int ent = d.body_joint_entry[e];
          int j = ent >> 1, side = ent & 1;
          int type = d.joint_type[j];
          if (!d.joint_enabled[j] || type == JT_FIXED || type == JT_CABLE) continue;
          int id_p = d.joint_parent[j], id_c = d.joint_child[j];
          int qs = d.joint_qd_start[j];
          if (type == JT_FREE || type == JT_DISTANCE) {
            f3 ft{a.joint_f[qs], a.joint_f[qs + 1], a.joint_f[qs + 2]};
            f3 tt{a.joint_f[qs + 3], a.joint_f[qs + 4], a.joint_f[qs + 5]};
            if (side == 1) { f.lin = f.lin + ft; f.ang = f.ang + tt; }
            else if (id_p >= 0) { f.lin = f.lin - ft; f.ang = f.ang - tt; }
            continue;
          }
          xform X_wp = load_xform(d.joint_X_p + 7 * j), pose_p = X_wp;
          f3 com_p{0.0f, 0.0f, 0.0f};
          if (id_p >= 0) {
            pose_p = lds_q(L, id_p - b0);
            X_wp = xmul(pose_p, X_wp);
            com_p = load3(L.com + 3 * (id_p - b0));
          }
          f3 r_p = X_wp.p - xpoint(pose_p, com_p);
          xform pose_c = lds_q(L, id_c - b0);
          xform X_wc = xmul(pose_c, load_xform(d.joint_X_c + 7 * j));
          f3 r_c = X_wc.p - xpoint(pose_c, load3(L.com + 3 * (id_c - b0)));
          int lin_n = d.joint_dof_dim[2 * j], ang_n = d.joint_dof_dim[2 * j + 1];
          f3 ft{0.0f, 0.0f, 0.0f}, tt{0.0f, 0.0f, 0.0f};
          if (type == JT_BALL) {
            tt = f3{a.joint_f[qs], a.joint_f[qs + 1], a.joint_f[qs + 2]};
          } else {
            for (int k = 0; k < lin_n && k < 3; ++k)
              ft = ft + a.joint_f[qs + k] * xvector(X_wp, load3(d.joint_axis + 3 * (qs + k)));
            for (int k = 0; k < ang_n && k < 3; ++k)
              tt = tt + a.joint_f[qs + lin_n + k] * xvector(X_wp, load3(d.joint_axis + 3 * (qs + lin_n + k)));
          }
          if (side == 0) { f.lin = f.lin - ft; f.ang = f.ang - (tt + cross(r_p, ft)); }
          else { f.lin = f.lin + ft; f.ang = f.ang + (tt + cross(r_c, ft)); }
        }
      }
      store3(fl + 6 * lb, f.lin);
      store3(fl + 6 * lb + 3, f.ang);
    }
  }
  // per-body (contact, side) counts for the ordering lists, or the contact masks
  unsigned* const cmask = reinterpret_cast<unsigned*>(contrib);
  if (masks) {
    for (int i = tid; i < 2 * nb * mask_words; i += BS) cmask[i] = 0u;
    __syncthreads();
  }
  for (int c = tid; c < nc; c += BS) {
    CReg r;
    if (c < CR) {
      r = creg[0];
#pragma unroll
      for (int jj = 1; jj < CPL; ++jj)
        if (c >= jj * BS) r = creg[jj];
    } else {
      r = load_contact(a, cstart + c, b0, nb);
    }
    if (r.valid()) {
      if (masks) {
        const unsigned bit = 1u << (c & 31);
        if (r.la() >= 0) atomicOr(&cmask[(2 * r.la()) * mask_words + (c >> 5)], bit);
        if (r.lb() >= 0) atomicOr(&cmask[(2 * r.lb() + 1) * mask_words + (c >> 5)], bit);
      } else {
        if (r.la() >= 0) atomicAdd(&cur[r.la()], 1);
        if (r.lb() >= 0) atomicAdd(&cur[r.lb()], 1);
      }
    }
  }
  __syncthreads();

  // ---- integrate in place (solver.py:98-157)
#pragma unroll
  for (int r = 0; r < BPL; ++r) {
    int lb = tid + r * BS;
    if (lb < nb) {
      int b = b0 + lb;
      if (!FAST)
        for (int k = 0; k < 9; ++k) L.I[9 * lb + k] = d.body_inertia[9 * b + k];
      if (MODE != XM_WARM && !(L.flags[lb] & BODY_KINEMATIC)) {
        int wi = d.body_world[b];
        int gi = wi > 0 ? wi : 0;
        Spatial f{load3(fl + 6 * lb), load3(fl + 6 * lb + 3)};
        BodyState o = integrate_body(lds_state(L, lb), f, load3(L.com + 3 * lb), load_m33(d.body_inertia + 9 * b),
                                     d.body_inv_mass[b],
                                     load_m33(d.body_inv_inertia + 9 * b), load3(d.gravity + 3 * gi),
                                     a.p.angular_damping, dt);
        lds_store_state(L, lb, o);
      }
    }
  }
  // ---- ordering lists: counting sort by body into per-body lists in contact order
  if (masks) {
    // counts from the masks, offsets, then each body's list in ascending (contact, side) order (a contact
    // never names the same body on both sides)
    for (int lb = tid; lb < nb; lb += BS) {
      int n = 0;
      for (int k = 0; k < 2 * mask_words; ++k) n += __popc(cmask[2 * lb * mask_words + k]);
      cur[lb] = n;
    }
    __syncthreads();
    block_exclusive_scan(cur, seg, nb, scan_tmp);
    for (int lb = tid; lb < nb; lb += BS) {
      int o = seg[lb];
      const unsigned* ma = cmask + 2 * lb * mask_words;
      const unsigned* mb = ma + mask_words;
      for (int wd = 0; wd < mask_words; ++wd) {
        const unsigned sa = ma[wd], sb = mb[wd];
        unsigned m = sa | sb;
        while (m) {
          const int bit = __ffs(m) - 1;
          m &= m - 1u;
          keys[o++] = static_cast<unsigned>(2 * ((wd << 5) + bit)) + ((sb >> bit) & 1u);
        }
      }
    }
  } else {
  block_exclusive_scan(cur, seg, nb, scan_tmp);
  for (int lb = tid; lb < nb; lb += BS) cur[lb] = seg[lb];
  __syncthreads();
  if (BS == 64 && nc > 256) {
    // one wave, many contacts: contacts in order, 64 at a time; per body touched in the chunk, a ballot
    // gives every touching lane its rank, so each list is filled in contact order (stable, no sort
    // afterwards; the insertion sort below is quadratic in a body's list length)
    for (int cb2 = 0; cb2 < nc; cb2 += 64) {
      const int c = cb2 + tid;
      int la = -1, lbb = -1;
      if (c < nc) {
        CReg r;
        if (c < CR) {
          r = creg[0];
#pragma unroll
          for (int jj = 1; jj < CPL; ++jj)
            if (c >= jj * BS) r = creg[jj];
        } else {
          r = load_contact(a, cstart + c, b0, nb);
        }
        if (r.valid()) {
          la = r.la();
          lbb = r.lb();
        }
      }
      bool pa = la >= 0, pb = lbb >= 0;
      const unsigned long long below = (1ull << tid) - 1ull;
      for (;;) {
        const unsigned long long pend = __ballot(pa || pb);
        if (pend == 0ull) break;
        const int leader = __ffsll(static_cast<long long>(pend)) - 1;
        const int x = __shfl(pa ? la : lbb, leader, 64);
        const bool ha = pa && la == x, hb = pb && lbb == x;
        const unsigned long long hit = __ballot(ha || hb);
        const int base = cur[x];
        if (ha || hb) keys[base + __popcll(hit & below)] = static_cast<unsigned>(2 * c + (ha ? 0 : 1));
        __builtin_amdgcn_wave_barrier();
        if (tid == leader) cur[x] = base + __popcll(hit);
        __builtin_amdgcn_wave_barrier();
        pa = pa && !ha;
        pb = pb && !hb;
      }
    }
  } else {
    for (int c = tid; c < nc; c += BS) {
      CReg r;
      if (c < CR) {
        r = creg[0];
#pragma unroll
        for (int jj = 1; jj < CPL; ++jj)
          if (c >= jj * BS) r = creg[jj];
      } else {
        r = load_contact(a, cstart + c, b0, nb);
      }
      if (r.valid()) {
        if (r.la() >= 0) keys[atomicAdd(&cur[r.la()], 1)] = static_cast<unsigned>(2 * c);
        if (r.lb() >= 0) keys[atomicAdd(&cur[r.lb()], 1)] = static_cast<unsigned>(2 * c + 1);
      }
    }
    __syncthreads();
    for (int lb = tid; lb < nb; lb += BS) {
      int s0 = seg[lb], s1 = seg[lb + 1];
      for (int i = s0 + 1; i < s1; ++i) {
        unsigned k = keys[i];
        int j = i - 1;
        while (j >= s0 && keys[j] > k) { keys[j + 1] = keys[j]; --j; }
        keys[j + 1] = k;
      }
    }
  }
  }
  __syncthreads();

  // ---- warm-start inputs at the integrated state
  if (MODE == XM_LOG || MODE == XM_PREP) {
    for (int c = tid; c < nc; c += BS) {
      CReg r;
      if (c < CR) {
        r = creg[0];
#pragma unroll
        for (int jj = 1; jj < CPL; ++jj)
          if (c >= jj * BS) r = creg[jj];
      } else {
        r = load_contact(a, cstart + c, b0, nb);
      }
      const int gc = cpos(a.c, cstart + c);
      if (gc < 0) continue;
      float lam_prev = 0.0f;
      if (lrn.prev_key && a.c.sort_key) {
        // the previous substep's contact with the same key in this world (keys ascend over the world's
        // canonical range; prev_order maps it to that buffer)
        const long long key = a.c.sort_key[gc];
        int lo = lrn.prev_start[w], hi = lo + lrn.prev_count[w] - 1;
        while (lo <= hi) {
          int mid = (lo + hi) >> 1;
          const int pm = lrn.prev_order ? lrn.prev_order[mid] : mid;
          if (pm < 0) break;
          long long km = lrn.prev_key[pm];
          if (km == key) { lam_prev = lrn.prev_lambda[pm]; break; }
          if (km < key) lo = mid + 1; else hi = mid - 1;
        }
      }
      int cnt_a = r.la() >= 0 ? seg[r.la() + 1] - seg[r.la()] : 0;
      int cnt_b = r.lb() >= 0 ? seg[r.lb() + 1] - seg[r.lb()] : 0;
      contact_features(L, r, cnt_a, cnt_b, lam_prev, dt, lrn.feat + 16ll * gc);
    }
  }
  if (MODE == XM_PREP) {
    for (int lb = tid; lb < nb; lb += BS) {
      int b = b0 + lb;
      for (int k = 0; k < 7; ++k) lrn.q_mid[7 * b + k] = L.q[7 * lb + k];
      for (int k = 0; k < 6; ++k) lrn.qd_mid[6 * b + k] = L.qd[6 * lb + k];
    }
    return;
  }
  if (FAST) {
    for (int lb = tid; lb < nb; lb += BS) frame_from_state(L, lb);
    __syncthreads();
  }

  const float relax = a.p.rigid_contact_relaxation;
  const bool weighted = a.p.rigid_contact_con_weighting != 0;
  const bool articulated = s_articulated != 0;
  constexpr bool ACC = MODE == XM_LOG || MODE == XM_WARM;  // accumulate normal multipliers
  float lamacc[CPL];
#pragma unroll
  for (int jj = 0; jj < CPL; ++jj) lamacc[jj] = 0.0f;
  if (ACC)
    for (int c = CR + tid; c < nc; c += BS) {
      const int pc = cpos(a.c, cstart + c);
      if (pc >= 0) lrn.lambda_sum[pc] = 0.0f;
    }

  // World-frame kernels, one body per lane: LPB lanes walk each body's list (entries k, k + LPB, ... in lane k
  // of the body's group) and the partial sums combine by a fixed shuffle tree, so small worlds keep their
  // lanes busy.  LPB depends on the world's body count only, so every launch shape sums in the same order
  // (deterministic, not the sequential order).  Contribution slots then hold (da.lin, active) as a float4 and
  // da.ang, db.ang in a second array, db.lin = -da.lin, zeros when inactive.
  constexpr bool SPLIT = SPLIT_SHAPE;
  const int lpb = SPLIT ? (nb <= 4 ? 16 : nb <= 8 ? 8 : nb <= 16 ? 4 : nb <= 32 ? 2 : 1) : 1;
  const int sub = SPLIT ? (tid & (lpb - 1)) : 0;
  const int walk_body = SPLIT ? tid / lpb : tid;
  float4* const contrib4 = reinterpret_cast<float4*>(contrib);  // split: (da.lin, active) per slot
  float* const contrib_ang = contrib + 4 * CH;                     // split: da.ang, db.ang per slot

  // pass 0: the step (XM_WARM: warm-start pass first, warm_iterations iterations); pass 1 (XM_WARM
  // only, when the world's residual exceeds the tolerance): cold solve from the integrated state
  constexpr int PASSES = MODE == XM_WARM ? 2 : 1;
  for (int pass = 0; pass < PASSES; ++pass) {
    int iters = a.p.iterations;
    if (MODE == XM_WARM && pass == 0) iters = lrn.warm_iterations;
    if (MODE == XM_WARM && pass == 1) {
      if (!s_fallback) break;
      for (int lb = tid; lb < nb; lb += BS) {
        int b = b0 + lb;
        for (int k = 0; k < 7; ++k) L.q[7 * lb + k] = lrn.q_mid[7 * b + k];
        for (int k = 0; k < 6; ++k) L.qd[6 * lb + k] = lrn.qd_mid[6 * b + k];
        if (FAST) frame_from_state(L, lb);
      }
#pragma unroll
      for (int jj = 0; jj < CPL; ++jj) lamacc[jj] = 0.0f;
      for (int c = CR + tid; c < nc; c += BS) {
      const int pc = cpos(a.c, cstart + c);
      if (pc >= 0) lrn.lambda_sum[pc] = 0.0f;
    }
      __syncthreads();
    }
    // it == -1: the warm-start pass (XM_WARM, pass 0)
    const int it0 = (MODE == XM_WARM && pass == 0) ? -1 : 0;
    for (int it = it0; it < iters; ++it) {
      const bool warm = MODE == XM_WARM && it < 0;
      // ---------------- contacts (kernels.py:2062-2299) + apply (kernels.py:809-879)
      if (has_c) {
        Spatial acc[BPL];
        float cnt[BPL];
        int ep[BPL], ee[BPL];
#pragma unroll
        for (int r = 0; r < BPL; ++r) {
          int lb = SPLIT ? walk_body : tid + r * BS;
          acc[r] = Spatial{{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};
          cnt[r] = 0.0f;
          ep[r] = lb < nb ? seg[lb] + sub : 0;
          ee[r] = lb < nb ? seg[lb + 1] : 0;
        }
        for (int cb = 0; cb < nc; cb += CH) {
#pragma unroll
          for (int jj = 0; jj < CH / BS; ++jj) {
            const int c = cb + jj * BS + tid;
            const int slot = jj * BS + tid;
            const int kreg = cb / BS + jj;  // the register set that holds contact c (c < CR)
            int act = 0;
            if (c < nc) {
              CReg r;
              if (c < CR) {
                r = creg[0];
#pragma unroll
                for (int k = 1; k < CPL; ++k)
                  if (kreg == k) r = creg[k];
              } else {
                r = load_contact(a, cstart + c, b0, nb);
              }
              Spatial da, db;
              bool on;
              float lam = 0.0f;
              if (warm) {
                const int pc = cpos(a.c, cstart + c);
                on = pc >= 0 && warm_contrib(L, r, lrn.lambda0[pc], da, db);
              } else {
                if (FAST)
                  on = r.valid() && contact_solve_fast(L, r, nb, relax, dt, da, db, lam);
                else
                  on = r.valid() && contact_solve(L, r.la(), r.lb(), r.p0, r.p1, r.o0, r.o1, r.n, r.thick, r.mu,
                                                r.mut, r.mur, relax, dt, da, db, lam);
              }
              if (MODE == XM_REPORT && a.rep.contact_impulse) {
                // this iteration's shape-0 side delta (zero when inactive) and the contact's bodies
                float* sc = a.rep.contact_scratch + 8ll * (cstart + c);
                const bool hit = on;
                sc[0] = hit ? da.lin.x : 0.0f; sc[1] = hit ? da.lin.y : 0.0f; sc[2] = hit ? da.lin.z : 0.0f;
                sc[3] = hit ? da.ang.x : 0.0f; sc[4] = hit ? da.ang.y : 0.0f; sc[5] = hit ? da.ang.z : 0.0f;
                sc[6] = __int_as_float(hit ? r.la() : -2);
                sc[7] = __int_as_float(r.lb());
              }
              if (SPLIT) {
                // every slot written (zeros when inactive): the walk needs no separate flag
                const f3 z{0.0f, 0.0f, 0.0f};
                contrib4[slot] = on ? make_float4(da.lin.x, da.lin.y, da.lin.z, 1.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                store3(contrib_ang + 6 * slot, on ? da.ang : z);
                store3(contrib_ang + 6 * slot + 3, on ? db.ang : z);
              }
              if (on) {
                act = 1;
                if (!SPLIT) {
                  float* o = contrib + 12 * slot;
                  store3(o, da.lin); store3(o + 3, da.ang); store3(o + 6, db.lin); store3(o + 9, db.ang);
                }
                if (ACC && !warm) {
                  if (c < CR) {
#pragma unroll
                    for (int k = 0; k < CPL; ++k)
                      if (kreg == k) lamacc[k] += lam;
                  } else {
                    lrn.lambda_sum[cpos(a.c, cstart + c)] += lam;
                  }
                }
              }
            }
            if (!SPLIT) active[slot] = act;
          }
          __syncthreads();
          const unsigned cend = static_cast<unsigned>(cb + CH);
          if (SPLIT) {
            while (ep[0] < ee[0]) {
              const unsigned ent = keys[ep[0]];
              const unsigned cc = ent >> 1;
              if (cc >= cend) break;
              const int slot = static_cast<int>(cc) - cb;
              const float4 u = contrib4[slot];
              const f3 v = load3(contrib_ang + 6 * slot + 3 * (ent & 1u));
              const float sgn = (ent & 1u) ? -1.0f : 1.0f;
              acc[0].lin = fastm::axpy(sgn, f3{u.x, u.y, u.z}, acc[0].lin);
              acc[0].ang = acc[0].ang + f3{v.x, v.y, v.z};
              cnt[0] += u.w;
              ep[0] += lpb;
            }
          } else {
#pragma unroll
            for (int r = 0; r < BPL; ++r) {
              while (ep[r] < ee[r]) {
                unsigned ent = keys[ep[r]];
                unsigned cc = ent >> 1;
                if (cc >= cend) break;
                int slot = static_cast<int>(cc) - cb;
                if (active[slot]) {
                  const float* o = contrib + 12 * slot + 6 * (ent & 1u);
                  acc[r].lin = acc[r].lin + load3(o);
                  acc[r].ang = acc[r].ang + load3(o + 3);
                  cnt[r] += 1.0f;
                }
                ep[r] += 1;
              }
            }
          }
          if (cb + CH < nc) __syncthreads();
        }
        if (SPLIT) {
          for (int off = lpb >> 1; off > 0; off >>= 1) {
            acc[0].lin.x += __shfl_xor(acc[0].lin.x, off, 64);
            acc[0].lin.y += __shfl_xor(acc[0].lin.y, off, 64);
            acc[0].lin.z += __shfl_xor(acc[0].lin.z, off, 64);
            acc[0].ang.x += __shfl_xor(acc[0].ang.x, off, 64);
            acc[0].ang.y += __shfl_xor(acc[0].ang.y, off, 64);
            acc[0].ang.z += __shfl_xor(acc[0].ang.z, off, 64);
            cnt[0] += __shfl_xor(cnt[0], off, 64);
          }
        }
        if (MODE == XM_REPORT && a.rep.contact_impulse) {
          // accumulate_weighted_contact_impulse (kernels.py:2301-2361): per contact, this iteration's
          // shape-0 delta scaled by 1/N of its dynamic body (harmonic mean 2/(N_a+N_b) for two)
#pragma unroll
          for (int r = 0; r < BPL; ++r) {
            int lb = SPLIT ? walk_body : tid + r * BS;
            if (lb < nb && sub == 0) fl[lb] = cnt[r];
          }
          __syncthreads();
          for (int c = tid; c < nc; c += BS) {
            const float* sc = a.rep.contact_scratch + 8ll * (cstart + c);
            const int la = __float_as_int(sc[6]), lbb = __float_as_int(sc[7]);
            if (la == -2) continue;  // inactive this iteration
            float wgt = 1.0f;
            if (weighted) {
              const float na = la >= 0 ? fl[la] : 0.0f, nbc = lbb >= 0 ? fl[lbb] : 0.0f;
              const float ns = na + nbc;
              if (ns > 0.0f) {
                if (na == 0.0f) wgt = 1.0f / nbc;
                else if (nbc == 0.0f) wgt = 1.0f / na;
                else wgt = 2.0f / ns;
              }
            }
            float* acc_c = a.rep.contact_impulse + 6ll * cpos(a.c, cstart + c);
#pragma unroll
            for (int k = 0; k < 6; ++k) acc_c[k] = acc_c[k] + sc[k] * wgt;
          }
          __syncthreads();
        }
#pragma unroll
        for (int r = 0; r < BPL; ++r) {
          int lb = SPLIT ? walk_body : tid + r * BS;
          // the warm pass leaves bodies without a warm contact untouched (no renormalisation / clamp)
          if (lb < nb && sub == 0 && !(warm && cnt[r] == 0.0f)) {
            if (FAST) {
              apply_fast(L, lb, acc[r], weighted, cnt[r], dt);
            } else {
              BodyState o = apply_deltas(lds_state(L, lb), load3(L.com + 3 * lb), lds_m33(L.I + 9 * lb),
                                         body_inv_m(L, lb), body_inv_I(L, lb), acc[r], weighted, cnt[r], dt);
              lds_store_state(L, lb, o);
            }
          }
        }
      }
      if (warm) {
        __syncthreads();
        continue;
      }
      // ---------------- joints (kernels.py:1441-1943) + apply without weights.  Without articulation the
      // pass applies zero deltas (renormalise, re-derive the origin, clamp): the reference-order kernel runs
      // it, the world-frame one skips it (it changes nothing beyond rounding after the contact apply).
      if (d.joint_count && (ART || !FAST)) {
        if (ART && articulated) {
          __syncthreads();
          for (int jl = tid; jl < nj; jl += BS) {
            int j = j0 + jl;
            int id_p = d.joint_parent[j], id_c = d.joint_child[j];
            Spatial dp, dc;
            bool ok = joint_solve(d, L, j, id_p >= 0 ? id_p - b0 : -1, id_c - b0, a.tpos, a.tvel,
                                  a.p.joint_linear_compliance, a.p.joint_angular_compliance,
                                  a.p.joint_angular_relaxation, a.p.joint_linear_relaxation, dt, dp, dc);
            jactive[jl] = ok ? 1 : 0;
            if (ok) {
              float* o = jcontrib + 12 * jl;
              store3(o, dp.lin); store3(o + 3, dp.ang); store3(o + 6, dc.lin); store3(o + 9, dc.ang);
              if (MODE == XM_REPORT && a.rep.body_parent_f) {  // child-side impulse (kernels.py:1937-1942)
                float* ji = a.rep.joint_impulse + 6ll * j;
                ji[0] = ji[0] + dc.lin.x; ji[1] = ji[1] + dc.lin.y; ji[2] = ji[2] + dc.lin.z;
                ji[3] = ji[3] + dc.ang.x; ji[4] = ji[4] + dc.ang.y; ji[5] = ji[5] + dc.ang.z;
              }
            }
          }
          __syncthreads();
        }
#pragma unroll
        for (int r = 0; r < BPL; ++r) {
          int lb = tid + r * BS;
          if (lb >= nb) continue;
          int b = b0 + lb;
          Spatial acc{{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};
          bool any = false;
          if (ART && articulated) {
            for (int e = d.body_joint_start[b]; e < d.body_joint_start[b + 1]; ++e) {
              int ent = d.body_joint_entry[e];
              int jl = (ent >> 1) - j0;
              if (!jactive[jl]) continue;
              const float* o = jcontrib + 12 * jl + 6 * (ent & 1);
              acc.lin = acc.lin + load3(o);
              acc.ang = acc.ang + load3(o + 3);
              any = true;
            }
          }
          if (FAST) {
            if (any) apply_fast(L, lb, acc, false, 0.0f, dt);
          } else {
            BodyState o = any ? apply_deltas(lds_state(L, lb), load3(L.com + 3 * lb), lds_m33(L.I + 9 * lb),
                                             body_inv_m(L, lb), body_inv_I(L, lb), acc, false, 0.0f, dt)
                              : apply_zero_delta(lds_state(L, lb), load3(L.com + 3 * lb), body_inv_m(L, lb));
            lds_store_state(L, lb, o);
          }
        }
      }
      __syncthreads();
    }
    // ---- per-world penetration residual: max over the world's contacts of max(0, -d)
    if (MODE == XM_WARM || lrn.residual) {
      float pen = 0.0f;
      for (int c = tid; c < nc; c += BS) {
        CReg r;
        if (c < CR) {
          r = creg[0];
#pragma unroll
          for (int jj = 1; jj < CPL; ++jj)
            if (c >= jj * BS) r = creg[jj];
        } else {
          r = load_contact(a, cstart + c, b0, nb);
        }
        pen = fmax2(pen, contact_penetration(L, r));
      }
      for (int off = 32; off > 0; off >>= 1) pen = fmax2(pen, __shfl_down(pen, off, 64));
      if ((tid & 63) == 0) s_red[tid >> 6] = pen;
      __syncthreads();
      if (tid == 0) {
        float m = 0.0f;
        for (int i = 0; i < BS / 64; ++i) m = fmax2(m, s_red[i]);
        if (MODE == XM_WARM && pass == 0 && m > lrn.residual_tol) s_fallback = 1;
        if (lrn.residual) lrn.residual[w] = m;
      }
      __syncthreads();
    }
  }
  if (ACC) {
#pragma unroll
    for (int jj = 0; jj < CPL; ++jj) {
      int c = jj * BS + tid;
      const int pc = c < nc ? cpos(a.c, cstart + c) : -1;
      if (pc >= 0) {
        float l0 = 0.0f;
        if (MODE == XM_WARM && !s_fallback) {
          float v = lrn.lambda0[pc];
          l0 = v > 0.0f ? v : 0.0f;
        }
        lrn.lambda_sum[pc] = l0 + lamacc[jj];
      }
    }
    if (MODE == XM_WARM && !s_fallback) {
      for (int c = CR + tid; c < nc; c += BS) {
        const int pc = cpos(a.c, cstart + c);
        if (pc < 0) continue;
        float v = lrn.lambda0[pc];
        lrn.lambda_sum[pc] = (v > 0.0f ? v : 0.0f) + lrn.lambda_sum[pc];
      }
    }
  }
  if (MODE == XM_WARM && lrn.fallback && tid == 0) lrn.fallback[w] = s_fallback;

  // ---- reported parent-joint wrench (convert_joint_impulse_to_parent_f, kernels.py:2397-2435)
  if (MODE == XM_REPORT && a.rep.body_parent_f) {
    for (int k = tid; k < 6 * nb; k += BS) a.rep.body_parent_f[6ll * b0 + k] = 0.0f;
    __syncthreads();
    const float inv_dt = 1.0f / dt;
    for (int jl = tid; jl < nj; jl += BS) {
      const int j = j0 + jl;
      if (!d.joint_enabled[j] || d.joint_type[j] == JT_FREE) continue;
      const int id_c = d.joint_child[j];
      if (id_c < 0) continue;
      for (int k = 0; k < 6; ++k) a.rep.body_parent_f[6ll * id_c + k] = a.rep.joint_impulse[6ll * j + k] * inv_dt;
    }
  }

  // ---- store (copy_kinematic_body_state_kernel, kernels.py:18-31: kinematic bodies pass through)
  for (int lb = tid; lb < nb; lb += BS) {
    int b = b0 + lb;
    if (L.flags[lb] & BODY_KINEMATIC) {
      for (int k = 0; k < 7; ++k) a.q_out[7 * b + k] = a.q_in[7 * b + k];
      for (int k = 0; k < 6; ++k) a.qd_out[6 * b + k] = a.qd_in[6 * b + k];
    } else {
      for (int k = 0; k < 7; ++k) a.q_out[7 * b + k] = L.q[7 * lb + k];
      for (int k = 0; k < 6; ++k) a.qd_out[6 * b + k] = L.qd[6 * lb + k];
    }
  }
}

// Launch shapes: <lanes, contacts per lane, bodies per lane>.  Worlds up to 64 bodies run on one
// wave (64 lanes, 128 contacts held in registers, barriers cost one wave's issue) or on 256 lanes
// (block_lanes 256: contact-heavy worlds with few bodies, e.g. triangle meshes with hundreds of contacts a
// world); larger worlds run one body per lane on 512 lanes (up to 512 bodies; two a lane up to 1024).  A
// world that large takes most of a CU's LDS, so its workgroup should bring as many waves as it can: on the
// uneven workload (2048 worlds of 8-262 bodies) 512 lanes in place of 256 for the worlds above 64 bodies took
// the tiered step from 1.56 to 1.30 ms per substep.  On 4096 worlds x 32 bodies: 64 lanes 0.377 ms per
// substep, 128 lanes 0.499.
struct XpbdShape { int bs, cpl, bpl; };

static XpbdShape xpbd_shape(int max_bodies, int block_lanes = 0) {
  if (max_bodies <= 64 && block_lanes != 256) return {64, 2, 1};
  if (max_bodies <= 64) return {256, 1, 1};
  if (max_bodies <= 512) return {512, 1, 1};
  return {512, 1, 2};
}

int xpbd_block_size(int max_bodies) { return xpbd_shape(max_bodies).bs; }

// LDS of the largest launch shape the model may use (both shapes when it has at most 64 bodies a world).
// LDS of one world in the kernel's layout (xpbd_world_kernel, in allocation order).
size_t xpbd_lds_bytes_shape(const XpbdShape& sh, int max_bodies, int max_joints, int key_cap, bool fast,
                            bool report) {
  const size_t mb = max_bodies > 0 ? max_bodies : 1, mj = max_joints > 0 ? max_joints : 0;
  const bool split = fast && sh.bpl == 1;
  const int ch = split ? sh.bs : sh.bs * sh.cpl;
  size_t b = sizeof(float4) * 6 * (mb + 1);                           // frames (+ static slot)
  b += sizeof(float) * (7 * mb + 6 * (mb + 1));                         // q, qd (+ static)
  if (!fast) b += sizeof(float) * (3 * mb + mb + 9 * mb + 9 * mb);      // com, inv_m, inv_I, I
  b += sizeof(int) * 3 * (mb + 1) + sizeof(int) * sh.bs;              // flags, seg, cur, scan
  b = (b + 15) & ~size_t(15);
  b += sizeof(float) * (size_t)xpbd_contrib_floats(ch, static_cast<int>(mb), split);
  if (report) b += sizeof(float) * 6 * mb;
  if (!split) b += sizeof(int) * ch;                                    // active flags
  b += sizeof(float) * 12 * mj + sizeof(int) * mj;                      // joint slots
  b += sizeof(unsigned) * (size_t)key_cap;
  return (b + 15) & ~size_t(15);
}

// The largest LDS any launch of the model may take (either launch shape, either arithmetic, any mode).
size_t xpbd_lds_bytes(int max_bodies, int max_joints, int key_cap) {
  size_t m = 0;
  for (int lanes : {0, 256})
    for (int fast = 0; fast < 2; ++fast) {
      const size_t v = xpbd_lds_bytes_shape(xpbd_shape(max_bodies, lanes), max_bodies, max_joints, key_cap,
                                            fast != 0, true);
      m = v > m ? v : m;
    }
  return m;
}

template <int MODE, bool FAST, bool ART>
static void launch_shape(const XpbdShape& sh, int worlds, size_t lds, hipStream_t stream, const XpbdArgs& a) {
  if (sh.bs == 64)
    hipLaunchKernelGGL((xpbd_world_kernel<64, 2, 1, MODE, FAST, ART>), dim3(worlds), dim3(64), lds, stream, a);
  else if (sh.bs == 256)
    hipLaunchKernelGGL((xpbd_world_kernel<256, 1, 1, MODE, FAST, ART>), dim3(worlds), dim3(256), lds, stream, a);
  else if (sh.bpl == 1)
    hipLaunchKernelGGL((xpbd_world_kernel<512, 1, 1, MODE, FAST, ART>), dim3(worlds), dim3(512), lds, stream, a);
  else
    hipLaunchKernelGGL((xpbd_world_kernel<512, 1, 2, MODE, FAST, ART>), dim3(worlds), dim3(512), lds, stream, a);
}

// reference_arithmetic picks the operation order; models whose joints are all FREE / DISTANCE (joint_lds 0)
// run the kernels compiled without the joint solve
template <int MODE>
static void launch_mode(const XpbdShape& sh, int worlds, size_t lds, hipStream_t stream, const XpbdArgs& a) {
  const bool art = a.joint_lds > 0;
  if (a.p.reference_arithmetic) {
    if (art) launch_shape<MODE, false, true>(sh, worlds, lds, stream, a);
    else launch_shape<MODE, false, false>(sh, worlds, lds, stream, a);
  } else {
    if (art) launch_shape<MODE, true, true>(sh, worlds, lds, stream, a);
    else launch_shape<MODE, true, false>(sh, worlds, lds, stream, a);
  }
}

hipError_t launch_xpbd(NmcModelImpl* m, const NmcXpbdParams* params, const float* q_in, const float* qd_in,
                       const float* f_in, const float* joint_f, const float* tpos, const float* tvel,
                       const NmcContacts* c, float dt, float* q_out, float* qd_out, const NmcLearnIO* lrn,
                       const NmcXpbdReport* rep, hipStream_t stream) {
  XpbdArgs a;
  a.d = m->d;
  a.p = *params;
  a.q_in = q_in;
  a.qd_in = qd_in;
  a.f_in = f_in;
  a.joint_f = joint_f;
  a.tpos = tpos;
  a.tvel = tvel;
  if (c) a.c = *c; else { a.c = NmcContacts{}; a.p.has_contacts = 0; }
  a.dt = dt;
  a.q_out = q_out;
  a.qd_out = qd_out;
  a.key_cap = m->xpbd_contact_cap;
  a.joint_lds = m->xpbd_joint_lds;
  a.keys_global = m->xpbd_keys;
  a.keys_global_cap = m->xpbd_keys_cap;
  a.counters = m->counters;
  if (lrn) a.lrn = *lrn; else a.lrn = NmcLearnIO{};
  if (rep) a.rep = *rep; else a.rep = NmcXpbdReport{};
  // launches: one per size tier of worlds (automatic shape: <= 64 bodies on one wave, larger on 512 lanes),
  // or all worlds in one launch of the requested shape, or (block_lanes 512, plain step) one launch over the
  // CSR bins, largest first, beside a launch of the worlds above 256 bodies when the model has any
  a.bins = nullptr;
  struct Launch { XpbdShape sh; int worlds, bodies; const int* list; const int* bins; int joints; };
  Launch ls[3];
  int nl = 0;
  const bool plain = rep == nullptr && a.lrn.mode == NMC_STEP_PLAIN;
  if (params->block_lanes == 512 && m->xpbd_bin_count > 0 && plain) {
    const int jl = m->xpbd_joint_lds > 0 ? std::max(m->xpbd_bin_max_joints, 1) : 0;
    // bins of up to 64 bodies on one wave, larger ones (at most 256 bodies) on 256 lanes, one body a lane
    const XpbdShape sh = m->xpbd_bin_max_bodies <= 64 ? xpbd_shape(m->xpbd_bin_max_bodies, 0) : XpbdShape{256, 1, 1};
    if (xpbd_lds_bytes_shape(sh, m->xpbd_bin_max_bodies, jl, m->xpbd_contact_cap, !params->reference_arithmetic,
                             false) <= 160 * 1024) {
      // the worlds above 256 bodies first (the longest), the bins beside them on the side stream
      if (m->xpbd_big_count > 0)
        ls[nl++] = Launch{xpbd_shape(m->xpbd_big_max, 0), m->xpbd_big_count, m->xpbd_big_max, m->xpbd_big, nullptr,
                          m->xpbd_joint_lds};
      ls[nl++] = Launch{sh, m->xpbd_bin_count, m->xpbd_bin_max_bodies, nullptr, m->xpbd_bins, jl};
    }
  }
  if (nl == 0 && params->block_lanes != 64 && params->block_lanes != 256 && m->xpbd_tiered) {
    for (int t = 2; t >= 0; --t)  // largest worlds first
      if (m->xpbd_tier_count[t] > 0)
        ls[nl++] = Launch{xpbd_shape(m->xpbd_tier_max[t], 0), m->xpbd_tier_count[t], m->xpbd_tier_max[t],
                          m->xpbd_tier_worlds[t], nullptr, m->xpbd_joint_lds};
  }
  if (nl == 0)
    ls[nl++] = Launch{xpbd_shape(m->d.max_world_bodies, params->block_lanes == 512 ? 0 : params->block_lanes),
                      m->d.world_count + 1, m->d.max_world_bodies, nullptr, nullptr, m->xpbd_joint_lds};
  // several tiers: launch them concurrently (tier i > 0 on a side stream forked from the caller's stream and
  // joined back into it), so that each size class's launch shares the chip with the others
  const bool fork = nl > 1;
  if (fork) {
    hipError_t e = hipEventRecord(m->tier_fork, stream);
    for (int i = 1; i < nl && e == hipSuccess; ++i) e = hipStreamWaitEvent(m->tier_stream[i - 1], m->tier_fork, 0);
    if (e != hipSuccess) return e;
  }
  for (int i = 0; i < nl; ++i) {
    const Launch& L = ls[i];
    hipStream_t s = i == 0 ? stream : m->tier_stream[i - 1];
    a.lds_bodies = L.bodies;
    a.world_list = L.list;
    a.bins = L.bins;
    a.joint_lds = L.joints;
    const size_t lds = xpbd_lds_bytes_shape(L.sh, L.bodies, L.joints, m->xpbd_contact_cap,
                                            !params->reference_arithmetic, rep != nullptr);
    if (rep) {
      launch_mode<XM_REPORT>(L.sh, L.worlds, lds, s, a);
      continue;
    }
    switch (a.lrn.mode) {
      case NMC_STEP_LOG: launch_mode<XM_LOG>(L.sh, L.worlds, lds, s, a); break;
      case NMC_STEP_PREP: launch_mode<XM_PREP>(L.sh, L.worlds, lds, s, a); break;
      case NMC_STEP_WARM: launch_mode<XM_WARM>(L.sh, L.worlds, lds, s, a); break;
      default: launch_mode<XM_PLAIN>(L.sh, L.worlds, lds, s, a); break;
    }
  }
  hipError_t e = hipGetLastError();
  if (fork) {
    for (int i = 1; i < nl && e == hipSuccess; ++i) {
      e = hipEventRecord(m->tier_join[i - 1], m->tier_stream[i - 1]);
      if (e == hipSuccess) e = hipStreamWaitEvent(stream, m->tier_join[i - 1], 0);
    }
  }
  return e;
}

}  // namespace nmc
