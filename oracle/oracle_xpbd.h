// CPU oracle -- TEST INFRASTRUCTURE ONLY.  SolverXPBD.step restatement (rigid bodies only).
//
// Follows:
//   SolverXPBD.step ordering                    newton/_src/solvers/xpbd/solver_xpbd.py:259-764
//   integrate_rigid_body / integrate_bodies     newton/_src/solvers/solver.py:50-157
//   _update_effective_inv_mass_inertia          newton/_src/solvers/solver.py:160-175
//   apply_joint_forces                          newton/_src/solvers/xpbd/kernels.py:890-1004
//   apply_body_deltas                           newton/_src/solvers/xpbd/kernels.py:809-879
//   solve_body_contact_positions                newton/_src/solvers/xpbd/kernels.py:2062-2299
//   compute_contact_constraint_delta            newton/_src/solvers/xpbd/kernels.py:1945-1980
//   solve_body_joints (+ positional/angular)    newton/_src/solvers/xpbd/kernels.py:1441-1943, 1982-2060
//   update_joint_axis_limits / weighted target  newton/_src/solvers/xpbd/kernels.py:1005-1030
//   copy_kinematic_body_state_kernel            newton/_src/solvers/xpbd/kernels.py:18-31
//   apply_rigid_restitution                     newton/_src/solvers/xpbd/kernels.py:2472-2615
//   apply_body_delta_velocities                 newton/_src/solvers/xpbd/kernels.py:882-887
// Per-body accumulations (deltas, contact counts, joint forces) are summed in the order contacts and
// joints are given; that order is the canonical one the HIP kernels reduce in.
#pragma once
#include <vector>

#include "oracle.h"
#include "oracle_math.h"

namespace oracle {

enum { KINEMATIC = 2 };
enum { J_PRISMATIC = 0, J_REVOLUTE = 1, J_BALL = 2, J_FIXED = 3, J_FREE = 4, J_DISTANCE = 5, J_D6 = 6, J_CABLE = 7 };

template <class T> struct Body {
  Xf<T> q;
  V3<T> v, w;  // qd = (v, w)
};
template <class T> struct Spatial { V3<T> lin, ang; };

template <class T> M3<T> load_m3(const float* a) {
  M3<T> r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.m[i][j] = T(a[3 * i + j]);
  return r;
}
template <class T> V3<T> load_v3(const float* a) { return V3<T>{T(a[0]), T(a[1]), T(a[2])}; }

template <class T> struct XModel {
  const OracleModel* m;
  std::vector<T> inv_m_eff;
  std::vector<M3<T>> inv_I_eff, I;
  std::vector<V3<T>> com;
};

template <class T> XModel<T> make_xmodel(const OracleModel* m) {
  XModel<T> x;
  x.m = m;
  int B = m->body_count;
  x.inv_m_eff.resize(B);
  x.inv_I_eff.resize(B);
  x.I.resize(B);
  x.com.resize(B);
  for (int b = 0; b < B; ++b) {
    bool kin = (m->body_flags[b] & KINEMATIC) != 0;
    x.inv_m_eff[b] = kin ? T(0) : T(m->body_inv_mass[b]);
    M3<T> z;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) z.m[i][j] = T(0);
    x.inv_I_eff[b] = kin ? z : load_m3<T>(m->body_inv_inertia + 9 * b);
    x.I[b] = load_m3<T>(m->body_inertia + 9 * b);
    x.com[b] = load_v3<T>(m->body_com + 3 * b);
  }
  return x;
}

// integrate_rigid_body (solver.py:50-95)
template <class T> Body<T> integrate_body(const Body<T>& s, const Spatial<T>& f, V3<T> com, const M3<T>& I, T inv_mass,
                                          const M3<T>& inv_I, V3<T> g, T ang_damp, T dt) {
  V3<T> x0 = s.q.p;
  Q<T> r0 = s.q.q;
  V3<T> w0 = s.w, v0 = s.v;
  V3<T> t0 = f.ang, f0 = f.lin;
  V3<T> x_com = x0 + qrot(r0, com);
  T nz = inv_mass != T(0) ? T(1) : T(0);
  V3<T> v1 = v0 + (f0 * inv_mass + g * nz) * dt;
  V3<T> x1 = x_com + v1 * dt;
  V3<T> wb = qrot_inv(r0, w0);
  V3<T> tb = qrot_inv(r0, t0) - cross(wb, mv(I, wb));
  V3<T> w1 = qrot(r0, wb + mv(inv_I, tb) * dt);
  Q<T> wq{w1.x, w1.y, w1.z, T(0)};
  Q<T> r1 = qnormalize(qadd(r0, qscale(qscale(qmul(wq, r0), T(0.5)), dt)));
  w1 = w1 * (T(1) - ang_damp * dt);
  Body<T> o;
  o.q = Xf<T>{x1 - qrot(r1, com), r1};
  o.v = v1;
  o.w = w1;
  return o;
}

// apply_body_deltas (kernels.py:809-879); weight <= 0 means "no weights array"
template <class T> Body<T> apply_deltas(const Body<T>& s, V3<T> com, const M3<T>& I, T inv_m, const M3<T>& inv_I,
                                        const Spatial<T>& delta, bool has_weights, T inv_weight, T dt) {
  if (inv_m == T(0)) return s;
  Xf<T> tf = s.q;
  V3<T> v0 = s.v, w0 = s.w;
  V3<T> p0 = tf.p;
  Q<T> q0 = tf.q;
  T weight = T(1);
  if (has_weights) {
    if (inv_weight > T(0)) weight = T(1) / inv_weight;
  }
  V3<T> dp = delta.lin * (inv_m * weight);
  V3<T> dq = delta.ang * weight;
  V3<T> wb = qrot_inv(q0, w0);
  V3<T> dwb = mv(inv_I, qrot_inv(q0, dq));
  V3<T> tb = cross(dwb, mv(I, wb + dwb)) + cross(wb, mv(I, dwb));
  V3<T> dw1 = qrot(q0, dwb - mv(mscale_l(dt, inv_I), tb));
  V3<T> dwdt = dw1 * dt;
  Q<T> dq4{dwdt.x, dwdt.y, dwdt.z, T(0)};
  Q<T> q1 = qadd(q0, qmul(qscale_l(T(0.5), dq4), q0));
  q1 = qnormalize(q1);
  V3<T> x_com = p0 + qrot(q0, com);
  V3<T> p1 = x_com + dp * dt;
  p1 = p1 - qrot(q1, com);
  Body<T> o;
  o.q = Xf<T>{p1, q1};
  V3<T> v1 = v0 + dp;
  V3<T> w1 = w0 + dw1;
  if (length(v1) < T(1e-4)) v1 = V3<T>{T(0), T(0), T(0)};
  if (length(w1) < T(1e-4)) w1 = V3<T>{T(0), T(0), T(0)};
  o.v = v1;
  o.w = w1;
  return o;
}

template <class T> T contact_delta(T err, const Xf<T>& tfa, const Xf<T>& tfb, T mia, T mib, const M3<T>& Ia,
                                   const M3<T>& Ib, V3<T> la, V3<T> lb, V3<T> aa, V3<T> ab, T relax, T dt) {
  T denom = T(0);
  denom += length_sq(la) * mia;
  denom += length_sq(lb) * mib;
  V3<T> ra = qrot_inv(tfa.q, aa);
  V3<T> rb = qrot_inv(tfb.q, ab);
  denom += dot(ra, mv(Ia, ra));
  denom += dot(rb, mv(Ib, rb));
  T dl = -err;
  if (denom > T(0)) dl /= dt * denom;
  return dl * relax;
}

template <class T> T positional_correction(T err, T derr, const Xf<T>& tfa, const Xf<T>& tfb, T mia, T mib,
                                           const M3<T>& Ia, const M3<T>& Ib, V3<T> la, V3<T> lb, V3<T> aa, V3<T> ab,
                                           T lambda_in, T compliance, T damping, T dt) {
  T denom = T(0);
  denom += length_sq(la) * mia;
  denom += length_sq(lb) * mib;
  V3<T> ra = qrot_inv(tfa.q, aa);
  V3<T> rb = qrot_inv(tfb.q, ab);
  denom += dot(ra, mv(Ia, ra));
  denom += dot(rb, mv(Ib, rb));
  T alpha = compliance;
  T gamma = compliance * damping;
  T dl = -(err + alpha * lambda_in + gamma * derr);
  if (denom + alpha > T(0)) dl /= (dt + gamma) * denom + alpha / dt;
  return dl;
}

template <class T> T angular_correction(T err, T derr, const Xf<T>& tfa, const Xf<T>& tfb, const M3<T>& Ia,
                                        const M3<T>& Ib, V3<T> aa, V3<T> ab, T lambda_in, T compliance, T damping,
                                        T dt) {
  T denom = T(0);
  V3<T> ra = qrot_inv(tfa.q, aa);
  V3<T> rb = qrot_inv(tfb.q, ab);
  denom += dot(ra, mv(Ia, ra));
  denom += dot(rb, mv(Ib, rb));
  T alpha = compliance;
  T gamma = compliance * damping;
  T dl = -(err + alpha * lambda_in + gamma * derr);
  if (denom + alpha > T(0)) dl /= (dt + gamma) * denom + alpha / dt;
  return dl;
}

template <class T> struct ContactIn {
  int shape0, shape1;
  V3<T> p0, p1, o0, o1, n;
  T m0, m1;
};

// solve_body_contact_positions (kernels.py:2062-2299); returns false when the contact is inactive
template <class T> bool contact_solve(const XModel<T>& X, const std::vector<Body<T>>& bodies, const ContactIn<T>& c,
                                      T relax, T dt, int& body_a, int& body_b, Spatial<T>& da, Spatial<T>& db,
                                      T* lambda_out = nullptr) {
  const OracleModel* m = X.m;
  int sa = c.shape0, sb = c.shape1;
  if (sa == sb) return false;
  body_a = sa >= 0 ? m->shape_body[sa] : -1;
  body_b = sb >= 0 ? m->shape_body[sb] : -1;
  if (body_a == body_b) return false;
  Xf<T> Xa = xf_identity<T>(), Xb = xf_identity<T>();
  if (body_a >= 0) Xa = bodies[body_a].q;
  if (body_b >= 0) Xb = bodies[body_b].q;
  V3<T> bx_a = xf_point(Xa, c.p0);
  V3<T> bx_b = xf_point(Xb, c.p1);
  T thickness = c.m0 + c.m1;
  V3<T> n = c.n;
  T d = dot(n, bx_b - bx_a) - thickness;
  if (d >= T(0)) return false;
  V3<T> zero{T(0), T(0), T(0)};
  T mia = T(0), mib = T(0);
  M3<T> Ia, Ib;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) { Ia.m[i][j] = T(0); Ib.m[i][j] = T(0); }
  V3<T> com_a = zero, com_b = zero, om_a = zero, om_b = zero;
  if (body_a >= 0) { com_a = X.com[body_a]; mia = X.inv_m_eff[body_a]; Ia = X.inv_I_eff[body_a]; om_a = bodies[body_a].w; }
  if (body_b >= 0) { com_b = X.com[body_b]; mib = X.inv_m_eff[body_b]; Ib = X.inv_I_eff[body_b]; om_b = bodies[body_b].w; }
  int nz = 0;
  T mu = T(0), mut = T(0), mur = T(0);
  if (sa >= 0) { nz += 1; mu += T(m->shape_mu[sa]); mut += T(m->shape_mu_torsional[sa]); mur += T(m->shape_mu_rolling[sa]); }
  if (sb >= 0) { nz += 1; mu += T(m->shape_mu[sb]); mut += T(m->shape_mu_torsional[sb]); mur += T(m->shape_mu_rolling[sb]); }
  if (nz > 0) { mu /= T(nz); mut /= T(nz); mur /= T(nz); }
  V3<T> r_a = bx_a - xf_point(Xa, com_a);
  V3<T> r_b = bx_b - xf_point(Xb, com_b);
  V3<T> ang_a = -cross(r_a, n);
  V3<T> ang_b = cross(r_b, n);
  T lambda_n = contact_delta(d, Xa, Xb, mia, mib, Ia, Ib, -n, n, ang_a, ang_b, relax, dt);
  if (lambda_out) *lambda_out = lambda_n;
  V3<T> lda = (-n) * lambda_n, ldb = n * lambda_n;
  V3<T> ada = ang_a * lambda_n, adb = ang_b * lambda_n;
  if (mu > T(0)) {
    bx_a = bx_a + xf_vector(Xa, c.o0);
    bx_b = bx_b + xf_vector(Xb, c.o1);
    V3<T> delta = bx_b - bx_a;
    V3<T> fd = delta - dot(n, delta) * n;
    r_a = bx_a - xf_point(Xa, com_a);
    r_b = bx_b - xf_point(Xb, com_b);
    V3<T> rel_v = zero;
    if (body_a >= 0 && (m->body_flags[body_a] & KINEMATIC)) {
      V3<T> va = bodies[body_a].v + cross(bodies[body_a].w, r_a);
      rel_v = rel_v - (va - dot(n, va) * n);
    }
    if (body_b >= 0 && (m->body_flags[body_b] & KINEMATIC)) {
      V3<T> vb = bodies[body_b].v + cross(bodies[body_b].w, r_b);
      rel_v = rel_v + (vb - dot(n, vb) * n);
    }
    fd = fd + rel_v * dt;
    V3<T> perp = normalize(fd);
    ang_a = -cross(r_a, perp);
    ang_b = cross(r_b, perp);
    T err = length(fd);
    if (err > T(0)) {
      T lf = contact_delta(err, Xa, Xb, mia, mib, Ia, Ib, -perp, perp, ang_a, ang_b, relax, dt);
      lf = tmax(lf, -lambda_n * mu);
      lda = lda - perp * lf;
      ldb = ldb + perp * lf;
      ada = ada + ang_a * lf;
      adb = adb + ang_b * lf;
    }
  }
  V3<T> dom = om_b - om_a;
  if (mut > T(0)) {
    T err = dot(dom, n) * dt;
    if (tabs(err) > T(0)) {
      T lt = contact_delta(err, Xa, Xb, mia, mib, Ia, Ib, zero, zero, -n, n, relax, dt);
      lt = tclamp(lt, -lambda_n * mut, lambda_n * mut);
      ada = ada - n * lt;
      adb = adb + n * lt;
    }
  }
  if (mur > T(0)) {
    dom = dom - dot(n, dom) * n;
    T err = length(dom) * dt;
    if (err > T(0)) {
      V3<T> rn = normalize(dom);
      T lr = contact_delta(err, Xa, Xb, mia, mib, Ia, Ib, zero, zero, -rn, rn, relax, dt);
      lr = tmax(lr, -lambda_n * mur);
      ada = ada - rn * lr;
      adb = adb + rn * lr;
    }
  }
  da = Spatial<T>{lda, ada};
  db = Spatial<T>{ldb, adb};
  return true;
}

template <class T> void axis_limits_update(V3<T> axis, T lo_l, T up_l, V3<T>& lo, V3<T>& up) {
  V3<T> lt = axis * lo_l, ut = axis * up_l;
  V3<T> l = vmin(lt, ut), u = vmax(lt, ut);
  lo = vmin(lo, l);
  up = vmax(up, u);
}
template <class T> void weighted_target(V3<T> axis, T target, T weight, V3<T>& targets, V3<T>& weights) {
  V3<T> wa = axis * weight;
  targets = targets + wa * target;
  weights = weights + V3<T>{tabs(wa.x), tabs(wa.y), tabs(wa.z)};
}

// solve_body_joints (kernels.py:1441-1943) for one joint; returns false if nothing is written
template <class T> bool joint_solve(const XModel<T>& X, const std::vector<Body<T>>& bodies, int j,
                                    const float* target_pos_arr, const float* target_vel_arr, T lin_compl,
                                    T ang_compl, T ang_relax, T lin_relax, T dt, Spatial<T>& dp_out,
                                    Spatial<T>& dc_out) {
  const OracleModel* m = X.m;
  int type = m->joint_type[j];
  if (!m->joint_enabled[j]) return false;
  if (type == J_FREE) return false;
  int id_c = m->joint_child[j], id_p = m->joint_parent[j];
  Xf<T> X_pj = load_xf<T>(m->joint_X_p + 7 * j);
  Xf<T> X_cj = load_xf<T>(m->joint_X_c + 7 * j);
  V3<T> zero{T(0), T(0), T(0)};
  Xf<T> X_wp = X_pj;
  T m_inv_p = T(0);
  M3<T> I_inv_p;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) I_inv_p.m[a][b] = T(0);
  Xf<T> pose_p = X_pj;
  V3<T> com_p = zero, vel_p = zero, omega_p = zero;
  if (id_p >= 0) {
    pose_p = bodies[id_p].q;
    X_wp = xf_mul(pose_p, X_wp);
    com_p = X.com[id_p];
    m_inv_p = X.inv_m_eff[id_p];
    I_inv_p = X.inv_I_eff[id_p];
    vel_p = bodies[id_p].v;
    omega_p = bodies[id_p].w;
  }
  Xf<T> pose_c = bodies[id_c].q;
  Xf<T> X_wc = xf_mul(pose_c, X_cj);
  V3<T> com_c = X.com[id_c];
  T m_inv_c = X.inv_m_eff[id_c];
  M3<T> I_inv_c = X.inv_I_eff[id_c];
  V3<T> vel_c = bodies[id_c].v, omega_c = bodies[id_c].w;
  if (m_inv_p == T(0) && m_inv_c == T(0)) return false;
  V3<T> ldp = zero, adp = zero, ldc = zero, adc = zero;
  Xf<T> rel_pose = xf_mul(xf_inverse(X_wp), X_wc);
  V3<T> rel_p = rel_pose.p;
  V3<T> x_c = X_wc.p;
  int axis_start = m->joint_qd_start[j];
  int lin_n = m->joint_dof_dim[2 * j], ang_n = m->joint_dof_dim[2 * j + 1];
  V3<T> world_com_p = xf_point(pose_p, com_p);
  if (type == J_DISTANCE) {
    return false;  // not part of the contact-step subset
  }
  {
    V3<T> lim_lo = zero, lim_up = zero, tpos = zero, kew = zero, tvel = zero, kdw = zero;
    for (int k = 0; k < lin_n; ++k) {
      int ai = axis_start + k;
      V3<T> axis = load_v3<T>(m->joint_axis + 3 * ai);
      if (k == 0) {
        V3<T> lt = axis * T(m->joint_limit_lower[ai]), ut = axis * T(m->joint_limit_upper[ai]);
        lim_lo = vmin(lt, ut);
        lim_up = vmax(lt, ut);
      } else {
        axis_limits_update(axis, T(m->joint_limit_lower[ai]), T(m->joint_limit_upper[ai]), lim_lo, lim_up);
      }
      T ke = T(m->joint_target_ke[ai]), kd = T(m->joint_target_kd[ai]);
      if (ke > T(0)) weighted_target(axis, T(target_pos_arr[ai]), ke, tpos, kew);
      if (kd > T(0)) weighted_target(axis, T(target_vel_arr[ai]), kd, tvel, kdw);
    }
    for (int i = 0; i < 3; ++i) {
      if (comp(kew, i) > T(0)) set_comp(tpos, i, comp(tpos, i) / comp(kew, i));
    }
    for (int i = 0; i < 3; ++i) {
      if (comp(kdw, i) > T(0)) set_comp(tvel, i, comp(tvel, i) / comp(kdw, i));
    }
    M3<T> frame_p = quat_to_matrix(X_wp.q);
    V3<T> r_p = x_c - world_com_p;
    V3<T> r_c = x_c - xf_point(pose_c, com_c);
    for (int dim = 0; dim < 3; ++dim) {
      T e = comp(rel_p, dim);
      V3<T> lc{frame_p.m[0][dim], frame_p.m[1][dim], frame_p.m[2][dim]};
      V3<T> lp = -lc;
      V3<T> ap = -cross(r_p, lc);
      V3<T> ac = cross(r_c, lc);
      T derr = dot(lp, vel_p) + dot(lc, vel_c) + dot(ap, omega_p) + dot(ac, omega_c);
      T err = T(0), compliance = lin_compl, damping = T(0);
      T target_vel = comp(tvel, dim);
      T derr_rel = derr - target_vel;
      T lower = comp(lim_lo, dim), upper = comp(lim_up, dim);
      if (e < lower) err = e - lower;
      else if (e > upper) err = e - upper;
      else {
        T target = tclamp(comp(tpos, dim), lower, upper);
        if (comp(kew, dim) > T(0)) { err = e - target; compliance = T(1) / comp(kew, dim); damping = comp(kdw, dim); }
        else if (comp(kdw, dim) > T(0)) { compliance = T(1) / comp(kdw, dim); damping = comp(kdw, dim); }
      }
      if (tabs(err) > T(1e-9) || tabs(derr_rel) > T(1e-9)) {
        T dl = positional_correction(err, derr_rel, pose_p, pose_c, m_inv_p, m_inv_c, I_inv_p, I_inv_c, lp, lc, ap, ac,
                                     T(0), compliance, damping, dt);
        ldp = ldp + lp * (dl * lin_relax);
        adp = adp + ap * (dl * ang_relax);
        ldc = ldc + lc * (dl * lin_relax);
        adc = adc + ac * (dl * ang_relax);
      }
    }
  }
  if (type == J_FIXED || type == J_PRISMATIC || type == J_REVOLUTE || type == J_D6) {
    Q<T> q_p = X_wp.q, q_c = X_wc.q;
    if (q_p.x * q_c.x + q_p.y * q_c.y + q_p.z * q_c.z + q_p.w * q_c.w < T(0)) q_c = qscale(q_c, T(-1));
    Q<T> rel_q = qmul(qinv(q_p), q_c);
    Q<T> qtwist = qnormalize(Q<T>{rel_q.x, T(0), T(0), rel_q.w});
    Q<T> qswing = qmul(rel_q, qinv(qtwist));
    T s = tsqrt(rel_q.x * rel_q.x + rel_q.w * rel_q.w);
    T invs = T(1) / s;
    T invscube = invs * invs * invs;
    T err_0 = T(2) * tasin(tclamp(qtwist.x, T(-1), T(1)));
    T err_1 = qswing.y, err_2 = qswing.z;
    Q<T> g0{invs - rel_q.x * rel_q.x * invscube, T(0), T(0), -(rel_q.w * rel_q.x) * invscube};
    Q<T> g1{-rel_q.w * (rel_q.w * rel_q.z + rel_q.x * rel_q.y) * invscube, rel_q.w * invs, -rel_q.x * invs,
            rel_q.x * (rel_q.w * rel_q.z + rel_q.x * rel_q.y) * invscube};
    Q<T> g2{rel_q.w * (rel_q.w * rel_q.y - rel_q.x * rel_q.z) * invscube, rel_q.x * invs, rel_q.w * invs,
            rel_q.x * (rel_q.z * rel_q.x - rel_q.w * rel_q.y) * invscube};
    g0 = qscale(g0, T(2) / tabs(qtwist.w));
    T swing_sq = qswing.w * qswing.w;
    const T angularEps = T(1.0e-4);
    if (swing_sq + angularEps < T(1)) {
      T d = tsqrt(T(1) - qswing.w * qswing.w);
      T theta = T(2) * tacos(tclamp(qswing.w, T(-1), T(1)));
      T scale = theta / d;
      err_1 *= scale;
      err_2 *= scale;
      g1 = qscale(g1, scale);
      g2 = qscale(g2, scale);
    }
    T errs[3] = {err_0, err_1, err_2};
    Q<T> grads[3] = {g0, g1, g2};
    V3<T> lim_lo = zero, lim_up = zero, tpos = zero, kew = zero, tvel = zero, kdw = zero;
    for (int k = 0; k < ang_n; ++k) {
      int ai = axis_start + lin_n + k;
      V3<T> axis = load_v3<T>(m->joint_axis + 3 * ai);
      if (k == 0) {
        V3<T> lt = axis * T(m->joint_limit_lower[ai]), ut = axis * T(m->joint_limit_upper[ai]);
        lim_lo = vmin(lt, ut);
        lim_up = vmax(lt, ut);
      } else {
        axis_limits_update(axis, T(m->joint_limit_lower[ai]), T(m->joint_limit_upper[ai]), lim_lo, lim_up);
      }
      T ke = T(m->joint_target_ke[ai]), kd = T(m->joint_target_kd[ai]);
      if (ke > T(0)) weighted_target(axis, T(target_pos_arr[ai]), ke, tpos, kew);
      if (kd > T(0)) weighted_target(axis, T(target_vel_arr[ai]), kd, tvel, kdw);
    }
    for (int i = 0; i < 3; ++i)
      if (comp(kew, i) > T(0)) set_comp(tpos, i, comp(tpos, i) / comp(kew, i));
    for (int i = 0; i < 3; ++i)
      if (comp(kdw, i) > T(0)) set_comp(tvel, i, comp(tvel, i) / comp(kdw, i));
    for (int dim = 0; dim < 3; ++dim) {
      T e = errs[dim];
      Q<T> grad = grads[dim];
      Q<T> quat_c = qmul(qmul(qscale_l(T(0.5), q_p), grad), qinv(q_c));
      V3<T> ac{quat_c.x, quat_c.y, quat_c.z};
      V3<T> ap = -ac;
      T derr = dot(ap, omega_p) + dot(ac, omega_c);
      T err = T(0), compliance = ang_compl, damping = T(0);
      T target_vel = comp(tvel, dim);
      T ac_len = length(ac);
      T derr_rel = derr - target_vel * ac_len;
      T lower = comp(lim_lo, dim), upper = comp(lim_up, dim);
      if (e < lower) err = e - lower;
      else if (e > upper) err = e - upper;
      else {
        T target = tclamp(comp(tpos, dim), lower, upper);
        if (comp(kew, dim) > T(0)) { err = e - target; compliance = T(1) / comp(kew, dim); damping = comp(kdw, dim); }
        else if (comp(kdw, dim) > T(0)) { damping = comp(kdw, dim); compliance = T(1) / comp(kdw, dim); }
      }
      T dl = angular_correction(err, derr_rel, pose_p, pose_c, I_inv_p, I_inv_c, ap, ac, T(0), compliance, damping,
                                dt) * ang_relax;
      adp = adp + ap * dl;
      adc = adc + ac * dl;
    }
  }
  dp_out = Spatial<T>{ldp, adp};
  dc_out = Spatial<T>{ldc, adc};
  (void)id_p;
  return true;
}

// apply_joint_forces (kernels.py:890-1004), accumulated in joint order
template <class T> void joint_forces(const XModel<T>& X, const std::vector<Body<T>>& bodies, const float* joint_f,
                                     std::vector<Spatial<T>>& body_f) {
  const OracleModel* m = X.m;
  V3<T> zero{T(0), T(0), T(0)};
  for (int j = 0; j < m->joint_count; ++j) {
    int type = m->joint_type[j];
    if (!m->joint_enabled[j]) continue;
    if (type == J_FIXED || type == J_CABLE) continue;
    int id_c = m->joint_child[j], id_p = m->joint_parent[j];
    Xf<T> X_pj = load_xf<T>(m->joint_X_p + 7 * j);
    Xf<T> X_cj = load_xf<T>(m->joint_X_c + 7 * j);
    Xf<T> X_wp = X_pj, pose_p = X_pj;
    V3<T> com_p = zero;
    if (id_p >= 0) {
      pose_p = bodies[id_p].q;
      X_wp = xf_mul(pose_p, X_wp);
      com_p = X.com[id_p];
    }
    V3<T> r_p = X_wp.p - xf_point(pose_p, com_p);
    Xf<T> pose_c = bodies[id_c].q;
    Xf<T> X_wc = xf_mul(pose_c, X_cj);
    V3<T> r_c = X_wc.p - xf_point(pose_c, X.com[id_c]);
    int qs = m->joint_qd_start[j];
    int lin_n = m->joint_dof_dim[2 * j], ang_n = m->joint_dof_dim[2 * j + 1];
    V3<T> t_total = zero, f_total = zero;
    if (type == J_FREE || type == J_DISTANCE) {
      f_total = V3<T>{T(joint_f[qs]), T(joint_f[qs + 1]), T(joint_f[qs + 2])};
      t_total = V3<T>{T(joint_f[qs + 3]), T(joint_f[qs + 4]), T(joint_f[qs + 5])};
      body_f[id_c].lin = body_f[id_c].lin + f_total;
      body_f[id_c].ang = body_f[id_c].ang + t_total;
      if (id_p >= 0) {
        body_f[id_p].lin = body_f[id_p].lin - f_total;
        body_f[id_p].ang = body_f[id_p].ang - t_total;
      }
      continue;
    } else if (type == J_BALL) {
      t_total = V3<T>{T(joint_f[qs]), T(joint_f[qs + 1]), T(joint_f[qs + 2])};
    } else if (type == J_REVOLUTE || type == J_PRISMATIC || type == J_D6) {
      for (int k = 0; k < lin_n && k < 3; ++k) {
        V3<T> a_p = xf_vector(X_wp, load_v3<T>(m->joint_axis + 3 * (qs + k)));
        f_total = f_total + T(joint_f[qs + k]) * a_p;
      }
      for (int k = 0; k < ang_n && k < 3; ++k) {
        V3<T> a_p = xf_vector(X_wp, load_v3<T>(m->joint_axis + 3 * (qs + lin_n + k)));
        t_total = t_total + T(joint_f[qs + lin_n + k]) * a_p;
      }
    }
    if (id_p >= 0) {
      body_f[id_p].lin = body_f[id_p].lin - f_total;
      body_f[id_p].ang = body_f[id_p].ang - (t_total + cross(r_p, f_total));
    }
    body_f[id_c].lin = body_f[id_c].lin + f_total;
    body_f[id_c].ang = body_f[id_c].ang + (t_total + cross(r_c, f_total));
  }
}

// integrate_bodies (solver.py:98-157) after the joint forces (kernels.py:890-1004); `in` keeps the
// start-of-step state for the kinematic pass-through
template <class T>
void integrate_all(const XModel<T>& X, const OracleModel* m, const OracleXpbdParams* p, std::vector<Body<T>>& state,
                   std::vector<Spatial<T>> body_f, const float* joint_f, T dt) {
  int B = m->body_count;
  std::vector<Body<T>> in = state;
  if (m->joint_count) joint_forces(X, in, joint_f, body_f);
  for (int b = 0; b < B; ++b) {
    if (m->body_flags[b] & KINEMATIC) continue;
    int wi = m->body_world[b];
    int gi = wi > 0 ? wi : 0;
    V3<T> g{T(m->gravity[3 * gi]), T(m->gravity[3 * gi + 1]), T(m->gravity[3 * gi + 2])};
    state[b] = integrate_body(in[b], body_f[b], X.com[b], X.I[b], T(m->body_inv_mass[b]),
                              load_m3<T>(m->body_inv_inertia + 9 * b), g, T(p->angular_damping), dt);
  }
}

// `iterations` x (contact solve + weighted apply, joint solve + apply) (solver_xpbd.py:518-650).  When
// lambda_sum is given, every contact's normal multiplier is added to it on each iteration it is active.
// contact_impulse (per contact) accumulates each iteration's shape-0 side delta scaled by the contact
// weighting (accumulate_weighted_contact_impulse, kernels.py:2301-2361); joint_impulse (per joint) the
// child side delta (kernels.py:1937-1942).
template <class T>
void iterate(const XModel<T>& X, const OracleModel* m, const OracleXpbdParams* p, std::vector<Body<T>>& state,
             const float* tpos, const float* tvel, const std::vector<ContactIn<T>>& contacts, bool has_contacts,
             T dt, int iterations, T* lambda_sum, Spatial<T>* contact_impulse = nullptr,
             Spatial<T>* joint_impulse = nullptr) {
  int B = m->body_count;
  V3<T> zero{T(0), T(0), T(0)};
  std::vector<Spatial<T>> deltas(B);
  std::vector<T> weights(B);
  std::vector<Spatial<T>> iter_imp(contact_impulse ? contacts.size() : 0);
  std::vector<char> iter_on(contact_impulse ? contacts.size() : 0);
  std::vector<int> iter_ba(contact_impulse ? contacts.size() : 0), iter_bb(contact_impulse ? contacts.size() : 0);
  for (int it = 0; it < iterations; ++it) {
    for (int b = 0; b < B; ++b) { deltas[b] = Spatial<T>{zero, zero}; weights[b] = T(0); }
    if (B && has_contacts) {
      for (size_t ci = 0; ci < contacts.size(); ++ci) {
        const ContactIn<T>& c = contacts[ci];
        int ba, bb;
        Spatial<T> da, db;
        T lam = T(0);
        if (contact_impulse) iter_on[ci] = 0;
        if (!contact_solve(X, state, c, T(p->rigid_contact_relaxation), dt, ba, bb, da, db, &lam)) continue;
        if (contact_impulse) {
          iter_on[ci] = 1;
          iter_imp[ci] = da;
          iter_ba[ci] = ba;
          iter_bb[ci] = bb;
        }
        if (lambda_sum) lambda_sum[ci] += lam;
        if (p->rigid_contact_con_weighting) {
          if (ba >= 0) weights[ba] += T(1);
          if (bb >= 0) weights[bb] += T(1);
        }
        if (ba >= 0) { deltas[ba].lin = deltas[ba].lin + da.lin; deltas[ba].ang = deltas[ba].ang + da.ang; }
        if (bb >= 0) { deltas[bb].lin = deltas[bb].lin + db.lin; deltas[bb].ang = deltas[bb].ang + db.ang; }
      }
      if (contact_impulse) {
        for (size_t ci = 0; ci < contacts.size(); ++ci) {
          if (!iter_on[ci]) continue;
          T w = T(1);
          if (p->rigid_contact_con_weighting) {
            T na = iter_ba[ci] >= 0 ? weights[iter_ba[ci]] : T(0);
            T nb = iter_bb[ci] >= 0 ? weights[iter_bb[ci]] : T(0);
            T ns = na + nb;
            if (ns > T(0)) {
              if (na == T(0)) w = T(1) / nb;
              else if (nb == T(0)) w = T(1) / na;
              else w = T(2) / ns;
            }
          }
          contact_impulse[ci].lin = contact_impulse[ci].lin + iter_imp[ci].lin * w;
          contact_impulse[ci].ang = contact_impulse[ci].ang + iter_imp[ci].ang * w;
        }
      }
      std::vector<Body<T>> next(B);
      for (int b = 0; b < B; ++b)
        next[b] = apply_deltas(state[b], X.com[b], X.I[b], X.inv_m_eff[b], X.inv_I_eff[b], deltas[b],
                               p->rigid_contact_con_weighting != 0, weights[b], dt);
      state = next;
    }
    if (m->joint_count) {
      for (int b = 0; b < B; ++b) deltas[b] = Spatial<T>{zero, zero};
      for (int j = 0; j < m->joint_count; ++j) {
        Spatial<T> dp, dc;
        if (!joint_solve(X, state, j, tpos, tvel, T(p->joint_linear_compliance), T(p->joint_angular_compliance),
                         T(p->joint_angular_relaxation), T(p->joint_linear_relaxation), dt, dp, dc))
          continue;
        int id_p = m->joint_parent[j], id_c = m->joint_child[j];
        if (joint_impulse) {
          joint_impulse[j].lin = joint_impulse[j].lin + dc.lin;
          joint_impulse[j].ang = joint_impulse[j].ang + dc.ang;
        }
        if (id_p >= 0) { deltas[id_p].lin = deltas[id_p].lin + dp.lin; deltas[id_p].ang = deltas[id_p].ang + dp.ang; }
        if (id_c >= 0) { deltas[id_c].lin = deltas[id_c].lin + dc.lin; deltas[id_c].ang = deltas[id_c].ang + dc.ang; }
      }
      std::vector<Body<T>> next(B);
      for (int b = 0; b < B; ++b)
        next[b] = apply_deltas(state[b], X.com[b], X.I[b], X.inv_m_eff[b], X.inv_I_eff[b], deltas[b], false, T(0), dt);
      state = next;
    }
  }
}

// apply_rigid_restitution (kernels.py:2472-2615) and apply_body_delta_velocities (kernels.py:882-887), run
// after the iterations when the solver is created with enable_restitution (solver_xpbd.py:719-764).  The
// contact points are taken on the start-of-step poses `prev` (point + offset, so on the surface), the
// approach velocity from the start-of-step velocities plus one step of gravity, the separation velocity from
// the solved ones in `state`; the velocity change along the normal that turns the separation into -e x the
// approach is split over both bodies by their effective inverse masses.  Per-body deltas are summed in
// contact order (the reference adds them atomically) and added to every body's velocity (+0 for bodies
// without an active contact); kinematic bodies, whose deltas are zero, are restored by the copy that follows.
template <class T>
void restitution(const XModel<T>& X, const OracleModel* m, const std::vector<Body<T>>& prev,
                 std::vector<Body<T>>& state, const std::vector<ContactIn<T>>& contacts, T dt) {
  const int B = m->body_count;
  const V3<T> zero{T(0), T(0), T(0)};
  std::vector<Spatial<T>> deltas(B, Spatial<T>{zero, zero});
  M3<T> mzero;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) mzero.m[i][j] = T(0);
  auto rest = [&](int s) { return m->shape_restitution ? T(m->shape_restitution[s]) : T(0); };
  for (const ContactIn<T>& c : contacts) {
    const int sa = c.shape0, sb = c.shape1;
    if (sa == sb) continue;
    int ba = -1, bb = -1, nz = 0;
    T e = T(0);
    if (sa >= 0) { nz += 1; e += rest(sa); ba = m->shape_body[sa]; }
    if (sb >= 0) { nz += 1; e += rest(sb); bb = m->shape_body[sb]; }
    if (nz > 0) e /= T(nz);
    if (ba == bb) continue;
    Xf<T> Xa = xf_identity<T>(), Xb = xf_identity<T>();
    V3<T> com_a = zero, com_b = zero;
    T mia = T(0), mib = T(0);
    M3<T> Ia = mzero, Ib = mzero;
    if (ba >= 0) { Xa = prev[ba].q; mia = X.inv_m_eff[ba]; Ia = X.inv_I_eff[ba]; com_a = X.com[ba]; }
    if (bb >= 0) { Xb = prev[bb].q; mib = X.inv_m_eff[bb]; Ib = X.inv_I_eff[bb]; com_b = X.com[bb]; }
    const V3<T> bx_a = xf_point(Xa, c.p0 + c.o0);
    const V3<T> bx_b = xf_point(Xb, c.p1 + c.o1);
    const T thickness = c.m0 + c.m1;
    const V3<T> n = c.n;
    const T d = dot(n, bx_b - bx_a) - thickness;
    if (d >= T(0)) continue;
    const V3<T> r_a = bx_a - xf_point(Xa, com_a);
    const V3<T> r_b = bx_b - xf_point(Xb, com_b);
    V3<T> va = zero, vb = zero, va_new = zero, vb_new = zero, rxn_a = zero, rxn_b = zero;
    T inv_mass = T(0);
    auto grav = [&](int b) {
      const int wi = m->body_world[b];
      const int gi = wi > 0 ? wi : 0;
      return V3<T>{T(m->gravity[3 * gi]), T(m->gravity[3 * gi + 1]), T(m->gravity[3 * gi + 2])};
    };
    if (ba >= 0) {
      va = (cross(prev[ba].w, r_a) + prev[ba].v) + grav(ba) * dt;
      va_new = cross(state[ba].w, r_a) + state[ba].v;
      rxn_a = qrot_inv(Xa.q, cross(r_a, n));
      inv_mass = inv_mass + (mia + dot(rxn_a, mv(Ia, rxn_a)));
    }
    if (bb >= 0) {
      vb = (cross(prev[bb].w, r_b) + prev[bb].v) + grav(bb) * dt;
      vb_new = cross(state[bb].w, r_b) + state[bb].v;
      rxn_b = qrot_inv(Xb.q, cross(r_b, n));
      inv_mass = inv_mass + (mib + dot(rxn_b, mv(Ib, rxn_b)));
    }
    if (inv_mass == T(0)) continue;
    const T rel_old = dot(n, vb - va);
    const T rel_new = dot(n, vb_new - va_new);
    if (rel_old >= T(0)) continue;
    const T dv = (-rel_new - e * rel_old) / inv_mass;
    if (ba >= 0) {
      const T dva = -dv;
      deltas[ba].lin = deltas[ba].lin + (n * mia) * dva;
      deltas[ba].ang = deltas[ba].ang + qrot(Xa.q, mv(Ia, rxn_a) * dva);
    }
    if (bb >= 0) {
      deltas[bb].lin = deltas[bb].lin + (n * mib) * dv;
      deltas[bb].ang = deltas[bb].ang + qrot(Xb.q, mv(Ib, rxn_b) * dv);
    }
  }
  for (int b = 0; b < B; ++b) {
    state[b].v = state[b].v + deltas[b].lin;
    state[b].w = state[b].w + deltas[b].ang;
  }
}

// contact_force (per contact, lin + ang) and body_parent_f (per body): update_contacts and the reported
// parent-joint wrench (solver_xpbd.py:632-655, 767-821; kernels.py:2364-2435), when given.
template <class T>
void xpbd_step(const OracleModel* m, const OracleXpbdParams* p, std::vector<Body<T>>& state,
               std::vector<Spatial<T>> body_f, const float* joint_f, const float* tpos, const float* tvel,
               const std::vector<ContactIn<T>>& contacts, bool has_contacts, T dt,
               Spatial<T>* contact_force = nullptr, Spatial<T>* body_parent_f = nullptr) {
  XModel<T> X = make_xmodel<T>(m);
  std::vector<Body<T>> in = state;
  V3<T> zero{T(0), T(0), T(0)};
  std::vector<Spatial<T>> cimp(contact_force ? contacts.size() : 0, Spatial<T>{zero, zero});
  std::vector<Spatial<T>> jimp(body_parent_f ? m->joint_count : 0, Spatial<T>{zero, zero});
  integrate_all(X, m, p, state, body_f, joint_f, dt);
  iterate(X, m, p, state, tpos, tvel, contacts, has_contacts, dt, p->iterations, static_cast<T*>(nullptr),
          contact_force ? cimp.data() : nullptr, body_parent_f && m->joint_count ? jimp.data() : nullptr);
  const T inv_dt = T(1) / dt;
  if (contact_force)
    for (size_t ci = 0; ci < contacts.size(); ++ci)
      contact_force[ci] = Spatial<T>{cimp[ci].lin * inv_dt, cimp[ci].ang * inv_dt};
  if (body_parent_f) {
    for (int b = 0; b < m->body_count; ++b) body_parent_f[b] = Spatial<T>{zero, zero};
    for (int j = 0; j < m->joint_count; ++j) {
      if (!m->joint_enabled[j] || m->joint_type[j] == 4 /* FREE */) continue;
      int id_c = m->joint_child[j];
      if (id_c < 0) continue;
      body_parent_f[id_c] = Spatial<T>{jimp[j].lin * inv_dt, jimp[j].ang * inv_dt};
    }
  }
  if (p->enable_restitution && has_contacts && m->body_count) restitution(X, m, in, state, contacts, dt);
  for (int b = 0; b < m->body_count; ++b)
    if (m->body_flags[b] & KINEMATIC) state[b] = in[b];
}

}  // namespace oracle
