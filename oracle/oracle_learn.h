// Test infrastructure: CPU restatement of the learned warm start of the XPBD contact solve and of its
// per-contact MLP.  Not part of the product; the tests check the HIP kernels against it.
//
// The reference XPBD contact solve keeps no per-contact multiplier (each iteration recomputes the
// normal multiplier from the current penetration, newton/_src/solvers/xpbd/kernels.py:1945-1980,
// 2192-2194).  The warm start adds one: before the plain iterations every contact applies a predicted
// normal multiplier lambda0 >= 0 along its normal with the reference's Jacobi weighting (1 / number of
// active contacts of the body, kernels.py:843-848), then `warm_iterations` plain iterations run.  If the
// world's penetration residual (largest max(0, -d) over its contacts) is still above the tolerance,
// the world falls back to a cold start: the integrated state and the full plain iteration count.
// With lambda0 = 0 everywhere and warm_iterations = iterations the step is the plain step, bit for bit.
#pragma once
#include <cmath>
#include <cstring>
#include <vector>

#include "oracle_xpbd.h"

namespace oracle {

constexpr int kFeatures = 16;
constexpr int kHidden = 128;

// Which world a contact belongs to (-1: between global shapes only).
template <class T> int contact_world(const OracleModel* m, const ContactIn<T>& c) {
  int ba = c.shape0 >= 0 ? m->shape_body[c.shape0] : -1;
  int bb = c.shape1 >= 0 ? m->shape_body[c.shape1] : -1;
  if (ba >= 0) return m->body_world[ba];
  if (bb >= 0) return m->body_world[bb];
  return -1;
}

template <class T> struct CGeom {
  bool valid;
  int ba, bb;
  T d;
  V3<T> r_a, r_b, ang_a, ang_b;
};

// Contact geometry at the current state (the normal part of solve_body_contact_positions).
template <class T> CGeom<T> contact_geom(const XModel<T>& X, const std::vector<Body<T>>& bodies, const ContactIn<T>& c) {
  const OracleModel* m = X.m;
  CGeom<T> g;
  g.valid = false;
  g.ba = c.shape0 >= 0 ? m->shape_body[c.shape0] : -1;
  g.bb = c.shape1 >= 0 ? m->shape_body[c.shape1] : -1;
  if (c.shape0 == c.shape1 || g.ba == g.bb) return g;
  g.valid = true;
  Xf<T> Xa = xf_identity<T>(), Xb = xf_identity<T>();
  if (g.ba >= 0) Xa = bodies[g.ba].q;
  if (g.bb >= 0) Xb = bodies[g.bb].q;
  V3<T> bx_a = xf_point(Xa, c.p0);
  V3<T> bx_b = xf_point(Xb, c.p1);
  g.d = dot(c.n, bx_b - bx_a) - (c.m0 + c.m1);
  V3<T> zero{T(0), T(0), T(0)};
  V3<T> com_a = g.ba >= 0 ? X.com[g.ba] : zero;
  V3<T> com_b = g.bb >= 0 ? X.com[g.bb] : zero;
  g.r_a = bx_a - xf_point(Xa, com_a);
  g.r_b = bx_b - xf_point(Xb, com_b);
  g.ang_a = -cross(g.r_a, c.n);
  g.ang_b = cross(g.r_b, c.n);
  return g;
}

// The 16 per-contact inputs of the warm-start MLP, at the integrated state:
//  0 d, 1 w_n (normal effective inverse mass), 2 first-iteration multiplier -d / (dt w_n) (0 if d >= 0),
//  3 previous substep's total normal multiplier of the matched contact, 4 relative normal velocity,
//  5 relative tangential speed, 6-7 inverse masses, 8-9 contact counts of the two bodies, 10 normal z,
//  11 friction mu, 12-13 static flags, 14-15 lever arms |r x n|.
template <class T>
void contact_features(const XModel<T>& X, const std::vector<Body<T>>& bodies, const std::vector<ContactIn<T>>& cs,
                      const T* lambda_prev, T dt, T* feat) {
  const OracleModel* m = X.m;
  std::vector<int> count(m->body_count, 0);
  for (const ContactIn<T>& c : cs) {
    int ba = c.shape0 >= 0 ? m->shape_body[c.shape0] : -1;
    int bb = c.shape1 >= 0 ? m->shape_body[c.shape1] : -1;
    if (c.shape0 == c.shape1 || ba == bb) continue;
    if (ba >= 0) count[ba] += 1;
    if (bb >= 0) count[bb] += 1;
  }
  V3<T> zero{T(0), T(0), T(0)};
  for (size_t i = 0; i < cs.size(); ++i) {
    const ContactIn<T>& c = cs[i];
    T* f = feat + kFeatures * i;
    for (int k = 0; k < kFeatures; ++k) f[k] = T(0);
    CGeom<T> g = contact_geom(X, bodies, c);
    if (!g.valid) continue;
    T mia = g.ba >= 0 ? X.inv_m_eff[g.ba] : T(0);
    T mib = g.bb >= 0 ? X.inv_m_eff[g.bb] : T(0);
    M3<T> Z;
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) Z.m[a][b] = T(0);
    M3<T> Ia = g.ba >= 0 ? X.inv_I_eff[g.ba] : Z;
    M3<T> Ib = g.bb >= 0 ? X.inv_I_eff[g.bb] : Z;
    Xf<T> Xa = g.ba >= 0 ? bodies[g.ba].q : xf_identity<T>();
    Xf<T> Xb = g.bb >= 0 ? bodies[g.bb].q : xf_identity<T>();
    T w = T(0);
    w += length_sq(-c.n) * mia;
    w += length_sq(c.n) * mib;
    V3<T> ra = qrot_inv(Xa.q, g.ang_a);
    V3<T> rb = qrot_inv(Xb.q, g.ang_b);
    w += dot(ra, mv(Ia, ra));
    w += dot(rb, mv(Ib, rb));
    V3<T> va = g.ba >= 0 ? bodies[g.ba].v + cross(bodies[g.ba].w, g.r_a) : zero;
    V3<T> vb = g.bb >= 0 ? bodies[g.bb].v + cross(bodies[g.bb].w, g.r_b) : zero;
    V3<T> rv = vb - va;
    T vn = dot(c.n, rv);
    V3<T> vt = rv - c.n * vn;
    int nz = 0;
    T mu = T(0);
    if (c.shape0 >= 0) { nz += 1; mu += T(m->shape_mu[c.shape0]); }
    if (c.shape1 >= 0) { nz += 1; mu += T(m->shape_mu[c.shape1]); }
    if (nz > 0) mu /= T(nz);
    f[0] = g.d;
    f[1] = w;
    f[2] = (g.d < T(0) && w > T(0)) ? -g.d / (dt * w) : T(0);
    f[3] = lambda_prev ? lambda_prev[i] : T(0);
    f[4] = vn;
    f[5] = length(vt);
    f[6] = mia;
    f[7] = mib;
    f[8] = g.ba >= 0 ? T(count[g.ba]) : T(0);
    f[9] = g.bb >= 0 ? T(count[g.bb]) : T(0);
    f[10] = c.n.z;
    f[11] = mu;
    f[12] = mia == T(0) ? T(1) : T(0);
    f[13] = mib == T(0) ? T(1) : T(0);
    f[14] = length(g.ang_a);
    f[15] = length(g.ang_b);
  }
}

// One Jacobi pass applying each contact's predicted normal multiplier (clamped at 0).
template <class T>
void warm_pass(const XModel<T>& X, const OracleModel* m, const OracleXpbdParams* p, std::vector<Body<T>>& state,
               const std::vector<ContactIn<T>>& cs, const T* lambda0, T dt) {
  int B = m->body_count;
  V3<T> zero{T(0), T(0), T(0)};
  std::vector<Spatial<T>> deltas(B, Spatial<T>{zero, zero});
  std::vector<T> weights(B, T(0));
  std::vector<char> touched(B, 0);  // bodies without a warm contact keep their state bit for bit
  for (size_t i = 0; i < cs.size(); ++i) {
    T lam = lambda0[i] > T(0) ? lambda0[i] : T(0);
    if (!(lam > T(0))) continue;
    CGeom<T> g = contact_geom(X, state, cs[i]);
    if (!g.valid) continue;
    const V3<T>& n = cs[i].n;
    if (g.ba >= 0) touched[g.ba] = 1;
    if (g.bb >= 0) touched[g.bb] = 1;
    if (p->rigid_contact_con_weighting) {
      if (g.ba >= 0) weights[g.ba] += T(1);
      if (g.bb >= 0) weights[g.bb] += T(1);
    }
    if (g.ba >= 0) { deltas[g.ba].lin = deltas[g.ba].lin + (-n) * lam; deltas[g.ba].ang = deltas[g.ba].ang + g.ang_a * lam; }
    if (g.bb >= 0) { deltas[g.bb].lin = deltas[g.bb].lin + n * lam; deltas[g.bb].ang = deltas[g.bb].ang + g.ang_b * lam; }
  }
  std::vector<Body<T>> next(B);
  for (int b = 0; b < B; ++b)
    next[b] = touched[b] ? apply_deltas(state[b], X.com[b], X.I[b], X.inv_m_eff[b], X.inv_I_eff[b], deltas[b],
                                        p->rigid_contact_con_weighting != 0, weights[b], dt)
                         : state[b];
  state = next;
}

// Per-world penetration residual: max over the world's contacts of max(0, -d).
template <class T>
void world_residual(const XModel<T>& X, const OracleModel* m, const std::vector<Body<T>>& state,
                    const std::vector<ContactIn<T>>& cs, T* res) {
  for (int w = 0; w <= m->world_count; ++w) res[w] = T(0);
  for (const ContactIn<T>& c : cs) {
    CGeom<T> g = contact_geom(X, state, c);
    if (!g.valid) continue;
    int w = g.ba >= 0 ? m->body_world[g.ba] : m->body_world[g.bb];
    int slot = w >= 0 ? w : m->world_count;
    T pen = -g.d > T(0) ? -g.d : T(0);
    if (pen > res[slot]) res[slot] = pen;
  }
}

// mode 0: plain step; 1: plain step that also logs features and per-contact normal multiplier sums;
// 2: warm start (lambda0) + warm_iterations, per-world residual check, cold fallback.
// lambda_sum (modes 1, 2): total normal multiplier applied per contact (warm value included).
template <class T>
void xpbd_step_learn(const OracleModel* m, const OracleXpbdParams* p, std::vector<Body<T>>& state,
                     std::vector<Spatial<T>> body_f, const float* joint_f, const float* tpos, const float* tvel,
                     const std::vector<ContactIn<T>>& cs, T dt, int mode, const T* lambda_prev, const T* lambda0,
                     int warm_iterations, T residual_tol, T* feat, T* lambda_sum, T* residual, int* fallback) {
  XModel<T> X = make_xmodel<T>(m);
  std::vector<Body<T>> in = state;
  integrate_all(X, m, p, state, body_f, joint_f, dt);
  if (feat) contact_features(X, state, cs, lambda_prev, dt, feat);
  size_t nc = cs.size();
  if (lambda_sum)
    for (size_t i = 0; i < nc; ++i) lambda_sum[i] = T(0);
  std::vector<T> res(m->world_count + 1, T(0));
  if (mode == 2) {
    std::vector<Body<T>> mid = state;
    warm_pass(X, m, p, state, cs, lambda0, dt);
    std::vector<T> lam(nc, T(0));
    iterate(X, m, p, state, tpos, tvel, cs, true, dt, warm_iterations, lam.data());
    world_residual(X, m, state, cs, res.data());
    std::vector<char> fb(m->world_count + 1, 0);
    bool any = false;
    for (int w = 0; w <= m->world_count; ++w)
      if (res[w] > residual_tol) { fb[w] = 1; any = true; }
    std::vector<T> lam_cold(nc, T(0));
    std::vector<T> res_cold(m->world_count + 1, T(0));
    if (any) {
      std::vector<Body<T>> cold = mid;
      iterate(X, m, p, cold, tpos, tvel, cs, true, dt, p->iterations, lam_cold.data());
      world_residual(X, m, cold, cs, res_cold.data());
      for (int b = 0; b < m->body_count; ++b) {
        int w = m->body_world[b];
        if (fb[w >= 0 ? w : m->world_count]) state[b] = cold[b];
      }
    }
    for (size_t i = 0; i < nc; ++i) {
      int w = contact_world(m, cs[i]);
      int slot = w >= 0 ? w : m->world_count;
      if (lambda_sum) {
        T l0 = lambda0[i] > T(0) ? lambda0[i] : T(0);
        lambda_sum[i] = fb[slot] ? lam_cold[i] : l0 + lam[i];
      }
    }
    for (int w = 0; w <= m->world_count; ++w) {
      if (fb[w]) res[w] = res_cold[w];
      if (fallback) fallback[w] = fb[w];
    }
  } else {
    iterate(X, m, p, state, tpos, tvel, cs, true, dt, p->iterations, lambda_sum);
    world_residual(X, m, state, cs, res.data());
    if (fallback)
      for (int w = 0; w <= m->world_count; ++w) fallback[w] = 0;
  }
  if (residual)
    for (int w = 0; w <= m->world_count; ++w) residual[w] = res[w];
  for (int b = 0; b < m->body_count; ++b)
    if (m->body_flags[b] & KINEMATIC) state[b] = in[b];
}

// ------------------------------------------------------------------ MLP 16 -> 128 -> 128 -> 1 (SiLU)
// Layers 1 and 2 take BF16 inputs and BF16 weights with FP32 accumulation (the MFMA kernels' numerics);
// biases, layer 3 and every gradient are FP32.

inline float bf16_round(float x) {  // round to nearest even, as the kernels' conversion
  unsigned u;
  std::memcpy(&u, &x, 4);
  unsigned r = ((u >> 16) & 1u) + 0x7FFFu;
  u = (u + r) & 0xFFFF0000u;
  float y;
  std::memcpy(&y, &u, 4);
  return y;
}

inline float silu(float z) { return z / (1.0f + std::exp(-z)); }
inline float silu_grad(float z) {
  float s = 1.0f / (1.0f + std::exp(-z));
  return s * (1.0f + z * (1.0f - s));
}

struct MlpParams {
  const float* w1;  // [128][16]
  const float* b1;  // [128]
  const float* w2;  // [128][128]
  const float* b2;  // [128]
  const float* w3;  // [128]
  const float* b3;  // [1]
};

// x: normalized features [n][16]; y: [n]
inline void mlp_forward(const MlpParams& P, const float* x, int n, float* y, float* z1 = nullptr, float* z2 = nullptr) {
  std::vector<float> h1(kHidden), a(kHidden);
  for (int r = 0; r < n; ++r) {
    float xb[kFeatures];
    for (int k = 0; k < kFeatures; ++k) xb[k] = bf16_round(x[kFeatures * r + k]);
    for (int o = 0; o < kHidden; ++o) {
      double s = 0.0;
      for (int k = 0; k < kFeatures; ++k) s += double(bf16_round(P.w1[kFeatures * o + k])) * xb[k];
      float z = float(s) + P.b1[o];
      if (z1) z1[kHidden * r + o] = z;
      h1[o] = bf16_round(silu(z));
    }
    double out = 0.0;
    for (int o = 0; o < kHidden; ++o) {
      double s = 0.0;
      for (int k = 0; k < kHidden; ++k) s += double(bf16_round(P.w2[kHidden * o + k])) * h1[k];
      float z = float(s) + P.b2[o];
      if (z2) z2[kHidden * r + o] = z;
      out += double(silu(z)) * P.w3[o];
    }
    y[r] = float(out) + P.b3[0];
  }
}

// Gradients of L = sum_r 0.5 * (y_r - t_r)^2 / n over the batch (FP32 restatement, FP64 sums).
inline void mlp_backward(const MlpParams& P, const float* x, const float* t, int n, float* gw1, float* gb1, float* gw2,
                         float* gb2, float* gw3, float* gb3, float* loss) {
  std::vector<float> z1(size_t(n) * kHidden), z2(size_t(n) * kHidden), y(n);
  mlp_forward(P, x, n, y.data(), z1.data(), z2.data());
  std::vector<double> dw1(kHidden * kFeatures, 0.0), db1(kHidden, 0.0), dw2(kHidden * kHidden, 0.0), db2(kHidden, 0.0),
      dw3(kHidden, 0.0);
  double db3 = 0.0, l = 0.0;
  std::vector<float> dz2(kHidden), dz1(kHidden), h1(kHidden);
  for (int r = 0; r < n; ++r) {
    float e = y[r] - t[r];
    l += 0.5 * double(e) * e / n;
    float dy = e / float(n);
    db3 += dy;
    for (int o = 0; o < kHidden; ++o) {
      float z = z2[size_t(kHidden) * r + o];
      dw3[o] += double(dy) * silu(z);
      dz2[o] = dy * P.w3[o] * silu_grad(z);
      db2[o] += dz2[o];
    }
    for (int k = 0; k < kHidden; ++k) h1[k] = bf16_round(silu(z1[size_t(kHidden) * r + k]));
    for (int o = 0; o < kHidden; ++o)
      for (int k = 0; k < kHidden; ++k) dw2[kHidden * o + k] += double(bf16_round(dz2[o])) * h1[k];
    for (int k = 0; k < kHidden; ++k) {
      double s = 0.0;
      for (int o = 0; o < kHidden; ++o) s += double(bf16_round(dz2[o])) * bf16_round(P.w2[kHidden * o + k]);
      dz1[k] = float(s) * silu_grad(z1[size_t(kHidden) * r + k]);
      db1[k] += dz1[k];
    }
    for (int o = 0; o < kHidden; ++o)
      for (int k = 0; k < kFeatures; ++k)
        dw1[kFeatures * o + k] += double(bf16_round(dz1[o])) * bf16_round(x[kFeatures * r + k]);
  }
  for (int i = 0; i < kHidden * kFeatures; ++i) gw1[i] = float(dw1[i]);
  for (int i = 0; i < kHidden * kHidden; ++i) gw2[i] = float(dw2[i]);
  for (int i = 0; i < kHidden; ++i) { gb1[i] = float(db1[i]); gb2[i] = float(db2[i]); gw3[i] = float(dw3[i]); }
  gb3[0] = float(db3);
  if (loss) *loss = float(l);
}

}  // namespace oracle
