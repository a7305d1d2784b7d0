// CPU oracle -- TEST INFRASTRUCTURE ONLY (tests/, __graft_entry__.smoke(), bench.py cpu_baseline).
// extern "C" entry points of the CPU restatement declared in oracle.h.
#include "oracle.h"

#include <vector>

#include "oracle_broad.h"
#include "oracle_collide.h"
#include "oracle_mesh.h"
#include "oracle_match.h"
#include "oracle_xpbd.h"

using namespace oracle;

template <class T, class S>
static int collide_impl(const OracleModel* m, const S* body_q, int capacity, int* shape0, int* shape1, S* point0,
                        S* point1, S* offset0, S* offset1, S* normal, S* margin0, S* margin1, int64_t* sort_key) {
  std::vector<Contact<T>> out;
  int n = collide<T, S>(m, body_q, out);
  int w = n < capacity ? n : capacity;
  for (int i = 0; i < w; ++i) {
    const Contact<T>& c = out[i];
    shape0[i] = c.shape0;
    shape1[i] = c.shape1;
    const V3<T>* vs[5] = {&c.point0, &c.point1, &c.offset0, &c.offset1, &c.normal};
    S* dst[5] = {point0, point1, offset0, offset1, normal};
    for (int k = 0; k < 5; ++k) {
      dst[k][3 * i] = S(vs[k]->x);
      dst[k][3 * i + 1] = S(vs[k]->y);
      dst[k][3 * i + 2] = S(vs[k]->z);
    }
    margin0[i] = S(c.margin0);
    margin1[i] = S(c.margin1);
    if (sort_key) sort_key[i] = c.key;
  }
  return n;
}

extern "C" int oracle_broadphase_f32(const OracleModel* m, const float* body_q, int capacity, int* pairs_out) {
  std::vector<ShapeWorld<float>> sw(m->shape_count);
  for (int s = 0; s < m->shape_count; ++s) sw[s] = compute_shape_world<float, float>(m, body_q, s);
  std::vector<std::pair<int, int>> pairs;
  sap_pairs(m, sw, pairs);
  int n = int(pairs.size());
  for (int i = 0; i < n && i < capacity; ++i) {
    pairs_out[2 * i] = pairs[i].first;
    pairs_out[2 * i + 1] = pairs[i].second;
  }
  return n;
}

extern "C" int oracle_collide_f32(const OracleModel* m, const float* body_q, int capacity, int* shape0, int* shape1,
                                  float* point0, float* point1, float* offset0, float* offset1, float* normal,
                                  float* margin0, float* margin1, int64_t* sort_key) {
  return collide_impl<float, float>(m, body_q, capacity, shape0, shape1, point0, point1, offset0, offset1, normal,
                                    margin0, margin1, sort_key);
}

extern "C" int oracle_collide_f64(const OracleModel* m, const double* body_q, int capacity, int* shape0, int* shape1,
                                  double* point0, double* point1, double* offset0, double* offset1, double* normal,
                                  double* margin0, double* margin1, int64_t* sort_key) {
  return collide_impl<double, double>(m, body_q, capacity, shape0, shape1, point0, point1, offset0, offset1, normal,
                                      margin0, margin1, sort_key);
}

template <class T, class S>
static void step_impl(const OracleModel* m, const OracleXpbdParams* p, const S* q_in, const S* qd_in, const S* f_in,
                      const float* joint_f, const float* tpos, const float* tvel, int nc, const int* shape0,
                      const int* shape1, const S* point0, const S* point1, const S* offset0, const S* offset1,
                      const S* normal, const S* margin0, const S* margin1, S dt, S* q_out, S* qd_out,
                      S* contact_force = nullptr, S* body_parent_f = nullptr) {
  int B = m->body_count;
  std::vector<Body<T>> state(B);
  std::vector<Spatial<T>> f(B);
  for (int b = 0; b < B; ++b) {
    state[b].q = load_xf_t<T, S>(q_in + 7 * b);
    state[b].v = V3<T>{T(qd_in[6 * b]), T(qd_in[6 * b + 1]), T(qd_in[6 * b + 2])};
    state[b].w = V3<T>{T(qd_in[6 * b + 3]), T(qd_in[6 * b + 4]), T(qd_in[6 * b + 5])};
    f[b].lin = V3<T>{T(f_in[6 * b]), T(f_in[6 * b + 1]), T(f_in[6 * b + 2])};
    f[b].ang = V3<T>{T(f_in[6 * b + 3]), T(f_in[6 * b + 4]), T(f_in[6 * b + 5])};
  }
  std::vector<ContactIn<T>> cs(nc > 0 ? nc : 0);
  auto ld = [](const S* a, int i) { return V3<T>{T(a[3 * i]), T(a[3 * i + 1]), T(a[3 * i + 2])}; };
  for (int i = 0; i < nc; ++i) {
    cs[i].shape0 = shape0[i];
    cs[i].shape1 = shape1[i];
    cs[i].p0 = ld(point0, i);
    cs[i].p1 = ld(point1, i);
    cs[i].o0 = ld(offset0, i);
    cs[i].o1 = ld(offset1, i);
    cs[i].n = ld(normal, i);
    cs[i].m0 = T(margin0[i]);
    cs[i].m1 = T(margin1[i]);
  }
  std::vector<Spatial<T>> cf(contact_force ? cs.size() : 0), pf(body_parent_f ? B : 0);
  xpbd_step<T>(m, p, state, f, joint_f, tpos, tvel, cs, nc >= 0, T(dt), contact_force ? cf.data() : nullptr,
               body_parent_f ? pf.data() : nullptr);
  auto put6 = [](S* o, const Spatial<T>& v) {
    o[0] = S(v.lin.x); o[1] = S(v.lin.y); o[2] = S(v.lin.z); o[3] = S(v.ang.x); o[4] = S(v.ang.y); o[5] = S(v.ang.z);
  };
  for (size_t i = 0; i < cf.size(); ++i) put6(contact_force + 6 * i, cf[i]);
  for (size_t b = 0; b < pf.size(); ++b) put6(body_parent_f + 6 * b, pf[b]);
  for (int b = 0; b < B; ++b) {
    const Body<T>& s = state[b];
    S* q = q_out + 7 * b;
    q[0] = S(s.q.p.x); q[1] = S(s.q.p.y); q[2] = S(s.q.p.z);
    q[3] = S(s.q.q.x); q[4] = S(s.q.q.y); q[5] = S(s.q.q.z); q[6] = S(s.q.q.w);
    S* qd = qd_out + 6 * b;
    qd[0] = S(s.v.x); qd[1] = S(s.v.y); qd[2] = S(s.v.z);
    qd[3] = S(s.w.x); qd[4] = S(s.w.y); qd[5] = S(s.w.z);
  }
}

extern "C" void oracle_xpbd_step_f32(const OracleModel* m, const OracleXpbdParams* p, const float* q_in,
                                     const float* qd_in, const float* f_in, const float* joint_f, const float* tpos,
                                     const float* tvel, int nc, const int* shape0, const int* shape1,
                                     const float* point0, const float* point1, const float* offset0,
                                     const float* offset1, const float* normal, const float* margin0,
                                     const float* margin1, float dt, float* q_out, float* qd_out) {
  step_impl<float, float>(m, p, q_in, qd_in, f_in, joint_f, tpos, tvel, nc, shape0, shape1, point0, point1, offset0,
                          offset1, normal, margin0, margin1, dt, q_out, qd_out);
}

extern "C" void oracle_xpbd_step_report_f32(const OracleModel* m, const OracleXpbdParams* p, const float* q_in,
                                            const float* qd_in, const float* f_in, const float* joint_f,
                                            const float* tpos, const float* tvel, int nc, const int* shape0,
                                            const int* shape1, const float* point0, const float* point1,
                                            const float* offset0, const float* offset1, const float* normal,
                                            const float* margin0, const float* margin1, float dt, float* q_out,
                                            float* qd_out, float* contact_force, float* body_parent_f) {
  step_impl<float, float>(m, p, q_in, qd_in, f_in, joint_f, tpos, tvel, nc, shape0, shape1, point0, point1, offset0,
                          offset1, normal, margin0, margin1, dt, q_out, qd_out, contact_force, body_parent_f);
}

extern "C" void oracle_xpbd_step_f64(const OracleModel* m, const OracleXpbdParams* p, const double* q_in,
                                     const double* qd_in, const double* f_in, const float* joint_f, const float* tpos,
                                     const float* tvel, int nc, const int* shape0, const int* shape1,
                                     const double* point0, const double* point1, const double* offset0,
                                     const double* offset1, const double* normal, const double* margin0,
                                     const double* margin1, double dt, double* q_out, double* qd_out) {
  step_impl<double, double>(m, p, q_in, qd_in, f_in, joint_f, tpos, tvel, nc, shape0, shape1, point0, point1,
                            offset0, offset1, normal, margin0, margin1, dt, q_out, qd_out);
}

extern "C" void oracle_asin_f32(const float* x, int n, float* out) {
  for (int i = 0; i < n; ++i) out[i] = tasin(x[i]);
}

extern "C" void oracle_acos_f32(const float* x, int n, float* out) {
  for (int i = 0; i < n; ++i) out[i] = tacos(x[i]);
}

// ---------------------------------------------------------------- learned warm start (oracle_learn.h)
#include "oracle_learn.h"

extern "C" void oracle_xpbd_step_learn_f32(const OracleModel* m, const OracleXpbdParams* p, const float* q_in,
                                           const float* qd_in, const float* f_in, const float* joint_f,
                                           const float* tpos, const float* tvel, int nc, const int* shape0,
                                           const int* shape1, const float* point0, const float* point1,
                                           const float* offset0, const float* offset1, const float* normal,
                                           const float* margin0, const float* margin1, float dt, int mode,
                                           const float* lambda_prev, const float* lambda0, int warm_iterations,
                                           float residual_tol, float* feat, float* lambda_sum, float* residual,
                                           int* fallback, float* q_out, float* qd_out) {
  int B = m->body_count;
  std::vector<Body<float>> state(B);
  std::vector<Spatial<float>> f(B);
  for (int b = 0; b < B; ++b) {
    state[b].q = load_xf_t<float, float>(q_in + 7 * b);
    state[b].v = V3<float>{qd_in[6 * b], qd_in[6 * b + 1], qd_in[6 * b + 2]};
    state[b].w = V3<float>{qd_in[6 * b + 3], qd_in[6 * b + 4], qd_in[6 * b + 5]};
    f[b].lin = V3<float>{f_in[6 * b], f_in[6 * b + 1], f_in[6 * b + 2]};
    f[b].ang = V3<float>{f_in[6 * b + 3], f_in[6 * b + 4], f_in[6 * b + 5]};
  }
  std::vector<ContactIn<float>> cs(nc > 0 ? nc : 0);
  auto ld = [](const float* a, int i) { return V3<float>{a[3 * i], a[3 * i + 1], a[3 * i + 2]}; };
  for (int i = 0; i < nc; ++i) {
    cs[i].shape0 = shape0[i];
    cs[i].shape1 = shape1[i];
    cs[i].p0 = ld(point0, i);
    cs[i].p1 = ld(point1, i);
    cs[i].o0 = ld(offset0, i);
    cs[i].o1 = ld(offset1, i);
    cs[i].n = ld(normal, i);
    cs[i].m0 = margin0[i];
    cs[i].m1 = margin1[i];
  }
  xpbd_step_learn<float>(m, p, state, f, joint_f, tpos, tvel, cs, dt, mode, lambda_prev, lambda0, warm_iterations,
                         residual_tol, feat, lambda_sum, residual, fallback);
  for (int b = 0; b < B; ++b) {
    const Body<float>& s = state[b];
    float* q = q_out + 7 * b;
    q[0] = s.q.p.x; q[1] = s.q.p.y; q[2] = s.q.p.z;
    q[3] = s.q.q.x; q[4] = s.q.q.y; q[5] = s.q.q.z; q[6] = s.q.q.w;
    float* qd = qd_out + 6 * b;
    qd[0] = s.v.x; qd[1] = s.v.y; qd[2] = s.v.z;
    qd[3] = s.w.x; qd[4] = s.w.y; qd[5] = s.w.z;
  }
}

extern "C" void oracle_mlp_forward(const float* w1, const float* b1, const float* w2, const float* b2, const float* w3,
                                   const float* b3, const float* x, int n, float* y) {
  MlpParams P{w1, b1, w2, b2, w3, b3};
  mlp_forward(P, x, n, y);
}

extern "C" void oracle_mlp_backward(const float* w1, const float* b1, const float* w2, const float* b2,
                                    const float* w3, const float* b3, const float* x, const float* t, int n,
                                    float* gw1, float* gb1, float* gw2, float* gb2, float* gw3, float* gb3,
                                    float* loss) {
  MlpParams P{w1, b1, w2, b2, w3, b3};
  mlp_backward(P, x, t, n, gw1, gb1, gw2, gb2, gw3, gb3, loss);
}

extern "C" void oracle_sdf_build_f32(const float* vertices, int vertex_count, const int* triangles, int triangle_count,
                                     const float lower[3], const float cell[3], const int n[3], float winding_threshold,
                                     float* out) {
  sdf_cook<float>(vertices, vertex_count, triangles, triangle_count, lower, cell, n, winding_threshold, out);
}

extern "C" void oracle_sdf_build_f64(const float* vertices, int vertex_count, const int* triangles, int triangle_count,
                                     const float lower[3], const float cell[3], const int n[3], float winding_threshold,
                                     double* out) {
  sdf_cook<double>(vertices, vertex_count, triangles, triangle_count, lower, cell, n, winding_threshold, out);
}

extern "C" int oracle_sdf_subgrids(const float* fine, const int dims[3], float band_lo, float band_hi, float lin_thr,
                                   float* coarse, int* occupied, int* required, int* linear, int* address) {
  return sdf_subgrids(fine, dims[0], dims[1], dims[2], band_lo, band_hi, lin_thr, coarse, occupied, required, linear,
                      address);
}

extern "C" void oracle_sdf_pack(const float* fine, const int dims[3], const int* required, const int* linear,
                                const int* address, int blocks, int fmt, float qmin, float qinv, uint32_t* slots,
                                unsigned char* texels) {
  sdf_pack(fine, dims[0], dims[1], dims[2], required, linear, address, blocks, fmt, qmin, qinv, slots, texels);
}

extern "C" void oracle_sdf_sample_f32(const OracleModel* m, int mesh, const float* p, int n, float* value,
                                      float* grad) {
  SdfG<float> g = sdf_grid<float>(m, mesh);
  for (int i = 0; i < n; ++i) {
    V3<float> x{p[3 * i], p[3 * i + 1], p[3 * i + 2]};
    V3<float> gr;
    value[i] = sdf_sample_grad(g, x, gr);
    grad[3 * i] = gr.x;
    grad[3 * i + 1] = gr.y;
    grad[3 * i + 2] = gr.z;
  }
}

extern "C" void oracle_mesh_distance_f64(const float* vertices, int vertex_count, const int* triangles,
                                         int triangle_count, const double* points, int n, double* out) {
  const double pi = 3.14159265358979323846;
  for (int i = 0; i < n; ++i) {
    V3<double> p{points[3 * i], points[3 * i + 1], points[3 * i + 2]};
    double best = 1e300, omega = 0.0;
    for (int t = 0; t < triangle_count; ++t) {
      int i0 = triangles[3 * t], i1 = triangles[3 * t + 1], i2 = triangles[3 * t + 2];
      V3<double> a{vertices[3 * i0], vertices[3 * i0 + 1], vertices[3 * i0 + 2]};
      V3<double> b{vertices[3 * i1], vertices[3 * i1 + 1], vertices[3 * i1 + 2]};
      V3<double> c{vertices[3 * i2], vertices[3 * i2 + 1], vertices[3 * i2 + 2]};
      double dd = length(p - closest_on_triangle(p, a, b, c));
      best = dd < best ? dd : best;
      omega += solid_angle(a - p, b - p, c - p);
    }
    out[i] = omega / (4.0 * pi) > 0.5 ? -best : best;
  }
}

extern "C" void oracle_closest_on_triangle_f64(const double* p, const double* a, const double* b, const double* c,
                                               int n, double* out) {
  for (int i = 0; i < n; ++i) {
    V3<double> q = closest_on_triangle(V3<double>{p[3 * i], p[3 * i + 1], p[3 * i + 2]},
                                       V3<double>{a[3 * i], a[3 * i + 1], a[3 * i + 2]},
                                       V3<double>{b[3 * i], b[3 * i + 1], b[3 * i + 2]},
                                       V3<double>{c[3 * i], c[3 * i + 1], c[3 * i + 2]});
    out[3 * i] = q.x;
    out[3 * i + 1] = q.y;
    out[3 * i + 2] = q.z;
  }
}

extern "C" void oracle_solid_angle_f64(const double* a, const double* b, const double* c, int n, double* out) {
  for (int i = 0; i < n; ++i)
    out[i] = solid_angle(V3<double>{a[3 * i], a[3 * i + 1], a[3 * i + 2]},
                         V3<double>{b[3 * i], b[3 * i + 1], b[3 * i + 2]},
                         V3<double>{c[3 * i], c[3 * i + 1], c[3 * i + 2]});
}

extern "C" void oracle_edge_search_f32(const OracleModel* m, int mesh, const float* v0, const float* v1, int n,
                                       float* value, float* point) {
  SdfG<float> g = sdf_grid<float>(m, mesh);
  for (int i = 0; i < n; ++i) {
    V3<float> a{v0[3 * i], v0[3 * i + 1], v0[3 * i + 2]}, b{v1[3 * i], v1[3 * i + 1], v1[3 * i + 2]};
    V3<float> mid = (a + b) * 0.5f, pt;
    value[i] = edge_sdf_search(g, a, b, sdf_sample(g, mid), pt);
    point[3 * i] = pt.x;
    point[3 * i + 1] = pt.y;
    point[3 * i + 2] = pt.z;
  }
}

extern "C" void* oracle_matcher_create(float pos_threshold, float normal_dot_threshold, int sticky) {
  MatcherState* S = new MatcherState();
  S->pos_threshold_sq = pos_threshold * pos_threshold;
  S->normal_dot_threshold = normal_dot_threshold;
  S->sticky = sticky != 0;
  return S;
}

extern "C" void oracle_matcher_destroy(void* h) { delete static_cast<MatcherState*>(h); }

extern "C" void oracle_matcher_reset(void* h) {
  MatcherState* S = static_cast<MatcherState*>(h);
  S->keys.clear();
}

// Returns the number of broken previous contacts written to broken_keys (at most broken_cap).
extern "C" int oracle_contact_match(void* h, const OracleModel* m, const float* body_q, int n, const int64_t* keys,
                                    const int* shape0, const int* shape1, float* point0, float* point1,
                                    float* offset0, float* offset1, float* normal, int* match_index,
                                    int64_t* matched_key, int64_t* broken_keys, int broken_cap) {
  std::vector<int64_t> broken;
  contact_match(*static_cast<MatcherState*>(h), m, body_q, n, keys, shape0, shape1, point0, point1, offset0, offset1,
                normal, match_index, matched_key, broken);
  int k = int(broken.size());
  for (int i = 0; i < k && i < broken_cap; ++i) broken_keys[i] = broken[i];
  return k;
}

extern "C" int oracle_broad_phase(int mode, const float* lower, const float* upper, const float* gap, const int* group,
                                  const int* world, const int* filter, int filter_count, const int* index_map,
                                  const int* slice_ends, int segments, const int* pairs, int pair_count, int* out,
                                  int capacity) {
  BroadInput in{lower, upper, gap, group, world, filter, filter_count};
  std::vector<int> res;
  if (mode == 0) broad_sap(in, index_map, slice_ends, segments, res);
  else if (mode == 1) broad_all_pairs(in, index_map, slice_ends, segments, res);
  else broad_explicit(in, pairs, pair_count, res);
  const int found = int(res.size() / 2);
  for (int k = 0; k < found && k < capacity; ++k) {
    out[2 * k] = res[2 * k];
    out[2 * k + 1] = res[2 * k + 1];
  }
  return found;
}

// ---------------------------------------------------------------- contact reduction primitives (oracle_reduce.h)
extern "C" void oracle_red_get_slot(const float* n, int count, int* out) {
  for (int i = 0; i < count; ++i) out[i] = red_get_slot(V3<float>{n[3 * i], n[3 * i + 1], n[3 * i + 2]});
}

extern "C" void oracle_red_project(const int* bin, const float* p, int count, float* out) {
  for (int i = 0; i < count; ++i) {
    V2<float> r = red_project(bin[i], V3<float>{p[3 * i], p[3 * i + 1], p[3 * i + 2]});
    out[2 * i] = r.x;
    out[2 * i + 1] = r.y;
  }
}

extern "C" void oracle_red_direction(float* out) {
  for (int i = 0; i < RED_DIRS; ++i) {
    V2<float> d = red_direction(i);
    out[2 * i] = d.x;
    out[2 * i + 1] = d.y;
  }
}

extern "C" void oracle_red_oct(const float* n, int count, float* enc, float* dec) {
  for (int i = 0; i < count; ++i) {
    V2<float> e = red_encode_oct(V3<float>{n[3 * i], n[3 * i + 1], n[3 * i + 2]});
    V3<float> d = red_decode_oct(e);
    enc[2 * i] = e.x;
    enc[2 * i + 1] = e.y;
    dec[3 * i] = d.x;
    dec[3 * i + 1] = d.y;
    dec[3 * i + 2] = d.z;
  }
}

extern "C" void oracle_red_value(const float* score, const int* fingerprint, const unsigned* id, int count, int det,
                                 unsigned long long* out) {
  for (int i = 0; i < count; ++i) out[i] = red_value(score[i], fingerprint[i], id[i], det != 0);
}

extern "C" void oracle_red_voxel_index(const float* p, int count, const float lo[3], const float hi[3],
                                       const int res[3], int* out) {
  const V3<float> l{lo[0], lo[1], lo[2]}, h{hi[0], hi[1], hi[2]};
  for (int i = 0; i < count; ++i) out[i] = red_voxel_index(V3<float>{p[3 * i], p[3 * i + 1], p[3 * i + 2]}, l, h, res);
}

extern "C" void oracle_red_voxel_resolution(const double lo[3], const double hi[3], int budget, int out[3]) {
  red_voxel_resolution(lo, hi, budget, out);
}

// The reduction over given candidates (position, normal used for the bin = normal, projected point = position,
// use_beta, voxel, depth, sub key); writes the winners' candidate indices, returns their number.
extern "C" int oracle_red_reduce(const float* proj_point, const float* bin_normal, const float* depth,
                                 const int* sub, const int* use_beta, const int* voxel, int count, int det,
                                 int* winners) {
  std::vector<RedCand<float>> c(count);
  for (int i = 0; i < count; ++i) {
    c[i].pos = V3<float>{proj_point[3 * i], proj_point[3 * i + 1], proj_point[3 * i + 2]};
    c[i].proj_point = c[i].pos;
    c[i].bin_normal = V3<float>{bin_normal[3 * i], bin_normal[3 * i + 1], bin_normal[3 * i + 2]};
    c[i].normal = c[i].bin_normal;
    c[i].depth = depth[i];
    c[i].sub = sub[i];
    c[i].use_beta = use_beta[i] != 0;
    c[i].voxel = voxel[i];
  }
  std::vector<int> w = red_reduce(c, det != 0);
  for (size_t k = 0; k < w.size(); ++k) winners[k] = w[k];
  return static_cast<int>(w.size());
}
