// Rigid restitution pass of SolverXPBD(enable_restitution=True) for gfx950: apply_rigid_restitution
// (newton/_src/solvers/xpbd/kernels.py:2472-2615) followed by apply_body_delta_velocities (kernels.py:882-887),
// run after the step's iterations (solver_xpbd.py:719-764).
//
// One workgroup per world (index world_count: the world -1 bodies).  The world's contacts are taken in chunks
// of BS, one contact per lane: each lane evaluates its contact on the start-of-step poses and velocities and
// the solved velocities and leaves the two velocity deltas (shape 0's body, shape 1's body) in LDS.  Then one
// lane per body walks the chunk in contact order and adds the entries that name its body, so every body's sum
// runs in ascending contact order (the reference adds them with float atomics) and equals a sequential sweep.
// The walk is a broadcast LDS read per entry; worlds above BS bodies repeat the chunks per group of BS bodies.
// Every contact reads the solved velocities, so the sums are held in LDS and added to the velocities only
// after the world's last chunk.
//
// Traffic per world: the contacts' records once per body group, the bodies' start and solved states once
// each, and the solved velocities written once; restitution changes no pose.
#include <hip/hip_runtime.h>

#include "nmc_internal.h"
#include "nmc_math.h"

namespace nmc {

struct RestArgs {
  NmcModelDesc d;
  NmcContacts c;
  const float* q_prev;   // start-of-step poses (state_in.body_q)
  const float* qd_prev;  // start-of-step velocities (state_in.body_qd)
  float* qd;             // solved velocities (state_out.body_qd), updated in place
  float dt;
};

struct RestEntry {
  int la, lb;  // local body of shape 0 / shape 1 that receives a delta, -1 when none
  f3 lin_a, ang_a, lin_b, ang_b;
};

// apply_rigid_restitution (kernels.py:2472-2615) for the contact at canonical position j, in the reference's
// operation order (the file is compiled without contraction).  Bodies are local to the world [b0, b0 + nb);
// a body outside it never appears in the world's segment.
__device__ RestEntry restitution_contact(const RestArgs& a, int j, int b0, int nb) {
  const NmcModelDesc& d = a.d;
  const f3 zero{0.0f, 0.0f, 0.0f};
  RestEntry o{-1, -1, zero, zero, zero, zero};
  const int gc = cpos(a.c, j);
  if (gc < 0) return o;
  const int sa = a.c.shape0[gc], sb = a.c.shape1[gc];
  if (sa == sb) return o;
  int ba = -1, bb = -1, nz = 0;
  float e = 0.0f;
  const float* rest = d.shape_material_restitution;
  if (sa >= 0) { nz += 1; e += rest ? rest[sa] : 0.0f; ba = d.shape_body[sa]; }
  if (sb >= 0) { nz += 1; e += rest ? rest[sb] : 0.0f; bb = d.shape_body[sb]; }
  if (nz > 0) e /= static_cast<float>(nz);
  if (ba == bb) return o;
  if (ba >= 0 && (ba < b0 || ba >= b0 + nb)) ba = -1;
  if (bb >= 0 && (bb < b0 || bb >= b0 + nb)) bb = -1;
  xform Xa = xident(), Xb = xident();
  f3 com_a = zero, com_b = zero;
  float mia = 0.0f, mib = 0.0f;
  m33 Ia{}, Ib{};
  // effective inverse mass and inertia: zero for kinematic bodies (solver.py:160-175)
  if (ba >= 0) {
    Xa = load_xform(a.q_prev + 7 * ba);
    const bool kin = (d.body_flags[ba] & BODY_KINEMATIC) != 0;
    mia = kin ? 0.0f : d.body_inv_mass[ba];
    if (!kin) Ia = load_m33(d.body_inv_inertia + 9 * ba);
    com_a = load3(d.body_com + 3 * ba);
  }
  if (bb >= 0) {
    Xb = load_xform(a.q_prev + 7 * bb);
    const bool kin = (d.body_flags[bb] & BODY_KINEMATIC) != 0;
    mib = kin ? 0.0f : d.body_inv_mass[bb];
    if (!kin) Ib = load_m33(d.body_inv_inertia + 9 * bb);
    com_b = load3(d.body_com + 3 * bb);
  }
  const f3 bx_a = xpoint(Xa, load3(a.c.point0 + 3 * gc) + load3(a.c.offset0 + 3 * gc));
  const f3 bx_b = xpoint(Xb, load3(a.c.point1 + 3 * gc) + load3(a.c.offset1 + 3 * gc));
  const float thickness = a.c.margin0[gc] + a.c.margin1[gc];
  const f3 n = load3(a.c.normal + 3 * gc);
  const float dist = dot(n, bx_b - bx_a) - thickness;
  if (dist >= 0.0f) return o;
  const f3 r_a = bx_a - xpoint(Xa, com_a);
  const f3 r_b = bx_b - xpoint(Xb, com_b);
  f3 va = zero, vb = zero, va_new = zero, vb_new = zero, rxn_a = zero, rxn_b = zero;
  float inv_mass = 0.0f;
  if (ba >= 0) {
    const int wi = d.body_world[ba];
    const f3 g = load3(d.gravity + 3 * (wi > 0 ? wi : 0));
    // velocity_at_point (newton/_src/math/spatial.py:54-78): v + w x r
    va = (cross(load3(a.qd_prev + 6 * ba + 3), r_a) + load3(a.qd_prev + 6 * ba)) + g * a.dt;
    va_new = cross(load3(a.qd + 6 * ba + 3), r_a) + load3(a.qd + 6 * ba);
    rxn_a = rot_inv(Xa.q, cross(r_a, n));
    inv_mass = inv_mass + (mia + dot(rxn_a, mv(Ia, rxn_a)));
  }
  if (bb >= 0) {
    const int wi = d.body_world[bb];
    const f3 g = load3(d.gravity + 3 * (wi > 0 ? wi : 0));
    vb = (cross(load3(a.qd_prev + 6 * bb + 3), r_b) + load3(a.qd_prev + 6 * bb)) + g * a.dt;
    vb_new = cross(load3(a.qd + 6 * bb + 3), r_b) + load3(a.qd + 6 * bb);
    rxn_b = rot_inv(Xb.q, cross(r_b, n));
    inv_mass = inv_mass + (mib + dot(rxn_b, mv(Ib, rxn_b)));
  }
  if (inv_mass == 0.0f) return o;
  const float rel_old = dot(n, vb - va);
  const float rel_new = dot(n, vb_new - va_new);
  if (rel_old >= 0.0f) return o;
  const float dv = (-rel_new - e * rel_old) / inv_mass;
  if (ba >= 0) {
    const float dva = -dv;
    o.la = ba - b0;
    o.lin_a = (n * mia) * dva;
    o.ang_a = rot(Xa.q, mv(Ia, rxn_a) * dva);
  }
  if (bb >= 0) {
    o.lb = bb - b0;
    o.lin_b = (n * mib) * dv;
    o.ang_b = rot(Xb.q, mv(Ib, rxn_b) * dv);
  }
  return o;
}

// MAXB: bodies a world may hold (the model's largest world, at most 512).  Every delta is computed from the
// solved velocities before any of them is updated: the sums wait in LDS until the last chunk of the last body
// group is done.
template <int BS, int MAXB>
__global__ void __launch_bounds__(BS) restitution_kernel(RestArgs a) {
  __shared__ int2 s_body[BS];
  __shared__ float s_delta[BS][12];  // lin_a, ang_a, lin_b, ang_b
  __shared__ float s_acc[MAXB][6];
  const NmcModelDesc& d = a.d;
  const int w = blockIdx.x;
  const int b0 = d.world_body_first[w];
  const int nb = d.world_body_count[w];
  if (nb == 0) return;  // the whole workgroup leaves together
  const int cstart = a.c.world_start[w];
  const int nc = a.c.world_count[w];
  const int tid = threadIdx.x;
  for (int g0 = 0; g0 < nb; g0 += BS) {
    const int lb = g0 + tid;
    f3 acc_v{0.0f, 0.0f, 0.0f}, acc_w{0.0f, 0.0f, 0.0f};
    for (int k0 = 0; k0 < nc; k0 += BS) {
      const int k = k0 + tid;
      if (k < nc) {
        const RestEntry r = restitution_contact(a, cstart + k, b0, nb);
        s_body[tid] = make_int2(r.la, r.lb);
        float* s = s_delta[tid];
        store3(s, r.lin_a);
        store3(s + 3, r.ang_a);
        store3(s + 6, r.lin_b);
        store3(s + 9, r.ang_b);
      }
      __syncthreads();
      if (lb < nb) {
        const int m = nc - k0 < BS ? nc - k0 : BS;
        for (int i = 0; i < m; ++i) {
          const int2 e = s_body[i];
          // a contact's two bodies differ, so at most one side names this body
          if (e.x == lb) {
            acc_v = acc_v + load3(s_delta[i]);
            acc_w = acc_w + load3(s_delta[i] + 3);
          } else if (e.y == lb) {
            acc_v = acc_v + load3(s_delta[i] + 6);
            acc_w = acc_w + load3(s_delta[i] + 9);
          }
        }
      }
      __syncthreads();
    }
    if (lb < nb) {
      store3(s_acc[lb], acc_v);
      store3(s_acc[lb] + 3, acc_w);
    }
  }
  __syncthreads();
  // apply_body_delta_velocities: every body's velocity += its delta (kinematic bodies keep their start
  // velocity, written by the step: copy_kinematic_body_state runs after this pass in the reference)
  for (int lb = tid; lb < nb; lb += BS) {
    const int b = b0 + lb;
    if (!(d.body_flags[b] & BODY_KINEMATIC)) {
      float* q = a.qd + 6 * b;
      store3(q, load3(q) + load3(s_acc[lb]));
      store3(q + 3, load3(q + 3) + load3(s_acc[lb] + 3));
    }
  }
}

hipError_t launch_restitution(NmcModelImpl* m, const float* q_prev, const float* qd_prev, const NmcContacts* c,
                              float dt, float* qd, hipStream_t stream) {
  RestArgs a;
  a.d = m->d;
  a.c = *c;
  a.q_prev = q_prev;
  a.qd_prev = qd_prev;
  a.qd = qd;
  a.dt = dt;
  const int worlds = m->d.world_count + 1;
  // the model's worlds hold at most 512 bodies (checked when it is created)
  if (m->d.max_world_bodies <= 64)
    hipLaunchKernelGGL((restitution_kernel<64, 64>), dim3(worlds), dim3(64), 0, stream, a);
  else if (m->d.max_world_bodies <= 512)
    hipLaunchKernelGGL((restitution_kernel<256, 512>), dim3(worlds), dim3(256), 0, stream, a);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace nmc
