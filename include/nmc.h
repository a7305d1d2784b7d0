/* C-ABI of libnmc.so: the gfx950 rigid-contact step (collision pipeline + XPBD solver).
 *
 * Drop-in boundary for Newton's rigid-contact hot path.  Every entry point takes device pointers and
 * plain sizes; no torch or Python types cross the boundary.  All launches are asynchronous on the
 * given hipStream_t and the library never synchronises the host inside nmc_collide / nmc_xpbd_step,
 * so the substep loop can be captured in a hipGraph.
 *
 * Reference interfaces replaced (newton @ nataliya-dev/newton_custom_contact):
 *   nmc_model_create / nmc_model_destroy  -- CollisionPipeline.__init__ buffers (newton/_src/sim/collide.py:483-856)
 *                                            and SolverXPBD.__init__ effective-mass state
 *                                            (newton/_src/solvers/xpbd/solver_xpbd.py:95-139, solver.py:199-216)
 *   nmc_collide        -- CollisionPipeline.collide(state, contacts)            (newton/_src/sim/collide.py:902-1230)
 *   nmc_xpbd_step      -- SolverXPBD.step(state_in, state_out, control, contacts, dt)
 *                                                                             (newton/_src/solvers/xpbd/solver_xpbd.py:259-764)
 *   nmc_get_counters   -- verify_narrow_phase_buffers overflow report        (newton/_src/geometry/narrow_phase.py:1334-1406)
 *   nmc_last_error     -- the ValueError/TypeError messages of the Python layer
 */
#ifndef NMC_H
#define NMC_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* nmc_stream_t; /* hipStream_t */

/* Device-resident model arrays (layouts of newton.Model; transforms are (p, q_xyzw)). */
typedef struct NmcModelDesc {
  int body_count, shape_count, joint_count, world_count;
  int joint_dof_count;
  /* bodies */
  const float* body_com;         /* B x 3 */
  const float* body_inertia;     /* B x 9 */
  const float* body_inv_mass;    /* B */
  const float* body_inv_inertia; /* B x 9 */
  const int* body_flags;         /* B */
  const int* body_world;         /* B */
  /* shapes */
  const float* shape_transform; /* S x 7 */
  const int* shape_body;
  const int* shape_type;
  const float* shape_scale; /* S x 3 */
  const float* shape_margin;
  const float* shape_gap;
  const float* shape_radius;
  const int* shape_flags;
  const int* shape_group;
  const int* shape_world;
  const float* shape_mu;
  const float* shape_mu_torsional;
  const float* shape_mu_rolling;
  /* broad-phase segments: segment w = world w shapes + global shapes (sorted by index);
   * the last segment holds only global shapes.  segment_count = world_count + 1. */
  int segment_count;
  const int* segment_start; /* segment_count + 1 */
  const int* segment_shape;
  int max_segment_shapes;
  /* sorted canonical (min,max) excluded shape pairs */
  int filter_count;
  const int* filter_pairs;
  /* gravity per world (max(world_count,1) x 3) */
  const float* gravity;
  /* per-world bodies and joints: world w owns bodies [world_body_first[w], +world_body_count[w])
   * and joints [world_joint_first[w], +world_joint_count[w]); index world_count = global entities */
  const int* world_body_first;
  const int* world_body_count;
  const int* world_joint_first;
  const int* world_joint_count;
  int max_world_bodies;
  int max_world_joints;
  /* body -> (joint, side) lists in joint order: entries [body_joint_start[b], body_joint_start[b+1]),
   * value = 2*joint + side (side 0 = parent, 1 = child) */
  const int* body_joint_start; /* B + 1 */
  const int* body_joint_entry;
  /* joints */
  const int* joint_type;
  const int* joint_enabled;
  const int* joint_parent;
  const int* joint_child;
  const float* joint_X_p; /* J x 7 */
  const float* joint_X_c; /* J x 7 */
  const int* joint_qd_start;
  const int* joint_dof_dim; /* J x 2 */
  const float* joint_axis;  /* D x 3 */
  const float* joint_limit_lower;
  const float* joint_limit_upper;
  const float* joint_target_ke;
  const float* joint_target_kd;
  /* capacities: nmc_xpbd_step accepts Contacts buffers of at most rigid_contact_max slots.
   * shape_pairs_max > 0 creates the collision buffers (each segment gets a fixed region for its
   * worst case n(n-1)/2 pairs, 5 contact slots per pair); 0 creates a solver-only model on which
   * nmc_collide fails. */
  int rigid_contact_max;
  int shape_pairs_max;
} NmcModelDesc;

/* Device-resident contact buffers (layout of newton.Contacts rigid part). */
typedef struct NmcContacts {
  int capacity;      /* rigid_contact_max */
  int* count;        /* [1] total contacts found (may exceed capacity; excess is dropped) */
  int* shape0;
  int* shape1;
  float* point0;  /* C x 3, body frame of shape0's body */
  float* point1;  /* C x 3, body frame of shape1's body */
  float* offset0; /* C x 3 */
  float* offset1; /* C x 3 */
  float* normal;  /* C x 3, world frame, shape0 -> shape1 */
  float* margin0;
  float* margin1;
  int64_t* sort_key; /* optional (NULL): (shape0<<43 | shape1<<23 | sub_key) */
  int* world_start;  /* segment_count: first contact of each segment */
  int* world_count;  /* segment_count: contacts of each segment stored in the buffer */
} NmcContacts;

typedef struct NmcXpbdParams {
  int iterations;
  float rigid_contact_relaxation;
  int rigid_contact_con_weighting;
  float angular_damping;
  float joint_linear_relaxation;
  float joint_angular_relaxation;
  float joint_linear_compliance;
  float joint_angular_compliance;
  int has_contacts; /* 0: step(..., contacts=None, ...) */
} NmcXpbdParams;

typedef struct NmcModel NmcModel;

/* Status codes: 0 ok, <0 error (see nmc_last_error()). */
int nmc_model_create(const NmcModelDesc* desc, NmcModel** out);
void nmc_model_destroy(NmcModel* model);

/* Shape AABBs -> SAP broad phase -> primitive / MPR-GJK narrow phase -> contacts written in
 * canonical order: segment-major, and inside a segment ascending by the 64-bit sort key. */
int nmc_collide(NmcModel* model, const float* body_q, const NmcContacts* contacts, nmc_stream_t stream);

/* One XPBD step: integrate, `iterations` x (contact solve + apply, joint solve + apply).  Fails when
 * contacts->capacity exceeds the model's rigid_contact_max or when state in/out arrays alias. */
int nmc_xpbd_step(NmcModel* model, const NmcXpbdParams* params, const float* body_q_in, const float* body_qd_in,
                  const float* body_f_in, const float* joint_f, const float* joint_target_pos,
                  const float* joint_target_vel, const NmcContacts* contacts, float dt, float* body_q_out,
                  float* body_qd_out, nmc_stream_t stream);

/* Copies the counters accumulated since model creation to host (synchronises the stream):
 * out[0] = broad-phase pairs dropped (per-segment pair buffer full), out[1] = reserved (0),
 * out[2] = contacts beyond the Contacts capacity, out[3] = worlds whose contacts did not fit the
 * solver's ordering buffer (cannot happen when contacts->capacity <= rigid_contact_max). */
int nmc_get_counters(NmcModel* model, int* out, nmc_stream_t stream);

const char* nmc_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
