/* CPU oracle -- TEST INFRASTRUCTURE ONLY (tests/, __graft_entry__.smoke(), bench.py cpu_baseline).
 *
 * C interface of the scalar CPU restatement of Newton's rigid-contact step
 * (newton/_src/sim/collide.py, geometry/{broad_phase_sap,narrow_phase,collision_*,mpr,simplex_solver,
 * multicontact}.py, solvers/solver.py, solvers/xpbd/{kernels,solver_xpbd}.py).
 * Built by oracle/Makefile into oracle/liboracle.so; the product library never links it.
 */
#ifndef NMC_ORACLE_H
#define NMC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct OracleModel {
  int body_count, shape_count, joint_count, world_count;
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
  /* broad-phase segments (world shapes + global shapes; last segment = global only) */
  int segment_count;
  const int* segment_start; /* segment_count + 1 */
  const int* segment_shape;
  /* sorted canonical excluded pairs */
  int filter_count;
  const int* filter_pairs; /* filter_count x 2 */
  /* gravity per world */
  int gravity_count;
  const float* gravity; /* gravity_count x 3 */
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
  /* triangle meshes: per-shape mesh index (-1 otherwise) and local AABB with the scale baked in;
   * concatenated vertices / unique edges; one texture SDF per mesh (coarse grid, slots, texels; per mesh the
   * offsets into the three arrays, box lower / upper / inverse cell size, quantisation min / range and
   * w, h, d, texture size, bytes per texel) */
  const int* shape_mesh;
  const float* shape_aabb_lower; /* S x 3 */
  const float* shape_aabb_upper; /* S x 3 */
  int mesh_count;
  const float* mesh_vertex;
  const int* mesh_vertex_start; /* M + 1 */
  const int* mesh_edge;         /* E x 2 */
  const int* mesh_edge_start;   /* M + 1 */
  const float* mesh_sdf_coarse;
  const uint32_t* mesh_sdf_slots;
  const unsigned char* mesh_sdf_texels;
  const int64_t* mesh_sdf_start; /* M x 3 */
  const float* mesh_sdf_box;     /* M x 9 */
  const float* mesh_sdf_quant;   /* M x 2 */
  const int* mesh_sdf_dims;      /* M x 5 */
  const int* mesh_sdf_flags;     /* M: bit 0 = scale baked into the grid */
  /* mesh contact reduction (reduce_contacts): 0 off (every contact), 1 fast packing, 2 deterministic packing;
   * per-shape voxel grid of the reduction (S x 3), builder.py:9936-10065 */
  int mesh_reduce;
  const int* shape_voxel_res;
  int broad_phase_mode;       /* 0 sap, 1 nxn (and explicit with the model's pairs), 2 explicit list */
  const int* explicit_pairs;  /* sorted (min, max) */
  int explicit_pair_count;
  const float* shape_restitution; /* S: shape_material_restitution (null: 0 for every shape) */
} OracleModel;

typedef struct OracleXpbdParams {
  int iterations;
  float rigid_contact_relaxation;
  int rigid_contact_con_weighting;
  float angular_damping;
  float joint_linear_relaxation;
  float joint_angular_relaxation;
  float joint_linear_compliance;
  float joint_angular_compliance;
  int enable_restitution; /* SolverXPBD(enable_restitution=True): apply_rigid_restitution after the iterations */
} OracleXpbdParams;

/* Broad phase only: writes canonical (min,max) pairs in discovery order, returns the pair count
 * (may exceed capacity; pairs past capacity are dropped).  body_q is B x 7. */
int oracle_broadphase_f32(const OracleModel* m, const float* body_q, int capacity, int* pairs_out);

/* Full collide: contacts sorted by the 64-bit key (shape_a<<43 | shape_b<<23 | sub_key).
 * Outputs per contact: shape0, shape1 (int), point0, point1, offset0, offset1, normal (x3),
 * margin0, margin1, sort_key (int64).  Returns the contact count (may exceed capacity). */
int oracle_collide_f32(const OracleModel* m, const float* body_q, int capacity, int* shape0, int* shape1,
                       float* point0, float* point1, float* offset0, float* offset1, float* normal, float* margin0,
                       float* margin1, int64_t* sort_key);
int oracle_collide_f64(const OracleModel* m, const double* body_q, int capacity, int* shape0, int* shape1,
                       double* point0, double* point1, double* offset0, double* offset1, double* normal,
                       double* margin0, double* margin1, int64_t* sort_key);

/* One SolverXPBD.step(state_in, state_out, control, contacts, dt) over contacts given in order. */
void oracle_xpbd_step_f32(const OracleModel* m, const OracleXpbdParams* p, const float* body_q_in,
                          const float* body_qd_in, const float* body_f_in, const float* joint_f,
                          const float* joint_target_pos, const float* joint_target_vel, int contact_count,
                          const int* shape0, const int* shape1, const float* point0, const float* point1,
                          const float* offset0, const float* offset1, const float* normal, const float* margin0,
                          const float* margin1, float dt, float* body_q_out, float* body_qd_out);
/* The same FP32 step, also writing the contact forces (update_contacts, per contact lin + ang, 6 floats;
 * solver_xpbd.py:767-821) and body_parent_f (per body, 6 floats; solver_xpbd.py:636-655). */
void oracle_xpbd_step_report_f32(const OracleModel* m, const OracleXpbdParams* p, const float* body_q_in,
                                 const float* body_qd_in, const float* body_f_in, const float* joint_f,
                                 const float* joint_target_pos, const float* joint_target_vel, int contact_count,
                                 const int* shape0, const int* shape1, const float* point0, const float* point1,
                                 const float* offset0, const float* offset1, const float* normal,
                                 const float* margin0, const float* margin1, float dt, float* body_q_out,
                                 float* body_qd_out, float* contact_force, float* body_parent_f);
void oracle_xpbd_step_f64(const OracleModel* m, const OracleXpbdParams* p, const double* body_q_in,
                          const double* body_qd_in, const double* body_f_in, const float* joint_f,
                          const float* joint_target_pos, const float* joint_target_vel, int contact_count,
                          const int* shape0, const int* shape1, const double* point0, const double* point1,
                          const double* offset0, const double* offset1, const double* normal,
                          const double* margin0, const double* margin1, double dt, double* body_q_out,
                          double* body_qd_out);

/* Signed distance of a triangle mesh at the vertices of a fine grid (distance to the nearest triangle,
 * negative where the winding number is above the threshold): vertex (i, j, k) at lower + (i, j, k) * cell,
 * n[0..2] vertices per axis, x fastest. */
void oracle_sdf_build_f32(const float* vertices, int vertex_count, const int* triangles, int triangle_count,
                          const float lower[3], const float cell[3], const int n[3], float winding_threshold,
                          float* out);
void oracle_sdf_build_f64(const float* vertices, int vertex_count, const int* triangles, int triangle_count,
                          const float lower[3], const float cell[3], const int n[3], float winding_threshold,
                          double* out);
/* The sparse texture derived from a fine grid of dims = (w, h, d) subgrids: coarse grid, occupied / required /
 * linear flags and addresses (returns the required count); then slots and texels (fmt bytes per texel). */
int oracle_sdf_subgrids(const float* fine, const int dims[3], float band_lo, float band_hi, float lin_thr,
                        float* coarse, int* occupied, int* required, int* linear, int* address);
void oracle_sdf_pack(const float* fine, const int dims[3], const int* required, const int* linear, const int* address,
                     int blocks, int fmt, float qmin, float qinv, uint32_t* slots, unsigned char* texels);

/* SDF value and gradient of mesh `mesh` of the model at local points p (n x 3). */
void oracle_sdf_sample_f32(const OracleModel* m, int mesh, const float* p, int n, float* value, float* grad);

/* FP64 signed distance of points to a triangle mesh (nearest triangle, winding-number sign). */
void oracle_mesh_distance_f64(const float* vertices, int vertex_count, const int* triangles, int triangle_count,
                              const double* points, int n, double* out);
/* Closest point on triangles (a, b, c) to points p and solid angles of triangles seen from the origin. */
void oracle_closest_on_triangle_f64(const double* p, const double* a, const double* b, const double* c, int n,
                                    double* out);
void oracle_solid_angle_f64(const double* a, const double* b, const double* c, int n, double* out);
/* Deepest point of segments v0 -> v1 in the SDF of mesh `mesh` (Brent search from the midpoint). */
void oracle_edge_search_f32(const OracleModel* m, int mesh, const float* v0, const float* v1, int n, float* value,
                            float* point);

/* Frame-to-frame contact matching (contact_match.py).  The matcher keeps the previous frame key-sorted.
 * Contacts are passed key-sorted; point0/1, offset0/1, normal are replayed in place when sticky.  Outputs
 * per contact the match index (into the previous key-sorted buffer, -1 not found, -2 broken) and the
 * matched previous key (-1 when none); returns the count of previous contacts no new contact won
 * (their keys, ascending, in broken_keys up to broken_cap). */
void* oracle_matcher_create(float pos_threshold, float normal_dot_threshold, int sticky);
void oracle_matcher_destroy(void* matcher);
void oracle_matcher_reset(void* matcher);
int oracle_contact_match(void* matcher, const OracleModel* m, const float* body_q, int n, const int64_t* keys,
                         const int* shape0, const int* shape1, float* point0, float* point1, float* offset0,
                         float* offset1, float* normal, int* match_index, int64_t* matched_key,
                         int64_t* broken_keys, int broken_cap);

/* Standalone broad phases over given AABBs (broad_phase_sap.py / broad_phase_nxn.py): mode 0 SAP, 1 all
 * pairs (both over the index map / slice ends of precompute_world_map), 2 explicit (pairs, pair_count).
 * gap and filter may be null.  Writes up to capacity pairs (x 2 ints) and returns the number found. */
int oracle_broad_phase(int mode, const float* lower, const float* upper, const float* gap, const int* group,
                       const int* world, const int* filter, int filter_count, const int* index_map,
                       const int* slice_ends, int segments, const int* pairs, int pair_count, int* out, int capacity);

// Elementwise FP32 asin / acos of the restatement (the revolute joint's twist and swing angles).
void oracle_asin_f32(const float* x, int n, float* out);
void oracle_acos_f32(const float* x, int n, float* out);

#ifdef __cplusplus
}
#endif
#endif
