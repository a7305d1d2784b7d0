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
  /* triangle meshes (GeoType MESH, newton/_src/geometry/types.py Mesh): shape_mesh[s] = mesh index or -1;
   * local AABB of each shape with its scale baked in (collide.py:215-245); mesh vertices and unique edges
   * (builder.py:10316-10351, edge order of Mesh.edges).  One sparse SDF per mesh, the reference's texture SDF
   * (TextureSDFData, sdf_texture.py:48-87, cooked by nmc_sdf_build / nmc_sdf_subgrids / nmc_sdf_pack):
   *   coarse: (d+1)(h+1)(w+1) floats, x fastest -- the field at the subgrid corners;
   *   slots: w*h*d uint32, z fastest (index (bx*h + by)*d + bz) -- texture block x | y << 10 | z << 20 of a
   *     narrow-band subgrid, 0xFFFFFFFE for a subgrid demoted to the coarse grid, 0xFFFFFFFF for far field;
   *   texels: ts^3 samples, x fastest, 9^3 per block, float32 or unorm16 / unorm8 (value = raw * range + min);
   * mesh_sdf_start[3m..3m+2] = offsets of mesh m into the coarse (floats), slot (uint32) and texel (bytes)
   * arrays; mesh_sdf_box[9m..] = sdf_box_lower, sdf_box_upper (the padded box), inv_sdf_dx;
   * mesh_sdf_quant[2m..] = subgrids_min_sdf_value, subgrids_sdf_value_range; mesh_sdf_dims[5m..] = w, h, d,
   * ts, bytes per texel; mesh_sdf_flags bit 0 set when the shape scale was multiplied into the vertices
   * before the cook (TextureSDFData.scale_baked, sdf_contact.py:918-930: the SDF side is then sampled with
   * unit scale). */
  const int* shape_mesh;
  const float* shape_aabb_lower; /* S x 3 */
  const float* shape_aabb_upper; /* S x 3 */
  int mesh_count;
  const float* mesh_vertex;      /* V x 3 */
  const int* mesh_vertex_start;  /* M + 1 */
  const int* mesh_edge;          /* E x 2, vertex indices local to the mesh */
  const int* mesh_edge_start;    /* M + 1 */
  const float* mesh_sdf_coarse;
  const uint32_t* mesh_sdf_slots;
  const uint8_t* mesh_sdf_texels;
  const int64_t* mesh_sdf_start; /* M x 3 */
  const float* mesh_sdf_box;     /* M x 9 */
  const float* mesh_sdf_quant;   /* M x 2 */
  const int* mesh_sdf_dims;      /* M x 5 */
  const int* mesh_sdf_flags;     /* M */
  /* CollisionPipeline(reduce_contacts, deterministic) for mesh pairs: 0 every contact the writer keeps
   * (narrow_phase.py:1090-1190, sdf_contact.py:830-1065); 1 the reference's normal-bin / voxel reduction with
   * the fast value packing, 2 with the deterministic packing (contact_reduction_global.py:232-1440); reduced,
   * a pair keeps at most 240 contacts. */
  int mesh_reduce;
  const int* shape_voxel_resolution; /* S x 3: the reduction's voxel grid per shape (builder.py:9924-10065) */
  /* CollisionPipeline broad_phase (collide.py:605-735): 0 "sap" -- a segment's shapes sorted by their
   * projection on the fixed axis, a pair tested only when the later one's lower bound is below the earlier
   * one's upper bound (broad_phase_sap.py:189-394); 1 "nxn" -- every pair of a segment (broad_phase_nxn.py:
   * 129-214; also "explicit" with the model's own shape_contact_pairs, the same set); 2 "explicit" with a
   * caller's list -- the pairs of explicit_pairs (sorted canonical (min, max)) whose AABBs overlap, without
   * the group and filter tests (broad_phase_nxn.py:28-65). */
  int broad_phase_mode;
  const int* explicit_pairs;
  int explicit_pair_count;
  /* S: shape_material_restitution (model.py), read by the restitution pass only; null = 0 for every shape */
  const float* shape_material_restitution;
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
  int* world_start;  /* segment_count: first canonical position of each segment's contacts */
  int* world_count;  /* segment_count: contacts of each segment (canonical positions below the capacity) */
  /* capacity: buffer position of the contact at each canonical position (-1: past the capacity, not stored).
   * Required with sort_key: the buffer is then in the reference's deterministic layout, the whole buffer
   * ascending by sort key (collide.py:1105-1124, contact_sort.py:194-445).  Null without sort keys: the buffer
   * is in the canonical order itself.  The canonical order is segment by segment (world order, the dedicated
   * world -1 segment last), by sort key inside a segment; a body's contacts are in the same relative order
   * in both. */
  int* order;
  /* [1] or NULL: Contacts.contact_generation, bumped once per nmc_collide (wrapping from INT32_MAX to 0 as
   * compute_shape_aabbs_and_prepare_kernel does, collide.py:181-191) */
  int* generation;
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
  int block_lanes;  /* launch shape: 0 = automatic (one 64-lane wave a world up to 64 bodies; models with
                     * larger worlds launch by size tier, worlds above 64 bodies on 512 lanes), 64 (worlds
                     * of up to 64 bodies), 256 (every world in one launch: 256 lanes up to 64 bodies a
                     * world, else 512) or 512 (plain step: one launch over CSR bins -- runs of consecutive
                     * worlds packed up to max(largest world, 64) bodies, or 256 when a world exceeds 256 (those
                     * worlds then run in a 512-lane launch beside the bins); bins of up to 64 bodies on one wave,
                     * larger ones on 256 lanes, largest first; other modes launch as 0).  The reference-order
                     * arithmetic gives the same result for every shape; the world-frame arithmetic's per-body
                     * sums follow the lanes a body gets, so bins (one lane a body) round like the 512-lane
                     * worlds (an autotuner picks per model) */
  int reference_arithmetic; /* 0 (default): contact rows and body updates in world-frame arithmetic (rotation
                             * matrices, world inverse inertia, FMA, hardware reciprocal), within rounding of
                             * the reference; 1: the reference's operation order, bit-identical to a CPU
                             * sweep compiled without contraction.  Same method, same iteration order. */
  int enable_restitution;   /* SolverXPBD(enable_restitution=True) (solver_xpbd.py:95-109): after the iterations,
                             * apply_rigid_restitution (kernels.py:2472-2615) on the start-of-step poses and
                             * velocities and the solved velocities, per-body deltas summed in contact order,
                             * then added to body_qd_out (kernels.py:882-887).  Needs contacts; in
                             * nmc_xpbd_step_learn every mode but NMC_STEP_PREP applies it. */
} NmcXpbdParams;

/* Learned warm start of the XPBD contact solve (no counterpart in the reference, whose contact solve
 * keeps no per-contact multiplier; added for the warm-started step).  Modes of nmc_xpbd_step_learn:
 *   NMC_STEP_PLAIN  the plain step (== nmc_xpbd_step); `residual` optional
 *   NMC_STEP_LOG    the plain step, also writing `feat` and `lambda_sum` (training data), `residual` optional
 *   NMC_STEP_PREP   joint forces + integration only: `feat` and the integrated state `q_mid`/`qd_mid` out
 *   NMC_STEP_WARM   from `q_mid`/`qd_mid`: apply `lambda0` (clamped at 0) along each contact normal with the
 *                   solver's 1/count weighting, run `warm_iterations` iterations; worlds whose residual
 *                   (largest penetration max(0, -d) over their contacts) exceeds `residual_tol` restart from
 *                   q_mid/qd_mid with the full iteration count.  `lambda_sum` = applied normal multipliers.
 * Per-contact arrays are indexed like the Contacts buffer; `feat` holds 16 floats per contact (see
 * newton_custom_contact_amd/learned.py FEATURES).  prev_* (optional): the previous substep's sort keys,
 * world_start / world_count, lambda_sum and order array, used to fill the "previous multiplier" feature
 * (requires contacts->sort_key). */
enum { NMC_STEP_PLAIN = 0, NMC_STEP_LOG = 1, NMC_STEP_PREP = 2, NMC_STEP_WARM = 3 };
typedef struct NmcLearnIO {
  int mode;
  int warm_iterations;
  float residual_tol;
  float* feat;             /* C x 16 */
  float* lambda_sum;       /* C */
  const float* lambda0;    /* C */
  float* residual;         /* segment_count, optional */
  int* fallback;           /* segment_count, optional */
  float* q_mid;            /* B x 7 */
  float* qd_mid;           /* B x 6 */
  const int64_t* prev_key; /* optional */
  const int* prev_start;
  const int* prev_count;
  const float* prev_lambda;
  const int* prev_order;   /* the previous Contacts' order array (null: canonical layout) */
} NmcLearnIO;

typedef struct NmcModel NmcModel;

/* Status codes: 0 ok, <0 error (see nmc_last_error()). */
int nmc_model_create(const NmcModelDesc* desc, NmcModel** out);
void nmc_model_destroy(NmcModel* model);

/* Shape AABBs -> SAP broad phase -> primitive / MPR-GJK / mesh narrow phase -> contacts.  With sort keys the
 * buffer is in global sort-key order (contacts->order maps canonical positions to it); without, in canonical
 * order (segment-major, ascending by the 64-bit sort key inside a segment). */
int nmc_collide(NmcModel* model, const float* body_q, const NmcContacts* contacts, nmc_stream_t stream);

/* One XPBD step: integrate, `iterations` x (contact solve + apply, joint solve + apply), then the restitution
 * pass when params->enable_restitution is set (a second launch).  Fails when contacts->capacity exceeds the
 * model's rigid_contact_max or when state in/out arrays alias. */
int nmc_xpbd_step(NmcModel* model, const NmcXpbdParams* params, const float* body_q_in, const float* body_qd_in,
                  const float* body_f_in, const float* joint_f, const float* joint_target_pos,
                  const float* joint_target_vel, const NmcContacts* contacts, float dt, float* body_q_out,
                  float* body_qd_out, nmc_stream_t stream);

/* nmc_xpbd_step with the learned warm-start modes above (mode NMC_STEP_PLAIN: identical to
 * nmc_xpbd_step).  No counterpart in the reference. */
int nmc_xpbd_step_learn(NmcModel* model, const NmcXpbdParams* params, const float* body_q_in,
                        const float* body_qd_in, const float* body_f_in, const float* joint_f,
                        const float* joint_target_pos, const float* joint_target_vel, const NmcContacts* contacts,
                        float dt, float* body_q_out, float* body_qd_out, const NmcLearnIO* learn,
                        nmc_stream_t stream);

/* Warm-start network 16 -> 128 -> 128 -> 1 (SiLU), BF16-MFMA layers 1-2 with FP32 accumulation (no
 * counterpart in the reference).  Weights row-major [out][in]; inputs normalized as
 * (feature - x_mean) * x_inv_std; output lambda0 = y * out_scale. */
typedef struct NmcMlp {
  const void* w1_bf16; /* 128 x 16 BF16 */
  const float* b1;     /* 128 */
  const void* w2_bf16; /* 128 x 128 BF16 */
  const float* b2;     /* 128 */
  const float* w3;     /* 128 */
  const float* b3;     /* 1 */
  const float* x_mean;    /* 16 */
  const float* x_inv_std; /* 16 */
  float out_scale;
} NmcMlp;

/* Device scratch of one training batch of n rows (FP32 unless noted). */
typedef struct NmcMlpTrainBuffers {
  float* y;    /* n */
  float* z1;   /* n x 128 */
  float* z2;   /* n x 128 */
  void* xb;    /* n x 16 BF16 */
  void* h1;    /* n x 128 BF16 */
  void* dz2b;  /* n x 128 BF16 */
  void* dz1b;  /* n x 128 BF16 */
  float* dz2f; /* n x 128 */
  float* dz1f; /* n x 128 */
  float* g3;   /* n x 128 */
  float* dy;   /* n */
} NmcMlpTrainBuffers;

/* Gradients (FP32) of 0.5 * mean((y - target)^2) with respect to the FP32 parameters. */
typedef struct NmcMlpGrads {
  float *w1, *b1, *w2, *b2, *w3, *b3;
} NmcMlpGrads;

/* lambda0[r] = network(feat[r]) for r < min(*count, capacity) (count read on the device). */
int nmc_mlp_forward(const NmcMlp* weights, const float* feat, const int* count, int capacity, float* lambda0,
                    nmc_stream_t stream);
/* One training batch: forward, backward, weight and bias gradients (hand-written MFMA kernels). */
int nmc_mlp_train_grad(const NmcMlp* weights, const float* feat, const float* target, int n,
                       const NmcMlpTrainBuffers* buffers, NmcMlpGrads* grads, nmc_stream_t stream);

/* The reference's sparse SDF texture of one triangle mesh (create_texture_sdf_from_mesh /
 * build_sparse_sdf_from_mesh, sdf_texture.py:959-1505), cooked in three calls; the host derives the grid
 * geometry and, between the second and third call, sizes the texture from the count of required subgrids
 * (the reference reads it back at the same point, :1163-1172).  No counterpart as C entry points in the
 * reference (Warp kernels there).
 *
 * nmc_sdf_build: the padded fine grid.  Vertex (ix, iy, iz) at lower + (ix cell[0], iy cell[1], iz cell[2])
 * gets the distance to the nearest triangle, negative inside; inside is decided by the generalized winding
 * number w against winding_threshold (get_distance_to_mesh, sdf_utils.py:640-651): w > threshold for
 * threshold >= 0, w <= threshold for a negative threshold (inward-wound meshes, -0.5 when the signed volume is
 * negative, sdf_utils.py:506-510).  out has vertices_per_axis[2] * [1] * [0] floats, x fastest; the texture
 * cook passes 8 w + 1, 8 h + 1, 8 d + 1 vertices (w, h, d subgrids of 8^3 cells). */
int nmc_sdf_build(const float* vertices, int vertex_count, const int* triangles, int triangle_count,
                  const float lower[3], const float cell[3], const int vertices_per_axis[3], float winding_threshold,
                  float* out, nmc_stream_t stream);
/* nmc_sdf_subgrids: from the fine grid of subgrids[0..2] = (w, h, d) subgrids, the coarse grid (coarse:
 * (d+1)(h+1)(w+1) floats, x fastest) and per subgrid s = (bz*h + by)*w + bx: flags[s] = occupied (the centre
 * value inside the band: below band_hi when not negative, above band_lo otherwise) | linear << 1 (occupied and
 * every sample within linear_threshold of the coarse interpolation; never when linear_threshold <= 0) |
 * required << 2 (occupied and not linear); address[s] = exclusive scan of required; *count = required
 * subgrids (device int). */
int nmc_sdf_subgrids(const float* fine, const int subgrids[3], float band_lo, float band_hi, float linear_threshold,
                     float* coarse, int* flags, int* address, int* count, nmc_stream_t stream);
/* nmc_sdf_pack: slots (w*h*d, z fastest) and the subgrid texture of blocks_per_axis^3 blocks of 9^3 texels
 * (ts = 9 blocks_per_axis texels per axis, x fastest; blocks_per_axis^3 >= the required count, the caller
 * zero-fills texels), texel_bytes 4 (float32), 2 (unorm16) or 1 (unorm8):
 * trunc(clamp((v - quant_min) * quant_range_inv, 0, 1) * (2^bits - 1)). */
int nmc_sdf_pack(const float* fine, const int subgrids[3], const int* flags, const int* address, int blocks_per_axis,
                 int texel_bytes, float quant_min, float quant_range_inv, uint32_t* slots, void* texels,
                 nmc_stream_t stream);

/* Reported forces of the XPBD step (SolverXPBD.update_contacts, newton/_src/solvers/xpbd/solver_xpbd.py:767-821,
 * and State.body_parent_f, :632-655).  contact_impulse (capacity x 6: lin, ang) accumulates, over the
 * iterations, each active contact's shape-0 side delta scaled by the contact weighting: 1/N of its dynamic
 * body, or 2/(N_a + N_b) for two (kernels.py:2301-2361); contact_scratch (capacity x 8) holds one iteration.
 * body_parent_f (body_count x 6) receives each enabled non-FREE joint's accumulated child-side impulse / dt
 * (joint_impulse: joint_count x 6 scratch), zero elsewhere (kernels.py:2397-2435). */
typedef struct NmcXpbdReport {
  float* contact_impulse; /* or null */
  float* contact_scratch;
  float* joint_impulse;
  float* body_parent_f; /* or null */
} NmcXpbdReport;
/* nmc_xpbd_step with the reported impulses. */
int nmc_xpbd_step_report(NmcModel* model, const NmcXpbdParams* params, const float* body_q_in,
                         const float* body_qd_in, const float* body_f_in, const float* joint_f,
                         const float* joint_target_pos, const float* joint_target_vel, const NmcContacts* contacts,
                         float dt, float* body_q_out, float* body_qd_out, const NmcXpbdReport* report,
                         nmc_stream_t stream);
/* convert_contact_impulse_to_force (kernels.py:2364-2394): force = impulse / dt for the first *count
 * contacts, zero for the rest of the capacity (contacts.force, 6 floats per contact). */
int nmc_contact_force(const float* contact_impulse, const int* count, int capacity, float dt, float* force,
                      nmc_stream_t stream);

/* Frame-to-frame contact matching (CollisionPipeline(contact_matching="latest" | "sticky", contact_report),
 * newton/_src/geometry/contact_match.py:570-992 ContactMatcher; collide.py:1084-1184).  The matcher keeps the
 * previous frame's contacts (keys, world midpoints, normals; body-frame points/offsets when sticky).
 * nmc_contact_match runs after nmc_collide on the same Contacts buffer (sort keys required): it writes
 * match_index (>= 0: index of the matched contact in the previous frame's buffer; -1 MATCH_NOT_FOUND: the
 * shape pair had no contact; -2 MATCH_BROKEN: none within pos_threshold / normal_dot_threshold, or a closer
 * new contact won it), replays the matched rows' previous point0/1, offset0/1 and normal when sticky, fills
 * the report when the matcher has one, and saves this frame for the next. */
typedef struct NmcMatcher NmcMatcher;
typedef struct NmcContactReport {
  int* new_indices;    /* capacity: buffer indices of this frame's contacts with match_index < 0 (canonical order;
                        * the reference appends through an atomic counter, in no fixed order) */
  int* new_count;      /* [1] */
  int* broken_indices; /* capacity: previous-frame buffer indices no new contact matched (its canonical order) */
  int* broken_count;   /* [1] */
} NmcContactReport;
int nmc_matcher_create(const NmcModel* model, int capacity, float pos_threshold, float normal_dot_threshold,
                       int sticky, int report, NmcMatcher** out);
void nmc_matcher_destroy(NmcMatcher* matcher);
/* ContactMatcher.reset (contact_match.py:745-758): the next frame matches nothing. */
int nmc_matcher_reset(NmcMatcher* matcher, nmc_stream_t stream);
int nmc_contact_match(NmcMatcher* matcher, const NmcModel* model, const float* body_q, const NmcContacts* contacts,
                      int* match_index, const NmcContactReport* report, nmc_stream_t stream);

/* Standalone broad phases over caller AABBs (newton.geometry BroadPhaseSAP, BroadPhaseAllPairs,
 * BroadPhaseExplicit; newton/_src/geometry/broad_phase_sap.py:397-677, broad_phase_nxn.py:217-472).
 * nmc_broad_phase_create replaces the constructors: mode 0 SAP or 1 all pairs take the host index map and
 * slice ends of precompute_world_map (broad_phase_common.py:183-300; the last slice is the dedicated
 * world -1 segment); mode 2 explicit takes the largest pair count it will be launched with.  Scratch is
 * allocated here, never per launch. */
typedef struct NmcBroadPhase NmcBroadPhase;
int nmc_broad_phase_create(int mode, const int* index_map, int index_count, const int* slice_ends, int segment_count,
                           int explicit_pair_capacity, NmcBroadPhase** out);
void nmc_broad_phase_destroy(NmcBroadPhase* broad_phase);
/* BroadPhaseSAP.launch / BroadPhaseAllPairs.launch (broad_phase_sap.py:510-677, broad_phase_nxn.py:305-391):
 * lower/upper S x 3, gap S or null (AABBs already expanded), group and world S, filter_pairs F x 2 sorted
 * (min, max) or null.  Writes candidate_pair (max x 2) and candidate_pair_count[0] = pairs found (may exceed
 * max; the pairs past it are dropped).  Deterministic order: segment by segment; all pairs by (row, column)
 * in index-map order, SAP by sweep origin in sorted order.  append = the reference's skip_count_zero
 * (broad_phase_sap.py:562-563, broad_phase_nxn.py:353-354): 0 counts from zero; non-zero keeps the
 * candidate_pair_count[0] already on the device, writes the new pairs after that many and adds their number
 * to it (a negative count on entry counts as 0). */
int nmc_broad_phase_launch(NmcBroadPhase* broad_phase, int shape_count, const float* lower, const float* upper,
                           const float* gap, const int* group, const int* world, const int* filter_pairs,
                           int filter_count, int* candidate_pair, int max_candidate_pair, int* candidate_pair_count,
                           int append, nmc_stream_t stream);
/* BroadPhaseExplicit.launch (broad_phase_nxn.py:408-472): the given pairs whose AABBs overlap, in input
 * order; a pair naming a shape outside [0, shape_count) is skipped.  append: skip_count_zero
 * (broad_phase_nxn.py:448-449), as for nmc_broad_phase_launch. */
int nmc_broad_phase_explicit(NmcBroadPhase* broad_phase, int shape_count, const float* lower, const float* upper,
                             const float* gap, const int* shape_pairs, int shape_pair_count, int* candidate_pair,
                             int max_candidate_pair, int* candidate_pair_count, int append, nmc_stream_t stream);
/* nmc_broad_phase_explicit without a broad-phase object and without allocating anything: no counterpart in the
 * reference; for callers that cannot call hipMalloc (code under graph capture, tight serving loops, harnesses
 * that forbid it).  Same arguments and results as nmc_broad_phase_explicit.  Its per-block scratch is one
 * static device array per GPU (128 KiB, held until the process ends), so one call takes at most 16 777 216
 * input pairs (more is refused with an error before any launch), and calls that overlap on one GPU race
 * (each call rewrites the whole part of the scratch it reads; nothing carries over between calls). */
int nmc_broad_phase_explicit_noalloc(int shape_count, const float* lower, const float* upper, const float* gap,
                                     const int* shape_pairs, int shape_pair_count, int* candidate_pair,
                                     int max_candidate_pair, int* candidate_pair_count, int append,
                                     nmc_stream_t stream);

/* Launch shape of nmc_collide's kernels, picked by the package's autotuner (no counterpart in the reference,
 * whose Warp launches have fixed shapes): segment_waves = segments of at most 64 shapes per workgroup of the
 * broad phase (1, 2 or 4, one wave each); narrow_max_blocks / mesh_max_blocks = grid caps of the grid-stride
 * narrow-phase and mesh kernels (256 lanes a workgroup).  0 keeps the default (4, 2048, 2048).  Every shape
 * gives bit-identical contacts; the setting holds for later nmc_collide calls on the model. */
typedef struct NmcLaunchConfig {
  int segment_waves;
  int narrow_max_blocks;
  int mesh_max_blocks;
} NmcLaunchConfig;
int nmc_model_set_launch(NmcModel* model, const NmcLaunchConfig* config);

/* Copies the counters accumulated since model creation to host (synchronises the stream):
 * out[0] = broad-phase pairs dropped (per-segment pair buffer full), out[1] = mesh-pair candidate
 * contacts dropped (candidate buffer full; cannot happen for models nmc_model_create accepted),
 * out[2] = contacts beyond the Contacts capacity, out[3] = worlds whose contacts did not fit the
 * solver's ordering buffer (cannot happen when contacts->capacity <= rigid_contact_max). */
int nmc_get_counters(NmcModel* model, int* out, nmc_stream_t stream);

const char* nmc_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
