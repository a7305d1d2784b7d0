// Internal declarations shared by the HIP translation units of libnmc.so.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/nmc.h"

namespace nmc {

// GeoType values (newton/_src/geometry/types.py:38-78)
enum GeoTypeId { GEO_PLANE = 1, GEO_SPHERE = 3, GEO_CAPSULE = 4, GEO_ELLIPSOID = 5, GEO_CYLINDER = 6, GEO_BOX = 7,
                 GEO_MESH = 8, GEO_CONE = 9 };
// ShapeFlags.COLLIDE_SHAPES (newton/_src/geometry/flags.py:25), BodyFlags.KINEMATIC (sim/enums.py:21)
enum { SHAPE_COLLIDE_SHAPES = 2, BODY_KINEMATIC = 2 };
// JointType (newton/_src/sim/enums.py:29-56)
enum { JT_PRISMATIC = 0, JT_REVOLUTE = 1, JT_BALL = 2, JT_FIXED = 3, JT_FREE = 4, JT_DISTANCE = 5, JT_D6 = 6,
       JT_CABLE = 7 };

constexpr float MAXVAL = 1.0e10f;  // newton/_src/core/types.py:72

// Broad-phase pairs and their staged contact records.  Every segment owns a fixed region of pair
// entries (its worst case n(n-1)/2, capped by the LDS pair buffer) at seg_pair_off[seg]; pair g owns
// contact slots [5g, 5g + 5), except mesh pairs: the i-th mesh pair of the call (route order) owns
// slots [mesh_slot_base + i K, mesh_slot_base + (i + 1) K), K = mesh_k (240 reduced, the largest candidate count
// of a pair unreduced).  Fixed regions need
// no global allocation counter.
// Mesh contact reduction parameters (reduce.h): the 6 scan directions (cos, sin) of angle k (2 pi / 6) taken in
// float32 (get_spatial_direction_2d, contact_reduction.py:334-345), correctly rounded and filled on the host,
// and the value packing: 0 no reduction, 1 fast (full score, sub key), 2 deterministic (22-bit score,
// 22-bit fingerprint = sub key, 20-bit id = sub key).
struct RedDirs {
  float c[6], s[6];
};
struct RedParams {
  RedDirs dirs;
  int mode;
};

constexpr int REC_F4 = 4;  // float4 per staging slot

struct StageBuffers {
  // one 48-byte record per contact slot: world-frame point on A 3, on B 3, unit normal 3, offset A,
  // offset B (radius + margin), sub_key (int bits) -- written and read as three 16-byte vectors at a
  // 64-byte stride (REC_F4 float4 per slot), so that a record fills one 64-byte memory sector
  float4* rec;
  unsigned short* mask;  // per pair: number of records in its slots (0-5, mesh pairs 0-K)
  int* pair_a;           // first shape (lower GeoType), global id
  int* pair_b;
  long long pair_capacity;  // total pair entries (sum of the segment regions)
  int* route;            // per segment region: local pair indices, MPR/GJK pairs, analytic, then mesh
  int* route_count;      // [3] analytic, MPR/GJK and mesh pairs over all segments
  int* mesh_index;       // per pair entry: mesh pair index (route order), written for mesh pairs only
  long long mesh_slot_base;  // first mesh slot (= 5 x pair_capacity)
  long long mesh_pair_capacity;  // mesh pairs the mesh slots hold (worst case over the segment regions)
  float* shape_wc;       // [shape_count x 20] world transform 7, scale 3, margin 1, AABB lo 3 / hi 3, pad 3
};

struct NmcModelImpl {
  NmcModelDesc d;
  int device;
  // collide launch shape (nmc_model_set_launch): segments (one wave each) per workgroup of the small-segment
  // broad phase, grid caps of the grid-stride narrow and mesh kernels
  int segment_waves = 4;
  bool all_types = false;  // the model holds ellipsoids, cylinders or cones (narrow phase with their colliders)
  int narrow_max_blocks = 2048;
  int mesh_max_blocks = 2048;
  // collide scratch
  StageBuffers stage;
  long long* seg_pair_off;  // [segments] first pair entry of each segment's region
  int* seg_pair_cap;        // [segments] size of the region
  int* seg_pair_count;      // [segments] pairs found this call (after the LDS cap)
  int* seg_conv_count;      // [segments] of which MPR/GJK pairs
  int* seg_conv_off;        // [segments] exclusive prefix of MPR/GJK pairs over segments
  int* seg_ana_off;         // [segments] exclusive prefix of analytic pairs over segments
  int* seg_mesh_count;      // [segments] mesh pairs (mesh-mesh, mesh-infinite plane)
  int* seg_mesh_off;        // [segments] exclusive prefix of mesh pairs over segments
  int mesh_cand_cap;        // candidate contacts one mesh pair gathers in LDS (power of two)
  int mesh_k;               // staging slots per mesh pair
  RedParams red;            // reduction of mesh-pair contacts
  size_t mesh_lds;
  int* seg_contact_count;   // [segments] contacts produced this call
  int* seg_small;           // segments of at most 64 shapes (one wave each), and the others (one workgroup)
  int* seg_big;
  int n_seg_small, n_seg_big;
  int seg_small_max_n;
  size_t seg_small_lds;     // dynamic LDS of one wave
  // global sort-key layout of deterministic Contacts buffers (collide.hip: order_count / order_scan / compaction).
  // Units are the W worlds' shape blocks and the G world -1 shapes, ordered by shape index: a contact's
  // buffer position is the contacts of lower shape0 (unit block bases, then the shape0's rank inside its
  // world block), then its rank among the contacts of its shape0 (for a world -1 shape0: by unit of shape1).
  int ord_G;              // world -1 shapes (-1: the layout is unavailable for this model)
  int ord_U;              // W + G units
  int* ord_gid;           // [S] index among the world -1 shapes, -1 for world shapes
  int* ord_unit_order;    // [U] unit ids (world w = w, world -1 shape g = W + g) in shape-index order
  int* ord_delta0;        // [S] world shape0: minus the world -1 shape0 contacts before its run in its segment
  int* ord_M;             // [segments x G] contacts of each world -1 shape0 in each segment
  int* ord_gstart;        // [segments x G] first segment-local contact of each world -1 shape0
  int* ord_L;             // [segments] contacts whose shape0 is a world shape
  int* ord_H;             // [G x G] last segment: contacts of each (world -1 shape0, world -1 shape1) pair
  int* ord_unit_off;      // [G x U] rank base of each unit among the contacts of one world -1 shape0
  int* ord_base;          // [U] first buffer position of each unit's shape0 block
  int* counters;         // [4] overflow counters (see nmc_get_counters)
  int pair_cap;          // per-segment pair buffer (LDS) capacity
  size_t collide_lds;
  size_t xpbd_lds;
  int xpbd_contact_cap;  // (contact, side) entries a world orders in LDS
  // worlds by size tier (<= 64, <= 256, <= 512 bodies) for the automatic launch shape; xpbd_tiered when the
  // model has worlds above 64 bodies and worlds of another tier (else one launch covers all)
  int* xpbd_tier_worlds[3];
  int xpbd_tier_count[3];
  int xpbd_tier_max[3];
  bool xpbd_tiered;
  // the tiers' launches run concurrently: tier t > 0 on tier_stream[t - 1], forked from and joined into the
  // caller's stream through events (graph-capturable); created with the model when it is tiered
  hipStream_t tier_stream[2];
  hipEvent_t tier_fork, tier_join[2];
  // CSR bins for the binned step (block_lanes 512): runs of consecutive worlds with adjacent body and joint
  // ranges, packed while their bodies fit max(largest world, 64) or 256 when a world exceeds 256 bodies;
  // xpbd_bins holds (first world, end world) per bin, largest bins first (the launch order)
  int* xpbd_bins;
  int xpbd_bin_count;
  int xpbd_bin_max_bodies;
  int xpbd_bin_max_joints;
  // worlds above 256 bodies stay out of the bins (which then launch on 256 lanes) and run in a launch of their
  // own beside them: their indices, count and largest body count
  int* xpbd_big;
  int xpbd_big_count;
  int xpbd_big_max;
  int xpbd_joint_lds;    // joints per world with LDS delta slots: max_world_joints when some enabled joint is
                         // not FREE / DISTANCE (the only joints the joint pass solves), else 0
  unsigned* xpbd_keys;   // [xpbd_keys_cap] ordering entries of denser worlds, at 2 x their contact offset
  long long xpbd_keys_cap;
};

// Explicit broad phase (broad.hip): argument checks, and the count / scan / write launches over caller-given
// per-block scratch (count and off: `blocks` ints each, total: 1 int).
int bp_explicit_check(int shape_count, const float* lower, const float* upper, const int* shape_pairs,
                      int shape_pair_count, const int* candidate_pair, int max_candidate_pair,
                      const int* candidate_pair_count, const char* what);
int bp_explicit_run(int shape_count, const float* lower, const float* upper, const float* gap,
                    const int* shape_pairs, int shape_pair_count, int blocks, int* count, int* off, int* total,
                    int* candidate_pair, int max_candidate_pair, int* candidate_pair_count, int append,
                    hipStream_t s, const char* what);

// Buffer position of the contact at canonical (segment-major, key-sorted per segment) position i: the identity
// for the segment-major layout, Contacts.order for the global key-order layout; -1 when not stored.
__device__ __forceinline__ int cpos(const NmcContacts& c, int i) { return c.order ? c.order[i] : i; }

void set_error(const std::string& msg);
int fail(const std::string& msg);               // sets the error string, returns -1
int hip_fail(hipError_t e, const char* what);  // same, with the HIP error text

}  // namespace nmc

struct NmcModel : nmc::NmcModelImpl {};
