// Internal declarations shared by the HIP translation units of libnmc.so.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/nmc.h"

namespace nmc {

// GeoType values (newton/_src/geometry/types.py:38-78)
enum GeoTypeId { GEO_PLANE = 1, GEO_SPHERE = 3, GEO_CAPSULE = 4, GEO_BOX = 7 };
// ShapeFlags.COLLIDE_SHAPES (newton/_src/geometry/flags.py:25), BodyFlags.KINEMATIC (sim/enums.py:21)
enum { SHAPE_COLLIDE_SHAPES = 2, BODY_KINEMATIC = 2 };
// JointType (newton/_src/sim/enums.py:29-56)
enum { JT_PRISMATIC = 0, JT_REVOLUTE = 1, JT_BALL = 2, JT_FIXED = 3, JT_FREE = 4, JT_DISTANCE = 5, JT_D6 = 6,
       JT_CABLE = 7 };

constexpr float MAXVAL = 1.0e10f;  // newton/_src/core/types.py:72

// Staged contact records produced by the narrow phase (5 slots per broad-phase pair).
struct StageBuffers {
  int* shape0;
  int* shape1;
  float* point0;
  float* point1;
  float* offset0;
  float* offset1;
  float* normal;
  float* margin0;
  float* margin1;
  int* sub_key;
  unsigned char* mask;  // per pair: number of records in its slots (0-5)
  long long slot_capacity;
  // broad-phase output, one entry per pair (slot_capacity / 5 entries)
  int* pair_a;      // first shape (lower GeoType), global id
  int* pair_b;
  int* pair_seg;    // segment the pair belongs to
  int* route;       // pair indices: analytic pairs from the front, MPR/GJK pairs from the back
  int* route_count; // [2] analytic, MPR/GJK
  float* shape_wc;  // [shape_count x 20] world transform 7, scale 3, margin 1, AABB lo 3 / hi 3, pad 3
};

struct NmcModelImpl {
  NmcModelDesc d;
  int device;
  // collide scratch
  StageBuffers stage;
  int* seg_pair_count;   // [segments] pairs found (may exceed the per-segment buffer)
  int* seg_pair_base;    // [segments] first staging pair of the segment
  int* seg_contact_count;  // [segments] valid staged contacts
  int* stage_alloc;      // [1] staging pairs handed out during the current collide
  int* counters;         // [4] overflow counters (see nmc_get_counters)
  int pair_cap;          // per-segment pair buffer (LDS) capacity
  size_t collide_lds;
  size_t xpbd_lds;
  int xpbd_contact_cap;  // (contact, side) entries a world orders in LDS
  unsigned* xpbd_keys;   // [xpbd_keys_cap] ordering entries of denser worlds, at 2 x their contact offset
  long long xpbd_keys_cap;
};

void set_error(const std::string& msg);

}  // namespace nmc

struct NmcModel : nmc::NmcModelImpl {};
