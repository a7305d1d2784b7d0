// extern "C" entry points of libnmc.so (declared in include/nmc.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <algorithm>
#include <array>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "nmc_internal.h"

namespace nmc {

size_t collide_lds_enum_bytes(int max_n);
size_t mesh_lds_bytes(int cand_cap, bool reduce);
size_t xpbd_lds_bytes(int max_bodies, int max_joints, int key_cap);
hipError_t launch_collide(NmcModelImpl* m, const float* body_q, const NmcContacts* c, hipStream_t stream);
hipError_t launch_xpbd(NmcModelImpl* m, const NmcXpbdParams* params, const float* q_in, const float* qd_in,
                       const float* f_in, const float* joint_f, const float* tpos, const float* tvel,
                       const NmcContacts* c, float dt, float* q_out, float* qd_out, const NmcLearnIO* lrn,
                       const NmcXpbdReport* rep, hipStream_t stream);
hipError_t launch_restitution(NmcModelImpl* m, const float* q_prev, const float* qd_prev, const NmcContacts* c,
                              float dt, float* qd, hipStream_t stream);

static thread_local std::string g_error;
void set_error(const std::string& msg) { g_error = msg; }

int fail(const std::string& msg) {
  set_error(msg);
  return -1;
}

int hip_fail(hipError_t e, const char* what) {
  return fail(std::string(what) + ": " + hipGetErrorString(e));
}

template <class T>
static hipError_t dmalloc(T** p, size_t count) {
  size_t bytes = sizeof(T) * (count > 0 ? count : 1);
  hipError_t e = hipMalloc(reinterpret_cast<void**>(p), bytes);
  if (e == hipSuccess) e = hipMemset(*p, 0, bytes);
  return e;
}

static bool valid_lanes(const NmcModelImpl* m, const NmcXpbdParams* p) {
  return p->block_lanes == 0 || p->block_lanes == 256 || (p->block_lanes == 64 && m->d.max_world_bodies <= 64) ||
         p->block_lanes == 512;
}

static int next_pow2(int n) {
  int p = 1;
  while (p < n) p <<= 1;
  return p;
}

// Tables of the global sort-key layout (collide.hip order_*).  The layout needs every world's shapes to form one
// contiguous index block (world -1 shapes lie between blocks) and at most 256 world -1 shapes; otherwise
// ord_G stays -1 and a collide that asks for the layout fails.
static hipError_t order_setup(NmcModelImpl* m) {
  const NmcModelDesc& d = m->d;
  const int S = d.shape_count, W = d.world_count, segs = d.segment_count;
  std::vector<int> sw(S > 0 ? S : 1);
  hipError_t e = hipSuccess;
  if (S > 0 && (e = hipMemcpy(sw.data(), d.shape_world, sizeof(int) * S, hipMemcpyDeviceToHost)) != hipSuccess) return e;
  std::vector<int> gid(S > 0 ? S : 1, -1), gshape;
  std::vector<int> lo(W > 0 ? W : 1, 0x7fffffff), hi(W > 0 ? W : 1, -1), n(W > 0 ? W : 1, 0);
  for (int s = 0; s < S; ++s) {
    if (sw[s] < 0) {
      gid[s] = static_cast<int>(gshape.size());
      gshape.push_back(s);
    } else if (sw[s] < W) {
      lo[sw[s]] = std::min(lo[sw[s]], s);
      hi[sw[s]] = std::max(hi[sw[s]], s);
      n[sw[s]] += 1;
    }
  }
  const int G = static_cast<int>(gshape.size());
  for (int w = 0; w < W; ++w)
    if (n[w] > 0 && hi[w] - lo[w] + 1 != n[w]) return hipSuccess;  // a world's shapes are not one block
  if (G > 256) return hipSuccess;
  // units by position: world w at its first shape, world -1 shape g at its index
  const int U = W + G;
  std::vector<std::pair<long long, int>> pos;
  pos.reserve(U);
  for (int w = 0; w < W; ++w) pos.push_back({n[w] > 0 ? lo[w] : (1ll << 40) + w, w});
  for (int g = 0; g < G; ++g) pos.push_back({gshape[g], W + g});
  std::stable_sort(pos.begin(), pos.end());
  std::vector<int> unit_order(U > 0 ? U : 1);
  for (int u = 0; u < U; ++u) unit_order[u] = pos[u].second;
  const size_t G1 = G > 0 ? G : 1;
  if ((e = dmalloc(&m->ord_gid, S)) != hipSuccess || (e = dmalloc(&m->ord_unit_order, U)) != hipSuccess ||
      (e = dmalloc(&m->ord_delta0, S)) != hipSuccess || (e = dmalloc(&m->ord_M, segs * G1)) != hipSuccess ||
      (e = dmalloc(&m->ord_gstart, segs * G1)) != hipSuccess || (e = dmalloc(&m->ord_L, segs)) != hipSuccess ||
      (e = dmalloc(&m->ord_H, G1 * G1)) != hipSuccess || (e = dmalloc(&m->ord_unit_off, G1 * (U > 0 ? U : 1))) != hipSuccess ||
      (e = dmalloc(&m->ord_base, U)) != hipSuccess)
    return e;
  if ((S > 0 && (e = hipMemcpy(m->ord_gid, gid.data(), sizeof(int) * S, hipMemcpyHostToDevice)) != hipSuccess) ||
      (U > 0 && (e = hipMemcpy(m->ord_unit_order, unit_order.data(), sizeof(int) * U, hipMemcpyHostToDevice)) != hipSuccess))
    return e;
  (void)segs;
  m->ord_G = G;
  m->ord_U = U;
  return hipSuccess;
}

}  // namespace nmc

using namespace nmc;

extern "C" const char* nmc_last_error(void) { return g_error.c_str(); }

extern "C" int nmc_model_create(const NmcModelDesc* desc, NmcModel** out) {
  if (!desc || !out) return fail("nmc_model_create: null argument");
  const NmcModelDesc& d = *desc;
  if (d.segment_count != d.world_count + 1) return fail("segment_count must equal world_count + 1");
  if (d.max_world_bodies > 512) return fail("worlds with more than 512 bodies are not supported by the XPBD kernel");
  if (d.max_segment_shapes > 65535) return fail("a broad-phase segment may hold at most 65535 shapes");
  if (d.mesh_count < 0) return fail("mesh_count must be >= 0");
  if (d.broad_phase_mode < 0 || d.broad_phase_mode > 2) return fail("broad_phase_mode must be 0 (sap), 1 (nxn) or 2 (explicit)");
  if (d.broad_phase_mode == 2 && (d.explicit_pair_count < 0 || (d.explicit_pair_count > 0 && !d.explicit_pairs)))
    return fail("an explicit broad phase needs its pair list");
  if (d.mesh_count > 0) {
    if (!d.shape_mesh || !d.shape_aabb_lower || !d.shape_aabb_upper || !d.mesh_vertex || !d.mesh_vertex_start ||
        !d.mesh_edge || !d.mesh_edge_start || !d.mesh_sdf_coarse || !d.mesh_sdf_slots || !d.mesh_sdf_texels ||
        !d.mesh_sdf_start || !d.mesh_sdf_box || !d.mesh_sdf_quant || !d.mesh_sdf_dims || !d.mesh_sdf_flags)
      return fail("a model with meshes needs every mesh table (vertices, edges, local AABBs, SDF textures)");
    if (d.mesh_reduce < 0 || d.mesh_reduce > 2) return fail("mesh_reduce must be 0, 1 or 2");
    if (!d.shape_voxel_resolution) return fail("a model with meshes needs shape_voxel_resolution");
  }
  NmcModel* m = new NmcModel();
  m->d = d;
  hipError_t e = hipGetDevice(&m->device);
  if (e != hipSuccess) { delete m; return hip_fail(e, "hipGetDevice"); }
  int segs = d.segment_count;
  // mesh pairs: candidates per pair are at most both meshes' edges (mesh-mesh) or the mesh's vertices
  // (mesh-plane); staging slots per pair: the reduction's 240, or every candidate unreduced
  m->mesh_k = 0;
  m->mesh_cand_cap = 0;
  m->red.mode = d.mesh_count > 0 ? d.mesh_reduce : 0;
  {
    // get_spatial_direction_2d: angle k (2 pi / 6) in float32, cosine and sine correctly rounded to float32
    const float step = (2.0f * 3.14159265358979323846f) / 6.0f;
    for (int k = 0; k < 6; ++k) {
      const float a = static_cast<float>(k) * step;
      m->red.dirs.c[k] = static_cast<float>(std::cos(static_cast<double>(a)));
      m->red.dirs.s[k] = static_cast<float>(std::sin(static_cast<double>(a)));
    }
  }
  if (d.mesh_count > 0) {
    std::vector<int> vstart(d.mesh_count + 1), estart(d.mesh_count + 1);
    if ((e = hipMemcpy(vstart.data(), d.mesh_vertex_start, sizeof(int) * (d.mesh_count + 1), hipMemcpyDeviceToHost)) != hipSuccess ||
        (e = hipMemcpy(estart.data(), d.mesh_edge_start, sizeof(int) * (d.mesh_count + 1), hipMemcpyDeviceToHost)) != hipSuccess) {
      delete m;
      return hip_fail(e, "nmc_model_create: mesh tables");
    }
    int need = 1;
    for (int k = 0; k < d.mesh_count; ++k) {
      const int nv = vstart[k + 1] - vstart[k], ne = estart[k + 1] - estart[k];
      need = std::max(need, std::max(nv, 2 * ne));
    }
    if (need > 4096) {
      delete m;
      return fail("mesh narrow phase: a pair may gather " + std::to_string(need) +
                  " candidate contacts (2 x edges or vertices of the largest mesh); the limit is 4096");
    }
    m->mesh_cand_cap = next_pow2(need);
    m->mesh_k = d.mesh_reduce ? 240 : need;
    m->mesh_lds = mesh_lds_bytes(m->mesh_cand_cap, d.mesh_reduce != 0);
  }
  // per-segment pair region: the segment's worst case n(n-1)/2, capped at 16384 pairs (a pair entry stages 5
  // contact records of 48 B)
  long long maxn = d.max_segment_shapes;
  long long worst = maxn * (maxn - 1) / 2;
  int pair_cap = static_cast<int>(worst < 16384 ? worst : 16384);
  if (pair_cap < 1) pair_cap = 1;
  m->pair_cap = next_pow2(pair_cap);
  m->collide_lds = collide_lds_enum_bytes(d.max_segment_shapes > 0 ? d.max_segment_shapes : 1);
  // solver ordering buffer: 2 entries per contact; worlds with up to 256 contacts order them in LDS,
  // denser worlds in a global buffer of 2 entries per contact slot.  Models with meshes (up to
  // mesh_k contacts per pair, few bodies per world) size the LDS buffer for their worst world,
  // up to 8192 entries, so that each body's contact list is walked out of LDS.
  m->xpbd_contact_cap = 512;
  if (d.shape_count > 0) {
    // ellipsoids, cylinders or cones present: the narrow phase compiled with their colliders and support maps
    std::vector<int> types(d.shape_count);
    if ((e = hipMemcpy(types.data(), d.shape_type, sizeof(int) * d.shape_count, hipMemcpyDeviceToHost)) != hipSuccess) {
      delete m;
      return hip_fail(e, "nmc_model_create: shape types");
    }
    for (int t : types) m->all_types = m->all_types || t == GEO_ELLIPSOID || t == GEO_CYLINDER || t == GEO_CONE;
  }
  if (d.mesh_count > 0 && segs > 1) {
    std::vector<int> starts(segs + 1), seg_shape, stype(d.shape_count > 0 ? d.shape_count : 1);
    if ((e = hipMemcpy(starts.data(), d.segment_start, sizeof(int) * (segs + 1), hipMemcpyDeviceToHost)) != hipSuccess) {
      delete m;
      return hip_fail(e, "nmc_model_create: segment_start");
    }
    seg_shape.resize(starts[segs] > 0 ? starts[segs] : 1);
    if ((starts[segs] > 0 &&
         (e = hipMemcpy(seg_shape.data(), d.segment_shape, sizeof(int) * starts[segs], hipMemcpyDeviceToHost)) != hipSuccess) ||
        (d.shape_count > 0 &&
         (e = hipMemcpy(stype.data(), d.shape_type, sizeof(int) * d.shape_count, hipMemcpyDeviceToHost)) != hipSuccess)) {
      delete m;
      return hip_fail(e, "nmc_model_create: shape tables");
    }
    long long worst_world = 0;
    for (int sg = 0; sg + 1 < segs; ++sg) {
      long long n = starts[sg + 1] - starts[sg], mm = 0;
      for (int k = starts[sg]; k < starts[sg + 1]; ++k) mm += stype[seg_shape[k]] == GEO_MESH;
      long long all = n * (n - 1) / 2, mesh_pairs = all - (n - mm) * (n - mm - 1) / 2;
      long long c = 5 * (all - mesh_pairs) + (long long)m->mesh_k * mesh_pairs;
      worst_world = c > worst_world ? c : worst_world;
    }
    long long want = (2 * worst_world + 63) / 64 * 64;
    m->xpbd_contact_cap = static_cast<int>(want < 512 ? 512 : (want > 8192 ? 8192 : want));
  }
  m->xpbd_keys_cap = 2ll * (d.rigid_contact_max > 0 ? d.rigid_contact_max : 1);
  m->xpbd_joint_lds = 0;
  if (d.joint_count > 0) {
    std::vector<int> jt(d.joint_count), je(d.joint_count);
    if ((e = hipMemcpy(jt.data(), d.joint_type, sizeof(int) * d.joint_count, hipMemcpyDeviceToHost)) != hipSuccess ||
        (e = hipMemcpy(je.data(), d.joint_enabled, sizeof(int) * d.joint_count, hipMemcpyDeviceToHost)) != hipSuccess) {
      delete m;
      return hip_fail(e, "nmc_model_create: joint tables");
    }
    for (int j = 0; j < d.joint_count; ++j)
      if (je[j] && jt[j] != JT_FREE && jt[j] != JT_DISTANCE) m->xpbd_joint_lds = d.max_world_joints > 0 ? d.max_world_joints : 1;
  }
  m->xpbd_lds = xpbd_lds_bytes(d.max_world_bodies, m->xpbd_joint_lds, m->xpbd_contact_cap);
  // world size tiers for the step kernel's automatic launch shape
  if (d.max_world_bodies > 64 && d.world_count >= 0) {
    const int W1 = d.world_count + 1;
    std::vector<int> nb(W1);
    if ((e = hipMemcpy(nb.data(), d.world_body_count, sizeof(int) * W1, hipMemcpyDeviceToHost)) != hipSuccess) {
      delete m;
      return hip_fail(e, "nmc_model_create: world_body_count");
    }
    std::vector<int> tier[3];
    for (int w = 0; w < W1; ++w) {
      const int t = nb[w] <= 64 ? 0 : (nb[w] <= 256 ? 1 : 2);
      tier[t].push_back(w);
      m->xpbd_tier_max[t] = std::max(m->xpbd_tier_max[t], std::max(nb[w], 1));
    }
    int used = 0;
    for (int t = 0; t < 3; ++t) used += tier[t].empty() ? 0 : 1;
    m->xpbd_tiered = used > 1;
    if (m->xpbd_tiered) {
      // side streams (default priority) and fork / join events for the concurrent tier launches
      e = hipEventCreateWithFlags(&m->tier_fork, hipEventDisableTiming);
      for (int k = 0; k < 2 && e == hipSuccess; ++k) {
        e = hipStreamCreateWithFlags(&m->tier_stream[k], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&m->tier_join[k], hipEventDisableTiming);
      }
      if (e != hipSuccess) {
        nmc_model_destroy(m);
        return hip_fail(e, "nmc_model_create: tier streams");
      }
      for (int t = 0; t < 3; ++t) {
        m->xpbd_tier_count[t] = static_cast<int>(tier[t].size());
        if (tier[t].empty()) continue;
        if ((e = dmalloc(&m->xpbd_tier_worlds[t], tier[t].size())) != hipSuccess ||
            (e = hipMemcpy(m->xpbd_tier_worlds[t], tier[t].data(), sizeof(int) * tier[t].size(),
                           hipMemcpyHostToDevice)) != hipSuccess) {
          for (int k = 0; k < 3; ++k) (void)hipFree(m->xpbd_tier_worlds[k]);
          delete m;
          return hip_fail(e, "nmc_model_create: world tiers");
        }
      }
    }
  }
  // CSR bins: consecutive worlds whose body and joint ranges follow each other, packed greedily while the
  // bin's bodies stay within max(largest world, 64) (so one LDS layout holds every bin); largest bins first.
  // With worlds above 256 bodies the bins stop at 256 (one 256-lane workgroup each, two resident per CU) and the
  // larger worlds run in a 512-lane launch of their own beside them.  Kept only when packing merges worlds.
  if (d.world_count >= 1) {
    const int W1 = d.world_count + 1;
    std::vector<int> nb(W1), bf(W1), jf(W1), jc(W1);
    if ((e = hipMemcpy(nb.data(), d.world_body_count, sizeof(int) * W1, hipMemcpyDeviceToHost)) != hipSuccess ||
        (e = hipMemcpy(bf.data(), d.world_body_first, sizeof(int) * W1, hipMemcpyDeviceToHost)) != hipSuccess ||
        (e = hipMemcpy(jf.data(), d.world_joint_first, sizeof(int) * W1, hipMemcpyDeviceToHost)) != hipSuccess ||
        (e = hipMemcpy(jc.data(), d.world_joint_count, sizeof(int) * W1, hipMemcpyDeviceToHost)) != hipSuccess) {
      nmc_model_destroy(m);
      return hip_fail(e, "nmc_model_create: world ranges");
    }
    const int cap = d.max_world_bodies <= 256 ? std::max(d.max_world_bodies, 64) : 256;
    std::vector<std::array<int, 4>> bins;  // first, end, bodies, joints
    std::vector<int> big;
    for (int w = 0; w < W1;) {
      if (nb[w] > cap) {
        big.push_back(w);
        m->xpbd_big_max = std::max(m->xpbd_big_max, nb[w]);
        ++w;
        continue;
      }
      int e1 = w + 1, bodies = nb[w], joints = jc[w];
      while (e1 < W1 && bodies + nb[e1] <= cap && bf[e1] == bf[e1 - 1] + nb[e1 - 1] &&
             jf[e1] == jf[e1 - 1] + jc[e1 - 1]) {
        bodies += nb[e1];
        joints += jc[e1];
        ++e1;
      }
      bins.push_back({w, e1, bodies, joints});
      w = e1;
    }
    // bins beside a launch of the large worlds need the side stream of the tiered launch
    if (static_cast<int>(bins.size() + big.size()) < W1 && !bins.empty() && (big.empty() || m->xpbd_tiered)) {
      std::stable_sort(bins.begin(), bins.end(),
                       [](const std::array<int, 4>& x, const std::array<int, 4>& y) { return x[2] > y[2]; });
      std::vector<int> flat;
      for (const auto& bn : bins) {
        flat.push_back(bn[0]);
        flat.push_back(bn[1]);
        m->xpbd_bin_max_bodies = std::max(m->xpbd_bin_max_bodies, std::max(bn[2], 1));
        m->xpbd_bin_max_joints = std::max(m->xpbd_bin_max_joints, bn[3]);
      }
      m->xpbd_bin_count = static_cast<int>(bins.size());
      m->xpbd_big_count = static_cast<int>(big.size());
      if ((e = dmalloc(&m->xpbd_bins, flat.size())) != hipSuccess ||
          (e = hipMemcpy(m->xpbd_bins, flat.data(), sizeof(int) * flat.size(), hipMemcpyHostToDevice)) != hipSuccess ||
          (!big.empty() &&
           ((e = dmalloc(&m->xpbd_big, big.size())) != hipSuccess ||
            (e = hipMemcpy(m->xpbd_big, big.data(), sizeof(int) * big.size(), hipMemcpyHostToDevice)) != hipSuccess))) {
        nmc_model_destroy(m);
        return hip_fail(e, "nmc_model_create: world bins");
      }
    }
  }
  if (m->collide_lds > 160 * 1024 || m->xpbd_lds > 160 * 1024) {
    delete m;
    return fail("per-world LDS requirement exceeds 160 KiB");
  }
  if ((e = dmalloc(&m->counters, 4)) != hipSuccess || (e = dmalloc(&m->xpbd_keys, m->xpbd_keys_cap)) != hipSuccess) {
    nmc_model_destroy(m);
    return hip_fail(e, "nmc_model_create: device allocation");
  }
  if (d.shape_pairs_max > 0) {
    // pair regions: segment s holds up to min(n_s (n_s - 1) / 2, LDS pair buffer) pairs
    std::vector<int> starts(segs + 1);
    e = hipMemcpy(starts.data(), d.segment_start, sizeof(int) * (segs + 1), hipMemcpyDeviceToHost);
    if (e != hipSuccess) { nmc_model_destroy(m); return hip_fail(e, "nmc_model_create: segment_start"); }
    std::vector<long long> off(segs);
    std::vector<int> cap(segs);
    long long total = 0;
    for (int sg = 0; sg < segs; ++sg) {
      long long n = starts[sg + 1] - starts[sg];
      long long w = n * (n - 1) / 2;
      cap[sg] = static_cast<int>(w < m->pair_cap ? w : m->pair_cap);
      off[sg] = total;
      total += cap[sg];
    }
    if (total >= (1ll << 31)) { nmc_model_destroy(m); return fail("broad-phase pair regions exceed 2^31 entries"); }
    // segments of at most 64 shapes run one wave each (their whole pair set fits the wave's LDS buffer)
    std::vector<int> small, big;
    int small_max = 1;
    for (int sg = 0; sg < segs; ++sg) {
      int n = starts[sg + 1] - starts[sg];
      if (n <= 64) {
        small.push_back(sg);
        small_max = n > small_max ? n : small_max;
      } else {
        big.push_back(sg);
      }
    }
    m->n_seg_small = static_cast<int>(small.size());
    m->n_seg_big = static_cast<int>(big.size());
    m->seg_small_max_n = small_max;
    m->seg_small_lds = collide_lds_enum_bytes(small_max);
    if ((e = dmalloc(&m->seg_small, small.size())) != hipSuccess || (e = dmalloc(&m->seg_big, big.size())) != hipSuccess ||
        (!small.empty() && (e = hipMemcpy(m->seg_small, small.data(), sizeof(int) * small.size(), hipMemcpyHostToDevice)) != hipSuccess) ||
        (!big.empty() && (e = hipMemcpy(m->seg_big, big.data(), sizeof(int) * big.size(), hipMemcpyHostToDevice)) != hipSuccess)) {
      nmc_model_destroy(m);
      return hip_fail(e, "nmc_model_create: segment lists");
    }
    long long pairs = total > 0 ? total : 1;
    long long slots = 5 * pairs;
    StageBuffers& s = m->stage;
    s.pair_capacity = pairs;
    s.mesh_slot_base = slots;
    s.mesh_pair_capacity = 0;
    if (d.mesh_count > 0) {
      // mesh pairs: per segment at most the pairs that involve one of its m mesh shapes,
      // n(n-1)/2 - (n-m)(n-m-1)/2, within the segment's region
      const int nss = starts[segs];
      std::vector<int> seg_shape(nss > 0 ? nss : 1), stype(d.shape_count > 0 ? d.shape_count : 1);
      if ((nss > 0 && (e = hipMemcpy(seg_shape.data(), d.segment_shape, sizeof(int) * nss, hipMemcpyDeviceToHost)) != hipSuccess) ||
          (d.shape_count > 0 && (e = hipMemcpy(stype.data(), d.shape_type, sizeof(int) * d.shape_count, hipMemcpyDeviceToHost)) != hipSuccess)) {
        nmc_model_destroy(m);
        return hip_fail(e, "nmc_model_create: mesh tables");
      }
      long long mesh_pairs = 0;
      for (int sg = 0; sg < segs; ++sg) {
        long long n = starts[sg + 1] - starts[sg], mm = 0;
        for (int k = starts[sg]; k < starts[sg + 1]; ++k) mm += stype[seg_shape[k]] == GEO_MESH;
        long long w = n * (n - 1) / 2 - (n - mm) * (n - mm - 1) / 2;
        mesh_pairs += w < cap[sg] ? w : cap[sg];
      }
      s.mesh_pair_capacity = mesh_pairs;
      slots += mesh_pairs * m->mesh_k;
    }
    if ((e = dmalloc(&s.rec, REC_F4 * slots)) != hipSuccess || (e = dmalloc(&s.mask, pairs)) != hipSuccess || (e = dmalloc(&s.pair_a, pairs)) != hipSuccess ||
        (e = dmalloc(&s.pair_b, pairs)) != hipSuccess || (e = dmalloc(&s.route, pairs)) != hipSuccess ||
        (e = dmalloc(&s.route_count, 3)) != hipSuccess ||
        (d.mesh_count > 0 && (e = dmalloc(&s.mesh_index, pairs)) != hipSuccess) ||
        (e = dmalloc(&m->seg_mesh_count, segs)) != hipSuccess || (e = dmalloc(&m->seg_mesh_off, segs)) != hipSuccess ||
        (e = dmalloc(&s.shape_wc, 20 * (size_t)(d.shape_count > 0 ? d.shape_count : 1))) != hipSuccess ||
        (e = dmalloc(&m->seg_pair_off, segs)) != hipSuccess || (e = dmalloc(&m->seg_pair_cap, segs)) != hipSuccess ||
        (e = dmalloc(&m->seg_pair_count, segs)) != hipSuccess || (e = dmalloc(&m->seg_conv_count, segs)) != hipSuccess ||
        (e = dmalloc(&m->seg_conv_off, segs)) != hipSuccess || (e = dmalloc(&m->seg_ana_off, segs)) != hipSuccess ||
        (e = dmalloc(&m->seg_contact_count, segs)) != hipSuccess ||
        (e = hipMemcpy(m->seg_pair_off, off.data(), sizeof(long long) * segs, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(m->seg_pair_cap, cap.data(), sizeof(int) * segs, hipMemcpyHostToDevice)) != hipSuccess) {
      nmc_model_destroy(m);
      return hip_fail(e, "nmc_model_create: device allocation");
    }
  }
  m->ord_G = -1;
  if (d.shape_pairs_max > 0) {
    e = order_setup(m);
    if (e != hipSuccess) {
      nmc_model_destroy(m);
      return hip_fail(e, "nmc_model_create: sort-key layout tables");
    }
  }
  *out = m;
  return 0;
}

extern "C" void nmc_model_destroy(NmcModel* m) {
  if (!m) return;
  StageBuffers& s = m->stage;
  void* ptrs[] = {s.rec, s.mask,
                  s.pair_a, s.pair_b, s.route, s.route_count, s.mesh_index, m->seg_mesh_count, m->seg_mesh_off,
                  m->seg_small, m->seg_big,
                  s.shape_wc, m->seg_pair_off, m->seg_pair_cap,
                  m->seg_pair_count, m->seg_conv_count, m->seg_conv_off, m->seg_ana_off, m->seg_contact_count,
                  m->counters, m->xpbd_keys, m->xpbd_tier_worlds[0], m->xpbd_tier_worlds[1], m->xpbd_tier_worlds[2],
                  m->xpbd_bins, m->xpbd_big,
                  m->ord_gid, m->ord_unit_order, m->ord_delta0, m->ord_M, m->ord_gstart, m->ord_L, m->ord_H,
                  m->ord_unit_off, m->ord_base};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (int k = 0; k < 2; ++k) {
    if (m->tier_stream[k]) (void)hipStreamDestroy(m->tier_stream[k]);
    if (m->tier_join[k]) (void)hipEventDestroy(m->tier_join[k]);
  }
  if (m->tier_fork) (void)hipEventDestroy(m->tier_fork);
  delete m;
}

extern "C" int nmc_collide(NmcModel* m, const float* body_q, const NmcContacts* c, nmc_stream_t stream) {
  if (!m || !c) return fail("nmc_collide: null argument");
  if (!m->stage.pair_a) return fail("nmc_collide: the model was created with shape_pairs_max = 0 (solver only)");
  if (m->d.body_count > 0 && !body_q) return fail("nmc_collide: body_q is null");
  if (!c->count || !c->shape0 || !c->shape1 || !c->point0 || !c->point1 || !c->offset0 || !c->offset1 ||
      !c->normal || !c->margin0 || !c->margin1 || !c->world_start || !c->world_count)
    return fail("nmc_collide: contacts buffer has a null array");
  if (c->sort_key && !c->order)
    return fail("nmc_collide: a contacts buffer with sort keys needs its order array (the global key layout)");
  if (c->order && m->ord_G < 0)
    return fail("nmc_collide: the global sort-key layout needs each world's shapes in one index block and at most "
                "256 world -1 shapes");
  hipError_t e = launch_collide(m, body_q, c, reinterpret_cast<hipStream_t>(stream));
  if (e != hipSuccess) return hip_fail(e, "nmc_collide launch");
  return 0;
}

// The restitution pass runs when asked for, with contacts and bodies (solver_xpbd.py:719-764: only with a
// contacts object).
static bool restitution_on(const NmcModel* m, const NmcXpbdParams* p) {
  return p->enable_restitution && p->has_contacts && m->d.body_count > 0;
}

extern "C" int nmc_xpbd_step(NmcModel* m, const NmcXpbdParams* p, const float* q_in, const float* qd_in,
                             const float* f_in, const float* joint_f, const float* tpos, const float* tvel,
                             const NmcContacts* c, float dt, float* q_out, float* qd_out, nmc_stream_t stream) {
  if (!m || !p) return fail("nmc_xpbd_step: null argument");
  if (m->d.body_count > 0 && (!q_in || !qd_in || !f_in || !q_out || !qd_out))
    return fail("nmc_xpbd_step: body state array is null");
  if (m->d.joint_count > 0 && (!joint_f || !tpos || !tvel)) return fail("nmc_xpbd_step: joint array is null");
  if (p->has_contacts && !c) return fail("nmc_xpbd_step: has_contacts set but contacts is null");
  if (q_in == q_out || qd_in == qd_out) return fail("nmc_xpbd_step: state_in and state_out must not alias");
  if (p->has_contacts && 2ll * c->capacity > m->xpbd_keys_cap)
    return fail("nmc_xpbd_step: contacts capacity exceeds the rigid_contact_max the model was created with");
  if (!valid_lanes(m, p)) return fail("nmc_xpbd_step: block_lanes must be 0, 256, 64 (worlds of up to 64 bodies) or 512 (CSR bins)");
  hipError_t e = launch_xpbd(m, p, q_in, qd_in, f_in, joint_f, tpos, tvel, p->has_contacts ? c : nullptr, dt, q_out,
                             qd_out, nullptr, nullptr, reinterpret_cast<hipStream_t>(stream));
  if (e != hipSuccess) return hip_fail(e, "nmc_xpbd_step launch");
  if (restitution_on(m, p)) {
    e = launch_restitution(m, q_in, qd_in, c, dt, qd_out, reinterpret_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "nmc_xpbd_step restitution launch");
  }
  return 0;
}

extern "C" int nmc_xpbd_step_learn(NmcModel* m, const NmcXpbdParams* p, const float* q_in, const float* qd_in,
                                   const float* f_in, const float* joint_f, const float* tpos, const float* tvel,
                                   const NmcContacts* c, float dt, float* q_out, float* qd_out, const NmcLearnIO* lrn,
                                   nmc_stream_t stream) {
  if (!m || !p || !lrn) return fail("nmc_xpbd_step_learn: null argument");
  if (m->d.body_count > 0 && (!q_in || !qd_in || !f_in))
    return fail("nmc_xpbd_step_learn: body state array is null");
  if (m->d.joint_count > 0 && (!joint_f || !tpos || !tvel)) return fail("nmc_xpbd_step_learn: joint array is null");
  if (p->has_contacts && !c) return fail("nmc_xpbd_step_learn: has_contacts set but contacts is null");
  if (p->has_contacts && 2ll * c->capacity > m->xpbd_keys_cap)
    return fail("nmc_xpbd_step_learn: contacts capacity exceeds the rigid_contact_max the model was created with");
  if (!valid_lanes(m, p))
    return fail("nmc_xpbd_step_learn: block_lanes must be 0, 256, 64 (worlds of up to 64 bodies) or 512 (CSR bins)");
  const int mode = lrn->mode;
  if (mode < NMC_STEP_PLAIN || mode > NMC_STEP_WARM) return fail("nmc_xpbd_step_learn: unknown mode");
  if (mode != NMC_STEP_PREP) {
    if (m->d.body_count > 0 && (!q_out || !qd_out)) return fail("nmc_xpbd_step_learn: output state is null");
    if (q_in == q_out || qd_in == qd_out) return fail("nmc_xpbd_step_learn: state_in and state_out must not alias");
  }
  if ((mode == NMC_STEP_LOG || mode == NMC_STEP_PREP) && !lrn->feat) return fail("nmc_xpbd_step_learn: feat is null");
  if ((mode == NMC_STEP_LOG || mode == NMC_STEP_WARM) && !lrn->lambda_sum)
    return fail("nmc_xpbd_step_learn: lambda_sum is null");
  if ((mode == NMC_STEP_PREP || mode == NMC_STEP_WARM) && (!lrn->q_mid || !lrn->qd_mid))
    return fail("nmc_xpbd_step_learn: q_mid / qd_mid is null");
  if (mode == NMC_STEP_WARM && !lrn->lambda0) return fail("nmc_xpbd_step_learn: lambda0 is null");
  if (mode == NMC_STEP_WARM && (lrn->warm_iterations < 0 || lrn->warm_iterations > p->iterations))
    return fail("nmc_xpbd_step_learn: warm_iterations must be in [0, iterations]");
  if (lrn->prev_key && (!lrn->prev_start || !lrn->prev_count || !lrn->prev_lambda))
    return fail("nmc_xpbd_step_learn: prev_key given without prev_start / prev_count / prev_lambda");
  hipError_t e = launch_xpbd(m, p, q_in, qd_in, f_in, joint_f, tpos, tvel, p->has_contacts ? c : nullptr, dt, q_out,
                             qd_out, lrn, nullptr, reinterpret_cast<hipStream_t>(stream));
  if (e != hipSuccess) return hip_fail(e, "nmc_xpbd_step_learn launch");
  if (mode != NMC_STEP_PREP && restitution_on(m, p)) {
    e = launch_restitution(m, q_in, qd_in, c, dt, qd_out, reinterpret_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "nmc_xpbd_step_learn restitution launch");
  }
  return 0;
}

extern "C" int nmc_xpbd_step_report(NmcModel* m, const NmcXpbdParams* p, const float* q_in, const float* qd_in,
                                    const float* f_in, const float* joint_f, const float* tpos, const float* tvel,
                                    const NmcContacts* c, float dt, float* q_out, float* qd_out,
                                    const NmcXpbdReport* rep, nmc_stream_t stream) {
  if (!m || !p || !rep) return fail("nmc_xpbd_step_report: null argument");
  if (m->d.body_count > 0 && (!q_in || !qd_in || !f_in || !q_out || !qd_out))
    return fail("nmc_xpbd_step_report: body state array is null");
  if (m->d.joint_count > 0 && (!joint_f || !tpos || !tvel)) return fail("nmc_xpbd_step_report: joint array is null");
  if (p->has_contacts && !c) return fail("nmc_xpbd_step_report: has_contacts set but contacts is null");
  if (q_in == q_out || qd_in == qd_out) return fail("nmc_xpbd_step_report: state_in and state_out must not alias");
  if (p->has_contacts && 2ll * c->capacity > m->xpbd_keys_cap)
    return fail("nmc_xpbd_step_report: contacts capacity exceeds the rigid_contact_max the model was created with");
  if (!valid_lanes(m, p))
    return fail("nmc_xpbd_step_report: block_lanes must be 0, 256, 64 (worlds of up to 64 bodies) or 512 (CSR bins)");
  if (rep->contact_impulse && !rep->contact_scratch)
    return fail("nmc_xpbd_step_report: contact_impulse needs contact_scratch");
  if (rep->body_parent_f && m->d.joint_count > 0 && !rep->joint_impulse)
    return fail("nmc_xpbd_step_report: body_parent_f needs joint_impulse");
  NmcXpbdReport r = *rep;
  if (!p->has_contacts) r.contact_impulse = nullptr;
  hipError_t e = launch_xpbd(m, p, q_in, qd_in, f_in, joint_f, tpos, tvel, p->has_contacts ? c : nullptr, dt, q_out,
                             qd_out, nullptr, &r, reinterpret_cast<hipStream_t>(stream));
  if (e != hipSuccess) return hip_fail(e, "nmc_xpbd_step_report launch");
  if (restitution_on(m, p)) {
    e = launch_restitution(m, q_in, qd_in, c, dt, qd_out, reinterpret_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "nmc_xpbd_step_report restitution launch");
  }
  return 0;
}

namespace nmc {
__global__ void __launch_bounds__(256) contact_force_kernel(const float* __restrict__ impulse,
                                                           const int* __restrict__ count, int capacity, float dt,
                                                           float* __restrict__ force) {
  // the buffers as flat float arrays of 6 x capacity, written 16 bytes per lane; float f belongs to contact f / 6
  const long long n6 = 6ll * static_cast<long long>(count[0] < capacity ? count[0] : capacity);
  const long long total = 6ll * capacity;
  const float inv_dt = 1.0f / dt;
  const long long nq = total >> 2;
  const float4* __restrict__ in4 = reinterpret_cast<const float4*>(impulse);
  float4* __restrict__ out4 = reinterpret_cast<float4*>(force);
  for (long long q = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; q < nq;
       q += static_cast<long long>(gridDim.x) * blockDim.x) {
    const long long f = 4 * q;
    float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (f < n6) {
      const float4 im = in4[q];
      v.x = im.x * inv_dt;
      v.y = f + 1 < n6 ? im.y * inv_dt : 0.0f;
      v.z = f + 2 < n6 ? im.z * inv_dt : 0.0f;
      v.w = f + 3 < n6 ? im.w * inv_dt : 0.0f;
    }
    out4[q] = v;
  }
  // 6 x capacity is even: at most two trailing floats
  if (blockIdx.x == 0 && threadIdx.x < static_cast<unsigned>(total - 4 * nq)) {
    const long long f = 4 * nq + threadIdx.x;
    force[f] = f < n6 ? impulse[f] * inv_dt : 0.0f;
  }
}
}  // namespace nmc

extern "C" int nmc_contact_force(const float* contact_impulse, const int* count, int capacity, float dt, float* force,
                                 nmc_stream_t stream) {
  if (!contact_impulse || !count || !force) return fail("nmc_contact_force: null argument");
  if (capacity <= 0) return 0;
  // flat float4 access needs 16-byte aligned buffers
  if ((reinterpret_cast<uintptr_t>(contact_impulse) | reinterpret_cast<uintptr_t>(force)) & 15)
    return fail("nmc_contact_force: impulse and force buffers must be 16-byte aligned");
  long long blocks = (6ll * capacity / 4 + 255) / 256;
  if (blocks < 1) blocks = 1;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(contact_force_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), contact_impulse, count, capacity, dt, force);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : hip_fail(e, "nmc_contact_force launch");
}

extern "C" int nmc_model_set_launch(NmcModel* m, const NmcLaunchConfig* cfg) {
  if (!m || !cfg) return fail("nmc_model_set_launch: null argument");
  const int sw = cfg->segment_waves ? cfg->segment_waves : 4;
  const int nb = cfg->narrow_max_blocks ? cfg->narrow_max_blocks : 2048;
  const int mb = cfg->mesh_max_blocks ? cfg->mesh_max_blocks : 2048;
  if (sw != 1 && sw != 2 && sw != 4) return fail("nmc_model_set_launch: segment_waves must be 1, 2 or 4");
  if (nb < 1 || nb > 65535 || mb < 1 || mb > 65535)
    return fail("nmc_model_set_launch: grid caps must lie in [1, 65535]");
  m->segment_waves = sw;
  m->narrow_max_blocks = nb;
  m->mesh_max_blocks = mb;
  return 0;
}

extern "C" int nmc_get_counters(NmcModel* m, int* out, nmc_stream_t stream) {
  if (!m || !out) return fail("nmc_get_counters: null argument");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  hipError_t e = hipMemcpyAsync(out, m->counters, 4 * sizeof(int), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_fail(e, "nmc_get_counters");
  return 0;
}
