// Triangle-mesh contacts on gfx950: the SDF cook and the mesh narrow phase.
//
// sdf_build_kernel: one lane per grid vertex; distance to the nearest triangle (closest point on each
// triangle), sign from the generalized winding number w (sum of solid angles / 4 pi): inside when
// w > threshold, or w <= threshold for a negative threshold (inward-wound meshes).
//
// mesh_narrow_kernel: one 256-lane workgroup per mesh pair (grid-stride over the call's mesh pairs).
//   mesh (shape0) - infinite plane (shape1), narrow_phase.py:1073-1190: every mesh vertex closer to the
//     plane than gap + margins is a contact at the midpoint of the vertex and its projection, normal
//     mesh -> plane (minus the plane normal), sub key = vertex index.
//   mesh - mesh, sdf_contact.py:813-1172: both ways (mode 0: edges of shape0 in the SDF of shape1,
//     mode 1: the reverse).  An edge is culled by its bounding sphere against the SDF box and the SDF at
//     its midpoint, searched for its deepest point (Brent), and becomes a contact at that point with the
//     SDF gradient as normal (negated in mode 0); sub key = (edge << 2) | (mode << 1).
// Pass 1 keeps (distance, sub key) of every candidate that passes the writer's gap test (collide.py:62-148)
// in LDS; the pair keeps the mesh_contact_max deepest (ordered by distance, then sub key) and pass 2
// evaluates those again and writes them in sub-key order into the pair's staging slots; the compaction
// then places them in canonical order.  The reference instead reduces a
// pair's contacts by normal bins in a global hash table (contact_reduction_global.py:671-1580).
#include <hip/hip_runtime.h>

#include "narrow.h"
#include "nmc_internal.h"
#include "sdf.h"

namespace nmc {

constexpr int MESH_BLOCK = 256;
constexpr int MESH_MAX_BLOCKS = 2048;

__device__ __forceinline__ unsigned mesh_sortable(float f) {
  if (f == 0.0f) f = 0.0f;
  unsigned u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

template <class K>
__device__ void mesh_bitonic(K* keys, int n2) {
  for (int k = 2; k <= n2; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < n2; i += blockDim.x) {
        int ixj = i ^ j;
        if (ixj > i) {
          K a = keys[i], b = keys[ixj];
          bool up = (i & k) == 0;
          if ((a > b) == up) {
            keys[i] = b;
            keys[ixj] = a;
          }
        }
      }
      __syncthreads();
    }
  }
}

// ------------------------------------------------------------------ SDF cook
__global__ void __launch_bounds__(256) sdf_build_kernel(const float* __restrict__ vtx, int nvtx,
                                                        const int* __restrict__ tri, int ntri, f3 lower, float voxel, int cx, int cy, int cz,
                                                        float threshold, float* __restrict__ out) {
  const long long nx = cx + 1, ny = cy + 1, nz = cz + 1;
  const long long total = nx * ny * nz;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    const int ix = (int)(i % nx), iy = (int)((i / nx) % ny), iz = (int)(i / (nx * ny));
    const f3 p{lower.x + voxel * (float)ix, lower.y + voxel * (float)iy, lower.z + voxel * (float)iz};
    float best = 3.0e38f;
    float omega = 0.0f;
    for (int t = 0; t < ntri; ++t) {
      const int i0 = tri[3 * t], i1 = tri[3 * t + 1], i2 = tri[3 * t + 2];
      if (static_cast<unsigned>(i0) >= static_cast<unsigned>(nvtx) || static_cast<unsigned>(i1) >= static_cast<unsigned>(nvtx) ||
          static_cast<unsigned>(i2) >= static_cast<unsigned>(nvtx))
        continue;  // the host validates indices; a bad one is skipped, never read
      const f3 a = load3(vtx + 3 * i0), b = load3(vtx + 3 * i1);
      const f3 c = load3(vtx + 3 * i2);
      const f3 q = closest_on_triangle(p, a, b, c);
      const float dd = len(p - q);
      best = dd < best ? dd : best;
      omega += solid_angle(a - p, b - p, c - p);
    }
    const float wn = omega / (4.0f * 3.14159265358979323846f);
    const bool above = wn > threshold;
    out[i] = (threshold >= 0.0f ? above : !above) ? -best : best;
  }
}

// ------------------------------------------------------------------ mesh narrow phase
// One side of a mesh pair: the edges of mesh T tested against the SDF of mesh S (mode 0: T = shape0).
struct EdgeSide {
  SdfGrid G;
  f3 ss, inv_s;  // SDF shape scale (clamped at 1e-10) and its inverse
  float min_s;
  xform X_ts;    // T frame -> S frame
  xform XS;      // S world transform
  f3 tscale;     // T scale
  float threshold, threshold_u;
  const int* edge;
  const float* vtx;
};

__device__ __forceinline__ EdgeSide edge_side(const NmcModelDesc& d, const ShapeW& T, const ShapeW& S, int tri_mesh,
                                              int sdf_mesh, float gap_sum) {
  EdgeSide e;
  e.G = sdf_grid(d, sdf_mesh);
  // a grid cooked from scaled vertices is sampled with unit scale (sdf_contact.py:918-930)
  const f3 s = (d.mesh_sdf_flags[sdf_mesh] & 1) ? f3{1.0f, 1.0f, 1.0f} : S.gscale;
  e.ss = f3{fmax2(s.x, 1e-10f), fmax2(s.y, 1e-10f), fmax2(s.z, 1e-10f)};
  e.inv_s = f3{1.0f / e.ss.x, 1.0f / e.ss.y, 1.0f / e.ss.z};
  e.min_s = fmin2(fmin2(e.ss.x, e.ss.y), e.ss.z);
  e.X_ts = xmul(xinverse(S.X), T.X);
  e.XS = S.X;
  e.tscale = T.gscale;
  e.threshold = gap_sum + T.margin + S.margin;
  e.threshold_u = e.threshold / e.min_s;
  e.edge = d.mesh_edge + 2 * d.mesh_edge_start[tri_mesh];
  e.vtx = d.mesh_vertex + 3 * d.mesh_vertex_start[tri_mesh];
  return e;
}

// Edge e of one side (sdf_contact.py:1010-1150): culling, Brent search, gradient normal.  Returns false
// when the edge yields no contact; else the contact's world point, its normal (from shape0 to shape1)
// and distance.
__device__ bool edge_contact(const EdgeSide& E, int mode, int e, f3& center, f3& normal, float& dist) {
  const int ia = E.edge[2 * e], ib = E.edge[2 * e + 1];
  f3 pa = load3(E.vtx + 3 * ia), pb = load3(E.vtx + 3 * ib);
  pa = f3{pa.x * E.tscale.x, pa.y * E.tscale.y, pa.z * E.tscale.z};
  pb = f3{pb.x * E.tscale.x, pb.y * E.tscale.y, pb.z * E.tscale.z};
  f3 ea = xpoint(E.X_ts, pa), eb = xpoint(E.X_ts, pb);
  ea = f3{ea.x * E.inv_s.x, ea.y * E.inv_s.y, ea.z * E.inv_s.z};
  eb = f3{eb.x * E.inv_s.x, eb.y * E.inv_s.y, eb.z * E.inv_s.z};
  const f3 mid = (ea + eb) * 0.5f;
  const float rad = len(eb - ea) * 0.5f;
  const float cull = rad + E.threshold_u;
  const SdfGrid& G = E.G;
  const f3 cl{clampf(mid.x, G.lower.x, G.upper.x), clampf(mid.y, G.lower.y, G.upper.y),
              clampf(mid.z, G.lower.z, G.upper.z)};
  if (len_sq(mid - cl) > cull * cull) return false;
  const float msdf = sdf_sample(G, mid);
  if (!(msdf <= cull)) return false;
  f3 pu;
  const float du0 = edge_sdf_search(G, ea, eb, msdf, pu);
  if (!(du0 * E.min_s < E.threshold)) return false;
  f3 gu;
  const float du = sdf_sample_grad(G, pu, gu);
  // scale_sdf_result_to_world (sdf_contact.py:51-82)
  dist = du * E.min_s;
  const f3 gsc{gu.x * E.inv_s.x, gu.y * E.inv_s.y, gu.z * E.inv_s.z};
  const float gl = len(gsc);
  const f3 dir = gl > 0.0f ? gsc / gl : gu;
  center = xpoint(E.XS, f3{pu.x * E.ss.x, pu.y * E.ss.y, pu.z * E.ss.z});
  f3 dw = xvector(E.XS, dir);
  const float dl = len(dw);
  if (dl > 0.0f) {
    dw = dw / dl;
  } else {
    const f3 fb = center - E.XS.p;
    const float fl = len(fb);
    dw = fl > 0.0f ? fb / fl : f3{0.0f, 1.0f, 0.0f};
  }
  normal = mode == 0 ? -dw : dw;
  return true;
}

// Vertex v of a mesh against an infinite plane (narrow_phase.py:1150-1190).
struct PlaneSide {
  xform XM, Xp, Xp_inv;
  f3 mscale, pn;
  float limit;  // gap_sum + both margins
  const float* vtx;
};

__device__ __forceinline__ bool vertex_contact(const PlaneSide& P, int v, f3& center, f3& normal, float& dist) {
  f3 vl = load3(P.vtx + 3 * v);
  vl = f3{vl.x * P.mscale.x, vl.y * P.mscale.y, vl.z * P.mscale.z};
  const f3 vw = xpoint(P.XM, vl);
  const f3 vp = xpoint(P.Xp_inv, vw);
  const f3 pp = xpoint(P.Xp, f3{vp.x, vp.y, 0.0f});
  dist = dot(vw - pp, P.pn);
  if (!(dist < P.limit)) return false;
  center = (vw + pp) * 0.5f;
  normal = -P.pn;
  return true;
}

// The writer's gap test (collide.py:62-148) for a contact of this pair.
__device__ __forceinline__ bool mesh_keeps(CData& c, f3 center, f3 normal, float dist) {
  c.center = center;
  c.normal = normal;
  c.dist = dist;
  return passes_gap(c);
}

__global__ void __launch_bounds__(MESH_BLOCK) mesh_narrow_kernel(NmcModelDesc d, const float* __restrict__ body_q,
                                                                 StageBuffers st,
                                                                 const long long* __restrict__ seg_pair_off,
                                                                 const int* __restrict__ seg_pair_count,
                                                                 const int* __restrict__ seg_mesh_count,
                                                                 const int* __restrict__ seg_mesh_off, int segments,
                                                                 int cand_cap, int* __restrict__ counters) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int s_n;
  // (sortable distance << 32) | sub key, then the kept sub keys
  unsigned long long* keys = reinterpret_cast<unsigned long long*>(smem);
  unsigned* sel = reinterpret_cast<unsigned*>(smem + sizeof(unsigned long long) * cand_cap);
  const int total = st.route_count[2];
  const int K = d.mesh_contact_max;
  for (int i = blockIdx.x; i < total; i += gridDim.x) {
    int lo = 0, hi = segments - 1;  // last segment whose mesh offset is <= i
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (seg_mesh_off[mid] <= i) lo = mid; else hi = mid - 1;
    }
    const long long poff = seg_pair_off[lo];
    const int r = seg_pair_count[lo] - seg_mesh_count[lo] + (i - seg_mesh_off[lo]);
    const long long g = poff + st.route[poff + r];
    if (i >= st.mesh_pair_capacity) {  // cannot happen: the capacity is the worst case of the regions
      if (threadIdx.x == 0) st.mask[g] = 0;
      continue;
    }
    const int sa = st.pair_a[g], sb = st.pair_b[g];
    const bool plane = d.shape_type[sb] == GEO_PLANE;
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    const ShapeW A = load_wc(st.shape_wc + 20ll * sa), B = load_wc(st.shape_wc + 20ll * sb);
    const float gap_sum = d.shape_gap[sa] + d.shape_gap[sb];
    CData c;
    c.r_a = 0.0f;
    c.r_b = 0.0f;
    c.m_a = A.margin;
    c.m_b = B.margin;
    c.shape_a = sa;
    c.shape_b = sb;
    c.gap_sum = gap_sum;
    c.sub_key = 0;
    const int ma = d.shape_mesh[sa];
    const int mb = plane ? -1 : d.shape_mesh[sb];
    PlaneSide P;
    if (plane) {
      P.XM = A.X;
      P.Xp = B.X;
      P.Xp_inv = xinverse(B.X);
      P.mscale = A.gscale;
      P.pn = xvector(B.X, f3{0.0f, 0.0f, 1.0f});
      P.limit = gap_sum + (A.margin + B.margin);
      P.vtx = d.mesh_vertex + 3 * d.mesh_vertex_start[ma];
    }
    // ---- pass 1: every candidate's distance, kept ones as keys (sub key: vertex, or edge and mode)
    auto push = [&](int sub, f3 center, f3 normal, float dist) {
      if (!mesh_keeps(c, center, normal, dist)) return;
      const int k = atomicAdd(&s_n, 1);
      if (k < cand_cap) keys[k] = (static_cast<unsigned long long>(mesh_sortable(dist)) << 32) | static_cast<unsigned>(sub);
    };
    if (plane) {
      const int nv = d.mesh_vertex_start[ma + 1] - d.mesh_vertex_start[ma];
      for (int v = threadIdx.x; v < nv; v += blockDim.x) {
        f3 center, normal;
        float dist;
        if (vertex_contact(P, v, center, normal, dist)) push(v, center, normal, dist);
      }
    } else {
      for (int mode = 0; mode < 2; ++mode) {
        const int tm = mode == 0 ? ma : mb;
        const EdgeSide E = mode == 0 ? edge_side(d, A, B, ma, mb, gap_sum) : edge_side(d, B, A, mb, ma, gap_sum);
        const int ne = d.mesh_edge_start[tm + 1] - d.mesh_edge_start[tm];
        for (int e = threadIdx.x; e < ne; e += blockDim.x) {
          f3 center, normal;
          float dist;
          if (edge_contact(E, mode, e, center, normal, dist)) push((e << 2) | (mode << 1), center, normal, dist);
        }
      }
    }
    __syncthreads();
    int n = s_n;
    if (n > cand_cap) {
      if (threadIdx.x == 0) atomicAdd(&counters[1], n - cand_cap);
      n = cand_cap;
    }
    // ---- keep the K deepest (distance, then sub key); order them by sub key
    const int kept = n < K ? n : K;
    if (n > K) {
      int n2 = 1;
      while (n2 < n) n2 <<= 1;
      for (int k = n + threadIdx.x; k < n2; k += blockDim.x) keys[k] = ~0ull;
      __syncthreads();
      mesh_bitonic(keys, n2);
    }
    int k2 = 1;
    while (k2 < kept) k2 <<= 1;
    for (int k = threadIdx.x; k < k2; k += blockDim.x)
      sel[k] = k < kept ? static_cast<unsigned>(keys[k] & 0xffffffffu) : 0xffffffffu;
    __syncthreads();
    mesh_bitonic(sel, k2);
    // ---- pass 2: the kept candidates again, written in sub-key order
    const WriteCtx w{d.shape_gap, st, st.mesh_slot_base + static_cast<long long>(i) * K, 0};
    for (int k = threadIdx.x; k < kept; k += blockDim.x) {
      const int sub = static_cast<int>(sel[k]);
      f3 center, normal;
      float dist;
      if (plane) {
        (void)vertex_contact(P, sub, center, normal, dist);
      } else {
        const int mode = (sub >> 1) & 1;
        const EdgeSide E = mode == 0 ? edge_side(d, A, B, ma, mb, gap_sum) : edge_side(d, B, A, mb, ma, gap_sum);
        (void)edge_contact(E, mode, sub >> 2, center, normal, dist);
      }
      CData o = c;
      o.center = center;
      o.normal = normal;
      o.dist = dist;
      o.sub_key = sub;
      f3 nn, aw, bw;
      (void)contact_d(o, nn, aw, bw);
      write_record(o, nn, aw, bw, w, w.slot0 + k);
    }
    if (threadIdx.x == 0) {
      st.mask[g] = static_cast<unsigned char>(kept);
      st.mesh_index[g] = i;
    }
    __syncthreads();
  }
}

size_t mesh_lds_bytes(int cand_cap) { return sizeof(unsigned long long) * (size_t)cand_cap + sizeof(unsigned) * 256; }

hipError_t launch_mesh_narrow(NmcModelImpl* m, const float* body_q, hipStream_t stream) {
  if (m->stage.mesh_pair_capacity <= 0) return hipSuccess;
  const long long cap = m->stage.mesh_pair_capacity;
  const int blocks = static_cast<int>(cap < MESH_MAX_BLOCKS ? cap : MESH_MAX_BLOCKS);
  hipLaunchKernelGGL(mesh_narrow_kernel, dim3(blocks), dim3(MESH_BLOCK), m->mesh_lds, stream, m->d, body_q, m->stage,
                     m->seg_pair_off, m->seg_pair_count, m->seg_mesh_count, m->seg_mesh_off, m->d.segment_count,
                     m->mesh_cand_cap, m->counters);
  return hipGetLastError();
}

}  // namespace nmc

using namespace nmc;

extern "C" int nmc_sdf_build(const float* vertices, int vertex_count, const int* triangles, int triangle_count,
                             const float lower[3], float voxel, const int cells[3], float winding_threshold,
                             float* out, nmc_stream_t stream) {
  if (!vertices || !triangles || !lower || !cells || !out) return nmc::fail("nmc_sdf_build: null argument");
  if (vertex_count <= 0 || triangle_count <= 0) return nmc::fail("nmc_sdf_build: empty mesh");
  if (!(voxel > 0.0f) || cells[0] <= 0 || cells[1] <= 0 || cells[2] <= 0)
    return nmc::fail("nmc_sdf_build: voxel and cells must be positive");
  const long long total = (long long)(cells[0] + 1) * (cells[1] + 1) * (cells[2] + 1);
  long long blocks = (total + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(sdf_build_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), vertices, vertex_count, triangles, triangle_count,
                     f3{lower[0], lower[1], lower[2]}, voxel, cells[0], cells[1], cells[2], winding_threshold, out);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : nmc::hip_fail(e, "nmc_sdf_build launch");
}
