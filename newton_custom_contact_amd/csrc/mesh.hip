// Triangle-mesh contacts on gfx950: the SDF cook and the mesh narrow phase.
//
// The cook of the reference's sparse SDF texture (sdf_texture.py:959-1310), in three entry points:
//   sdf_build_kernel: one lane per vertex of the padded fine grid; distance to the nearest triangle (closest
//     point on each triangle), sign from the generalized winding number w (sum of solid angles / 4 pi):
//     inside when w > threshold, or w <= threshold for a negative threshold (inward-wound meshes);
//   sdf_coarse_kernel, sdf_subgrid_kernel, sdf_address_kernel: the coarse grid, narrow-band occupancy,
//     linearity demotion and texture addresses of the subgrids;
//   sdf_pack_kernel: indirection slots and the (quantised) subgrid texture.
//
// mesh_narrow_kernel: one 256-lane workgroup per mesh pair (grid-stride over the call's mesh pairs).
//   mesh (shape0) - infinite plane (shape1), narrow_phase.py:1073-1326: every mesh vertex closer to the
//     plane than gap + margins is a candidate at the midpoint of the vertex and its projection, normal
//     mesh -> plane (minus the plane normal), sub key = vertex index.
//   mesh - mesh, sdf_contact.py:813-1514: both ways (mode 0: edges of shape0 in the SDF of shape1,
//     mode 1: the reverse).  An edge is culled by its bounding sphere against the SDF box and the SDF at
//     its midpoint, searched for its deepest point (Brent), and becomes a candidate at that point with the
//     SDF gradient as normal (negated in mode 0) when the search distance is below gap + margins;
//     sub key = (edge << 2) | (mode << 1).
// reduce_contacts=False: every candidate the writer keeps (collide.py:62-148) is a contact.  Pass 1 lists
//   their sub keys in LDS, they are sorted, pass 2 evaluates each again and writes it.
// reduce_contacts=True: pass 1 enters every candidate into the pair's 240-slot reduction table in LDS
//   (reduce.h); the distinct winners are sorted by sub key, evaluated again, given the octahedral round trip
//   of their normal the reference's reducer buffer applies, kept when the writer keeps them and written.
// Contacts go to the pair's staging slots in sub-key order; the compaction places them in canonical order.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "narrow.h"
#include "nmc_internal.h"
#include "reduce.h"
#include "scan.h"
#include "sdf.h"

namespace nmc {

constexpr int MESH_BLOCK = 256;
// Slot tables of up to this many subgrids (12^3: up to 96 cells a side) are staged in LDS per mesh pair, so a
// sample's indirection is an LDS read; larger textures are read from global memory.
constexpr int MESH_SLOT_LDS = 2048;
// Edges culled per round of the mesh-mesh pass (the survivors' list in LDS).
constexpr int MESH_EDGE_LIST = 1024;

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
// Fine vertex (gx, gy, gz) at lower + (gx cx, gy cy, gz cz): the reference queries the mesh only at fine
// vertices (subgrid corners, subgrid centres, subgrid samples), so the cook evaluates the padded fine grid
// once and every later pass reads it.
__global__ void __launch_bounds__(256) sdf_build_kernel(const float* __restrict__ vtx, int nvtx,
                                                        const int* __restrict__ tri, int ntri, f3 lower, f3 cell,
                                                        int nx, int ny, int nz, float threshold,
                                                        float* __restrict__ out) {
  const long long total = (long long)nx * ny * nz;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    const int ix = (int)(i % nx), iy = (int)((i / nx) % ny), iz = (int)(i / ((long long)nx * ny));
    const f3 p{lower.x + (float)ix * cell.x, lower.y + (float)iy * cell.y, lower.z + (float)iz * cell.z};
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

struct FineGrid {
  const float* v;
  int nx, ny;  // fine vertices along x and y (8 w + 1, 8 h + 1)
  __device__ __forceinline__ float at(int x, int y, int z) const {
    return v[((long long)z * ny + y) * nx + x];
  }
};

// The coarse grid: the field at every subgrid corner (_build_coarse_sdf_from_mesh_kernel,
// sdf_texture.py:329-360), (d+1)(h+1)(w+1) values, x fastest.
__global__ void __launch_bounds__(256) sdf_coarse_kernel(FineGrid F, int w, int h, int d, float* __restrict__ coarse) {
  const int bx = w + 1, by = h + 1, total = bx * by * (d + 1);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
    const int z = i / (bx * by), r = i - z * bx * by, y = r / bx, x = r - y * bx;
    coarse[i] = F.at(x * SDF_SUBGRID, y * SDF_SUBGRID, z * SDF_SUBGRID);
  }
}

// Trilinear value of the coarse grid at fine sample (lx, ly, lz) of subgrid (bx, by, bz)
// (_interp_coarse_sdf, sdf_texture.py:119-161: coarse coordinates clamped to the last cell, weights to [0, 1]).
__device__ __forceinline__ float coarse_interp(const FineGrid& F, int bx, int by, int bz, int lx, int ly, int lz,
                                               int w, int h, int d) {
  const float cfx = (float)bx + (float)lx * 0.125f;
  const float cfy = (float)by + (float)ly * 0.125f;
  const float cfz = (float)bz + (float)lz * 0.125f;
  const int x0 = max(0, min((int)floorf(cfx), w - 1));
  const int y0 = max(0, min((int)floorf(cfy), h - 1));
  const int z0 = max(0, min((int)floorf(cfz), d - 1));
  const float tx = clampf(cfx - (float)x0, 0.0f, 1.0f);
  const float ty = clampf(cfy - (float)y0, 0.0f, 1.0f);
  const float tz = clampf(cfz - (float)z0, 0.0f, 1.0f);
  const int X0 = x0 * SDF_SUBGRID, X1 = X0 + SDF_SUBGRID, Y0 = y0 * SDF_SUBGRID, Y1 = Y0 + SDF_SUBGRID;
  const int Z0 = z0 * SDF_SUBGRID, Z1 = Z0 + SDF_SUBGRID;
  const float v000 = F.at(X0, Y0, Z0), v100 = F.at(X1, Y0, Z0), v010 = F.at(X0, Y1, Z0), v110 = F.at(X1, Y1, Z0);
  const float v001 = F.at(X0, Y0, Z1), v101 = F.at(X1, Y0, Z1), v011 = F.at(X0, Y1, Z1), v111 = F.at(X1, Y1, Z1);
  const float c00 = v000 * (1.0f - tx) + v100 * tx;
  const float c10 = v010 * (1.0f - tx) + v110 * tx;
  const float c01 = v001 * (1.0f - tx) + v101 * tx;
  const float c11 = v011 * (1.0f - tx) + v111 * tx;
  const float c0 = c00 * (1.0f - ty) + c10 * ty;
  const float c1 = c01 * (1.0f - ty) + c11 * ty;
  return c0 * (1.0f - tz) + c1 * tz;
}

// One workgroup per subgrid (grid-stride): narrow-band occupancy from the value at its centre
// (_check_subgrid_occupied_kernel, :198-224, _is_in_narrow_band, :164-169), then for occupied subgrids the
// largest deviation of its 9^3 samples from the coarse interpolation (_accumulate_subgrid_linearity_error_kernel,
// :227-304); a subgrid below linear_threshold is demoted to the coarse grid (_apply_subgrid_linearity_kernel,
// :307-326).  flags[s] = occupied | linear << 1 | required << 2.
__global__ void __launch_bounds__(256) sdf_subgrid_kernel(FineGrid F, int w, int h, int d, float band_lo,
                                                          float band_hi, float linear_threshold,
                                                          int* __restrict__ flags) {
  __shared__ float s_err[256 / 64];
  const int total = w * h * d;
  for (int s = blockIdx.x; s < total; s += gridDim.x) {
    const int bz = s / (w * h), r = s - bz * w * h, by = r / w, bx = r - by * w;
    const float c = F.at(bx * SDF_SUBGRID + 4, by * SDF_SUBGRID + 4, bz * SDF_SUBGRID + 4);
    const bool occupied = !(c < 0.0f) ? c < band_hi : c > band_lo;
    float err = 0.0f;
    if (occupied && linear_threshold > 0.0f) {
      for (int k = threadIdx.x; k < 729; k += blockDim.x) {
        const int lz = k / 81, rr = k - lz * 81, ly = rr / 9, lx = rr - ly * 9;
        const float v = F.at(bx * SDF_SUBGRID + lx, by * SDF_SUBGRID + ly, bz * SDF_SUBGRID + lz);
        const float e = fabsf(v - coarse_interp(F, bx, by, bz, lx, ly, lz, w, h, d));
        err = e > err ? e : err;
      }
      for (int off = 32; off > 0; off >>= 1) {
        const float o = __shfl_xor(err, off, 64);
        err = o > err ? o : err;
      }
      if ((threadIdx.x & 63) == 0) s_err[threadIdx.x >> 6] = err;
      __syncthreads();
      if (threadIdx.x == 0) {
        for (int q = 1; q < (int)(blockDim.x >> 6); ++q) err = s_err[q] > err ? s_err[q] : err;
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      const bool linear = occupied && linear_threshold > 0.0f && err < linear_threshold;
      flags[s] = (occupied ? 1 : 0) | (linear ? 2 : 0) | ((occupied && !linear) ? 4 : 0);
    }
  }
}

// Texture addresses of the required subgrids: exclusive scan (one workgroup); count[0] = number required.
__global__ void __launch_bounds__(SCAN_BLOCK) sdf_address_kernel(const int* __restrict__ flags, int n,
                                                                 int* __restrict__ address, int* __restrict__ count) {
  int total[1];
  segment_scan_fn<1>(
      n, [&](int i, int (&v)[1]) { v[0] = (flags[i] >> 2) & 1; },
      [&](int i, const int (&ex)[1], const int (&)[1]) { address[i] = ex[0]; }, total);
  if (threadIdx.x == 0) count[0] = total[0];
}

// One lane per subgrid sample (_populate_subgrid_texture_*_kernel, :363-566, _write_subgrid_slot, :172-195):
// a required subgrid's samples go to texture block `address` (x fastest in a blocks^3 arrangement of 9^3
// blocks), quantised for unorm formats as trunc(clamp((v - qmin) qinv, 0, 1) * (2^bits - 1)); sample 0 of
// every subgrid writes its slot (block coordinates, SLOT_LINEAR for a demoted subgrid, SLOT_EMPTY otherwise).
__global__ void __launch_bounds__(256) sdf_pack_kernel(FineGrid F, int w, int h, int d, const int* __restrict__ flags,
                                                       const int* __restrict__ address, int blocks, int ts, int fmt,
                                                       float qmin, float qinv, unsigned* __restrict__ slots,
                                                       void* __restrict__ texels) {
  const long long total = (long long)w * h * d * 729;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    const int s = (int)(t / 729), k = (int)(t - (long long)s * 729);
    const int bz = s / (w * h), r = s - bz * w * h, by = r / w, bx = r - by * w;
    const int f = flags[s];
    const int a = address[s];
    const int az = a / (blocks * blocks), ar = a - az * blocks * blocks, ay = ar / blocks, ax = ar - ay * blocks;
    if (k == 0) {
      const unsigned slot = (f & 4) ? ((unsigned)ax | ((unsigned)ay << 10) | ((unsigned)az << 20))
                                    : ((f & 2) ? SDF_SLOT_LINEAR : SDF_SLOT_EMPTY);
      slots[((long long)bx * h + by) * d + bz] = slot;
    }
    if (!(f & 4)) continue;
    const int lz = k / 81, rr = k - lz * 81, ly = rr / 9, lx = rr - ly * 9;
    const float v = F.at(bx * SDF_SUBGRID + lx, by * SDF_SUBGRID + ly, bz * SDF_SUBGRID + lz);
    const long long ti = ((long long)(az * 9 + lz) * ts + (ay * 9 + ly)) * ts + (ax * 9 + lx);
    if (fmt == 4) {
      static_cast<float*>(texels)[ti] = v;
    } else {
      const float vn = clampf((v - qmin) * qinv, 0.0f, 1.0f);
      if (fmt == 2)
        static_cast<unsigned short*>(texels)[ti] = (unsigned short)(vn * 65535.0f);
      else
        static_cast<unsigned char*>(texels)[ti] = (unsigned char)(vn * 255.0f);
    }
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
  xform XT;      // T world transform
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
  e.XT = T.X;
  e.tscale = T.gscale;
  e.threshold = gap_sum + T.margin + S.margin;
  e.threshold_u = e.threshold / e.min_s;
  e.edge = d.mesh_edge + 2 * d.mesh_edge_start[tri_mesh];
  e.vtx = d.mesh_vertex + 3 * d.mesh_vertex_start[tri_mesh];
  return e;
}

// The edge's end points in the SDF side's unscaled frame.
__device__ __forceinline__ void edge_points(const EdgeSide& E, int e, f3& ea, f3& eb) {
  const int ia = E.edge[2 * e], ib = E.edge[2 * e + 1];
  f3 pa = load3(E.vtx + 3 * ia), pb = load3(E.vtx + 3 * ib);
  pa = f3{pa.x * E.tscale.x, pa.y * E.tscale.y, pa.z * E.tscale.z};
  pb = f3{pb.x * E.tscale.x, pb.y * E.tscale.y, pb.z * E.tscale.z};
  ea = xpoint(E.X_ts, pa);
  eb = xpoint(E.X_ts, pb);
  ea = f3{ea.x * E.inv_s.x, ea.y * E.inv_s.y, ea.z * E.inv_s.z};
  eb = f3{eb.x * E.inv_s.x, eb.y * E.inv_s.y, eb.z * E.inv_s.z};
}

// The culling of edge e (sdf_contact.py:1010-1045): its bounding sphere against the SDF box, then the SDF at
// its midpoint.  True when the edge goes on to the search; msdf is the midpoint value the search starts from.
__device__ __forceinline__ bool edge_cull(const EdgeSide& E, int e, float& msdf) {
  f3 ea, eb;
  edge_points(E, e, ea, eb);
  const f3 mid = (ea + eb) * 0.5f;
  const float rad = len(eb - ea) * 0.5f;
  const float cull = rad + E.threshold_u;
  const SdfGrid& G = E.G;
  const f3 cl{clampf(mid.x, G.lower.x, G.upper.x), clampf(mid.y, G.lower.y, G.upper.y),
              clampf(mid.z, G.lower.z, G.upper.z)};
  if (len_sq(mid - cl) > cull * cull) return false;
  msdf = sdf_sample(G, mid);
  return msdf <= cull;
}

// The rest of edge e once it passed the culling (sdf_contact.py:1045-1150): Brent search from the midpoint
// value, threshold, gradient normal.  Returns false when the edge yields no contact; else the contact's world
// point, its normal (from shape0 to shape1) and distance.
__device__ bool edge_finish(const EdgeSide& E, int mode, int e, float msdf, f3& center, f3& normal, float& dist) {
  f3 ea, eb;
  edge_points(E, e, ea, eb);
  const SdfGrid& G = E.G;
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

// Edge e of one side (sdf_contact.py:1010-1150): culling, Brent search, gradient normal.
__device__ __forceinline__ bool edge_contact(const EdgeSide& E, int mode, int e, f3& center, f3& normal,
                                             float& dist) {
  float msdf;
  if (!edge_cull(E, e, msdf)) return false;
  return edge_finish(E, mode, e, msdf, center, normal, dist);
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

__device__ __forceinline__ int mesh_block_scan(bool f, int* s_w, int& total) {
  const unsigned long long b = __ballot(f);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int pre = __popcll(b & ((1ull << lane) - 1ull));
  if (lane == 0) s_w[w] = __popcll(b);
  __syncthreads();
  int base = 0;
  total = 0;
  for (int q = 0; q < MESH_BLOCK / 64; ++q) {
    base += q < w ? s_w[q] : 0;
    total += s_w[q];
  }
  __syncthreads();
  return base + pre;
}

// The reduction frame of one side of a pair: where its candidates compete (contact_reduction_global.py
// :968-1059 for mesh-plane, :1092-1227 for mesh-mesh).
struct RedFrame {
  bool centered;  // mesh-mesh: project relative to the shapes' midpoint, fixed beta
  f3 midpoint;
  float beta;     // depth below which a candidate competes in the scan slots
  xform X_inv;    // world -> frame of the voxel shape
  f3 lo, hi;      // voxel shape's local collision AABB
  int res[3];
};

__device__ __forceinline__ RedFrame red_frame(const NmcModelDesc& d, int shape, const xform& X, bool centered,
                                              f3 midpoint) {
  RedFrame f;
  f.centered = centered;
  f.midpoint = midpoint;
  f.lo = load3(d.shape_aabb_lower + 3 * shape);
  f.hi = load3(d.shape_aabb_upper + 3 * shape);
  f.beta = centered ? RED_BETA : RED_BETA * len(f.hi - f.lo);
  f.X_inv = xinverse(X);
  f.res[0] = d.shape_voxel_resolution[3 * shape];
  f.res[1] = d.shape_voxel_resolution[3 * shape + 1];
  f.res[2] = d.shape_voxel_resolution[3 * shape + 2];
  return f;
}

__device__ __forceinline__ void red_candidate(unsigned long long* slots, const RedParams& rp, const RedFrame& f,
                                              f3 center, f3 normal, float dist, int sub) {
  // mesh-plane candidates are binned by their stored (octahedral) normal, mesh-mesh ones by the computed one
  const f3 bn = f.centered ? normal : red_oct_roundtrip(normal);
  const f3 proj = f.centered ? center - f.midpoint : center;
  const int vox = red_voxel(xpoint(f.X_inv, center), f.lo, f.hi, f.res[0], f.res[1], f.res[2]);
  red_insert(slots, rp, bn, proj, dist, dist < f.beta, vox, sub);
}

template <bool REDUCE>
__global__ void __launch_bounds__(MESH_BLOCK) mesh_narrow_kernel(NmcModelDesc d, const float* __restrict__ body_q,
                                                                 StageBuffers st,
                                                                 const long long* __restrict__ seg_pair_off,
                                                                 const int* __restrict__ seg_pair_count,
                                                                 const int* __restrict__ seg_mesh_count,
                                                                 const int* __restrict__ seg_mesh_off, int segments,
                                                                 int cand_cap, int mesh_k, RedParams rp,
                                                                 int* __restrict__ counters) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int s_n;
  __shared__ int s_w[MESH_BLOCK / 64];
  __shared__ unsigned s_slot[2][MESH_SLOT_LDS];  // [mode]: slots of the SDF that side samples
  __shared__ int s_nlist;
  __shared__ int s_list[MESH_EDGE_LIST];           // edges of one side that passed the culling
  __shared__ float s_lmsdf[MESH_EDGE_LIST];        // and their midpoint SDF values
  // reduced: the 240 slots, then 256 winner sub keys; unreduced: cand_cap candidate sub keys
  unsigned long long* slots = reinterpret_cast<unsigned long long*>(smem);
  unsigned* sel = reinterpret_cast<unsigned*>(smem + (rp.mode ? sizeof(unsigned long long) * 256 : 0));
  const int total = st.route_count[2];
  constexpr bool reduce = REDUCE;
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
    if (reduce)
      for (int k = threadIdx.x; k < RED_SLOTS; k += blockDim.x) slots[k] = 0ull;
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
    int staged = 0;  // bit m: side m's slot table is in LDS
    if (!plane) {
      for (int mode = 0; mode < 2; ++mode) {
        const int sm = mode == 0 ? mb : ma;
        const int* dm = d.mesh_sdf_dims + 5 * sm;
        const int ns = dm[0] * dm[1] * dm[2];
        if (ns <= MESH_SLOT_LDS) {
          staged |= 1 << mode;
          const unsigned* src = d.mesh_sdf_slots + d.mesh_sdf_start[3 * sm + 1];
          for (int k = threadIdx.x; k < ns; k += blockDim.x) s_slot[mode][k] = src[k];
        }
      }
      __syncthreads();
    }
    // the edges of one side against the other side's SDF, its slot table in LDS when staged
    auto side = [&](int mode) {
      EdgeSide E = mode == 0 ? edge_side(d, A, B, ma, mb, gap_sum) : edge_side(d, B, A, mb, ma, gap_sum);
      if ((staged >> mode) & 1) E.G.slots = s_slot[mode];
      return E;
    };
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
    // ---- pass 1: every candidate into the reduction table, or its sub key into the list when the writer keeps it
    auto push = [&](int sub, f3 center, f3 normal, float dist) {
      c.center = center;
      c.normal = normal;
      c.dist = dist;
      if (!passes_gap(c)) return;
      const int k = atomicAdd(&s_n, 1);
      if (k < cand_cap) sel[k] = static_cast<unsigned>(sub);
    };
    if (plane) {
      const int nv = d.mesh_vertex_start[ma + 1] - d.mesh_vertex_start[ma];
      RedFrame F;
      if (reduce) F = red_frame(d, sa, A.X, false, f3{0.0f, 0.0f, 0.0f});
      for (int v = threadIdx.x; v < nv; v += blockDim.x) {
        f3 center, normal;
        float dist;
        if (!vertex_contact(P, v, center, normal, dist)) continue;
        if (reduce) red_candidate(slots, rp, F, center, normal, dist, v);
        else push(v, center, normal, dist);
      }
    } else {
      // Edges that pass the culling are listed in LDS first and searched from the list, so that the lanes
      // of a wave run the (long) Brent search together instead of waiting on the few of them that do.
      for (int mode = 0; mode < 2; ++mode) {
        const int tm = mode == 0 ? ma : mb;
        const EdgeSide E = side(mode);
        RedFrame F;
        if (reduce) F = red_frame(d, mode == 0 ? sa : sb, E.XT, true, (E.XT.p + E.XS.p) * 0.5f);
        const int ne = d.mesh_edge_start[tm + 1] - d.mesh_edge_start[tm];
        for (int base = 0; base < ne; base += MESH_EDGE_LIST) {
          const int end = ne < base + MESH_EDGE_LIST ? ne : base + MESH_EDGE_LIST;
          if (threadIdx.x == 0) s_nlist = 0;
          __syncthreads();
          for (int e = base + threadIdx.x; e < end; e += blockDim.x) {
            float msdf;
            const bool pass = edge_cull(E, e, msdf);
            const unsigned long long bal = __ballot(pass);
            const int lane = threadIdx.x & 63;
            int at = 0;
            if (lane == 0 && bal) at = atomicAdd(&s_nlist, __popcll(bal));
            at = __shfl(at, 0, 64);
            if (pass) {
              const int k = at + __popcll(bal & ((1ull << lane) - 1ull));
              s_list[k] = e;
              s_lmsdf[k] = msdf;
            }
          }
          __syncthreads();
          const int nl = s_nlist;
          for (int k = threadIdx.x; k < nl; k += blockDim.x) {
            const int e = s_list[k];
            f3 center, normal;
            float dist;
            if (!edge_finish(E, mode, e, s_lmsdf[k], center, normal, dist)) continue;
            const int sub = (e << 2) | (mode << 1);
            if (reduce) red_candidate(slots, rp, F, center, normal, dist, sub);
            else push(sub, center, normal, dist);
          }
          __syncthreads();
        }
      }
    }
    __syncthreads();
    // ---- the sub keys to write, sorted: the list, or the table's winners (each once)
    int n, n2;
    if (reduce) {
      for (int k = threadIdx.x; k < 256; k += blockDim.x) {
        const unsigned long long v = k < RED_SLOTS ? slots[k] : 0ull;
        sel[k] = v ? static_cast<unsigned>(red_winner(v, rp.mode)) : 0xffffffffu;
      }
      n = 256;
      n2 = 256;
    } else {
      n = s_n;
      if (n > cand_cap) {
        if (threadIdx.x == 0) atomicAdd(&counters[1], n - cand_cap);
        n = cand_cap;
      }
      n2 = 1;
      while (n2 < n) n2 <<= 1;
      for (int k = n + threadIdx.x; k < n2; k += blockDim.x) sel[k] = 0xffffffffu;
    }
    __syncthreads();
    mesh_bitonic(sel, n2);
    // ---- pass 2: evaluate each again; reduced: first of equal sub keys, octahedral normal, writer test
    const WriteCtx w{d.shape_gap, st, st.mesh_slot_base + static_cast<long long>(i) * mesh_k, 0};
    int written = 0;
    for (int k0 = 0; k0 < n; k0 += blockDim.x) {
      const int k = k0 + threadIdx.x;
      const unsigned sub_u = k < n ? sel[k] : 0xffffffffu;
      bool live = sub_u != 0xffffffffu && (!reduce || k == 0 || sel[k - 1] != sub_u);
      CData o = c;
      f3 nn, aw, bw;
      if (live) {
        const int sub = static_cast<int>(sub_u);
        f3 center, normal;
        float dist;
        if (plane) {
          (void)vertex_contact(P, sub, center, normal, dist);
        } else {
          const int mode = (sub >> 1) & 1;
          const EdgeSide E = side(mode);
          (void)edge_contact(E, mode, sub >> 2, center, normal, dist);
        }
        o.center = center;
        o.normal = reduce ? red_oct_roundtrip(normal) : normal;
        o.dist = dist;
        o.sub_key = sub;
        const float dd = contact_d(o, nn, aw, bw);
        if (reduce) live = dd <= o.gap_sum;
      }
      int cnt;
      const int at = mesh_block_scan(live, s_w, cnt);
      if (live && written + at < mesh_k) write_record(o, nn, aw, bw, w, w.slot0 + written + at);
      written += cnt;
    }
    if (threadIdx.x == 0) {
      st.mask[g] = static_cast<unsigned short>(written < mesh_k ? written : mesh_k);
      st.mesh_index[g] = i;
    }
    __syncthreads();
  }
}

size_t mesh_lds_bytes(int cand_cap, bool reduce) {
  return reduce ? sizeof(unsigned long long) * 256 + sizeof(unsigned) * 256 : sizeof(unsigned) * (size_t)cand_cap;
}

hipError_t launch_mesh_narrow(NmcModelImpl* m, const float* body_q, hipStream_t stream) {
  if (m->stage.mesh_pair_capacity <= 0) return hipSuccess;
  const long long cap = m->stage.mesh_pair_capacity;
  const int blocks = static_cast<int>(cap < m->mesh_max_blocks ? cap : m->mesh_max_blocks);
  if (m->red.mode != 0)
    hipLaunchKernelGGL(mesh_narrow_kernel<true>, dim3(blocks), dim3(MESH_BLOCK), m->mesh_lds, stream, m->d, body_q,
                       m->stage, m->seg_pair_off, m->seg_pair_count, m->seg_mesh_count, m->seg_mesh_off,
                       m->d.segment_count, m->mesh_cand_cap, m->mesh_k, m->red, m->counters);
  else
    hipLaunchKernelGGL(mesh_narrow_kernel<false>, dim3(blocks), dim3(MESH_BLOCK), m->mesh_lds, stream, m->d, body_q,
                       m->stage, m->seg_pair_off, m->seg_pair_count, m->seg_mesh_count, m->seg_mesh_off,
                       m->d.segment_count, m->mesh_cand_cap, m->mesh_k, m->red, m->counters);
  return hipGetLastError();
}

}  // namespace nmc

using namespace nmc;

extern "C" int nmc_sdf_build(const float* vertices, int vertex_count, const int* triangles, int triangle_count,
                             const float lower[3], const float cell[3], const int vertices_per_axis[3],
                             float winding_threshold, float* out, nmc_stream_t stream) {
  if (!vertices || !triangles || !lower || !cell || !vertices_per_axis || !out)
    return nmc::fail("nmc_sdf_build: null argument");
  if (vertex_count <= 0 || triangle_count <= 0) return nmc::fail("nmc_sdf_build: empty mesh");
  if (!(cell[0] > 0.0f) || !(cell[1] > 0.0f) || !(cell[2] > 0.0f) || vertices_per_axis[0] < 2 ||
      vertices_per_axis[1] < 2 || vertices_per_axis[2] < 2)
    return nmc::fail("nmc_sdf_build: cell sizes must be positive and every axis needs two vertices");
  const long long total = (long long)vertices_per_axis[0] * vertices_per_axis[1] * vertices_per_axis[2];
  long long blocks = (total + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(sdf_build_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), vertices, vertex_count, triangles, triangle_count,
                     f3{lower[0], lower[1], lower[2]}, f3{cell[0], cell[1], cell[2]}, vertices_per_axis[0],
                     vertices_per_axis[1], vertices_per_axis[2], winding_threshold, out);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : nmc::hip_fail(e, "nmc_sdf_build launch");
}

extern "C" int nmc_sdf_subgrids(const float* fine, const int subgrids[3], float band_lo, float band_hi,
                                float linear_threshold, float* coarse, int* flags, int* address, int* count,
                                nmc_stream_t stream) {
  if (!fine || !subgrids || !coarse || !flags || !address || !count) return nmc::fail("nmc_sdf_subgrids: null argument");
  const int w = subgrids[0], h = subgrids[1], d = subgrids[2];
  if (w <= 0 || h <= 0 || d <= 0 || w > 1024 || h > 1024 || d > 1024 || (long long)w * h * d > (1LL << 30))
    return nmc::fail("nmc_sdf_subgrids: subgrid counts must lie in [1, 1024]");
  const hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const FineGrid F{fine, w * SDF_SUBGRID + 1, h * SDF_SUBGRID + 1};
  const int nc = (w + 1) * (h + 1) * (d + 1), n = w * h * d;
  hipLaunchKernelGGL(sdf_coarse_kernel, dim3(static_cast<unsigned>(std::min(4096, (nc + 255) / 256))), dim3(256), 0,
                     st, F, w, h, d, coarse);
  hipLaunchKernelGGL(sdf_subgrid_kernel, dim3(static_cast<unsigned>(std::min(n, 65535))), dim3(256), 0, st, F, w,
                     h, d, band_lo, band_hi, linear_threshold, flags);
  hipLaunchKernelGGL(sdf_address_kernel, dim3(1), dim3(SCAN_BLOCK), 0, st, flags, n, address, count);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : nmc::hip_fail(e, "nmc_sdf_subgrids launch");
}

extern "C" int nmc_sdf_pack(const float* fine, const int subgrids[3], const int* flags, const int* address,
                            int blocks_per_axis, int texel_bytes, float quant_min, float quant_range_inv,
                            uint32_t* slots, void* texels, nmc_stream_t stream) {
  if (!fine || !subgrids || !flags || !address || !slots || !texels) return nmc::fail("nmc_sdf_pack: null argument");
  const int w = subgrids[0], h = subgrids[1], d = subgrids[2];
  if (w <= 0 || h <= 0 || d <= 0 || w > 1024 || h > 1024 || d > 1024 || (long long)w * h * d > (1LL << 30))
    return nmc::fail("nmc_sdf_pack: subgrid counts must lie in [1, 1024]");
  if (texel_bytes != 1 && texel_bytes != 2 && texel_bytes != 4) return nmc::fail("nmc_sdf_pack: texel_bytes must be 1, 2 or 4");
  if (blocks_per_axis < 1 || blocks_per_axis > 1024) return nmc::fail("nmc_sdf_pack: blocks_per_axis must lie in [1, 1024]");
  const FineGrid F{fine, w * SDF_SUBGRID + 1, h * SDF_SUBGRID + 1};
  const long long total = (long long)w * h * d * 729;
  const long long blocks = std::min<long long>((total + 255) / 256, 16384);
  hipLaunchKernelGGL(sdf_pack_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), F, w, h, d, flags, address, blocks_per_axis,
                     blocks_per_axis * 9, texel_bytes, quant_min, quant_range_inv, slots, texels);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : nmc::hip_fail(e, "nmc_sdf_pack launch");
}
