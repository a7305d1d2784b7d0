// Standalone broad phases over caller AABBs for gfx950: newton.geometry BroadPhaseSAP, BroadPhaseAllPairs
// and BroadPhaseExplicit (newton/_src/geometry/broad_phase_sap.py:397-677, broad_phase_nxn.py:217-472).
//
// Output: candidate_pair (max x 2 int32, canonical (min, max) shape ids) and candidate_pair_count[0] = the
// number of pairs found, which may exceed max (pairs past it are dropped, as write_pair does,
// broad_phase_common.py:116-128).  The reference appends through one global atomic counter, so its order
// depends on scheduling; every launch here writes one fixed order, with no atomics:
//   all pairs  segment by segment; inside a segment row i, then column j (i < j, index-map order)
//   SAP        segment by segment; inside a segment as all pairs, except segments of BP_ENUM_MAX..BP_LDS_MAX
//              shapes: by sweep origin in sorted order, then sweep position
//   explicit   input order
// Each call is three launches: count (one workgroup per segment, or per block of input pairs), an exclusive
// scan of the counts by one workgroup, and write (the workgroup counts again, scans its lanes' counts and
// writes at its offset).
//
// Pair tests, in the reference's order of meaning (broad_phase_nxn.py:161-214, broad_phase_sap.py:300-394):
// distinct shapes; both world -1 only in the last (dedicated) segment; test_world_and_group_pair; AABB overlap
// with cutoff gap_a + gap_b (check_aabb_overlap, broad_phase_common.py:18-35); not in the sorted filter list.
// SAP adds its sweep: each segment is sorted by the projection of its AABBs on the fixed axis (ties keep
// index-map order, the stable segmented sort) and element a sweeps the elements b after it while
// lower(b) < upper(a) (broad_phase_sap.py:189-250).  A pair is therefore found iff, with a the earlier of
// the two in that order, lower(b) < upper(a) and the tests pass.  Segments of up to BP_ENUM_MAX shapes
// evaluate that predicate on every pair (all 256 lanes busy, no sort: faster than a sort and a serial sweep
// at these sizes); segments up to BP_LDS_MAX sort in LDS and sweep; larger ones enumerate again, reading
// global memory (the same set, O(n^2) work).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "broad_common.h"
#include "nmc_internal.h"
#include "nmc_math.h"
#include "scan.h"

namespace nmc {

constexpr int BP_LDS_MAX = 2048;         // shapes per segment held in LDS (56 B each for SAP)
constexpr int BP_ENUM_MAX = 256;         // SAP segments up to this size enumerate their pairs (no sort)
constexpr int BP_MAX_BLOCKS = 4096;

struct BpArgs {
  int lds_n;           // segments of up to lds_n shapes run out of LDS (a power of two, <= BP_LDS_MAX)
  const float* lower;  // S x 3
  const float* upper;  // S x 3
  const float* gap;    // S, or null (0)
  const int* group;
  const int* world;
  const int* filter;   // F x 2, lexicographically sorted (min, max) pairs, or null
  int filter_count;
  const int* index_map;
  const int* seg_start;  // segments + 1
  int segments;
  int sap;
  int* seg_count;
  int* seg_off;
  int* total;
  int* out_pairs;
  int max_pairs;
  int* out_count;
};

struct BpShape {
  f3 lo, hi;
  float gap;
  int world, group, id;
};

__device__ __forceinline__ BpShape bp_load(const BpArgs& a, int id) {
  BpShape s;
  s.lo = load3(a.lower + 3 * id);
  s.hi = load3(a.upper + 3 * id);
  s.gap = a.gap ? a.gap[id] : 0.0f;
  s.world = a.world[id];
  s.group = a.group[id];
  s.id = id;
  return s;
}

// The tests every broad phase applies to the pair of two segment members.
__device__ __forceinline__ bool bp_pair_ok(const BpArgs& a, const BpShape& p, const BpShape& q, bool dedicated,
                                           int& s1, int& s2) {
  if (p.id == q.id) return false;
  const BpShape& A = p.id < q.id ? p : q;
  const BpShape& B = p.id < q.id ? q : p;
  s1 = A.id;
  s2 = B.id;
  if (A.world == -1 && B.world == -1 && !dedicated) return false;
  if (A.world != -1 && B.world != -1 && A.world != B.world) return false;
  if (!group_pair(A.group, B.group)) return false;
  const float c = A.gap + B.gap;
  const bool ov = A.lo.x <= B.hi.x + c && A.hi.x >= B.lo.x - c && A.lo.y <= B.hi.y + c && A.hi.y >= B.lo.y - c &&
                  A.lo.z <= B.hi.z + c && A.hi.z >= B.lo.z - c;
  if (!ov) return false;
  if (a.filter_count > 0 && pair_listed(a.filter, a.filter_count, s1, s2)) return false;
  return true;
}

// Projection of one AABB on the sweep axis (_sap_project_aabb, broad_phase_sap.py:43-63).
__device__ __forceinline__ void bp_project(const BpShape& s, f3 dir, float& plo, float& phi) {
  f3 half = 0.5f * (s.hi - s.lo);
  half = f3{half.x + s.gap, half.y + s.gap, half.z + s.gap};
  const float radius = dot(dir, half);
  const float center = dot(dir, 0.5f * (s.lo + s.hi));
  plo = center - radius;
  phi = center + radius;
}

__device__ __forceinline__ unsigned long long bp_order_key(float plo, int l) {
  return (static_cast<unsigned long long>(sortable_float(plo)) << 32) | static_cast<unsigned>(l);
}

// Exclusive scan of one value per lane over the CBLOCK-lane workgroup; returns the workgroup total.
__device__ __forceinline__ int bp_block_scan(int v, int& excl, int* s_wave) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int inc = v;
  for (int off = 1; off < 64; off <<= 1) {
    const int t = __shfl_up(inc, off, 64);
    if (lane >= off) inc += t;
  }
  if (lane == 63) s_wave[w] = inc;
  __syncthreads();
  int before = 0, total = 0;
  for (int i = 0; i < CBLOCK / 64; ++i) {
    if (i < w) before += s_wave[i];
    total += s_wave[i];
  }
  __syncthreads();
  excl = before + inc - v;
  return total;
}

__device__ __forceinline__ void bp_emit(const BpArgs& a, long long at, int s1, int s2) {
  if (at < a.max_pairs) {
    a.out_pairs[2 * at] = s1;
    a.out_pairs[2 * at + 1] = s2;
  }
}

// ---------------------------------------------------------------------------------------- pair enumeration
// Lane L takes the contiguous run [L c, (L + 1) c) of the segment's n(n-1)/2 pairs in row order.  SAP adds
// the sweep predicate: the earlier of the two in sort order (projected lower bound, then index-map position)
// sweeps the later one, so the pair counts when lower(later) < upper(earlier).  Shapes (and projections)
// come from LDS when the segment fits, else from global memory.
__device__ __forceinline__ long long bp_row_start(long long i, long long n) { return i * (2 * n - i - 1) / 2; }

struct BpMember {
  BpShape s;
  float lo, hi;  // projection (SAP)
};

template <bool SAP>
__device__ __forceinline__ BpMember bp_member(const BpArgs& a, int s0, int l, bool in_lds, const BpShape* sh,
                                              const float* plo, const float* phi, f3 dir) {
  BpMember m;
  if (in_lds) {
    m.s = sh[l];
    if (SAP) {
      m.lo = plo[l];
      m.hi = phi[l];
    }
  } else {
    m.s = bp_load(a, a.index_map[s0 + l]);
    if (SAP) bp_project(m.s, dir, m.lo, m.hi);
  }
  return m;
}

template <bool WRITE, bool SAP>
__device__ int bp_enum_segment(const BpArgs& a, int seg, const BpShape* sh, const float* plo, const float* phi,
                               int* s_wave, long long out_base) {
  const int s0 = a.seg_start[seg];
  const int n = a.seg_start[seg + 1] - s0;
  const bool dedicated = seg == a.segments - 1;
  const bool in_lds = n <= a.lds_n;
  const f3 dir = normalize(f3{0.5935f, 0.7790f, 0.1235f});
  const long long P = static_cast<long long>(n) * (n - 1) / 2;
  const long long chunk = (P + CBLOCK - 1) / CBLOCK;
  const long long p0 = chunk * threadIdx.x;
  const long long p1 = p0 + chunk < P ? p0 + chunk : P;
  int i = 0, j = 0;
  if (p0 < p1) {
    int lo = 0, hi = n - 1;  // last row whose start is <= p0
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (bp_row_start(mid, n) <= p0) lo = mid; else hi = mid - 1;
    }
    i = lo;
    j = static_cast<int>(p0 - bp_row_start(i, n)) + i + 1;
  }
  int found = 0;
  for (int pass = 0; pass < (WRITE ? 2 : 1); ++pass) {
    int ii = i, jj = j, k = 0;
    long long at = 0;
    if (WRITE && pass == 1) {
      int ex;
      bp_block_scan(found, ex, s_wave);
      at = out_base + ex;
    }
    if (p0 < p1) {
      BpMember A = bp_member<SAP>(a, s0, ii, in_lds, sh, plo, phi, dir);
      for (long long p = p0; p < p1; ++p) {
        const BpMember B = bp_member<SAP>(a, s0, jj, in_lds, sh, plo, phi, dir);
        bool take = true;
        if (SAP) {
          const bool a_first = bp_order_key(A.lo, ii) < bp_order_key(B.lo, jj);
          take = a_first ? (B.lo < A.hi) : (A.lo < B.hi);
        }
        int s1, s2;
        if (take && bp_pair_ok(a, A.s, B.s, dedicated, s1, s2)) {
          if (WRITE && pass == 1) bp_emit(a, at + k, s1, s2);
          ++k;
        }
        if (++jj == n) {
          ++ii;
          jj = ii + 1;
          if (ii < n - 1) A = bp_member<SAP>(a, s0, ii, in_lds, sh, plo, phi, dir);
        }
      }
    }
    found = k;
  }
  return found;
}

// ---------------------------------------------------------------------------------------- SAP
// LDS path: sorted order in keys[]; cnt[] = pairs found by each sweep origin.
template <bool WRITE>
__device__ int bp_sap_segment_lds(const BpArgs& a, int seg, const BpShape* sh, const float* plo, const float* phi,
                                  const unsigned long long* keys, int* s_wave, long long out_base) {
  const int s0 = a.seg_start[seg];
  const int n = a.seg_start[seg + 1] - s0;
  (void)s0;
  const bool dedicated = seg == a.segments - 1;
  if (!WRITE) {
    // lane L sweeps origins L, L + CBLOCK, ...
    int mine = 0;
    for (int r = threadIdx.x; r < n; r += CBLOCK) {
      const int la = static_cast<int>(keys[r] & 0xffffffffu);
      const float up = phi[la];
      for (int q = r + 1; q < n; ++q) {
        const int lb = static_cast<int>(keys[q] & 0xffffffffu);
        if (!(plo[lb] < up)) break;
        int s1, s2;
        if (bp_pair_ok(a, sh[la], sh[lb], dedicated, s1, s2)) ++mine;
      }
    }
    return mine;
  }
  // the output order is origin order: scan the lanes' counts per round of CBLOCK origins
  long long base = out_base;
  for (int r0 = 0; r0 < n; r0 += CBLOCK) {
    const int r = r0 + threadIdx.x;
    int c = 0;
    int la = 0;
    float up = 0.0f;
    if (r < n) {
      la = static_cast<int>(keys[r] & 0xffffffffu);
      up = phi[la];
      for (int q = r + 1; q < n; ++q) {
        const int lb = static_cast<int>(keys[q] & 0xffffffffu);
        if (!(plo[lb] < up)) break;
        int s1, s2;
        if (bp_pair_ok(a, sh[la], sh[lb], dedicated, s1, s2)) ++c;
      }
    }
    int ex;
    const int round_total = bp_block_scan(c, ex, s_wave);
    if (r < n && c > 0) {
      long long at = base + ex;
      for (int q = r + 1; q < n; ++q) {
        const int lb = static_cast<int>(keys[q] & 0xffffffffu);
        if (!(plo[lb] < up)) break;
        int s1, s2;
        if (bp_pair_ok(a, sh[la], sh[lb], dedicated, s1, s2)) bp_emit(a, at++, s1, s2);
      }
    }
    base += round_total;
  }
  return 0;
}

// One workgroup per segment (grid-stride).  WRITE = false: pairs per segment into seg_count; WRITE = true:
// pairs at seg_off[seg].
template <bool WRITE>
__global__ void __launch_bounds__(CBLOCK) bp_segment_kernel(BpArgs a) {
  extern __shared__ __align__(16) unsigned char bp_smem[];
  __shared__ int s_wave[CBLOCK / 64];
  __shared__ int s_sum;
  // LDS: keys (SAP), shapes, projections
  unsigned long long* keys = reinterpret_cast<unsigned long long*>(bp_smem);
  BpShape* sh = reinterpret_cast<BpShape*>(keys + (a.sap ? a.lds_n : 0));
  float* plo = reinterpret_cast<float*>(sh + a.lds_n);
  float* phi = plo + (a.sap ? a.lds_n : 0);
  const f3 dir = normalize(f3{0.5935f, 0.7790f, 0.1235f});
  for (int seg = blockIdx.x; seg < a.segments; seg += gridDim.x) {
    const int s0 = a.seg_start[seg];
    const int n = a.seg_start[seg + 1] - s0;
    const bool in_lds = n <= a.lds_n;
    // SAP segments above BP_ENUM_MAX shapes held in LDS sort and sweep; every other segment enumerates pairs
    const bool sweep = a.sap && in_lds && n > BP_ENUM_MAX;
    if (in_lds) {
      for (int l = threadIdx.x; l < n; l += CBLOCK) {
        const BpShape s = bp_load(a, a.index_map[s0 + l]);
        sh[l] = s;
        if (a.sap) {
          float lo, hi;
          bp_project(s, dir, lo, hi);
          plo[l] = lo;
          phi[l] = hi;
          keys[l] = bp_order_key(lo, l);
        }
      }
      if (sweep) {
        int n2 = 1;
        while (n2 < n) n2 <<= 1;
        for (int l = n + threadIdx.x; l < n2; l += CBLOCK) keys[l] = ~0ull;
        __syncthreads();
        bitonic_sort(keys, n2);
      }
      __syncthreads();
    }
    const long long base = WRITE ? a.seg_off[seg] : 0;
    int c;
    if (sweep) {
      c = bp_sap_segment_lds<WRITE>(a, seg, sh, plo, phi, keys, s_wave, base);
    } else if (a.sap) {
      c = bp_enum_segment<WRITE, true>(a, seg, sh, plo, phi, s_wave, base);
    } else {
      c = bp_enum_segment<WRITE, false>(a, seg, sh, plo, phi, s_wave, base);
    }
    if (!WRITE) {
      if (threadIdx.x == 0) s_sum = 0;
      __syncthreads();
      atomicAdd(&s_sum, c);  // LDS only
      __syncthreads();
      if (threadIdx.x == 0) a.seg_count[seg] = s_sum;
    }
    __syncthreads();
  }
}

size_t bp_lds_bytes(int lds_n, bool sap) {
  return static_cast<size_t>(lds_n) * (sizeof(BpShape) + (sap ? 2 * sizeof(float) + sizeof(unsigned long long) : 0));
}

// Exclusive scan of the per-segment (or per-block) counts into write offsets.  append != 0: the pairs go after
// the out_count[0] pairs already in the output (a caller that skips the zeroing of the count and chains
// launches into one buffer), and the count becomes that base plus the pairs found; a negative count on entry
// counts as 0.  The base is read into LDS before any lane writes out_count.
__global__ void __launch_bounds__(SCAN_BLOCK) bp_scan_kernel(const int* __restrict__ count, int n,
                                                             int* __restrict__ off, int* __restrict__ total,
                                                             int* __restrict__ out_count, int append) {
  __shared__ int s_base;
  if (threadIdx.x == 0) {
    const int b = append ? out_count[0] : 0;
    s_base = b > 0 ? b : 0;
  }
  __syncthreads();
  const int base = s_base;
  int t[1];
  segment_scan_fn<1>(
      n, [&](int i, int (&v)[1]) { v[0] = count[i]; },
      [&](int i, const int (&ex)[1], const int (&)[1]) { off[i] = base + ex[0]; }, t);
  if (threadIdx.x == 0) {
    const long long sum = static_cast<long long>(base) + t[0];
    const int tot = sum > 0x7fffffffLL ? 0x7fffffff : static_cast<int>(sum);
    total[0] = tot;
    out_count[0] = tot;
  }
}

// ---------------------------------------------------------------------------------------- explicit
// Block b owns input pairs [b B, (b + 1) B), B = CBLOCK x BP_EXPLICIT_ITEMS; lane L its items
// [b B + L I, b B + (L + 1) I).
struct BpExplicitArgs {
  int shape_count;
  const float* lower;
  const float* upper;
  const float* gap;
  const int* pairs;
  int pair_count;
  int* block_count;
  int* block_off;
  int* total;
  int* out_pairs;
  int max_pairs;
  int* out_count;
};

__device__ __forceinline__ bool bp_explicit_ok(const BpExplicitArgs& a, int e, int& s1, int& s2) {
  s1 = a.pairs[2 * e];
  s2 = a.pairs[2 * e + 1];
  // an index outside the AABB arrays is never read; such a pair is not a candidate
  if (static_cast<unsigned>(s1) >= static_cast<unsigned>(a.shape_count) ||
      static_cast<unsigned>(s2) >= static_cast<unsigned>(a.shape_count))
    return false;
  const f3 lo1 = load3(a.lower + 3 * s1), hi1 = load3(a.upper + 3 * s1);
  const f3 lo2 = load3(a.lower + 3 * s2), hi2 = load3(a.upper + 3 * s2);
  const float c = (a.gap ? a.gap[s1] : 0.0f) + (a.gap ? a.gap[s2] : 0.0f);
  return lo1.x <= hi2.x + c && hi1.x >= lo2.x - c && lo1.y <= hi2.y + c && hi1.y >= lo2.y - c &&
         lo1.z <= hi2.z + c && hi1.z >= lo2.z - c;
}

template <bool WRITE>
__global__ void __launch_bounds__(CBLOCK) bp_explicit_kernel(BpExplicitArgs a) {
  __shared__ int s_wave[CBLOCK / 64];
  const int blocks = (a.pair_count + CBLOCK * BP_EXPLICIT_ITEMS - 1) / (CBLOCK * BP_EXPLICIT_ITEMS);
  for (int b = blockIdx.x; b < blocks; b += gridDim.x) {
    const int e0 = b * CBLOCK * BP_EXPLICIT_ITEMS + threadIdx.x * BP_EXPLICIT_ITEMS;
    unsigned hit = 0;
    int c = 0;
#pragma unroll
    for (int t = 0; t < BP_EXPLICIT_ITEMS; ++t) {
      int s1, s2;
      if (e0 + t < a.pair_count && bp_explicit_ok(a, e0 + t, s1, s2)) {
        hit |= 1u << t;
        ++c;
      }
    }
    int ex;
    const int total = bp_block_scan(c, ex, s_wave);
    if (!WRITE) {
      if (threadIdx.x == 0) a.block_count[b] = total;
    } else {
      long long at = static_cast<long long>(a.block_off[b]) + ex;
#pragma unroll
      for (int t = 0; t < BP_EXPLICIT_ITEMS; ++t) {
        if (hit & (1u << t)) {
          if (at < a.max_pairs) {
            a.out_pairs[2 * at] = a.pairs[2 * (e0 + t)];
            a.out_pairs[2 * at + 1] = a.pairs[2 * (e0 + t) + 1];
          }
          ++at;
        }
      }
    }
  }
}

}  // namespace nmc

using namespace nmc;

struct NmcBroadPhase {
  int device;
  int mode;  // 0 SAP, 1 all pairs, 2 explicit
  int segments;
  int* index_map;
  int* seg_start;
  int* count;   // per segment (modes 0, 1) or per block of input pairs (mode 2)
  int* off;
  int* total;
  int capacity;  // entries of count / off
  int max_segment;
  int max_index;  // largest shape index in the index map
  int lds_n;
};

static void bp_free(NmcBroadPhase* b) {
  if (!b) return;
  hipFree(b->index_map);
  hipFree(b->seg_start);
  hipFree(b->count);
  hipFree(b->off);
  hipFree(b->total);
  delete b;
}

extern "C" int nmc_broad_phase_create(int mode, const int* index_map, int index_count, const int* slice_ends,
                                      int segment_count, int explicit_pair_capacity, NmcBroadPhase** out) {
  if (!out) return fail("nmc_broad_phase_create: null output");
  *out = nullptr;
  if (mode < 0 || mode > 2) return fail("nmc_broad_phase_create: mode must be 0 (SAP), 1 (all pairs) or 2 (explicit)");
  NmcBroadPhase* b = new NmcBroadPhase();
  b->mode = mode;
  hipGetDevice(&b->device);
  int cap = 1;
  if (mode == 2) {
    if (explicit_pair_capacity < 0) {
      delete b;
      return fail("nmc_broad_phase_create: explicit_pair_capacity must be >= 0");
    }
    const long long per = static_cast<long long>(CBLOCK) * BP_EXPLICIT_ITEMS;
    cap = static_cast<int>((explicit_pair_capacity + per - 1) / per);
    if (cap < 1) cap = 1;
  } else {
    if (!index_map || !slice_ends || segment_count < 1 || index_count < 0)
      return delete b, fail("nmc_broad_phase_create: index map and at least one segment required");
    // slice_ends: exclusive ends of each segment, non-decreasing, last == index_count
    std::vector<int> start(segment_count + 1, 0);
    int prev = 0, maxn = 0;
    for (int s = 0; s < segment_count; ++s) {
      if (slice_ends[s] < prev || slice_ends[s] > index_count)
        return delete b, fail("nmc_broad_phase_create: slice_ends must be non-decreasing and within the index map");
      maxn = std::max(maxn, slice_ends[s] - prev);
      prev = slice_ends[s];
      start[s + 1] = prev;
    }
    if (prev != index_count) return delete b, fail("nmc_broad_phase_create: the last slice end must equal index_count");
    int mx = -1;
    for (int k = 0; k < index_count; ++k) {
      if (index_map[k] < 0) return delete b, fail("nmc_broad_phase_create: negative shape index in the index map");
      mx = std::max(mx, index_map[k]);
    }
    b->segments = segment_count;
    b->max_segment = maxn;
    b->max_index = mx;
    int l = 1;
    while (l < maxn && l < BP_LDS_MAX) l <<= 1;
    b->lds_n = l;
    cap = segment_count;
    hipError_t e = hipMalloc(&b->index_map, sizeof(int) * std::max(index_count, 1));
    if (e == hipSuccess) e = hipMalloc(&b->seg_start, sizeof(int) * (segment_count + 1));
    if (e == hipSuccess && index_count > 0)
      e = hipMemcpy(b->index_map, index_map, sizeof(int) * index_count, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b->seg_start, start.data(), sizeof(int) * (segment_count + 1), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      bp_free(b);
      return hip_fail(e, "nmc_broad_phase_create");
    }
  }
  b->capacity = cap;
  hipError_t e = hipMalloc(&b->count, sizeof(int) * cap);
  if (e == hipSuccess) e = hipMalloc(&b->off, sizeof(int) * cap);
  if (e == hipSuccess) e = hipMalloc(&b->total, sizeof(int));
  if (e != hipSuccess) {
    bp_free(b);
    return hip_fail(e, "nmc_broad_phase_create");
  }
  *out = b;
  return 0;
}

extern "C" void nmc_broad_phase_destroy(NmcBroadPhase* bp) { bp_free(bp); }

extern "C" int nmc_broad_phase_launch(NmcBroadPhase* bp, int shape_count, const float* lower, const float* upper,
                                      const float* gap, const int* group, const int* world, const int* filter_pairs,
                                      int filter_count, int* candidate_pair, int max_candidate_pair,
                                      int* candidate_pair_count, int append, nmc_stream_t stream) {
  if (!bp || bp->mode == 2) return fail("nmc_broad_phase_launch: needs a SAP or all-pairs broad phase");
  if (bp->max_index >= shape_count)
    return fail("nmc_broad_phase_launch: the index map names shapes beyond shape_count");
  if (!lower || !upper || !group || !world || !candidate_pair_count || (max_candidate_pair > 0 && !candidate_pair))
    return fail("nmc_broad_phase_launch: null argument");
  if (max_candidate_pair < 0 || filter_count < 0 || (filter_count > 0 && !filter_pairs))
    return fail("nmc_broad_phase_launch: bad candidate capacity or filter list");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  BpArgs a;
  a.lds_n = bp->lds_n;
  a.lower = lower;
  a.upper = upper;
  a.gap = gap;
  a.group = group;
  a.world = world;
  a.filter = filter_pairs;
  a.filter_count = filter_count;
  a.index_map = bp->index_map;
  a.seg_start = bp->seg_start;
  a.segments = bp->segments;
  a.sap = bp->mode == 0;
  a.seg_count = bp->count;
  a.seg_off = bp->off;
  a.total = bp->total;
  a.out_pairs = candidate_pair;
  a.max_pairs = max_candidate_pair;
  a.out_count = candidate_pair_count;
  const int grid = bp->segments < BP_MAX_BLOCKS ? bp->segments : BP_MAX_BLOCKS;
  const size_t lds = bp_lds_bytes(bp->lds_n, bp->mode == 0);
  hipLaunchKernelGGL(bp_segment_kernel<false>, dim3(grid), dim3(CBLOCK), lds, s, a);
  hipLaunchKernelGGL(bp_scan_kernel, dim3(1), dim3(SCAN_BLOCK), 0, s, bp->count, bp->segments, bp->off, bp->total,
                     candidate_pair_count, append);
  hipLaunchKernelGGL(bp_segment_kernel<true>, dim3(grid), dim3(CBLOCK), lds, s, a);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : hip_fail(e, "nmc_broad_phase_launch");
}

// The three launches of an explicit broad phase over per-block scratch (count, off: `blocks` entries each,
// total: 1).  Every entry of count is written by the first launch before the scan reads it, so the scratch
// needs no clearing.
int nmc::bp_explicit_run(int shape_count, const float* lower, const float* upper, const float* gap,
                           const int* shape_pairs, int shape_pair_count, int blocks, int* count, int* off,
                           int* total, int* candidate_pair, int max_candidate_pair, int* candidate_pair_count,
                           int append, hipStream_t s, const char* what) {
  BpExplicitArgs a;
  a.shape_count = shape_count;
  a.lower = lower;
  a.upper = upper;
  a.gap = gap;
  a.pairs = shape_pairs;
  a.pair_count = shape_pair_count;
  a.block_count = count;
  a.block_off = off;
  a.total = total;
  a.out_pairs = candidate_pair;
  a.max_pairs = max_candidate_pair;
  a.out_count = candidate_pair_count;
  const int grid = blocks < 1 ? 1 : (blocks < BP_MAX_BLOCKS ? blocks : BP_MAX_BLOCKS);
  hipLaunchKernelGGL(bp_explicit_kernel<false>, dim3(grid), dim3(CBLOCK), 0, s, a);
  hipLaunchKernelGGL(bp_scan_kernel, dim3(1), dim3(SCAN_BLOCK), 0, s, count, blocks, off, total,
                     candidate_pair_count, append);
  hipLaunchKernelGGL(bp_explicit_kernel<true>, dim3(grid), dim3(CBLOCK), 0, s, a);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : hip_fail(e, what);
}

int nmc::bp_explicit_check(int shape_count, const float* lower, const float* upper, const int* shape_pairs,
                             int shape_pair_count, const int* candidate_pair, int max_candidate_pair,
                             const int* candidate_pair_count, const char* what) {
  if (!lower || !upper || !candidate_pair_count || (shape_pair_count > 0 && !shape_pairs) ||
      (max_candidate_pair > 0 && !candidate_pair))
    return fail(std::string(what) + ": null argument");
  if (shape_pair_count < 0 || max_candidate_pair < 0 || shape_count < 0)
    return fail(std::string(what) + ": negative size");
  return 0;
}

extern "C" int nmc_broad_phase_explicit(NmcBroadPhase* bp, int shape_count, const float* lower, const float* upper,
                                        const float* gap, const int* shape_pairs, int shape_pair_count,
                                        int* candidate_pair, int max_candidate_pair, int* candidate_pair_count,
                                        int append, nmc_stream_t stream) {
  if (!bp || bp->mode != 2) return fail("nmc_broad_phase_explicit: needs an explicit broad phase");
  if (bp_explicit_check(shape_count, lower, upper, shape_pairs, shape_pair_count, candidate_pair,
                        max_candidate_pair, candidate_pair_count, "nmc_broad_phase_explicit"))
    return -1;
  const long long per = static_cast<long long>(CBLOCK) * BP_EXPLICIT_ITEMS;
  const int blocks = static_cast<int>((shape_pair_count + per - 1) / per);
  if (blocks > bp->capacity)
    return fail("nmc_broad_phase_explicit: more pairs than the broad phase was created for");
  return bp_explicit_run(shape_count, lower, upper, gap, shape_pairs, shape_pair_count, blocks, bp->count, bp->off,
                         bp->total, candidate_pair, max_candidate_pair, candidate_pair_count, append,
                         reinterpret_cast<hipStream_t>(stream), "nmc_broad_phase_explicit");
}

