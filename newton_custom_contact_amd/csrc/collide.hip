// Collision pipeline kernels for gfx950.
//
// One wave (segments of up to 64 shapes) or one workgroup owns one broad-phase segment (a world's shapes
// plus the global shapes, reference precompute_world_map, newton/_src/geometry/broad_phase_common.py:183-300)
// and runs, out of LDS: shape world transforms + AABBs (collide.py:150-275) and the pair tests of the
// configured broad phase (broad_phase_sap.py:43-394 sweep on the fixed axis, broad_phase_nxn.py every pair
// or listed pairs) over an enumeration of the segment's pairs in contact-sort-key order, so the pairs come
// out sorted with no sort (segment_broad_phase_enum).  The narrow phase then runs over all pairs of all segments at once, one pair per lane, with the
// MPR/GJK pairs grouped apart from the analytic ones.  Each segment owns a fixed region of pair
// entries and each pair five contact slots, so no kernel uses a global atomic counter (same-address
// atomics serialise across the chip's eight L2 slices); the two cross-segment prefix sums are done by
// one-workgroup scan kernels.  The Contacts buffer comes out in canonical order (segment-major, sort
// key inside a segment) with no float atomics and no global sort.
//
// Launch sequence of nmc_collide (all on the caller's stream, graph-capturable):
//   collide_segment_wave_kernel <<<segments/4, 256>>> one wave per segment of <= 64 shapes:
//                                                 AABB + pair tests in key order -> pair region, route list
//   collide_segment_kernel  <<<big segments, 256>>> the same for larger segments, one workgroup each
//   route_scan_kernel       <<<1, 1024>>>         route offsets over segments
//   narrow_kernel           <<<<=2048, 256>>>     narrow phase -> contact slots, per-pair counts
//   mesh_narrow_kernel      <<<<=2048, 256>>>     mesh pairs, one workgroup per pair (models with meshes)
//   contact_count_kernel    <<<segments, 64>>>    contacts per segment
//   contact_scan_kernel     <<<1, 1024>>>         Contacts.world_start / world_count / count, generation bump
//   contact_compact_kernel  <<<segments, 256>>>   slots -> Contacts arrays
#include <hip/hip_runtime.h>

#include <algorithm>

#include "broad_common.h"
#include "narrow.h"
static_assert(nmc::kNarrowBlock == nmc::CBLOCK, "the narrow kernel's clip buffers are laid out for CBLOCK lanes");
#include "nmc_internal.h"
#include "scan.h"

namespace nmc {


// Inclusive scan of one value per lane over a CBLOCK-lane workgroup (wave shuffles + 4 wave totals).
__device__ __forceinline__ int block_inclusive_scan(int v, int* s_wave) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int off = 1; off < 64; off <<= 1) {
    int t = __shfl_up(v, off, 64);
    if (lane >= off) v += t;
  }
  if (lane == 63) s_wave[w] = v;
  __syncthreads();
  int add = 0;
  for (int i = 0; i < w; ++i) add += s_wave[i];
  __syncthreads();
  return v + add;
}

struct SegLds {
  float* lo;   // 3 per shape, AABB expanded by margin + gap
  float* hi;   // 3
  float* plo;  // 1, SAP projection interval
  float* phi;  // 1
  int* type;
  int* gid;
  int* world;
  int* group;
  int* fbeg;   // filter-pair entries with this shape first: [fbeg, fend)
  int* fend;
  unsigned long long* skeys;  // sort keys (n2)
};

// Exclusive scan over a lane group of four counters per lane (x: all pairs, y/z/w: route classes 0/1/2);
// returns this lane's exclusive prefix, `total` gets the group's sums.  s_tmp: CBLOCK / 64 entries (block).
template <class G>
__device__ __forceinline__ int4 group_scan4(const G g, int4 v, int4& total, int4* s_tmp) {
  const int lane = threadIdx.x & 63;
  int4 inc = v;
  for (int off = 1; off < 64; off <<= 1) {
    const int tx = __shfl_up(inc.x, off, 64), ty = __shfl_up(inc.y, off, 64);
    const int tz = __shfl_up(inc.z, off, 64), tw = __shfl_up(inc.w, off, 64);
    if (lane >= off) {
      inc.x += tx;
      inc.y += ty;
      inc.z += tz;
      inc.w += tw;
    }
  }
  int4 before = make_int4(0, 0, 0, 0);
  if (G::size == 64) {
    total = make_int4(__shfl(inc.x, 63, 64), __shfl(inc.y, 63, 64), __shfl(inc.z, 63, 64), __shfl(inc.w, 63, 64));
  } else {
    const int w = threadIdx.x >> 6;
    if (lane == 63) s_tmp[w] = inc;
    g.sync();
    total = make_int4(0, 0, 0, 0);
    for (int i = 0; i < G::size / 64; ++i) {
      const int4 t = s_tmp[i];
      if (i < w) {
        before.x += t.x;
        before.y += t.y;
        before.z += t.z;
        before.w += t.w;
      }
      total.x += t.x;
      total.y += t.y;
      total.z += t.z;
      total.w += t.w;
    }
    g.sync();
  }
  return make_int4(before.x + inc.x - v.x, before.y + inc.y - v.y, before.z + inc.z - v.z, before.w + inc.w - v.w);
}

// Broad phase of one segment by lane group g, enumerating pairs instead of sorting and sweeping: the
// segment's shapes (ranks ordered like global ids) form an n x n matrix of (first, second) candidates,
// row-major, where `first` is the shape that leads the pair in the narrow phase (lower GeoType, then lower
// rank; a mesh before an infinite plane, narrow_phase.py:212-215, 303-313).  Row-major order over that
// matrix IS the pair key order, so each lane takes a contiguous run, counts its accepted pairs per route
// class, and after one group scan writes them and their route entries at their final positions -- no
// atomics, no sort, no LDS pair buffer.  The tests are the sweep's (broad_phase_sap.py:189-394): "sap"
// keeps a pair when the later of the two in (projected lower bound, rank) order starts below the earlier
// one's upper bound; "nxn" and "explicit" test every pair.  Pairs past the region's capacity are dropped,
// and counted, in key order.
struct EnumRun {
  long long p0, p1;  // this lane's run of the n x n matrix
  int n;
  bool regular, sweep, listed;
};

// Whether shape a (segment rank) leads the pair {a, b} in the narrow phase: lower GeoType, then lower rank; a
// mesh before an infinite plane.
__device__ __forceinline__ bool pair_lead(int a, int b, int ta, int tb) {
  bool first = a != b && (ta < tb || (ta == tb && a < b));
  if (ta == GEO_PLANE && tb == GEO_MESH) first = false;
  if (ta == GEO_MESH && tb == GEO_PLANE) first = true;
  return first;
}

// The broad phase's tests of a pair other than the sweep's (ranks la < lb): world, collision group, AABB
// overlap, filter or explicit list.
__device__ __forceinline__ bool pair_accepted(const NmcModelDesc& d, const SegLds& L, int la, int lb, bool regular,
                                              bool listed) {
  const int w1 = L.world[la], w2 = L.world[lb];
  bool keep = !(w1 == -1 && w2 == -1 && regular);
  if (keep && !listed) keep = !(w1 != -1 && w2 != -1 && w1 != w2) && group_pair(L.group[la], L.group[lb]);
  if (keep) {
    const float c = 0.0f + 0.0f;
    keep = L.lo[3 * la] <= L.hi[3 * lb] + c && L.hi[3 * la] >= L.lo[3 * lb] - c &&
           L.lo[3 * la + 1] <= L.hi[3 * lb + 1] + c && L.hi[3 * la + 1] >= L.lo[3 * lb + 1] - c &&
           L.lo[3 * la + 2] <= L.hi[3 * lb + 2] + c && L.hi[3 * la + 2] >= L.lo[3 * lb + 2] - c;
  }
  if (keep) {
    if (listed) keep = pair_listed(d.explicit_pairs, d.explicit_pair_count, L.gid[la], L.gid[lb]);
    else keep = !excluded(L.gid[lb], d.filter_pairs, L.fbeg[la], L.fend[la]);
  }
  return keep;
}

// Visits the accepted pairs of the lane's run in order: f(a, b, route class).
template <class F>
__device__ __forceinline__ void enum_pairs(const NmcModelDesc& d, const SegLds& L, const EnumRun& r, F&& f) {
  const int n = r.n;
  if (r.p0 >= r.p1) return;
  int a = static_cast<int>(r.p0 / n), b = static_cast<int>(r.p0 % n);
  int ta = L.type[a];
  for (long long q = r.p0; q < r.p1; ++q) {
    const int tb = L.type[b];
    if (pair_lead(a, b, ta, tb)) {
      const int la = a < b ? a : b, lb = a < b ? b : a;  // lower rank first for the tests
      bool keep = true;
      if (r.sweep) {
        const bool a_early = L.skeys[a] < L.skeys[b];
        keep = a_early ? (L.plo[b] < L.phi[a]) : (L.plo[a] < L.phi[b]);
      }
      if (keep) keep = pair_accepted(d, L, la, lb, r.regular, r.listed);
      if (keep) f(a, b, route_class(ta, tb));
    }
    if (++b == n) {
      b = 0;
      ++a;
      if (a < n) ta = L.type[a];
    }
  }
}

// Ascending bitonic sort of a[0, npow2) (npow2 a power of two) by the whole workgroup in LDS.
template <class T>
__device__ void block_bitonic_sort(T* a, int npow2) {
  for (int k = 2; k <= npow2; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < npow2; i += blockDim.x) {
        const int ixj = i ^ j;
        if (ixj > i) {
          const T x = a[i], y = a[ixj];
          if ((x > y) == ((i & k) == 0)) {
            a[i] = y;
            a[ixj] = x;
          }
        }
      }
      __syncthreads();
    }
}

constexpr int SWEEP_MAX_SHAPES = 512;  // segments the sweep path takes (above 64 shapes, up to this many)
constexpr int SWEEP_MAX_PAIRS = 4096;  // accepted pairs its LDS buffer holds; more: the enumeration

// SAP broad phase of one large segment by a 256-lane workgroup (broad_phase_sap.py:99-141, 189-394): the
// shapes sorted by (projected lower bound, rank) in LDS, each swept against the later ones that start below
// its upper bound, the accepted pairs put in key order (first * n + second over the segment's ranks, the
// enumeration's row-major order: a counting sort by the leading shape, then each pair placed in its bucket by
// the number of smaller keys there) and written like the enumeration writes them.  The pair set is the enumeration's: the
// same tests, each pair visited once.  Returns false, having written nothing, when the segment finds more
// than SWEEP_MAX_PAIRS pairs.
__device__ bool segment_pairs_sweep(const BlockGroup g, const NmcModelDesc& d, const SegLds& L, int n, bool regular,
                                    const StageBuffers& st, long long poff, int pcap, int* __restrict__ seg_pair_count,
                                    int* __restrict__ seg_conv_count, int* __restrict__ seg_mesh_count,
                                    int* __restrict__ counters, int seg, int4* s_tmp) {
  __shared__ unsigned long long s_sk[SWEEP_MAX_SHAPES];
  __shared__ unsigned s_pk[SWEEP_MAX_PAIRS];
  __shared__ unsigned s_po[SWEEP_MAX_PAIRS];
  __shared__ int s_cnt[SWEEP_MAX_SHAPES + 1];
  __shared__ int s_off[SWEEP_MAX_SHAPES + 1];
  __shared__ int s_wv[CBLOCK / 64];
  __shared__ int s_np;
  int npow = 1;
  while (npow < n) npow <<= 1;
  for (int i = g.lane; i < npow; i += CBLOCK) s_sk[i] = i < n ? L.skeys[i] : ~0ull;
  for (int i = g.lane; i < n; i += CBLOCK) s_cnt[i] = 0;
  if (g.lane == 0) {
    s_cnt[SWEEP_MAX_SHAPES] = 0;
    s_np = 0;
  }
  g.sync();
  block_bitonic_sort(s_sk, npow);
  // sweep: the earlier of two in (projected lower bound, rank) order keeps each later one that starts below
  // its upper bound.  The lower bounds ascend in sort order (NaN bounds sit at the ends, where the test fails),
  // so shape k's candidates are the run k+1 .. end_k - 1, end_k found by binary search; the runs are laid end
  // to end and split evenly over the lanes (one shape -- an infinite plane -- may span the whole segment).
  for (int base = 0; base < n; base += CBLOCK) {
    const int k = base + g.lane;
    int len = 0;
    if (k < n) {
      const float hk = L.phi[static_cast<int>(s_sk[k] & 0xFFFFFFFFull)];
      int lo = k + 1, hi = n;  // first j in (k, n] whose lower bound does not start below hk
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (L.plo[static_cast<int>(s_sk[mid] & 0xFFFFFFFFull)] < hk) lo = mid + 1; else hi = mid;
      }
      len = lo - k - 1;
    }
    const int incl = block_inclusive_scan(len, s_wv);
    const int carry = base == 0 ? 0 : s_off[base];
    g.sync();
    if (k < n) s_off[k] = carry + incl - len;
    if (g.lane == CBLOCK - 1) s_off[(base + CBLOCK < n ? base + CBLOCK : n)] = carry + incl;
    g.sync();
  }
  const int total = s_off[n];
  const int per = (total + CBLOCK - 1) / CBLOCK;
  int c = g.lane * per < total ? g.lane * per : total;
  const int c_end = c + per < total ? c + per : total;
  if (c < c_end) {
    int k = 0, hi = n - 1;  // the last shape whose run starts at or before c
    while (k < hi) {
      const int mid = (k + hi + 1) >> 1;
      if (s_off[mid] <= c) k = mid; else hi = mid - 1;
    }
    int j = k + 1 + (c - s_off[k]);
    for (; c < c_end; ++c) {
      while (c >= s_off[k + 1]) {  // past shape k's run: the next shape with a non-empty run
        ++k;
        j = k + 1;
      }
      const int rk = static_cast<int>(s_sk[k] & 0xFFFFFFFFull);
      const int rj = static_cast<int>(s_sk[j] & 0xFFFFFFFFull);
      ++j;
      const int la = rk < rj ? rk : rj, lb = rk < rj ? rj : rk;
      if (!pair_accepted(d, L, la, lb, regular, false)) continue;
      const bool lead = pair_lead(rk, rj, L.type[rk], L.type[rj]);
      const unsigned key = static_cast<unsigned>(lead ? rk * n + rj : rj * n + rk);
      const int pos = atomicAdd(&s_np, 1);
      if (pos < SWEEP_MAX_PAIRS) {
        s_pk[pos] = key;
        atomicAdd(&s_cnt[lead ? rk : rj], 1);
      }
    }
  }
  g.sync();
  const int np_all = s_np;
  if (np_all > SWEEP_MAX_PAIRS) return false;
  // key order: offsets of the leading shapes' buckets (exclusive scan of the counts), scatter, then each
  // bucket sorted by its shape's lane (a shape leads a handful of pairs)
  for (int base = 0; base < n; base += CBLOCK) {
    const int i = base + g.lane;
    const int v = i < n ? s_cnt[i] : 0;
    const int incl = block_inclusive_scan(v, s_wv);
    const int carry = s_cnt[SWEEP_MAX_SHAPES];  // total of the earlier passes (slot past the shapes)
    g.sync();
    if (i < n) s_cnt[i] = carry + incl - v;
    g.sync();
    if (g.lane == CBLOCK - 1) s_cnt[SWEEP_MAX_SHAPES] = carry + incl;
    g.sync();
  }
  for (int q = g.lane; q < np_all; q += CBLOCK) {
    const unsigned key = s_pk[q];
    const int pos = atomicAdd(&s_cnt[key / static_cast<unsigned>(n)], 1);
    s_po[pos] = key;
  }
  g.sync();
  // each pair's place in its bucket: the number of smaller keys there (keys are unique); s_cnt[a] is now the
  // end of a's bucket, its start the previous shape's end (0 for the first)
  for (int q = g.lane; q < np_all; q += CBLOCK) {
    const unsigned key = s_po[q];
    const int a = static_cast<int>(key / static_cast<unsigned>(n));
    const int b0 = a > 0 ? s_cnt[a - 1] : 0, b1 = s_cnt[a];
    int rank = 0;
    for (int i = b0; i < b1; ++i) rank += s_po[i] < key ? 1 : 0;
    s_pk[b0 + rank] = key;
  }
  g.sync();
  // the first pcap pairs in key order are kept (the rest dropped and counted, as the enumeration does)
  int np = np_all;
  if (np > pcap) {
    if (g.lane == 0) atomicAdd(&counters[0], np - pcap);
    np = pcap;
  }
  const int chunk = (np + CBLOCK - 1) / CBLOCK;
  const int q0 = g.lane * chunk < np ? g.lane * chunk : np;
  const int q1 = q0 + chunk < np ? q0 + chunk : np;
  int4 mine = make_int4(0, 0, 0, 0);
  for (int q = q0; q < q1; ++q) {
    const unsigned key = s_pk[q];
    const int a = static_cast<int>(key / static_cast<unsigned>(n)), b = static_cast<int>(key % static_cast<unsigned>(n));
    const int cls = route_class(L.type[a], L.type[b]);
    mine.x += 1;
    if (cls == 0) mine.y += 1; else if (cls == 1) mine.z += 1; else mine.w += 1;
  }
  int4 tot;
  const int4 at = group_scan4(g, mine, tot, s_tmp);
  // route order: MPR/GJK (class 1), then analytic (class 0), then mesh (class 2)
  const int base0 = tot.z, base1 = 0, base2 = tot.z + tot.y;
  int c0 = at.y, c1 = at.z, c2 = at.w;
  for (int q = q0; q < q1; ++q) {
    const unsigned key = s_pk[q];
    const int a = static_cast<int>(key / static_cast<unsigned>(n)), b = static_cast<int>(key % static_cast<unsigned>(n));
    const int cls = route_class(L.type[a], L.type[b]);
    st.pair_a[poff + q] = L.gid[a];
    st.pair_b[poff + q] = L.gid[b];
    const int slot = cls == 1 ? base1 + c1++ : (cls == 0 ? base0 + c0++ : base2 + c2++);
    st.route[poff + slot] = q;
  }
  if (g.lane == 0) {
    seg_pair_count[seg] = np;
    seg_conv_count[seg] = tot.z;
    seg_mesh_count[seg] = tot.w;
  }
  return true;
}

template <class G>
__device__ void segment_broad_phase_enum(const G g, const NmcModelDesc& d, const float* __restrict__ body_q,
                                         const StageBuffers& st, const long long* __restrict__ seg_pair_off,
                                         const int* __restrict__ seg_pair_cap, int* __restrict__ seg_pair_count,
                                         int* __restrict__ seg_conv_count, int* __restrict__ seg_mesh_count,
                                         int* __restrict__ counters, int max_n, int seg, unsigned char* smem,
                                         int4* s_tmp) {
  const int s0 = d.segment_start[seg];
  const int n = d.segment_start[seg + 1] - s0;
  const bool regular = seg < d.segment_count - 1;

  SegLds L;
  unsigned char* p = smem;
  L.skeys = reinterpret_cast<unsigned long long*>(p); p += sizeof(unsigned long long) * max_n;
  L.lo = reinterpret_cast<float*>(p); p += sizeof(float) * 3 * max_n;
  L.hi = reinterpret_cast<float*>(p); p += sizeof(float) * 3 * max_n;
  L.plo = reinterpret_cast<float*>(p); p += sizeof(float) * max_n;
  L.phi = reinterpret_cast<float*>(p); p += sizeof(float) * max_n;
  L.type = reinterpret_cast<int*>(p); p += sizeof(int) * max_n;
  L.gid = reinterpret_cast<int*>(p); p += sizeof(int) * max_n;
  L.world = reinterpret_cast<int*>(p); p += sizeof(int) * max_n;
  L.group = reinterpret_cast<int*>(p); p += sizeof(int) * max_n;
  L.fbeg = reinterpret_cast<int*>(p); p += sizeof(int) * max_n;
  L.fend = reinterpret_cast<int*>(p);

  // ---- shape world transforms, AABBs, SAP projections (collide.py:150-275, broad_phase_sap.py:43-63)
  const f3 dir = normalize(f3{0.5935f, 0.7790f, 0.1235f});
  for (int l = g.lane; l < n; l += G::size) {
    int s = d.segment_shape[s0 + l];
    ShapeW w = shape_world(s, body_q, d);
    // a world's shapes are written by its own segment, global shapes by the dedicated last segment
    if (d.shape_world[s] != -1 || !regular) store_wc(st.shape_wc + 20ll * s, w);
    store3(L.lo + 3 * l, w.lo);
    store3(L.hi + 3 * l, w.hi);
    L.type[l] = d.shape_type[s];
    L.gid[l] = s;
    L.world[l] = d.shape_world[s];
    L.group[l] = d.shape_group[s];
    int fb = 0, fe = 0;
    if (d.filter_count > 0) {
      fb = filter_lower_bound(d.filter_pairs, d.filter_count, s);
      fe = filter_lower_bound(d.filter_pairs, d.filter_count, s + 1);
    }
    L.fbeg[l] = fb;
    L.fend[l] = fe;
    f3 half = 0.5f * (w.hi - w.lo);
    half = f3{half.x + 0.0f, half.y + 0.0f, half.z + 0.0f};
    float radius = dot(dir, half);
    float center = dot(dir, 0.5f * (w.lo + w.hi));
    float plo = center - radius;
    L.plo[l] = plo;
    L.phi[l] = center + radius;
    // the stable sort's order: projected lower bound, then rank
    L.skeys[l] = (static_cast<unsigned long long>(sortable_float(plo)) << 32) | static_cast<unsigned>(l);
  }
  g.sync();

  EnumRun r;
  r.n = n;
  r.regular = regular;
  r.sweep = d.broad_phase_mode == 0;
  r.listed = d.broad_phase_mode == 2;
  const long long poff_s = seg_pair_off[seg];
  const int pcap_s = seg_pair_cap[seg];
  if constexpr (G::size == CBLOCK) {
    // large segments under SAP: sort and sweep (the enumeration is quadratic in the shape count)
    if (r.sweep && n > 64 && n <= SWEEP_MAX_SHAPES &&
        segment_pairs_sweep(g, d, L, n, regular, st, poff_s, pcap_s, seg_pair_count, seg_conv_count, seg_mesh_count,
                            counters, seg, s_tmp))
      return;
  }
  const long long N2 = static_cast<long long>(n) * n;
  const long long chunk = (N2 + G::size - 1) / G::size;
  r.p0 = chunk * g.lane;
  r.p1 = r.p0 + chunk < N2 ? r.p0 + chunk : N2;
  if (r.p0 > N2) r.p0 = N2;
  const long long poff = seg_pair_off[seg];
  const int pcap = seg_pair_cap[seg];

  // pass 0: accepted pairs per route class
  int4 mine = make_int4(0, 0, 0, 0);
  enum_pairs(d, L, r, [&](int, int, int cls) {
    mine.x += 1;
    if (cls == 0) mine.y += 1; else if (cls == 1) mine.z += 1; else mine.w += 1;
  });
  int4 tot;
  int4 at = group_scan4(g, mine, tot, s_tmp);
  int np = tot.x;
  if (np > pcap) {
    // rare: the region is full; recount the classes over the first pcap pairs only
    if (g.lane == 0) atomicAdd(&counters[0], np - pcap);
    np = pcap;
    int4 kept = make_int4(0, 0, 0, 0);
    int pi = at.x;
    enum_pairs(d, L, r, [&](int, int, int cls) {
      if (pi < pcap) {
        kept.x += 1;
        if (cls == 0) kept.y += 1; else if (cls == 1) kept.z += 1; else kept.w += 1;
      }
      ++pi;
    });
    const int ax = at.x;
    at = group_scan4(g, kept, tot, s_tmp);
    at.x = ax;  // pair positions stay those of the full order
  }
  // route order: MPR/GJK (class 1), then analytic (class 0), then mesh (class 2)
  const int base0 = tot.z, base1 = 0, base2 = tot.z + tot.y;
  int pi = at.x, c0 = at.y, c1 = at.z, c2 = at.w;
  enum_pairs(d, L, r, [&](int a, int b, int cls) {
    if (pi < pcap) {
      st.pair_a[poff + pi] = L.gid[a];
      st.pair_b[poff + pi] = L.gid[b];
      const int slot = cls == 1 ? base1 + c1++ : (cls == 0 ? base0 + c0++ : base2 + c2++);
      st.route[poff + slot] = pi;
    }
    ++pi;
  });
  if (g.lane == 0) {
    seg_pair_count[seg] = np;
    seg_conv_count[seg] = tot.z;
    seg_mesh_count[seg] = tot.w;
  }
}

// One 256-lane workgroup per listed segment (segments above 64 shapes), pairs by enumeration.
__global__ void __launch_bounds__(CBLOCK) collide_segment_kernel(NmcModelDesc d, const float* __restrict__ body_q,
                                                                  StageBuffers st,
                                                                  const long long* __restrict__ seg_pair_off,
                                                                  const int* __restrict__ seg_pair_cap,
                                                                  int* __restrict__ seg_pair_count,
                                                                  int* __restrict__ seg_conv_count,
                                                                  int* __restrict__ seg_mesh_count,
                                                                  int* __restrict__ counters, int max_n,
                                                                  const int* __restrict__ seg_list) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int4 s_tmp[CBLOCK / 64];
  segment_broad_phase_enum(BlockGroup{static_cast<int>(threadIdx.x)}, d, body_q, st, seg_pair_off, seg_pair_cap,
                           seg_pair_count, seg_conv_count, seg_mesh_count, counters, max_n, seg_list[blockIdx.x],
                           smem, s_tmp);
}

// One wave per listed segment (segments of at most 64 shapes), blockDim / 64 segments per workgroup (1, 2 or
// 4, a tuned launch parameter): the enumeration synchronises one wave instead of the workgroup.
// lds_per_wave: the dynamic LDS of one wave.
__global__ void __launch_bounds__(CBLOCK) collide_segment_wave_kernel(NmcModelDesc d, const float* __restrict__ body_q,
                                                                       StageBuffers st,
                                                                       const long long* __restrict__ seg_pair_off,
                                                                       const int* __restrict__ seg_pair_cap,
                                                                       int* __restrict__ seg_pair_count,
                                                                       int* __restrict__ seg_conv_count,
                                                                       int* __restrict__ seg_mesh_count,
                                                                       int* __restrict__ counters, int max_n,
                                                                       const int* __restrict__ seg_list,
                                                                       int n_list, int lds_per_wave) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int wave = threadIdx.x >> 6;
  const int idx = blockIdx.x * static_cast<int>(blockDim.x >> 6) + wave;
  if (idx >= n_list) return;  // whole waves only; the kernel has no workgroup barrier
  segment_broad_phase_enum(WaveGroup{static_cast<int>(threadIdx.x & 63)}, d, body_q, st, seg_pair_off, seg_pair_cap,
                           seg_pair_count, seg_conv_count, seg_mesh_count, counters, max_n, seg_list[idx],
                           smem + static_cast<size_t>(wave) * lds_per_wave, nullptr);
}

// Route offsets: exclusive prefix of MPR/GJK, analytic and mesh pair counts over segments.
__global__ void __launch_bounds__(SCAN_BLOCK) route_scan_kernel(const int* __restrict__ seg_pair_count,
                                                               const int* __restrict__ seg_conv_count,
                                                               const int* __restrict__ seg_mesh_count,
                                                               int* __restrict__ seg_conv_off,
                                                               int* __restrict__ seg_ana_off,
                                                               int* __restrict__ seg_mesh_off,
                                                               int* __restrict__ route_count, int segments) {
  // scans MPR/GJK, analytic (all pairs minus MPR/GJK minus mesh) and mesh counts in one pass
  int total[3];
  segment_scan_fn<3>(
      segments,
      [&](int i, int (&v)[3]) {
        v[0] = seg_conv_count[i];
        v[2] = seg_mesh_count ? seg_mesh_count[i] : 0;
        v[1] = seg_pair_count[i] - v[0] - v[2];
      },
      [&](int i, const int (&ex)[3], const int (&)[3]) {
        seg_conv_off[i] = ex[0];
        seg_ana_off[i] = ex[1];
        if (seg_mesh_off) seg_mesh_off[i] = ex[2];
      },
      total);
  if (threadIdx.x == 0) {
    route_count[0] = total[1];
    route_count[1] = total[0];
    route_count[2] = total[2];
  }
}

// Narrow phase over every staged pair, one pair per lane: MPR/GJK pairs first (the long ones), then
// the analytic pairs, so that a wave's lanes mostly run the same collider.  Lane i finds its segment
// by binary search over the route offsets.
// ALL: the model holds ellipsoids, cylinders or cones (their colliders and support maps compiled in)
template <bool ALL>
__global__ void __launch_bounds__(CBLOCK) narrow_kernel(NmcModelDesc d, const float* __restrict__ body_q,
                                                        StageBuffers st, const long long* __restrict__ seg_pair_off,
                                                        const int* __restrict__ seg_conv_count,
                                                        const int* __restrict__ seg_conv_off,
                                                        const int* __restrict__ seg_ana_off, int segments) {
  const int n_conv = st.route_count[1];
  const int total = st.route_count[0] + n_conv;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
    const bool conv = i < n_conv;
    const int key = conv ? i : i - n_conv;
    const int* __restrict__ offs = conv ? seg_conv_off : seg_ana_off;
    int lo = 0, hi = segments - 1;  // last segment whose offset is <= key
    while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (offs[mid] <= key) lo = mid; else hi = mid - 1;
    }
    const long long poff = seg_pair_off[lo];
    const int r = key - offs[lo] + (conv ? 0 : seg_conv_count[lo]);
    const long long g = poff + st.route[poff + r];
    int sa = st.pair_a[g], sb = st.pair_b[g];
    ShapeW A = load_wc(st.shape_wc + 20ll * sa), B = load_wc(st.shape_wc + 20ll * sb);
    float gap_sum = d.shape_gap[sa] + d.shape_gap[sb];
    WriteCtx w{d.shape_gap, st, 5ll * g, 0};
    narrow_pair<ALL>(sa, sb, d.shape_type[sa], d.shape_type[sb], A, B, gap_sum, w);
    st.mask[g] = static_cast<unsigned short>(w.n);
  }
}

// Contacts per segment (sum of its pairs' record counts).
__global__ void __launch_bounds__(64) contact_count_kernel(StageBuffers st, const long long* __restrict__ seg_pair_off,
                                                           const int* __restrict__ seg_pair_count,
                                                           int* __restrict__ seg_contact_count) {
  const int seg = blockIdx.x;
  const int np = seg_pair_count[seg];
  const long long poff = seg_pair_off[seg];
  int local = 0;
  for (int pi = threadIdx.x; pi < np; pi += 64) local += st.mask[poff + pi];
  for (int off = 32; off > 0; off >>= 1) local += __shfl_down(local, off, 64);
  if (threadIdx.x == 0) seg_contact_count[seg] = local;
}

// Global sort-key layout, counting pass (one wave per segment, pairs in key order): per-segment counts for
// the scan below.  The segment's contacts with a world shape0 form one run per shape0, all inside this
// segment; for each run's shape0 it records minus the number of world -1 shape0 contacts before the run
// (ord_delta0).  For each world -1 shape0 it records the segment's contacts and the first one's segment-local
// position (ord_M, ord_gstart; LDS accumulation), and in the last (world -1 only) segment each pair's count by
// (shape0, shape1) (ord_H).
struct OrderArgs {
  int G, U, W;
  const int* gid;
  const int* unit_order;
  int* delta0;
  int* M;
  int* gstart;
  int* L;
  int* H;
  int* unit_off;
  int* base;
};

__global__ void __launch_bounds__(64) order_count_kernel(StageBuffers st, const long long* __restrict__ seg_pair_off,
                                                         const int* __restrict__ seg_pair_count, OrderArgs o) {
  extern __shared__ int s_gm[];  // [G] contacts, [G] first position
  const int seg = blockIdx.x;
  const int np = seg_pair_count[seg];
  const long long poff = seg_pair_off[seg];
  const int G = o.G;
  const bool last = seg == o.W;
  for (int k = threadIdx.x; k < G; k += 64) {
    s_gm[k] = 0;
    s_gm[G + k] = 0x7fffffff;
  }
  if (last)
    for (int k = threadIdx.x; k < G * G; k += 64) o.H[k] = 0;
  __syncthreads();
  const int lane = threadIdx.x;
  int carry = 0, carry_g = 0;
  for (int b0 = 0; b0 < np; b0 += 64) {
    const int p = b0 + lane;
    const bool valid = p < np;
    const int cnt = valid ? static_cast<int>(st.mask[poff + p]) : 0;
    const int a = valid ? st.pair_a[poff + p] : -1;
    const int gi = valid ? o.gid[a] : -1;
    const int gc = gi >= 0 ? cnt : 0;
    int ic = cnt, ig = gc;
    for (int off = 1; off < 64; off <<= 1) {
      const int t = __shfl_up(ic, off, 64), tg = __shfl_up(ig, off, 64);
      if (lane >= off) {
        ic += t;
        ig += tg;
      }
    }
    const int ex = carry + ic - cnt, exg = carry_g + ig - gc;
    if (valid) {
      const bool run_start = p == 0 || st.pair_a[poff + p - 1] != a;
      if (gi < 0) {
        if (run_start) o.delta0[a] = -exg;
      } else {
        if (cnt > 0) {
          atomicAdd(&s_gm[gi], cnt);  // LDS
          atomicMin(&s_gm[G + gi], ex);
        }
        if (last) {
          const int gj = o.gid[st.pair_b[poff + p]];
          if (gj >= 0) o.H[gi * G + gj] = cnt;
        }
      }
    }
    carry += __shfl(ic, 63, 64);
    carry_g += __shfl(ig, 63, 64);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < G; k += 64) {
    o.M[static_cast<long long>(seg) * G + k] = s_gm[k];
    o.gstart[static_cast<long long>(seg) * G + k] = s_gm[G + k];
  }
  if (threadIdx.x == 0) o.L[seg] = carry - carry_g;
}

// Global sort-key layout, scans (one workgroup): for each world -1 shape0 g, the exclusive prefix of its
// contacts over the units in shape-index order (world w: segment w's contacts with shape0 g; world -1 shape t:
// the last segment's (g, t) contacts) -> ord_unit_off, total T[g]; then the shape0 blocks' first positions
// over the units (world w: its world-shape0 contacts; world -1 shape g: T[g]) -> ord_base.
__global__ void __launch_bounds__(SCAN_BLOCK) order_scan_kernel(OrderArgs o) {
  __shared__ int s_T[256];
  const int G = o.G, U = o.U, W = o.W;
  for (int g = 0; g < G; ++g) {
    int tot[1];
    segment_scan_fn<1>(
        U,
        [&](int i, int (&v)[1]) {
          const int u = o.unit_order[i];
          v[0] = u < W ? o.M[static_cast<long long>(u) * G + g] : o.H[g * G + (u - W)];
        },
        [&](int i, const int (&ex)[1], const int (&)[1]) { o.unit_off[static_cast<long long>(g) * U + o.unit_order[i]] = ex[0]; },
        tot);
    if (threadIdx.x == 0) s_T[g] = tot[0];
    __syncthreads();
  }
  int tot[1];
  segment_scan_fn<1>(
      U,
      [&](int i, int (&v)[1]) {
        const int u = o.unit_order[i];
        v[0] = u < W ? o.L[u] : s_T[u - W];
      },
      [&](int i, const int (&ex)[1], const int (&)[1]) { o.base[o.unit_order[i]] = ex[0]; }, tot);
}

// Contact offsets: exclusive prefix over segments -> Contacts.world_start / world_count / count.
__global__ void __launch_bounds__(SCAN_BLOCK) contact_scan_kernel(const int* __restrict__ seg_contact_count,
                                                                 NmcContacts c, int segments,
                                                                 int* __restrict__ counters) {
  // world_start, and world_count clipped to the capacity, written as the scan emits each segment
  int tot[1];
  segment_scan_fn<1>(
      segments, [&](int i, int (&v)[1]) { v[0] = seg_contact_count[i]; },
      [&](int i, const int (&ex)[1], const int (&v)[1]) {
        c.world_start[i] = ex[0];
        const int room = c.capacity - ex[0];
        const int stored = v[0] < room ? v[0] : room;
        c.world_count[i] = stored > 0 ? stored : 0;
      },
      tot);
  const int total = tot[0];
  if (threadIdx.x == 0) {
    c.count[0] = total;
    if (total > c.capacity) atomicAdd(&counters[2], total - c.capacity);
    if (c.generation) {
      const int g = c.generation[0];
      c.generation[0] = g == 2147483647 ? 0 : g + 1;
    }
  }
}

// Staging -> Contacts at canonical indices, one lane per output contact (coalesced stores).  MESH:
// the model has mesh shapes, whose pairs (mesh first) keep their records in the mesh slots.
// ORDER: the global sort-key layout; each contact goes to its buffer position (below the capacity) and
// Contacts.order records it.
template <bool MESH, bool ORDER>
__global__ void __launch_bounds__(CBLOCK) contact_compact_kernel(StageBuffers st,
                                                                  const long long* __restrict__ seg_pair_off,
                                                                  const int* __restrict__ seg_pair_count,
                                                                  const int* __restrict__ shape_type,
                                                                  const int* __restrict__ shape_body,
                                                                  const float* __restrict__ body_q,
                                                                  int mesh_k, NmcContacts c, OrderArgs ord,
                                                                  const int* __restrict__ seg_total) {
  __shared__ int s_scan[CBLOCK];
  __shared__ int s_wave[CBLOCK / 64];
  const int seg = blockIdx.x;
  const int np = seg_pair_count[seg];
  const long long poff = seg_pair_off[seg];
  // ORDER: every contact of the segment is placed (those whose buffer position is below the capacity are
  // stored, so the buffer holds the first contacts of the key order); else the canonical range only
  const int start = c.world_start[seg], stored = ORDER ? seg_total[seg] : c.world_count[seg];
  int run = 0;  // contacts of the earlier chunks
  for (int b0 = 0; b0 < np && run < stored; b0 += CBLOCK) {
    const int pi = b0 + threadIdx.x;
    const int cnt = pi < np ? static_cast<int>(st.mask[poff + pi]) : 0;
    const int incl = block_inclusive_scan(cnt, s_wave);
    s_scan[threadIdx.x] = incl;
    __syncthreads();
    const int chunk = s_scan[CBLOCK - 1];
    for (int t = threadIdx.x; t < chunk; t += CBLOCK) {
      const int local_idx = run + t;
      if (local_idx >= stored) break;
      int lo = 0, hi = CBLOCK - 1;  // first pair of the chunk whose inclusive count exceeds t
      while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (s_scan[mid] > t) hi = mid; else lo = mid + 1;
      }
      const int k = t - (lo > 0 ? s_scan[lo - 1] : 0);
      const long long g = poff + b0 + lo;
      const int s0 = st.pair_a[g], s1 = st.pair_b[g];
      int o = start + local_idx;
      if (ORDER) {
        // buffer position: shape0 block base + rank inside it
        const int gi = ord.gid[s0];
        long long pos;
        if (gi < 0) {
          pos = static_cast<long long>(ord.base[seg]) + ord.delta0[s0] + local_idx;
        } else if (seg < ord.W) {
          pos = static_cast<long long>(ord.base[ord.W + gi]) + ord.unit_off[static_cast<long long>(gi) * ord.U + seg] +
                (local_idx - ord.gstart[static_cast<long long>(seg) * ord.G + gi]);
        } else {
          pos = static_cast<long long>(ord.base[ord.W + gi]) +
                ord.unit_off[static_cast<long long>(gi) * ord.U + ord.W + ord.gid[s1]] + k;
        }
        const bool kept = pos < c.capacity;
        if (o < c.capacity) c.order[o] = kept ? static_cast<int>(pos) : -1;
        if (!kept) continue;
        o = static_cast<int>(pos);
      }
      long long s = 5 * g + k;
      if (MESH && shape_type[s0] == GEO_MESH) s = st.mesh_slot_base + static_cast<long long>(st.mesh_index[g]) * mesh_k + k;
      const float4* r = st.rec + REC_F4 * s;
      const float4 r0 = r[0], r1 = r[1], r2 = r[2];
      // write_contact (collide.py:62-148), second half: body-frame points and offsets
      const f3 aw{r0.x, r0.y, r0.z}, bw{r0.w, r1.x, r1.y}, n{r1.z, r1.w, r2.x};
      const float off_a = r2.y, off_b = r2.z;
      const int b0 = shape_body[s0], b1 = shape_body[s1];
      const xform Xa = b0 == -1 ? xident() : xinverse(load_xform(body_q + 7 * b0));
      const xform Xb = b1 == -1 ? xident() : xinverse(load_xform(body_q + 7 * b1));
      const f3 p0 = xpoint(Xa, aw), p1 = xpoint(Xb, bw);
      const f3 o0 = xvector(Xa, off_a * n), o1 = xvector(Xb, (-off_b) * n);
      c.shape0[o] = s0;
      c.shape1[o] = s1;
      c.point0[3 * o] = p0.x; c.point0[3 * o + 1] = p0.y; c.point0[3 * o + 2] = p0.z;
      c.point1[3 * o] = p1.x; c.point1[3 * o + 1] = p1.y; c.point1[3 * o + 2] = p1.z;
      c.offset0[3 * o] = o0.x; c.offset0[3 * o + 1] = o0.y; c.offset0[3 * o + 2] = o0.z;
      c.offset1[3 * o] = o1.x; c.offset1[3 * o + 1] = o1.y; c.offset1[3 * o + 2] = o1.z;
      c.normal[3 * o] = n.x; c.normal[3 * o + 1] = n.y; c.normal[3 * o + 2] = n.z;
      c.margin0[o] = off_a;
      c.margin1[o] = off_b;
      if (c.sort_key) {
        c.sort_key[o] = ((static_cast<long long>(s0) & 0xFFFFF) << 43) | ((static_cast<long long>(s1) & 0xFFFFF) << 23) |
                        (static_cast<long long>(__float_as_int(r2.w)) & 0x7FFFFF);
      }
    }
    run += chunk;
    __syncthreads();
  }
}

// LDS of segment_broad_phase_enum: order key 8 B, AABB 24 B, projection 8 B, six ints per shape.
size_t collide_lds_enum_bytes(int max_n) {
  size_t b = (sizeof(unsigned long long) + sizeof(float) * 8 + sizeof(int) * 6) * (size_t)max_n;
  return (b + 15) & ~size_t(15);
}


hipError_t launch_mesh_narrow(NmcModelImpl* m, const float* body_q, hipStream_t stream);

hipError_t launch_collide(NmcModelImpl* m, const float* body_q, const NmcContacts* c, hipStream_t stream) {
  const NmcModelDesc& d = m->d;
  int segs = d.segment_count;
  if (segs <= 0) return hipSuccess;
  if (m->n_seg_small > 0) {
    const int per = m->segment_waves;
    hipLaunchKernelGGL(collide_segment_wave_kernel, dim3((m->n_seg_small + per - 1) / per), dim3(64 * per),
                       static_cast<size_t>(per) * m->seg_small_lds, stream, d, body_q, m->stage, m->seg_pair_off,
                       m->seg_pair_cap, m->seg_pair_count, m->seg_conv_count, m->seg_mesh_count, m->counters,
                       m->seg_small_max_n, m->seg_small, m->n_seg_small,
                       static_cast<int>(m->seg_small_lds));
  }
  if (m->n_seg_big > 0)
    hipLaunchKernelGGL(collide_segment_kernel, dim3(m->n_seg_big), dim3(CBLOCK), m->collide_lds, stream, d, body_q,
                       m->stage, m->seg_pair_off, m->seg_pair_cap, m->seg_pair_count, m->seg_conv_count,
                       m->seg_mesh_count, m->counters, d.max_segment_shapes, m->seg_big);
  const bool mesh = d.mesh_count > 0;
  hipLaunchKernelGGL(route_scan_kernel, dim3(1), dim3(SCAN_BLOCK), 0, stream, m->seg_pair_count, m->seg_conv_count,
                     mesh ? m->seg_mesh_count : nullptr, m->seg_conv_off, m->seg_ana_off,
                     mesh ? m->seg_mesh_off : nullptr, m->stage.route_count, segs);
  long long pairs = m->stage.pair_capacity;
  long long nblk = (pairs + CBLOCK - 1) / CBLOCK;
  int narrow_blocks = static_cast<int>(nblk < m->narrow_max_blocks ? nblk : m->narrow_max_blocks);
  if (narrow_blocks < 1) narrow_blocks = 1;
  if (m->all_types)
    hipLaunchKernelGGL(narrow_kernel<true>, dim3(narrow_blocks), dim3(CBLOCK), 0, stream, d, body_q, m->stage,
                       m->seg_pair_off, m->seg_conv_count, m->seg_conv_off, m->seg_ana_off, segs);
  else
    hipLaunchKernelGGL(narrow_kernel<false>, dim3(narrow_blocks), dim3(CBLOCK), 0, stream, d, body_q, m->stage,
                       m->seg_pair_off, m->seg_conv_count, m->seg_conv_off, m->seg_ana_off, segs);
  if (mesh) {
    hipError_t e = launch_mesh_narrow(m, body_q, stream);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(contact_count_kernel, dim3(segs), dim3(64), 0, stream, m->stage, m->seg_pair_off,
                     m->seg_pair_count, m->seg_contact_count);
  hipLaunchKernelGGL(contact_scan_kernel, dim3(1), dim3(SCAN_BLOCK), 0, stream, m->seg_contact_count, *c, segs,
                     m->counters);
  const bool order = c->order != nullptr;
  OrderArgs oa{};
  if (order) {
    oa = OrderArgs{m->ord_G, m->ord_U, d.world_count, m->ord_gid, m->ord_unit_order, m->ord_delta0, m->ord_M,
                   m->ord_gstart, m->ord_L, m->ord_H, m->ord_unit_off, m->ord_base};
    hipLaunchKernelGGL(order_count_kernel, dim3(segs), dim3(64), sizeof(int) * 2 * (m->ord_G > 0 ? m->ord_G : 1),
                       stream, m->stage, m->seg_pair_off, m->seg_pair_count, oa);
    hipLaunchKernelGGL(order_scan_kernel, dim3(1), dim3(SCAN_BLOCK), 0, stream, oa);
  }
  auto compact = [&](auto kernel) {
    hipLaunchKernelGGL(kernel, dim3(segs), dim3(CBLOCK), 0, stream, m->stage, m->seg_pair_off, m->seg_pair_count,
                       d.shape_type, d.shape_body, body_q, m->mesh_k, *c, oa, m->seg_contact_count);
  };
  if (mesh) {
    if (order) compact(contact_compact_kernel<true, true>); else compact(contact_compact_kernel<true, false>);
  } else {
    if (order) compact(contact_compact_kernel<false, true>); else compact(contact_compact_kernel<false, false>);
  }
  return hipGetLastError();
}

}  // namespace nmc
