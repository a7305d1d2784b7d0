// Frame-to-frame contact matching (CollisionPipeline(contact_matching="latest" | "sticky"),
// newton/_src/geometry/contact_match.py:1-992, collide.py:1084-1184).
//
// For each contact of this frame: among the previous frame's contacts of the same shape pair (sort key
// without the 23 sub-key bits), the one whose world-space midpoint 0.5 (world(point0) + world(point1)) is
// closest, within `pos_threshold` and with normals whose dot product reaches `normal_dot_threshold`
// (contact_match.py:250-324).  Several new contacts may pick the same previous one; the reference keeps
// the one with the smallest packed claim (float_flip(dist^2) << 32 | low 32 bits of its sort key,
// `atomic_min`, :160-214) and demotes the others to MATCH_BROKEN (:341-370).
//
// Here the contacts are walked in canonical order (segment-major, sort key inside a segment; Contacts.order maps
// a canonical position to the buffer), and a shape pair lives in exactly one segment, so both the previous
// pair range and the competing new contacts of a pair are contiguous runs inside one segment.  One 64-lane workgroup per segment:
//   match_kernel    candidate + its squared distance per contact (binary search over the segment's
//                   previous keys, linear scan of the pair range)
//   resolve_kernel  claims settled by scanning the pair's own new contacts (no atomics), sticky replay of
//                   the winners' previous body-frame points/offsets and normal, report counts per segment
//   report: one-workgroup scan of the per-segment counts, then the ordered index lists
//   save_kernel     this frame's keys, midpoints, normals (and sticky columns) for the next frame
// Match indices and report entries are buffer positions (the previous frame's for a match, as in the reference).
#include <hip/hip_runtime.h>

#include "nmc_internal.h"
#include "nmc_math.h"
#include "scan.h"

namespace nmc {

constexpr int MATCH_NOT_FOUND = -1;
constexpr int MATCH_BROKEN = -2;
constexpr long long SUB_KEY_MASK = 0x7FFFFF;

struct MatchState {
  int capacity, segments;
  float pos_threshold_sq, normal_dot_threshold;
  int sticky, report;
  // previous frame (canonical order)
  long long* keys;
  float* pos;     // 3 per contact: world midpoint
  float* normal;  // 3
  float* p0;      // sticky: body-frame point0 / point1 / offset0 / offset1
  float* p1;
  float* o0;
  float* o1;
  int* seg_start;
  int* seg_count;
  int* bufpos;       // buffer position of each previous contact in the previous Contacts
  int* count;        // [1]
  // this frame's scratch
  int* cand;
  float* dist_sq;
  int* prev_matched;  // report: per previous contact
  int* seg_new;       // report: per-segment counts and offsets
  int* seg_broken;
  int* seg_new_off;
  int* seg_broken_off;
};

__device__ __forceinline__ f3 body_point(const float* __restrict__ body_q, const int* __restrict__ shape_body,
                                         int shape, f3 p) {
  const int b = shape_body[shape];
  return b == -1 ? p : xpoint(load_xform(body_q + 7 * b), p);
}

__device__ __forceinline__ f3 midpoint(const NmcContacts& c, const float* __restrict__ body_q,
                                       const int* __restrict__ shape_body, int i) {
  const f3 p0w = body_point(body_q, shape_body, c.shape0[i], load3(c.point0 + 3 * i));
  const f3 p1w = body_point(body_q, shape_body, c.shape1[i], load3(c.point1 + 3 * i));
  return (p0w + p1w) * 0.5f;
}

// First index in keys[lo, hi) whose value is >= target.
__device__ __forceinline__ int lower_bound64(const long long* __restrict__ keys, int lo, int hi, long long target) {
  while (lo < hi) {
    const int mid = lo + ((hi - lo) >> 1);
    if (keys[mid] < target) lo = mid + 1; else hi = mid;
  }
  return lo;
}

__global__ void __launch_bounds__(64) match_kernel(MatchState m, NmcContacts c, const float* __restrict__ body_q,
                                                   const int* __restrict__ shape_body) {
  const int seg = blockIdx.x;
  const int ws = c.world_start[seg], wc = c.world_count[seg];
  const int ps = m.seg_start[seg], pe = ps + m.seg_count[seg];
  if (m.report)
    for (int p = ps + static_cast<int>(threadIdx.x); p < pe; p += 64) m.prev_matched[p] = 0;
  for (int i = ws + static_cast<int>(threadIdx.x); i < ws + wc; i += 64) {
    int best = MATCH_NOT_FOUND;
    float best_d2 = m.pos_threshold_sq;
    const int bi = cpos(c, i);
    if (pe > ps && bi >= 0) {
      const long long key = c.sort_key[bi];
      const long long prefix = key & ~SUB_KEY_MASK;
      const int lo = lower_bound64(m.keys, ps, pe, prefix);
      const int hi = lower_bound64(m.keys, lo, pe, prefix + SUB_KEY_MASK + 1);
      if (lo < hi) {
        const f3 pw = midpoint(c, body_q, shape_body, bi);
        const f3 nn = load3(c.normal + 3 * bi);
        best = MATCH_BROKEN;
        int found = -1;
        for (int j = lo; j < hi; ++j) {
          const f3 diff = pw - load3(m.pos + 3 * j);
          const float d2 = dot(diff, diff);
          const float nd = dot(nn, load3(m.normal + 3 * j));
          if (d2 <= best_d2 && nd >= m.normal_dot_threshold) {
            best_d2 = d2;
            found = j;
          }
        }
        if (found >= 0) best = found;
      }
    }
    m.cand[i] = best;
    m.dist_sq[i] = best_d2;
  }
}

__device__ __forceinline__ unsigned float_flip(float f) {
  const unsigned u = __float_as_uint(f);
  return u ^ ((u >> 31) ? 0xFFFFFFFFu : 0x80000000u);
}

// (flip(d2), low 32 key bits) lexicographic, as the reference's packed int64 claim orders them.
__device__ __forceinline__ bool claim_less(float d2a, long long ka, float d2b, long long kb) {
  const unsigned fa = float_flip(d2a), fb = float_flip(d2b);
  if (fa != fb) return fa < fb;
  return static_cast<unsigned>(ka & 0xFFFFFFFFll) < static_cast<unsigned>(kb & 0xFFFFFFFFll);
}

__global__ void __launch_bounds__(64) resolve_kernel(MatchState m, NmcContacts c, int* __restrict__ match_index) {
  __shared__ int s_new;
  const int seg = blockIdx.x;
  const int ws = c.world_start[seg], wc = c.world_count[seg];
  if (threadIdx.x == 0) s_new = 0;
  __syncthreads();
  int my_new = 0;
  // the sort key at canonical position j, or -1 when that contact is not stored (no pair prefix matches it)
  auto ckey = [&](int j) -> long long {
    const int b = cpos(c, j);
    return b >= 0 ? c.sort_key[b] : -1ll;
  };
  for (int i = ws + static_cast<int>(threadIdx.x); i < ws + wc; i += 64) {
    const int bi = cpos(c, i);
    if (bi < 0) continue;
    int r = m.cand[i];
    if (r >= 0) {
      // competitors: the pair's other new contacts (a contiguous run around i) with the same candidate
      const long long key = c.sort_key[bi];
      const long long prefix = key & ~SUB_KEY_MASK;
      const float d2 = m.dist_sq[i];
      bool win = true;
      for (int j = i - 1; j >= ws && (ckey(j) & ~SUB_KEY_MASK) == prefix && win; --j)
        if (m.cand[j] == r && claim_less(m.dist_sq[j], ckey(j), d2, key)) win = false;
      for (int j = i + 1; j < ws + wc && (ckey(j) & ~SUB_KEY_MASK) == prefix && win; ++j)
        if (m.cand[j] == r && claim_less(m.dist_sq[j], ckey(j), d2, key)) win = false;
      if (!win) {
        r = MATCH_BROKEN;
      } else {
        if (m.report) m.prev_matched[r] = 1;
        if (m.sticky) {
          store3(c.point0 + 3 * bi, load3(m.p0 + 3 * r));
          store3(c.point1 + 3 * bi, load3(m.p1 + 3 * r));
          store3(c.offset0 + 3 * bi, load3(m.o0 + 3 * r));
          store3(c.offset1 + 3 * bi, load3(m.o1 + 3 * r));
          store3(c.normal + 3 * bi, load3(m.normal + 3 * r));
        }
      }
    }
    match_index[bi] = r >= 0 ? m.bufpos[r] : r;
    my_new += r < 0 ? 1 : 0;
  }
  if (!m.report) return;
  for (int off = 32; off > 0; off >>= 1) my_new += __shfl_down(my_new, off, 64);
  __syncthreads();  // prev_matched of this segment written by all lanes
  int my_broken = 0;
  const int ps = m.seg_start[seg], pe = ps + m.seg_count[seg];
  for (int p = ps + static_cast<int>(threadIdx.x); p < pe; p += 64) my_broken += m.prev_matched[p] ? 0 : 1;
  for (int off = 32; off > 0; off >>= 1) my_broken += __shfl_down(my_broken, off, 64);
  if (threadIdx.x == 0) {
    m.seg_new[seg] = my_new;
    m.seg_broken[seg] = my_broken;
  }
}

__global__ void __launch_bounds__(SCAN_BLOCK) report_scan_kernel(MatchState m, int* __restrict__ new_count,
                                                                  int* __restrict__ broken_count) {
  const int* const in[2] = {m.seg_new, m.seg_broken};
  int* const out[2] = {m.seg_new_off, m.seg_broken_off};
  int tot[2];
  segment_scan<2>(in, out, m.segments, tot);
  if (threadIdx.x == 0) {
    new_count[0] = tot[0];
    broken_count[0] = tot[1];
  }
}

// Index lists (buffer positions): new contacts (match index < 0) in this frame's canonical order, broken
// previous contacts (not won by any new contact) in the previous frame's canonical order.  The reference
// fills both through atomic counters, in no fixed order.  One wave per segment, ballot ranks.
__global__ void __launch_bounds__(64) report_write_kernel(MatchState m, NmcContacts c,
                                                         const int* __restrict__ match_index,
                                                         int* __restrict__ new_indices,
                                                         int* __restrict__ broken_indices) {
  const int seg = blockIdx.x;
  const unsigned long long below = (1ull << threadIdx.x) - 1ull;
  const int ws = c.world_start[seg], wc = c.world_count[seg];
  int base = m.seg_new_off[seg];
  for (int i0 = ws; i0 < ws + wc; i0 += 64) {
    const int i = i0 + static_cast<int>(threadIdx.x);
    const int bi = i < ws + wc ? cpos(c, i) : -1;
    const bool is_new = bi >= 0 && match_index[bi] < 0;
    const unsigned long long bal = __ballot(is_new);
    if (is_new) new_indices[base + __popcll(bal & below)] = bi;
    base += __popcll(bal);
  }
  const int ps = m.seg_start[seg], pe = ps + m.seg_count[seg];
  base = m.seg_broken_off[seg];
  for (int p0 = ps; p0 < pe; p0 += 64) {
    const int p = p0 + static_cast<int>(threadIdx.x);
    const bool brk = p < pe && !m.prev_matched[p];
    const unsigned long long bal = __ballot(brk);
    if (brk) broken_indices[base + __popcll(bal & below)] = m.bufpos[p];
    base += __popcll(bal);
  }
}

// This frame's contacts become the next frame's previous contacts (contact_match.py:432-480).
__global__ void __launch_bounds__(256) save_kernel(MatchState m, NmcContacts c, const float* __restrict__ body_q,
                                                   const int* __restrict__ shape_body) {
  const int segs = m.segments;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < segs; s += gridDim.x * blockDim.x) {
    m.seg_start[s] = c.world_start[s];
    m.seg_count[s] = c.world_count[s];
  }
  const int n0 = c.count[0];
  const int n = n0 < c.capacity ? n0 : c.capacity;
  if (blockIdx.x == 0 && threadIdx.x == 0) m.count[0] = n;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int bi = cpos(c, i);
    m.bufpos[i] = bi;
    if (bi < 0) {  // not stored: a key no contact carries, so nothing matches it
      m.keys[i] = -1ll;
      continue;
    }
    m.keys[i] = c.sort_key[bi];
    store3(m.pos + 3 * i, midpoint(c, body_q, shape_body, bi));
    store3(m.normal + 3 * i, load3(c.normal + 3 * bi));
    if (m.sticky) {
      store3(m.p0 + 3 * i, load3(c.point0 + 3 * bi));
      store3(m.p1 + 3 * i, load3(c.point1 + 3 * bi));
      store3(m.o0 + 3 * i, load3(c.offset0 + 3 * bi));
      store3(m.o1 + 3 * i, load3(c.offset1 + 3 * bi));
    }
  }
}

__global__ void zero_kernel(int* __restrict__ a, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] = 0;
}

}  // namespace nmc

using namespace nmc;

struct NmcMatcher {
  MatchState s;
  int device;
};

template <class T>
static hipError_t mmalloc(T** p, size_t count) {
  return hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * (count > 0 ? count : 1));
}

extern "C" void nmc_matcher_destroy(NmcMatcher* mt) {
  if (!mt) return;
  MatchState& s = mt->s;
  void* ptrs[] = {s.keys, s.pos, s.normal, s.p0, s.p1, s.o0, s.o1, s.seg_start, s.seg_count, s.bufpos, s.count, s.cand,
                  s.dist_sq, s.prev_matched, s.seg_new, s.seg_broken, s.seg_new_off, s.seg_broken_off};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete mt;
}

extern "C" int nmc_matcher_create(const NmcModel* model, int capacity, float pos_threshold, float normal_dot_threshold,
                                  int sticky, int report, NmcMatcher** out) {
  if (!model || !out) return nmc::fail("nmc_matcher_create: null argument");
  if (capacity <= 0) return nmc::fail("nmc_matcher_create: capacity must be positive");
  if (!(pos_threshold >= 0.0f)) return nmc::fail("nmc_matcher_create: pos_threshold must be non-negative");
  if (!(normal_dot_threshold >= -1.0f && normal_dot_threshold <= 1.0f))
    return nmc::fail("nmc_matcher_create: normal_dot_threshold must be in [-1, 1]");
  NmcMatcher* mt = new NmcMatcher();
  MatchState& s = mt->s;
  s = MatchState{};
  s.capacity = capacity;
  s.segments = model->d.segment_count;
  s.pos_threshold_sq = pos_threshold * pos_threshold;
  s.normal_dot_threshold = normal_dot_threshold;
  s.sticky = sticky ? 1 : 0;
  s.report = report ? 1 : 0;
  (void)hipGetDevice(&mt->device);
  const size_t C = static_cast<size_t>(capacity), S = static_cast<size_t>(s.segments);
  hipError_t e = hipSuccess;
  if ((e = mmalloc(&s.keys, C)) != hipSuccess || (e = mmalloc(&s.pos, 3 * C)) != hipSuccess ||
      (e = mmalloc(&s.normal, 3 * C)) != hipSuccess || (e = mmalloc(&s.seg_start, S)) != hipSuccess ||
      (e = mmalloc(&s.seg_count, S)) != hipSuccess || (e = mmalloc(&s.count, 1)) != hipSuccess ||
      (e = mmalloc(&s.bufpos, C)) != hipSuccess ||
      (e = mmalloc(&s.cand, C)) != hipSuccess || (e = mmalloc(&s.dist_sq, C)) != hipSuccess ||
      (s.sticky && ((e = mmalloc(&s.p0, 3 * C)) != hipSuccess || (e = mmalloc(&s.p1, 3 * C)) != hipSuccess ||
                    (e = mmalloc(&s.o0, 3 * C)) != hipSuccess || (e = mmalloc(&s.o1, 3 * C)) != hipSuccess)) ||
      (s.report && ((e = mmalloc(&s.prev_matched, C)) != hipSuccess || (e = mmalloc(&s.seg_new, S)) != hipSuccess ||
                    (e = mmalloc(&s.seg_broken, S)) != hipSuccess || (e = mmalloc(&s.seg_new_off, S)) != hipSuccess ||
                    (e = mmalloc(&s.seg_broken_off, S)) != hipSuccess))) {
    nmc_matcher_destroy(mt);
    return nmc::hip_fail(e, "nmc_matcher_create: device allocation");
  }
  // no previous frame: empty segment ranges
  if ((e = hipMemset(s.seg_start, 0, sizeof(int) * S)) != hipSuccess ||
      (e = hipMemset(s.seg_count, 0, sizeof(int) * S)) != hipSuccess ||
      (e = hipMemset(s.count, 0, sizeof(int))) != hipSuccess || (e = hipDeviceSynchronize()) != hipSuccess) {
    nmc_matcher_destroy(mt);
    return nmc::hip_fail(e, "nmc_matcher_create: initialisation");
  }
  *out = mt;
  return 0;
}

extern "C" int nmc_matcher_reset(NmcMatcher* mt, nmc_stream_t stream) {
  if (!mt) return nmc::fail("nmc_matcher_reset: null argument");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const int S = mt->s.segments;
  hipLaunchKernelGGL(zero_kernel, dim3((S + 255) / 256), dim3(256), 0, st, mt->s.seg_count, S);
  hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, st, mt->s.count, 1);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : nmc::hip_fail(e, "nmc_matcher_reset launch");
}

extern "C" int nmc_contact_match(NmcMatcher* mt, const NmcModel* model, const float* body_q, const NmcContacts* c,
                                 int* match_index, const NmcContactReport* report, nmc_stream_t stream) {
  if (!mt || !model || !c || !match_index) return nmc::fail("nmc_contact_match: null argument");
  const MatchState& s = mt->s;
  if (!c->sort_key || !c->order)
    return nmc::fail("nmc_contact_match: the contacts buffer needs sort keys and its order array (deterministic pipeline)");
  if (c->capacity > s.capacity) return nmc::fail("nmc_contact_match: contacts capacity exceeds the matcher's");
  if (model->d.segment_count != s.segments) return nmc::fail("nmc_contact_match: model does not match the matcher");
  if (model->d.body_count > 0 && !body_q) return nmc::fail("nmc_contact_match: body_q is null");
  if (s.report && (!report || !report->new_indices || !report->new_count || !report->broken_indices ||
                   !report->broken_count))
    return nmc::fail("nmc_contact_match: the matcher was created with a report; pass its arrays");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const int S = s.segments;
  if (S <= 0) return 0;
  hipLaunchKernelGGL(match_kernel, dim3(S), dim3(64), 0, st, s, *c, body_q, model->d.shape_body);
  hipLaunchKernelGGL(resolve_kernel, dim3(S), dim3(64), 0, st, s, *c, match_index);
  if (s.report) {
    hipLaunchKernelGGL(report_scan_kernel, dim3(1), dim3(SCAN_BLOCK), 0, st, s, report->new_count,
                       report->broken_count);
    hipLaunchKernelGGL(report_write_kernel, dim3(S), dim3(64), 0, st, s, *c, match_index, report->new_indices,
                       report->broken_indices);
  }
  const long long work = c->capacity > S ? c->capacity : S;
  long long blocks = (work + 255) / 256;
  if (blocks > 2048) blocks = 2048;
  hipLaunchKernelGGL(save_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, st, s, *c, body_q,
                     model->d.shape_body);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : nmc::hip_fail(e, "nmc_contact_match launch");
}
