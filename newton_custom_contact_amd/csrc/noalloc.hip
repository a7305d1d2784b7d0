// Explicit broad phase that allocates nothing (nmc_broad_phase_explicit_noalloc): the launches of
// nmc_broad_phase_explicit with their per-block scratch in a file-scope device array instead of a
// broad-phase object, for callers that cannot call hipMalloc (code under graph capture, tight serving
// loops).  No counterpart in the reference.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "broad_common.h"
#include "nmc_internal.h"

namespace nmc {

// count[NOALLOC_BLOCKS], off[NOALLOC_BLOCKS], total[1] (+ padding).  Each device holds its own copy for the
// life of the process.  In host code `scratch` only names the symbol: its device address on the current
// device comes from hipGetSymbolAddress, taken once per device and kept in g_scratch_dev, and only that
// address is handed to kernels.
constexpr int NOALLOC_BLOCKS = 16384;  // input pairs per call: NOALLOC_BLOCKS x CBLOCK x BP_EXPLICIT_ITEMS
static __device__ __align__(256) int scratch[2 * NOALLOC_BLOCKS + 64];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mutex;
static int* g_scratch_dev[NOALLOC_MAX_DEVICES] = {};

// The device address of `scratch` on the current device (looked up on the first call per device).
static int noalloc_scratch(int** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "nmc_broad_phase_explicit_noalloc: hipGetDevice");
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
    return fail("nmc_broad_phase_explicit_noalloc: device index beyond the supported range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "nmc_broad_phase_explicit_noalloc: hipGetSymbolAddress");
    if (!p) return fail("nmc_broad_phase_explicit_noalloc: hipGetSymbolAddress returned no address");
    g_scratch_dev[dev] = static_cast<int*>(p);
  }
  *out = g_scratch_dev[dev];
  return 0;
}

}  // namespace nmc

using namespace nmc;

extern "C" int nmc_broad_phase_explicit_noalloc(int shape_count, const float* lower, const float* upper,
                                                const float* gap, const int* shape_pairs, int shape_pair_count,
                                                int* candidate_pair, int max_candidate_pair,
                                                int* candidate_pair_count, int append, nmc_stream_t stream) {
  if (bp_explicit_check(shape_count, lower, upper, shape_pairs, shape_pair_count, candidate_pair,
                        max_candidate_pair, candidate_pair_count, "nmc_broad_phase_explicit_noalloc"))
    return -1;
  const long long per = static_cast<long long>(CBLOCK) * BP_EXPLICIT_ITEMS;
  const long long blocks = (shape_pair_count + per - 1) / per;
  if (blocks > NOALLOC_BLOCKS)
    return fail("nmc_broad_phase_explicit_noalloc: " + std::to_string(shape_pair_count) +
                " input pairs exceed the limit of " + std::to_string(NOALLOC_BLOCKS * per));
  int* base = nullptr;
  if (noalloc_scratch(&base)) return -1;
  return bp_explicit_run(shape_count, lower, upper, gap, shape_pairs, shape_pair_count, static_cast<int>(blocks),
                         base, base + NOALLOC_BLOCKS, base + 2 * NOALLOC_BLOCKS, candidate_pair,
                         max_candidate_pair, candidate_pair_count, append, reinterpret_cast<hipStream_t>(stream),
                         "nmc_broad_phase_explicit_noalloc");
}
