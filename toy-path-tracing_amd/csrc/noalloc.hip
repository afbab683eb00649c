// tpt_trace_noalloc: the trace queries of tpt_trace_closest / tpt_trace_any
// for callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops).  The persistent trace kernel's scratch -- its
// per-region work counters and the stack-overflow status word -- lives in a
// file-scope device array instead of the device handle's allocations, so a
// call allocates nothing: it launches a small clear kernel and the trace
// kernel, both on the device's stream.
//
// Every GPU holds its own copy of `scratch` at its own address;
// tpt_device_create looks that address up once per device with the device
// current (noalloc_setup), and the host uses only that address.  All calls on
// one GPU share the array: calls that overlap on one GPU race on it.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "tpt_device.h"
#include "tpt_internal.h"
#include "trace_kernel.h"
#include "traverse.h"

namespace tpt {
namespace {

// Work counters (one 256-byte line per region plus the finished-wave count)
// followed by one 256-byte line for the status word: 16,896 bytes.
constexpr size_t kScratchStatusOffset = sizeof(WorkCounters);
constexpr size_t kScratchBytes = sizeof(WorkCounters) + 256;
static_assert(kScratchStatusOffset % 256 == 0, "status line must start on a 256-byte boundary");

}  // namespace
}  // namespace tpt

static __device__ __align__(256) unsigned char scratch[tpt::kScratchBytes];

namespace tpt {
namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};  // per device: the address of its `scratch`

// Zeroes the counters and the status word (a kernel, not a memset, so a
// caller may capture the call in a graph).
__global__ __launch_bounds__(256) void noalloc_clear(uint32_t* p, uint32_t words) {
  for (uint32_t i = threadIdx.x; i < words; i += blockDim.x) p[i] = 0u;
}

// The traversal stack keeps the fixed kLdsDepth entries per lane in the
// kernels' static LDS here (not the device's tuned depth), so the resident
// blocks looked up at set-up stay valid; no launch-sized LDS.
constexpr size_t kNoallocLds = 0;

// Kernel variants: query 0 closest, 1 early-exit shadow, 2 exact-rule shadow;
// FP32 or 8-bit nodes; one or two levels.
constexpr int kVariants = 12;
int variant_index(int query, bool exact, bool q8, bool two) {
  const int qm = query == 0 ? 0 : (exact ? 2 : 1);
  return qm * 4 + (q8 ? 2 : 0) + (two ? 1 : 0);
}
const void* variant_kernel(int v) {
#define TPT_NK(Q8, M, E, T) reinterpret_cast<const void*>(&trace_kernel<Q8, M, E, false, T, kLdsDepth>)
  static const void* const table[kVariants] = {
      TPT_NK(false, kClosest, false, false), TPT_NK(false, kClosest, false, true),
      TPT_NK(true, kClosest, false, false),  TPT_NK(true, kClosest, false, true),
      TPT_NK(false, kAny, false, false),     TPT_NK(false, kAny, false, true),
      TPT_NK(true, kAny, false, false),      TPT_NK(true, kAny, false, true),
      TPT_NK(false, kAny, true, false),      TPT_NK(false, kAny, true, true),
      TPT_NK(true, kAny, true, false),       TPT_NK(true, kAny, true, true),
  };
#undef TPT_NK
  return table[v];
}

}  // namespace

// Looks up this device's copy of `scratch` (the device must be current) and
// the resident blocks per CU of every kernel variant, once per device handle.
bool noalloc_setup(TptDevice* d) {
  void* ptr = nullptr;
  if (!hip_ok(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)")) return false;
  if (!ptr) {
    set_error("hipGetSymbolAddress(scratch): null address");
    return false;
  }
  d->noalloc_scratch = ptr;
  for (int v = 0; v < kVariants; ++v) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, variant_kernel(v), kBlock, kNoallocLds) != hipSuccess ||
        per_cu < 1)
      per_cu = 1;
    d->noalloc_per_cu[v] = per_cu;
  }
  if (d->index >= 0 && d->index < kMaxDevices) {
    std::lock_guard<std::mutex> lock(g_mu);
    g_scratch[d->index] = ptr;
  }
  return hip_ok(hipMemset(ptr, 0, kScratchBytes), "hipMemset(scratch)");
}

}  // namespace tpt

using namespace tpt;

extern "C" int32_t tpt_trace_noalloc(TptScene* s, uint32_t query, const TptRay* rays, size_t n,
                                     const uint32_t* target, const float* distance, TptHit* hits,
                                     uint8_t* unoccluded, uint32_t flags) {
  try {
    if (!s) {
      set_error("tpt_trace_noalloc: null scene");
      return TPT_ERROR_BAD_ARGUMENT;
    }
    if (query > 1) {
      set_error("tpt_trace_noalloc: query must be 0 (closest) or 1 (shadow)");
      return TPT_ERROR_BAD_ARGUMENT;
    }
    if (!(flags & TPT_FLAG_DEVICE_POINTERS)) {
      set_error("tpt_trace_noalloc: needs TPT_FLAG_DEVICE_POINTERS (host arrays would need a staging allocation)");
      return TPT_ERROR_BAD_ARGUMENT;
    }
    if (n == 0) return TPT_OK;
    if (!rays || (query == 0 && !hits) || (query == 1 && !unoccluded)) {
      set_error("tpt_trace_noalloc: null rays or output");
      return TPT_ERROR_BAD_ARGUMENT;
    }
    if (n > 0x7FFFFFFFull) {
      set_error("tpt_trace_noalloc: too many rays for one call (at most 2^31 - 1)");
      return TPT_ERROR_BAD_ARGUMENT;
    }
    TptDevice* d = s->dev;
    // the scratch holds kMaxRegions counter lines; the device's region count
    // must fit them
    if (d->regions < 1 || d->regions > (uint32_t)kMaxRegions) {
      set_error("tpt_trace_noalloc: work regions beyond the scratch's " + std::to_string(kMaxRegions) + " counters");
      return TPT_ERROR_BAD_ARGUMENT;
    }
    if (!d->noalloc_scratch) {
      set_error("tpt_trace_noalloc: the device's scratch was not set up");
      return TPT_ERROR_DEVICE;
    }
    if (!hip_ok(hipSetDevice(d->index), "hipSetDevice")) return TPT_ERROR_DEVICE;
    char* base = static_cast<char*>(d->noalloc_scratch);
    const bool q8 = flags & TPT_FLAG_QUANTIZED;
    const bool exact = flags & TPT_FLAG_ANY_EXACT;
    const bool two = !(s->instance_count == 1 && s->tlas_node_count == 0);
    const int v = variant_index((int)query, exact, q8, two);
    hipLaunchKernelGGL(noalloc_clear, dim3(1), dim3(256), 0, d->stream, reinterpret_cast<uint32_t*>(base),
                       (uint32_t)(kScratchBytes / 4));
    const SceneView sv = view_of(s);
    Query qr;
    qr.hits = query == 0 ? hits : nullptr;
    qr.target = query == 1 ? target : nullptr;
    qr.distance = query == 1 ? distance : nullptr;
    qr.out = query == 1 ? unoccluded : nullptr;
    qr.counts = nullptr;
    qr.status = reinterpret_cast<uint32_t*>(base + kScratchStatusOffset);
    qr.work = reinterpret_cast<WorkCounters*>(base);
    qr.n_dev = nullptr;
    qr.n = (uint32_t)n;
    qr.refill_below = d->refill_below;
    qr.grab = d->refill_grab;
    qr.regions = d->regions;
    qr.lds_depth = kLdsDepth;
    int per_cu = d->noalloc_per_cu[v];
    if (d->blocks_per_cu && (int)d->blocks_per_cu < per_cu) per_cu = (int)d->blocks_per_cu;
    const size_t want = (n + kBlock - 1) / kBlock;
    const size_t cap = (size_t)per_cu * (size_t)(d->cu_count > 0 ? d->cu_count : 1);
    const unsigned blocks = (unsigned)(want < cap ? want : cap);
    void* args[] = {(void*)&sv, (void*)&rays, (void*)&qr};
    if (!hip_ok(hipLaunchKernel(variant_kernel(v), dim3(blocks), dim3(kBlock), args, kNoallocLds, d->stream),
                "hipLaunchKernel(trace)"))
      return TPT_ERROR_DEVICE;
    return hip_ok(hipGetLastError(), "kernel launch") ? TPT_OK : TPT_ERROR_DEVICE;
  } catch (const std::exception& e) {
    set_error(e.what());
    return TPT_ERROR;
  }
}
