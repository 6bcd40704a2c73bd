// The no-allocation kNN entry point (nabo_hip_knn_device_noalloc).
//
// nabo_hip_knn_device takes its temporaries (chunk counters, the query
// order's keys, histograms, scan space and permutation) from a per-device
// pool that allocates device memory when it has no block of the right size.
// This entry point takes them from `scratch`, a static device array of
// kNoAllocScratchBytes that every GPU the library runs on holds from the
// library's first use on it until the process ends, so that the call makes
// no device allocation at all (for callers that may not allocate). The
// kernels are the same; results equal nabo_hip_knn_device's.
//
// The array's host-side name is not a device address: host code uses only
// the address hipGetSymbolAddress returns, fetched once per device with that
// device current and kept in g_scratch_ptr.
//
// Every call on one GPU shares that GPU's array: calls that overlap on a
// GPU (two host threads, or two streams in flight) race on it and must be
// serialized by the caller. Nothing in it needs to start at a given value:
// each call zeroes its chunk counters with a kernel and writes every other
// temporary before reading it.

#include <mutex>
#include <string>

#include "internal.h"

namespace nabo_hip {
namespace {

// 64 MiB: the query order's temporaries (about 24 bytes per query) of calls
// up to about 2.6 million queries.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on device `dev` (the current device).
ExternalScratch device_scratch(int dev) {
  if (dev < 0 || dev >= kMaxDevices)
    throw Error(NABO_HIP_ERR_UNSUPPORTED,
                "the no-allocation entry point serves devices 0 to " +
                    std::to_string(kMaxDevices - 1));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      throw Error(NABO_HIP_ERR_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") +
                                        hipGetErrorString(e));
    g_scratch_ptr[dev] = ptr;
  }
  ExternalScratch ext;
  ext.base = g_scratch_ptr[dev];
  ext.bytes = kNoAllocScratchBytes;
  return ext;
}

}  // namespace
}  // namespace nabo_hip

extern "C" int nabo_hip_knn_device_noalloc(const nabo_hip_index* index, const float* d_query,
                                           int32_t qrows, int64_t qstride, int64_t m, int32_t k,
                                           float epsilon, uint32_t option_flags,
                                           float max_radius, const float* d_max_radii,
                                           int32_t* d_indices, float* d_dists2,
                                           uint64_t* d_touched, void* hip_stream) {
  return nabo_hip::knn_device_call(index, d_query, qrows, qstride, m, k, epsilon, option_flags,
                                   max_radius, d_max_radii, d_indices, d_dists2, d_touched,
                                   hip_stream, nabo_hip::device_scratch);
}
