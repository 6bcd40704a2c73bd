// Internal declarations shared by the HIP translation units of libnabo_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/nabo_hip.h"

namespace nabo_hip {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& msg) : std::runtime_error(msg), code(c) {}
};

[[noreturn]] void throw_hip(hipError_t e, const char* what, const char* file, int line);
// Sets the message returned by nabo_hip_last_error() on this thread.
void set_last_error(const std::string& msg);

#define NH_CHECK(expr)                                              \
  do {                                                              \
    hipError_t nh_e_ = (expr);                                      \
    if (nh_e_ != hipSuccess) throw_hip(nh_e_, #expr, __FILE__, __LINE__); \
  } while (0)

#define NH_LAUNCH_CHECK() NH_CHECK(hipGetLastError())

// Stream-ordered device buffer (hipMallocAsync / hipFreeAsync).
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t count = 0;
  hipStream_t s = nullptr;
  DevBuf() = default;
  DevBuf(size_t n, hipStream_t stream) { alloc(n, stream); }
  void alloc(size_t n, hipStream_t stream) {
    release();
    s = stream;
    count = n;
    if (n) {
      hipError_t e = hipMallocAsync(reinterpret_cast<void**>(&p), n * sizeof(T), stream);
      if (e != hipSuccess) {
        p = nullptr;
        throw Error(NABO_HIP_ERR_MEMORY, "device allocation of " + std::to_string(n * sizeof(T)) +
                                             " bytes failed: " + hipGetErrorString(e));
      }
    }
  }
  void release() {
    if (p) (void)hipFreeAsync(p, s);
    p = nullptr;
    count = 0;
  }
  ~DevBuf() { release(); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
};

// A device array carved from a Scratch block.
template <class T>
struct Span {
  T* p = nullptr;
  size_t count = 0;
};

// Device scratch of one call (a build, a knn call): one block carved into
// 256-byte aligned arrays. Blocks come from a per-device pool; a call gives
// its block back with an event recorded on its stream. A later call on the
// same stream reuses it at once (stream order); a call on another stream
// takes it only once that event has completed (and still orders itself
// after it on the device), else it allocates: calls on different streams
// never wait for each other through the pool, and nothing waits on the
// host. Calls thus make no device allocation once the pool holds blocks of
// their size (stream-ordered frees of a call's dozens of arrays cost up to
// milliseconds each). The pool keeps at most kScratchKeep blocks per device,
// each at most kScratchKeepBytes; larger ones are freed after the call.
constexpr int kScratchKeep = 8;
constexpr size_t kScratchKeepBytes = size_t(1) << 30;

class Scratch {
 public:
  explicit Scratch(hipStream_t s) : s_(s) {}
  // Carves the arrays from caller-owned device memory [base, base + bytes)
  // instead of a pool block: commit() then allocates nothing and throws
  // NABO_HIP_ERR_UNSUPPORTED when the arrays do not fit; nothing is given
  // back on destruction.
  // (base null: the pool, as the one-argument form.)
  Scratch(hipStream_t s, void* base, size_t bytes) : s_(s), ext_(base), ext_bytes_(bytes) {}
  Scratch(const Scratch&) = delete;
  Scratch& operator=(const Scratch&) = delete;
  ~Scratch();

  template <class T>
  void want(Span<T>& span, size_t count) {
    span.count = count;
    reqs_.push_back({reinterpret_cast<void**>(&span.p), (count * sizeof(T) + 255) / 256 * 256});
  }
  // Takes a block holding every array asked for and points the spans into it.
  void commit(int dev);
  // Bytes the arrays asked for so far take (each 256-byte aligned).
  size_t bytes() const {
    size_t t = 0;
    for (const Req& r : reqs_) t += r.bytes;
    return t;
  }

  struct Block {
    int dev;
    void* p;
    size_t bytes;
    hipStream_t last;  // stream of the last call that used it
    hipEvent_t done;   // recorded after that call's work
  };

 private:
  struct Req {
    void** slot;
    size_t bytes;
  };
  hipStream_t s_;
  void* ext_ = nullptr;
  size_t ext_bytes_ = 0;
  Block block_{-1, nullptr, 0, nullptr, nullptr};
  std::vector<Req> reqs_;
};

// A built search structure resident in HBM.
struct Index {
  int search_type = NABO_HIP_SEARCH_KDTREE;
  int device = 0;
  int rows = 0;
  int dim = 0;
  int64_t n = 0;
  uint32_t creation_flags = 0;
  nabo_hip_params params{};

  // kd-tree (points in leaves, implicit bounds)
  uint32_t bucket_size = 8;
  uint32_t dim_bits = 0;
  uint32_t dim_mask = 0;
  int depth = 0;
  int build_levels = 0;  // levels the build's level loop ran (deeper parts finished in LDS)
  int64_t node_count = 0;
  uint2* d_nodes = nullptr;   // {dim | child << dim_bits, cut bits or bucket start},
                              // node_count entries + one padding leaf
  uint4* d_links = nullptr;   // per node: the right child's node word (internal
                              // nodes) and {cut dim, cut bits} of the parent's split
  float* d_pts = nullptr;     // leaf-ordered points, pts_stride floats each,
                              // slot `dim` holds the point index bits
  int pts_stride = 4;

  // brute force
  float* d_cloud = nullptr;   // dim x n, column stride bf_stride
  int bf_stride = 0;
  uint16_t* d_cloud_lp = nullptr;  // 16-bit rows for the MFMA shortlist
  int lp_kp = 0;                   // 16-bit row length (dim + 2 norm columns, padded)
  int64_t lp_n_pad = 0;            // rows, padded to a multiple of 32
  float* d_norms = nullptr;        // |p|^2 per point (low-precision path)

  // Batched trees: clouds > 1 trees stored one after the other in d_nodes /
  // d_pts; cloud c holds points [cloud_offsets[c], cloud_offsets[c+1]) of
  // the concatenated input and its root is node d_roots[c] (node 0 when
  // clouds == 1, d_roots null).
  int32_t clouds = 1;
  std::vector<int64_t> cloud_offsets;
  int32_t* d_roots = nullptr;
  float* d_boxes = nullptr;  // batched 2-D / 3-D trees: per cloud {min bound[3],
                             // 1 / extent[3] (0 if none)}, for the query order

  std::vector<float> min_bound, max_bound;  // clouds x dim
  int64_t device_bytes = 0;

  ~Index();
};

// Caller-owned device scratch for one knn call (the no-allocation entry
// point): every temporary of the call is carved from it; calls whose
// temporaries do not fit, or that need memory sized by the grid (admission
// logs of the logged IndexHeapSTL search, HBM stacks of trees deeper than
// the LDS stacks), are refused with NABO_HIP_ERR_UNSUPPORTED.
struct ExternalScratch {
  void* base = nullptr;
  size_t bytes = 0;
};

struct KnnArgs {
  const float* query = nullptr;  // device
  int qrows = 0;
  int64_t qstride = 0;
  int64_t m = 0;
  int k = 1;
  float epsilon = 0.f;
  uint32_t flags = 0;
  float max_radius = 0.f;
  const float* radii = nullptr;  // device, nullable
  int32_t* out_idx = nullptr;    // device
  float* out_d2 = nullptr;       // device
  uint64_t* touched = nullptr;   // device, nullable
  const int32_t* qcloud = nullptr;  // device, per-query cloud id (batched trees)
  const ExternalScratch* scratch = nullptr;  // null: the per-device scratch pool
  const int32_t* ok = nullptr;  // device, nullable: the search is skipped while *ok == 0
  size_t* measure = nullptr;    // set: kd_knn_launch only computes the call's scratch bytes
};

// Host memory copies on the library's pool of copying threads (up to 16),
// and the first-touch faults of a destination taken ahead of a copy into it.
void host_copy(void* dst, const void* src, size_t bytes);
void host_touch(void* dst, size_t bytes);

// Makes `dev` the current device for the scope, and restores the previous
// one after.
struct DeviceScope {
  int prev = -1;
  explicit DeviceScope(int dev) {
    NH_CHECK(hipGetDevice(&prev));
    if (dev != prev) NH_CHECK(hipSetDevice(dev));
  }
  ~DeviceScope() {
    int cur = -1;
    if (hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};

// Runs f, mapping its exceptions to a status code and the thread's message
// (the C ABI's entry points).
int run_guarded(const std::function<void()>& f);
// Argument checks of a create / knn call (nabo/nabo.cpp:68-80, 102-131).
void check_create(int64_t n, int32_t rows, int64_t stride, int32_t dim, int32_t type,
                  const nabo_hip_params& p);
void check_knn(const Index& ix, int32_t qrows, int64_t qstride, int64_t m, int32_t k,
               uint32_t flags);
// Bits of the node word holding a cut dimension (getStorageBitCount,
// nabo/kdtree_cpu.cpp:60-70).
uint32_t storage_bits(uint32_t v);

// Supplies the caller-owned scratch of a knn call on device `dev` (current
// when called); throws Error on failure.
typedef ExternalScratch (*ScratchSource)(int dev);

}  // namespace nabo_hip

// The opaque handle of the C ABI.
struct nabo_hip_index {
  nabo_hip::Index ix;
};

namespace nabo_hip {

// Body of nabo_hip_knn_device (capi.hip): argument checks, then the search
// enqueued on hip_stream. With `source`, the call's temporaries come from the
// memory it returns and kd-tree indices only are served; without, from the
// per-device scratch pool. Returns a status code and sets the thread's
// message.
int knn_device_call(const nabo_hip_index* index, const float* d_query, int32_t qrows,
                    int64_t qstride, int64_t m, int32_t k, float epsilon, uint32_t option_flags,
                    float max_radius, const float* d_max_radii, int32_t* d_indices,
                    float* d_dists2, uint64_t* d_touched, void* hip_stream,
                    ScratchSource source);

// A kd-tree build level's state in device memory: its segment count and
// the BFS id of its first segment.
struct BuildLevel {
  int32_t S;
  int32_t pad;
  int64_t base;
};

// What a build is built from and into.
struct KdBuildShape {
  int64_t n;
  int dim;
  int pts_stride;
  int32_t clouds;
  uint32_t bucket;
  uint32_t dim_bits, dim_mask;
};
struct KdBuildOut {
  float* pts = nullptr;       // (n + kPtsPad) x pts_stride floats, leaf-ordered on return
  uint2* nodes = nullptr;     // eager: allocated by the build; recorded: 2n + 1 nodes
  uint4* links = nullptr;     // eager: allocated by the build; recorded: 2n
  int32_t* roots = nullptr;   // per-cloud roots (nullable)
  float* boxes = nullptr;     // per-cloud query-order boxes, 6 floats each (nullable)
  int32_t* ok = nullptr;      // recorded builds: 1 when the levels sufficed, else 0
                              // (node arrays then unwritten)
};

class Launcher;

// Scratch bytes of a recorded build of this shape.
size_t kd_build_record_bytes(const KdBuildShape& sh);
// Records the build's launches (`levels` levels of the level loop) through
// L, with its scratch in [scratch, scratch + scratch_bytes), which must live
// as long as the recording is replayed; copies the host offsets into it
// first. A replay is complete (out.ok) when the levels finished the loop and
// the tree is at most max_depth levels deep.
void kd_build_record(const KdBuildShape& sh, const float* d_cloud, int64_t stride,
                     const int64_t* h_offsets, void* scratch, size_t scratch_bytes,
                     KdBuildOut& out, Launcher& L, int levels, int max_depth);

// Builds the kd-tree of d_cloud on the GPU (kdtree_build.hip).
// offsets: clouds + 1 point offsets into d_cloud ({0, n} for one cloud).
void kd_build(Index& ix, const float* d_cloud, int64_t stride,
              const std::vector<int64_t>& offsets, hipStream_t s);
// Fills d_boxes from the per-cloud bounds (batched 2-D / 3-D trees).
void kd_make_boxes(Index& ix, hipStream_t s);
// Cloud id of each of m queries from clouds + 1 device query offsets.
__global__ void k_query_clouds(int64_t m, int32_t clouds, const int64_t* __restrict__ offsets,
                               int32_t* __restrict__ qcloud);

// Searches a built kd-tree (kdtree_search.hip); the _launch form launches
// or records through a Launcher (a recorded search needs a.scratch, which
// must live as long as the recording is replayed).
void kd_knn(const Index& ix, const KnnArgs& a, hipStream_t s);
void kd_knn_launch(const Index& ix, const KnnArgs& a, Launcher& L);
// Prepares and searches a brute-force index (brute_force.hip).
void bf_prepare(Index& ix, const float* d_cloud, int64_t stride, hipStream_t s);
void bf_knn(const Index& ix, const KnnArgs& a, hipStream_t s);

// Measurement hook: brackets a kernel launch with events when enabled.
struct KernelTimer {
  hipEvent_t start = nullptr, stop = nullptr;
  hipStream_t s = nullptr;
  explicit KernelTimer(hipStream_t stream);  // records start if timing is on
  ~KernelTimer();                            // records stop and files the pair
};
bool kernel_timing_enabled();

// kd-tree clouds (and batches) hold fewer points: the build numbers the
// nodes of its LDS-finished subtrees from 2n up, below 4n, in int32 words.
constexpr int64_t kKdPointsLimit = int64_t(1) << 29;
// Query counts stay below this: the query order holds int32 positions.
constexpr int64_t kQueryLimit = int64_t(1) << 31;

// Largest k the register-resident k-best arrays support.
constexpr int kMaxK = 128;
// Zeroed entries after the leaf-ordered points, so a leaf scan can load a
// batch of up to kPtsPad entries past its bucket's end.
constexpr int kPtsPad = 8;

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace nabo_hip
