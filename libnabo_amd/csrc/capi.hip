// C ABI of libnabo_hip (declared in include/nabo_hip.h).
//
// Argument checks and their messages follow the reference:
//   create      nabo/nabo.cpp:68-80, 134-155; nabo/kdtree_cpu.cpp:227-251
//   knn         nabo/nabo.cpp:102-131 (checkSizesKnn)
// Exceptions never cross the ABI: each entry point maps them to a status
// code and keeps the message for nabo_hip_last_error().

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "internal.h"

namespace nabo_hip {
namespace {

thread_local std::string g_last_error;

// Every entry point first clears the thread's HIP error state: the launch
// checks read it with hipGetLastError, which also returns an error that an
// earlier HIP call of the caller's, or of another library on the thread,
// left unread (a graph library's teardown, for instance), and that error is
// not this call's.
template <typename F>
int guarded(F&& f) {
  (void)hipGetLastError();
  try {
    f();
    g_last_error.clear();
    return NABO_HIP_OK;
  } catch (const Error& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return NABO_HIP_ERR_MEMORY;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return NABO_HIP_ERR_ARGUMENT;
  }
}

[[noreturn]] void arg_error(const std::string& msg) { throw Error(NABO_HIP_ERR_ARGUMENT, msg); }

// Stream-ordered allocations are served from the device's default pool; keep
// freed blocks in the pool instead of returning them at every synchronize.
void prepare_device(int dev) {
  static std::mutex mu;
  static std::vector<int> done;
  std::lock_guard<std::mutex> lock(mu);
  if (std::find(done.begin(), done.end(), dev) != done.end()) return;
  hipMemPool_t pool;
  if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
    uint64_t threshold = UINT64_MAX;
    (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &threshold);
  }
  done.push_back(dev);
}

// One non-blocking stream per host thread and device for the host-buffer
// entry points (default priority).
hipStream_t thread_stream(int dev) {
  thread_local std::vector<std::pair<int, hipStream_t>> streams;
  for (auto& e : streams)
    if (e.first == dev) return e.second;
  hipStream_t s;
  NH_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  streams.emplace_back(dev, s);
  return s;
}

void build_index(nabo_hip_index* h, const float* d_cloud, int64_t n, int32_t rows, int64_t stride,
                 int32_t dim, int32_t type, uint32_t cflags, const nabo_hip_params& p,
                 hipStream_t s, const std::vector<int64_t>* offsets = nullptr) {
  Index& ix = h->ix;
  ix.search_type = type;
  ix.rows = rows;
  ix.dim = std::min(dim, rows);
  ix.n = n;
  ix.creation_flags = cflags;
  ix.params = p;
  ix.bucket_size = p.bucket_size;
  ix.dim_bits = storage_bits(uint32_t(ix.dim));
  ix.dim_mask = (1u << ix.dim_bits) - 1u;
  if (type == NABO_HIP_SEARCH_KDTREE) {
    kd_build(ix, d_cloud, stride, offsets ? *offsets : std::vector<int64_t>{0, n}, s);
    ix.device_bytes = int64_t(sizeof(uint2)) * (ix.node_count + 1) +
                      int64_t(sizeof(uint4)) * ix.node_count +
                      int64_t(sizeof(float)) * ix.pts_stride * (ix.n + kPtsPad);
  } else {
    bf_prepare(ix, d_cloud, stride, s);
    ix.device_bytes = int64_t(sizeof(float)) * ix.n * ix.dim;
  }
  NH_CHECK(hipStreamSynchronize(s));
}

// Serialized index: a fixed header, then the device arrays and the host
// vectors, each at a 256-byte aligned offset of the buffer.
constexpr char kSerialMagic[8] = {'N', 'A', 'B', 'O', 'H', 'I', 'P', '1'};
constexpr uint32_t kSerialVersion = 2;  // 2: shortlist rows in MFMA operand order

struct SerialHeader {
  char magic[8];
  uint32_t version;
  uint32_t header_bytes;
  int32_t search_type, rows, dim, clouds;
  int64_t n;
  uint32_t creation_flags, bucket_size, dim_bits, dim_mask;
  nabo_hip_params params;
  int32_t depth, pts_stride, bf_stride, lp_kp;
  int64_t node_count, lp_n_pad, device_bytes;
  // byte offset and size of each array (size 0: absent)
  uint64_t off[9], len[9];
  uint64_t total;
};
enum { kArNodes, kArLinks, kArPts, kArRoots, kArCloud, kArCloudLp, kArOffsets, kArMin, kArMax };

uint64_t align256(uint64_t v) { return (v + 255) / 256 * 256; }

SerialHeader serial_layout(const Index& ix) {
  SerialHeader h{};
  std::memcpy(h.magic, kSerialMagic, 8);
  h.version = kSerialVersion;
  h.header_bytes = sizeof(SerialHeader);
  h.search_type = ix.search_type;
  h.rows = ix.rows;
  h.dim = ix.dim;
  h.clouds = ix.clouds;
  h.n = ix.n;
  h.creation_flags = ix.creation_flags;
  h.bucket_size = ix.bucket_size;
  h.dim_bits = ix.dim_bits;
  h.dim_mask = ix.dim_mask;
  h.params = ix.params;
  h.depth = ix.depth;
  h.pts_stride = ix.pts_stride;
  h.bf_stride = ix.bf_stride;
  h.lp_kp = ix.lp_kp;
  h.node_count = ix.node_count;
  h.lp_n_pad = ix.lp_n_pad;
  h.device_bytes = ix.device_bytes;
  const bool kd = ix.search_type == NABO_HIP_SEARCH_KDTREE;
  h.len[kArNodes] = kd ? sizeof(uint2) * uint64_t(ix.node_count + 1) : 0;
  h.len[kArLinks] = kd ? sizeof(uint4) * uint64_t(ix.node_count) : 0;
  h.len[kArPts] = kd ? sizeof(float) * uint64_t(ix.n + kPtsPad) * ix.pts_stride : 0;
  h.len[kArRoots] = ix.d_roots ? sizeof(int32_t) * uint64_t(ix.clouds) : 0;
  h.len[kArCloud] = ix.d_cloud ? sizeof(float) * uint64_t(ix.n) * ix.dim : 0;
  h.len[kArCloudLp] = ix.d_cloud_lp ? sizeof(uint16_t) * uint64_t(ix.lp_n_pad) * ix.lp_kp : 0;
  h.len[kArOffsets] = sizeof(int64_t) * ix.cloud_offsets.size();
  h.len[kArMin] = sizeof(float) * ix.min_bound.size();
  h.len[kArMax] = sizeof(float) * ix.max_bound.size();
  uint64_t at = align256(sizeof(SerialHeader));
  for (int a = 0; a < 9; ++a) {
    h.off[a] = at;
    at = align256(at + h.len[a]);
  }
  h.total = at;
  return h;
}

__global__ void k_set_u64(uint64_t* p, uint64_t v) { *p = v; }


// Stream capture is refused: replays of captured knn calls (which allocate
// their scratch stream-ordered) made illegal memory accesses on the GPU,
// intermittently, after other captured graphs had run in the process; the
// cause was not found.
// knn calls allocate from the scratch pool and query its events, which a
// stream capture forbids: they refuse capture before doing anything.
// Calls that take temporaries from the per-device pool refuse stream
// capture: a graph would keep a pool block's address after the pool hands
// the block to the next call or frees it. The no-allocation entry point
// (temporaries in a static device array, kernels only) may be captured.
void refuse_capture(hipStream_t s) {
  hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
  NH_CHECK(hipStreamIsCapturing(s, &capture));
  if (capture != hipStreamCaptureStatusNone)
    throw Error(NABO_HIP_ERR_UNSUPPORTED, "knn calls cannot be captured in a HIP graph");
}

void run_knn(const Index& ix, const KnnArgs& a, hipStream_t s) {
  if (!a.scratch) refuse_capture(s);
  if (a.touched) {
    k_set_u64<<<1, 1, 0, s>>>(a.touched, 0);
    NH_LAUNCH_CHECK();
  }
  if (a.k == 0 || a.m == 0) return;
  if (ix.search_type == NABO_HIP_SEARCH_KDTREE) {
    KnnArgs b = a;
    if (!(ix.creation_flags & NABO_HIP_TOUCH_STATISTICS)) b.touched = nullptr;
    kd_knn(ix, b, s);
  } else {
    bf_knn(ix, a, s);
  }
}

}  // namespace


void check_create(int64_t n, int32_t rows, int64_t stride, int32_t dim, int32_t type,
                  const nabo_hip_params& p) {
  if (dim <= 0) arg_error("Your space must have at least one dimension");
  if (n == 0) arg_error("Cloud has no points");
  if (rows == 0) arg_error("Cloud has 0 dimensions");
  if (n < 0 || rows < 0) arg_error("Cloud has a negative size");
  if (stride < rows) arg_error("Cloud column stride is smaller than its number of rows");
  if (type >= 0 && type <= 5)
    throw Error(NABO_HIP_ERR_UNSUPPORTED,
                "search type " + std::to_string(type) +
                    " is a CPU search of the reference; the HIP library offers KDTREE_HIP (6) "
                    "and BRUTE_FORCE_HIP (7)");
  if (type != NABO_HIP_SEARCH_KDTREE && type != NABO_HIP_SEARCH_BRUTE_FORCE)
    arg_error("Unknown search type");
  if (n >= (int64_t(1) << 30))
    throw Error(NABO_HIP_ERR_UNSUPPORTED, "clouds of 2^30 points or more are not supported");
  if (type == NABO_HIP_SEARCH_KDTREE && n >= kKdPointsLimit)
    throw Error(NABO_HIP_ERR_UNSUPPORTED,
                "kd-tree clouds of 2^29 points or more are not supported");
  if (p.heap != NABO_HIP_HEAP_LINEAR && p.heap != NABO_HIP_HEAP_STL)
    arg_error("Unknown heap type " + std::to_string(p.heap));
  if (p.wave_queries != 0 && p.wave_queries != 64 && p.wave_queries != 128 &&
      p.wave_queries != 256 && p.wave_queries != 512)
    arg_error("wave_queries must be 0, 64, 128, 256 or 512, not " +
              std::to_string(p.wave_queries));
  if (p.stack_place != NABO_HIP_STACK_AUTO && p.stack_place != NABO_HIP_STACK_HBM)
    arg_error("Unknown stack placement " + std::to_string(p.stack_place));
  if (p.stl_log != 0 && (p.stl_log < 16 || p.stl_log > 4096))
    arg_error("stl_log must be 0 or 16 to 4096, not " + std::to_string(p.stl_log));
  if (type == NABO_HIP_SEARCH_KDTREE) {
    const uint32_t bucket = p.bucket_size;
    if (bucket < 2)
      arg_error("Requested bucket size " + std::to_string(bucket) + ", but must be larger than 2");
    const uint32_t used_dim = uint32_t(std::min(dim, rows));
    const uint32_t bits = storage_bits(used_dim);
    if (n > int64_t(bucket)) {
      const uint64_t max_nodes = (uint64_t(1) << (32 - bits)) - 1;
      const uint64_t est = uint64_t(n) / (bucket / 2);
      if (est > max_nodes)
        arg_error("Cloud has a risk to have more nodes (" + std::to_string(est) +
                  ") than the kd-tree allows (" + std::to_string(max_nodes) +
                  "). The kd-tree has " + std::to_string(bits) + " bits for dimensions and " +
                  std::to_string(32 - bits) + " bits for node indices");
    }
  }
}

void check_knn(const Index& ix, int32_t qrows, int64_t qstride, int64_t m, int32_t k,
               uint32_t flags) {
  const bool allow_self = flags & NABO_HIP_ALLOW_SELF_MATCH;
  // Batched trees: every cloud must satisfy the reference's k rule.
  int64_t npts = ix.n;
  for (int32_t c = 0; c + 1 < int32_t(ix.cloud_offsets.size()); ++c)
    npts = std::min(npts, ix.cloud_offsets[c + 1] - ix.cloud_offsets[c]);
  if (allow_self) {
    if (k > npts)
      arg_error("Requesting more points (" + std::to_string(k) + ") than available in cloud (" +
                std::to_string(npts) + ")");
  } else if (k > npts - 1) {
    arg_error("Requesting more points (" + std::to_string(k) +
              ") than available in cloud minus 1 (" + std::to_string(npts - 1) +
              ") (as self match is forbidden)");
  }
  if (qrows < ix.dim)
    arg_error("Query has less dimensions (" + std::to_string(qrows) +
              ") than requested for cloud (" + std::to_string(ix.dim) + ")");
  if (flags > 3u)
    arg_error("OR-ed value of option flags (" + std::to_string(flags) +
              ") is larger than maximal valid value (3)");
  if (k < 0) arg_error("k must not be negative");
  if (m < 0) arg_error("Query has a negative number of columns");
  if (m >= kQueryLimit)
    throw Error(NABO_HIP_ERR_UNSUPPORTED, "2^31 queries or more in one call are not supported");
  if (m > 0 && qstride < qrows) arg_error("Query column stride is smaller than its number of rows");
}

int run_guarded(const std::function<void()>& f) { return guarded(f); }

uint32_t storage_bits(uint32_t v) {  // getStorageBitCount, nabo/kdtree_cpu.cpp:60-70
  uint32_t b = 0;
  while (v) {
    ++b;
    v >>= 1;
  }
  return b;
}

namespace {
std::mutex g_scratch_mu;
std::vector<Scratch::Block> g_scratch;
}  // namespace

void Scratch::commit(int dev) {
  size_t total = 256;
  for (const Req& r : reqs_) total += r.bytes;
  if (ext_) {
    // caller-owned memory: carve, or refuse; no pool, no event, no allocation
    if (total - 256 > ext_bytes_)
      throw Error(NABO_HIP_ERR_UNSUPPORTED,
                  "the call needs " + std::to_string(total - 256) +
                      " bytes of scratch; the caller's scratch holds " +
                      std::to_string(ext_bytes_));
    char* cur = static_cast<char*>(ext_);
    for (const Req& r : reqs_) {
      *r.slot = cur;
      cur += r.bytes;
    }
    return;
  }
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    // A block last used on this stream is reused at once (stream order); a
    // block of another stream only once its work is done, so that calls on
    // different streams never wait for each other through the pool.
    int best = -1;
    for (int i = 0; i < int(g_scratch.size()); ++i) {
      const Block& b = g_scratch[i];
      if (b.dev != dev || b.bytes < total) continue;
      if (b.last != s_ && hipEventQuery(b.done) != hipSuccess) {
        (void)hipGetLastError();  // not ready: not an error of this call
        continue;
      }
      if (best < 0) {
        best = i;
        continue;
      }
      const Block& c = g_scratch[best];
      // same stream first, then the smallest
      const bool b_same = b.last == s_, c_same = c.last == s_;
      if (b_same != c_same ? b_same : b.bytes < c.bytes) best = i;
    }
    if (best >= 0) {
      block_ = g_scratch[best];
      g_scratch.erase(g_scratch.begin() + best);
    }
  }
  if (block_.p) {
    // Also on the same stream: a destroyed stream's handle can come back
    // for a new stream that does not follow the old one's work.
    NH_CHECK(hipStreamWaitEvent(s_, block_.done, 0));
  } else {
    // Power-of-two sizes from 2 MiB: blocks then fit calls of similar sizes.
    size_t bytes = size_t(2) << 20;
    while (bytes < total) bytes <<= 1;
    const hipError_t e = hipMallocAsync(&block_.p, bytes, s_);
    if (e != hipSuccess) {
      block_.p = nullptr;
      throw Error(NABO_HIP_ERR_MEMORY, "device allocation of " + std::to_string(bytes) +
                                           " bytes failed: " + hipGetErrorString(e));
    }
    block_.dev = dev;
    block_.bytes = bytes;
    if (hipEventCreateWithFlags(&block_.done, hipEventDisableTiming) != hipSuccess) {
      (void)hipFreeAsync(block_.p, s_);
      block_.p = nullptr;
      throw Error(NABO_HIP_ERR_HIP, "hipEventCreateWithFlags failed");
    }
  }
  char* cur = static_cast<char*>(block_.p);
  for (const Req& r : reqs_) {
    *r.slot = cur;
    cur += r.bytes;
  }
}

Scratch::~Scratch() {
  if (ext_ || !block_.p) return;
  auto drop = [this](const Block& b) {
    // stream-ordered after the block's last use, on this call's stream
    (void)hipStreamWaitEvent(s_, b.done, 0);
    (void)hipFreeAsync(b.p, s_);
    (void)hipEventDestroy(b.done);
  };
  block_.last = s_;
  if (hipEventRecord(block_.done, s_) != hipSuccess || block_.bytes > kScratchKeepBytes) {
    drop(block_);
    return;
  }
  std::vector<Block> dropped;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    g_scratch.push_back(block_);
    for (;;) {
      int held = 0, smallest = -1;
      for (int i = 0; i < int(g_scratch.size()); ++i) {
        if (g_scratch[i].dev != block_.dev) continue;
        ++held;
        if (smallest < 0 || g_scratch[i].bytes < g_scratch[smallest].bytes) smallest = i;
      }
      if (held <= kScratchKeep) break;
      dropped.push_back(g_scratch[smallest]);
      g_scratch.erase(g_scratch.begin() + smallest);
    }
  }
  for (const Block& b : dropped) drop(b);
}

namespace {

// Pinned host staging for the host-buffer entry points. A copy between
// pageable user memory and the device goes through a pinned buffer in
// chunks: the DMA of one chunk overlaps the host copy of the next (a
// pageable hipMemcpy moved ~10 GB/s here, pinned DMA ~55 GB/s and a host
// memcpy ~30 GB/s). Buffers are pooled (at most kStageKeep, each at most
// kStageKeepBytes kept); every call that uses one synchronizes its stream
// before giving it back.
constexpr size_t kStageChunk = size_t(8) << 20;

// Host copies between the caller's pageable memory and the pinned stage run
// on a small persistent pool of threads (at most 16, the calling thread
// included): a copy into freshly allocated memory is bound by its
// first-touch page faults, which the operating system serves in parallel
// (1M queries x k = 16: 128 MB of results). Several calls may copy at once;
// each waits for its own pieces only.
class CopyPool {
 public:
  static CopyPool& get() {
    static CopyPool* pool = new CopyPool();  // never destroyed: idle workers wait
    return *pool;
  }
  void copy(void* dst, const void* src, size_t len) {
    if (workers_ == 0 || len < kParallelMin) {
      std::memcpy(dst, src, len);
      return;
    }
    run(static_cast<char*>(dst), static_cast<const char*>(src), len);
  }
  // Writes a zero into every page of [dst, dst + len): pages not yet backed
  // by memory take their first-touch faults now, on all workers, instead of
  // one piece at a time during the copies into them.
  void touch(void* dst, size_t len) {
    if (len == 0) return;
    if (workers_ == 0 || len < kParallelMin) {
      touch_range(static_cast<char*>(dst), len);
      return;
    }
    run(static_cast<char*>(dst), nullptr, len);
  }

 private:
  static constexpr size_t kParallelMin = size_t(2) << 20;
  static constexpr size_t kPieceMin = size_t(256) << 10;
  static constexpr size_t kPage = 4096;
  struct Job {
    char* dst = nullptr;
    const char* src = nullptr;  // null: touch the pages of dst
    size_t len = 0, piece = 0, count = 0;
    std::atomic<size_t> next{0}, done{0};
    std::atomic<int> users{0};
  };
  static void touch_range(char* p, size_t len) {
    for (size_t off = 0; off < len; off += kPage) *reinterpret_cast<volatile char*>(p + off) = 0;
    *reinterpret_cast<volatile char*>(p + len - 1) = 0;
  }
  void run(char* dst, const char* src, size_t len) {
    Job j;
    j.dst = dst;
    j.src = src;
    j.len = len;
    j.piece = std::max(kPieceMin, (len + 2 * (workers_ + 1) - 1) / (2 * (workers_ + 1)));
    j.count = (len + j.piece - 1) / j.piece;
    {
      std::lock_guard<std::mutex> lock(mu_);
      jobs_.push_back(&j);
    }
    cv_.notify_all();
    while (run_piece(j)) {
    }
    {
      // after this no worker can join the job; wait for those that did
      std::lock_guard<std::mutex> lock(mu_);
      auto it = std::find(jobs_.begin(), jobs_.end(), &j);
      if (it != jobs_.end()) jobs_.erase(it);
    }
    while (j.done.load(std::memory_order_acquire) < j.count ||
           j.users.load(std::memory_order_acquire) > 0)
      std::this_thread::yield();
  }
  CopyPool() {
    const unsigned hw = std::thread::hardware_concurrency();
    workers_ = hw > 1 ? std::min(hw, 16u) - 1 : 0;
    for (unsigned i = 0; i < workers_; ++i) {
      try {
        std::thread(&CopyPool::work, this).detach();
      } catch (const std::exception&) {
        workers_ = i;  // run with the workers that started
        break;
      }
    }
  }
  static bool run_piece(Job& j) {
    const size_t i = j.next.fetch_add(1, std::memory_order_relaxed);
    if (i >= j.count) return false;
    const size_t off = i * j.piece;
    if (j.src)
      std::memcpy(j.dst + off, j.src + off, std::min(j.piece, j.len - off));
    else
      touch_range(j.dst + off, std::min(j.piece, j.len - off));
    j.done.fetch_add(1, std::memory_order_release);
    return true;
  }
  void work() {
    for (;;) {
      Job* j = nullptr;
      {
        std::unique_lock<std::mutex> lock(mu_);
        cv_.wait(lock, [&] {
          while (!jobs_.empty() &&
                 jobs_.front()->next.load(std::memory_order_relaxed) >= jobs_.front()->count)
            jobs_.pop_front();
          return !jobs_.empty();
        });
        j = jobs_.front();
        j->users.fetch_add(1, std::memory_order_relaxed);
      }
      while (run_piece(*j)) {
      }
      j->users.fetch_sub(1, std::memory_order_release);
    }
  }
  unsigned workers_ = 0;
  std::mutex mu_;
  std::condition_variable cv_;
  std::deque<Job*> jobs_;
};
constexpr int kStageEvents = 32;  // chunks per direction and call, at most
constexpr int kStageKeep = 2;
constexpr size_t kStageKeepBytes = size_t(256) << 20;

struct StageBuf {
  char* p = nullptr;
  size_t bytes = 0;
  hipEvent_t ev[kStageEvents] = {};
};
std::mutex g_stage_mu;
std::vector<StageBuf> g_stage;

class HostStage {
 public:
  HostStage(size_t bytes, hipStream_t s) : s_(s) {
    {
      std::lock_guard<std::mutex> lock(g_stage_mu);
      int best = -1;
      for (int i = 0; i < int(g_stage.size()); ++i)
        if (g_stage[i].bytes >= bytes && (best < 0 || g_stage[i].bytes < g_stage[best].bytes))
          best = i;
      if (best >= 0) {
        b_ = g_stage[best];
        g_stage.erase(g_stage.begin() + best);
      }
    }
    if (!b_.p) {
      size_t sz = kStageChunk;
      while (sz < bytes) sz <<= 1;
      if (hipHostMalloc(reinterpret_cast<void**>(&b_.p), sz, hipHostMallocDefault) != hipSuccess) {
        b_.p = nullptr;
        throw Error(NABO_HIP_ERR_MEMORY, "pinned host allocation of " + std::to_string(sz) +
                                             " bytes failed");
      }
      b_.bytes = sz;
      for (hipEvent_t& e : b_.ev) {
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
          release_events();
          (void)hipHostFree(b_.p);
          b_.p = nullptr;
          throw Error(NABO_HIP_ERR_HIP, "hipEventCreateWithFlags failed");
        }
      }
    }
  }
  HostStage(const HostStage&) = delete;
  HostStage& operator=(const HostStage&) = delete;
  ~HostStage() {
    if (!b_.p) return;
    if (hipStreamSynchronize(s_) != hipSuccess || b_.bytes > kStageKeepBytes) {
      free_buf(b_);
      return;
    }
    std::vector<StageBuf> drop;
    {
      std::lock_guard<std::mutex> lock(g_stage_mu);
      g_stage.push_back(b_);
      while (int(g_stage.size()) > kStageKeep) {
        int smallest = 0;
        for (int i = 1; i < int(g_stage.size()); ++i)
          if (g_stage[i].bytes < g_stage[smallest].bytes) smallest = i;
        drop.push_back(g_stage[smallest]);
        g_stage.erase(g_stage.begin() + smallest);
      }
    }
    for (StageBuf& d : drop) free_buf(d);
  }

  // Host to device: chunk i is copied into the stage, then its DMA is
  // enqueued; the next chunk's host copy overlaps it.
  void to_device(void* dst, const void* src, size_t bytes, size_t stage_off) {
    for (size_t off = 0; off < bytes; off += kStageChunk) {
      const size_t len = std::min(kStageChunk, bytes - off);
      CopyPool::get().copy(b_.p + stage_off + off, static_cast<const char*>(src) + off, len);
      NH_CHECK(hipMemcpyAsync(static_cast<char*>(dst) + off, b_.p + stage_off + off, len,
                              hipMemcpyHostToDevice, s_));
    }
  }
  // Device to host, for several arrays: all DMAs are enqueued (one event
  // per chunk), the destination pages are faulted in while the GPU works,
  // then each chunk is copied out as soon as it has landed.
  struct Out {
    void* dst;
    const void* src;
    size_t bytes;
  };
  void to_host(const Out* outs, int count) {
    size_t chunk = kStageChunk;
    for (;;) {
      size_t pieces = 0;
      for (int i = 0; i < count; ++i) pieces += (outs[i].bytes + chunk - 1) / chunk;
      if (pieces <= size_t(kStageEvents)) break;
      chunk <<= 1;
    }
    struct Piece {
      char* dst;
      size_t stage_off, len;
    };
    Piece pieces[kStageEvents];
    int n = 0;
    size_t stage_off = 0;
    for (int i = 0; i < count; ++i) {
      for (size_t off = 0; off < outs[i].bytes; off += chunk) {
        const size_t len = std::min(chunk, outs[i].bytes - off);
        NH_CHECK(hipMemcpyAsync(b_.p + stage_off, static_cast<const char*>(outs[i].src) + off, len,
                                hipMemcpyDeviceToHost, s_));
        NH_CHECK(hipEventRecord(b_.ev[n], s_));
        pieces[n++] = {static_cast<char*>(outs[i].dst) + off, stage_off, len};
        stage_off += len;
      }
    }
    for (int i = 0; i < count; ++i) CopyPool::get().touch(outs[i].dst, outs[i].bytes);
    for (int j = 0; j < n; ++j) {
      NH_CHECK(hipEventSynchronize(b_.ev[j]));
      CopyPool::get().copy(pieces[j].dst, b_.p + pieces[j].stage_off, pieces[j].len);
    }
  }

 private:
  void release_events() {
    for (hipEvent_t& e : b_.ev)
      if (e) (void)hipEventDestroy(e);
  }
  static void free_buf(StageBuf& b) {
    for (hipEvent_t& e : b.ev)
      if (e) (void)hipEventDestroy(e);
    (void)hipHostFree(b.p);
  }
  hipStream_t s_;
  StageBuf b_;
};

}  // namespace

void set_last_error(const std::string& msg) { g_last_error = msg; }

void host_copy(void* dst, const void* src, size_t bytes) {
  if (bytes) CopyPool::get().copy(dst, src, bytes);
}
void host_touch(void* dst, size_t bytes) { CopyPool::get().touch(dst, bytes); }

void throw_hip(hipError_t e, const char* what, const char* file, int line) {
  const int code = (e == hipErrorOutOfMemory) ? NABO_HIP_ERR_MEMORY : NABO_HIP_ERR_HIP;
  throw Error(code, std::string(what) + " failed: " + hipGetErrorString(e) + " (" + file + ":" +
                        std::to_string(line) + ")");
}

namespace {
std::mutex g_timing_mu;
bool g_timing = false;
std::vector<std::pair<hipEvent_t, hipEvent_t>> g_timing_events;
}  // namespace

bool kernel_timing_enabled() {
  std::lock_guard<std::mutex> lock(g_timing_mu);
  return g_timing;
}

KernelTimer::KernelTimer(hipStream_t stream) : s(stream) {
  if (!kernel_timing_enabled()) return;
  // a graph being captured gets no event nodes: the pair would never be
  // recorded outside the graph
  hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &capture) != hipSuccess || capture != hipStreamCaptureStatusNone)
    return;
  NH_CHECK(hipEventCreate(&start));
  NH_CHECK(hipEventCreate(&stop));
  NH_CHECK(hipEventRecord(start, s));
}

KernelTimer::~KernelTimer() {
  if (!start) return;
  if (hipEventRecord(stop, s) != hipSuccess) return;
  std::lock_guard<std::mutex> lock(g_timing_mu);
  g_timing_events.emplace_back(start, stop);
}

Index::~Index() {
  if (d_nodes) (void)hipFree(d_nodes);
  if (d_links) (void)hipFree(d_links);
  if (d_pts) (void)hipFree(d_pts);
  if (d_cloud) (void)hipFree(d_cloud);
  if (d_cloud_lp) (void)hipFree(d_cloud_lp);
  if (d_norms) (void)hipFree(d_norms);
  if (d_roots) (void)hipFree(d_roots);
  if (d_boxes) (void)hipFree(d_boxes);
}

}  // namespace nabo_hip

using namespace nabo_hip;

extern "C" {

const char* nabo_hip_last_error(void) { return g_last_error.c_str(); }

void nabo_hip_params_default(nabo_hip_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->bucket_size = 8;
  p->query_sort = 1;
  p->bf_precision = NABO_HIP_BF_FP32;
  p->bf_extra = 32;
  p->device = -1;
}

int nabo_hip_kernel_timing(int32_t enable) {
  return guarded([&] {
    std::lock_guard<std::mutex> lock(g_timing_mu);
    g_timing = enable != 0;
  });
}

int nabo_hip_kernel_timing_collect(double* milliseconds, int64_t* launches) {
  return guarded([&] {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
    {
      std::lock_guard<std::mutex> lock(g_timing_mu);
      ev.swap(g_timing_events);
    }
    double total = 0.0;
    for (auto& e : ev) {
      NH_CHECK(hipEventSynchronize(e.second));
      float ms = 0.f;
      NH_CHECK(hipEventElapsedTime(&ms, e.first, e.second));
      total += ms;
      (void)hipEventDestroy(e.first);
      (void)hipEventDestroy(e.second);
    }
    if (milliseconds) *milliseconds = total;
    if (launches) *launches = int64_t(ev.size());
  });
}

int nabo_hip_device_count(int32_t* count) {
  return guarded([&] {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    if (count) *count = c;
  });
}

int nabo_hip_create(const float* cloud, int64_t n, int32_t rows, int64_t stride, int32_t dim,
                    int32_t search_type, uint32_t creation_flags, const nabo_hip_params* params,
                    nabo_hip_index** out) {
  return guarded([&] {
    if (!out) arg_error("output handle pointer is null");
    *out = nullptr;
    nabo_hip_params p;
    nabo_hip_params_default(&p);
    if (params) p = *params;
    check_create(n, rows, stride, dim, search_type, p);
    if (!cloud) arg_error("cloud pointer is null");
    int dev = p.device;
    if (dev < 0) NH_CHECK(hipGetDevice(&dev));
    DeviceScope scope(dev);
    prepare_device(dev);
    hipStream_t s = thread_stream(dev);
    const int64_t floats = stride * (n - 1) + rows;
    Scratch scratch(s);
    Span<float> d_cloud;
    scratch.want(d_cloud, size_t(floats));
    scratch.commit(dev);
    {
      HostStage stage(sizeof(float) * size_t(floats), s);
      stage.to_device(d_cloud.p, cloud, sizeof(float) * size_t(floats), 0);
    }
    auto* h = new nabo_hip_index();
    h->ix.device = dev;
    try {
      build_index(h, d_cloud.p, n, rows, stride, dim, search_type, creation_flags, p, s);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int nabo_hip_create_device(const float* d_cloud, int64_t n, int32_t rows, int64_t stride,
                           int32_t dim, int32_t search_type, uint32_t creation_flags,
                           const nabo_hip_params* params, void* hip_stream,
                           nabo_hip_index** out) {
  return guarded([&] {
    if (!out) arg_error("output handle pointer is null");
    *out = nullptr;
    nabo_hip_params p;
    nabo_hip_params_default(&p);
    if (params) p = *params;
    check_create(n, rows, stride, dim, search_type, p);
    if (!d_cloud) arg_error("cloud pointer is null");
    int dev = p.device;
    if (dev < 0) NH_CHECK(hipGetDevice(&dev));
    DeviceScope scope(dev);
    prepare_device(dev);
    hipStream_t s = static_cast<hipStream_t>(hip_stream);
    auto* h = new nabo_hip_index();
    h->ix.device = dev;
    try {
      build_index(h, d_cloud, n, rows, stride, dim, search_type, creation_flags, p, s);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int nabo_hip_destroy(nabo_hip_index* index) {
  return guarded([&] {
    if (!index) return;
    DeviceScope scope(index->ix.device);
    delete index;
  });
}

int nabo_hip_knn_device(const nabo_hip_index* index, const float* d_query, int32_t qrows,
                        int64_t qstride, int64_t m, int32_t k, float epsilon,
                        uint32_t option_flags, float max_radius, const float* d_max_radii,
                        int32_t* d_indices, float* d_dists2, uint64_t* d_touched,
                        void* hip_stream) {
  return knn_device_call(index, d_query, qrows, qstride, m, k, epsilon, option_flags, max_radius,
                         d_max_radii, d_indices, d_dists2, d_touched, hip_stream, nullptr);
}

}  // extern "C"

int nabo_hip::knn_device_call(const nabo_hip_index* index, const float* d_query, int32_t qrows,
                              int64_t qstride, int64_t m, int32_t k, float epsilon,
                              uint32_t option_flags, float max_radius, const float* d_max_radii,
                              int32_t* d_indices, float* d_dists2, uint64_t* d_touched,
                              void* hip_stream, ScratchSource source) {
  return guarded([&] {
    if (!index) arg_error("index is null");
    const Index& ix = index->ix;
    check_knn(ix, qrows, qstride, m, k, option_flags);
    if (m > 0 && k > 0 && (!d_query || !d_indices || !d_dists2))
      arg_error("query or output pointer is null");
    if (source && ix.search_type != NABO_HIP_SEARCH_KDTREE)
      throw Error(NABO_HIP_ERR_UNSUPPORTED,
                  "the no-allocation knn entry point serves kd-tree indices only");
    DeviceScope scope(ix.device);
    hipStream_t s = static_cast<hipStream_t>(hip_stream);
    ExternalScratch ext;
    if (source) ext = source(ix.device);
    KnnArgs a;
    a.scratch = source ? &ext : nullptr;
    a.query = d_query;
    a.qrows = qrows;
    a.qstride = qstride;
    a.m = m;
    a.k = k;
    a.epsilon = epsilon;
    a.flags = option_flags;
    a.max_radius = max_radius;
    a.radii = d_max_radii;
    a.out_idx = d_indices;
    a.out_d2 = d_dists2;
    a.touched = d_touched;
    run_knn(ix, a, s);
    if (d_touched && ix.search_type == NABO_HIP_SEARCH_BRUTE_FORCE &&
        (ix.creation_flags & NABO_HIP_TOUCH_STATISTICS)) {
      k_set_u64<<<1, 1, 0, s>>>(d_touched, uint64_t(m) * uint64_t(ix.n));
      NH_LAUNCH_CHECK();
    }
  });
}

extern "C" {

int nabo_hip_knn(const nabo_hip_index* index, const float* query, int32_t qrows, int64_t qstride,
                 int64_t m, int32_t k, float epsilon, uint32_t option_flags, float max_radius,
                 const float* max_radii, int32_t* indices, float* dists2, uint64_t* touched) {
  return guarded([&] {
    if (!index) arg_error("index is null");
    const Index& ix = index->ix;
    check_knn(ix, qrows, qstride, m, k, option_flags);
    if (touched) *touched = 0;
    if (m == 0 || k == 0) return;
    if (!query || !indices || !dists2) arg_error("query or output pointer is null");
    DeviceScope scope(ix.device);
    hipStream_t s = thread_stream(ix.device);
    const int64_t qfloats = qstride * (m - 1) + qrows;
    Scratch scratch(s);
    Span<float> dq, dr, dd;
    Span<int32_t> di;
    Span<uint64_t> dt;
    scratch.want(dq, size_t(qfloats));
    scratch.want(dr, max_radii ? size_t(m) : 0);
    scratch.want(dd, size_t(m) * k);
    scratch.want(di, size_t(m) * k);
    scratch.want(dt, 1);
    scratch.commit(ix.device);
    const size_t q_bytes = sizeof(float) * size_t(qfloats);
    const size_t r_bytes = max_radii ? sizeof(float) * size_t(m) : 0;
    const size_t o_bytes = sizeof(float) * size_t(m) * size_t(k);
    HostStage stage(std::max(q_bytes + r_bytes, 2 * o_bytes + sizeof(uint64_t)), s);
    stage.to_device(dq.p, query, q_bytes, 0);
    if (max_radii) stage.to_device(dr.p, max_radii, r_bytes, q_bytes);
    KnnArgs a;
    a.query = dq.p;
    a.qrows = qrows;
    a.qstride = qstride;
    a.m = m;
    a.k = k;
    a.epsilon = epsilon;
    a.flags = option_flags;
    a.max_radius = max_radius;
    a.radii = max_radii ? dr.p : nullptr;
    a.out_idx = di.p;
    a.out_d2 = dd.p;
    a.touched = dt.p;
    run_knn(ix, a, s);
    uint64_t t = 0;
    const HostStage::Out outs[3] = {{indices, di.p, o_bytes}, {dists2, dd.p, o_bytes},
                                    {&t, dt.p, sizeof(uint64_t)}};
    stage.to_host(outs, 3);
    NH_CHECK(hipStreamSynchronize(s));
    if (touched) {
      if (!(ix.creation_flags & NABO_HIP_TOUCH_STATISTICS))
        *touched = 0;
      else if (ix.search_type == NABO_HIP_SEARCH_BRUTE_FORCE)
        *touched = uint64_t(m) * uint64_t(ix.n);
      else
        *touched = t;
    }
  });
}

int nabo_hip_batch_create(const float* d_points, const int64_t* offsets, int32_t clouds,
                          int32_t rows, int64_t stride, int32_t dim, uint32_t creation_flags,
                          const nabo_hip_params* params, void* hip_stream,
                          nabo_hip_index** out) {
  return guarded([&] {
    if (!out) arg_error("output handle pointer is null");
    *out = nullptr;
    if (clouds < 1 || !offsets) arg_error("a batch needs at least one cloud");
    nabo_hip_params p;
    nabo_hip_params_default(&p);
    if (params) p = *params;
    std::vector<int64_t> off(offsets, offsets + clouds + 1);
    if (off[0] != 0) arg_error("cloud offsets must start at 0");
    for (int32_t c = 0; c < clouds; ++c) {
      if (off[c + 1] < off[c]) arg_error("cloud offsets must not decrease");
      check_create(off[c + 1] - off[c], rows, stride, dim, NABO_HIP_SEARCH_KDTREE, p);
    }
    const int64_t n = off[clouds];
    if (n >= kKdPointsLimit)
      throw Error(NABO_HIP_ERR_UNSUPPORTED, "batches of 2^29 points or more are not supported");
    if (!d_points) arg_error("points pointer is null");
    int dev = p.device;
    if (dev < 0) NH_CHECK(hipGetDevice(&dev));
    DeviceScope scope(dev);
    prepare_device(dev);
    hipStream_t s = static_cast<hipStream_t>(hip_stream);
    auto* h = new nabo_hip_index();
    h->ix.device = dev;
    try {
      build_index(h, d_points, n, rows, stride, dim, NABO_HIP_SEARCH_KDTREE, creation_flags, p, s,
                  &off);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int nabo_hip_batch_knn(const nabo_hip_index* index, const float* d_query, int32_t qrows,
                       int64_t qstride, int64_t m, const int64_t* d_query_offsets, int32_t k,
                       float epsilon, uint32_t option_flags, float max_radius,
                       int32_t* d_indices, float* d_dists2, uint64_t* d_touched,
                       void* hip_stream) {
  return guarded([&] {
    if (!index) arg_error("index is null");
    const Index& ix = index->ix;
    check_knn(ix, qrows, qstride, m, k, option_flags);
    if (m > 0 && k > 0 && (!d_query || !d_indices || !d_dists2 || !d_query_offsets))
      arg_error("query, offsets or output pointer is null");
    DeviceScope scope(ix.device);
    hipStream_t s = static_cast<hipStream_t>(hip_stream);
    refuse_capture(s);
    Scratch scratch(s);
    Span<int32_t> qcloud;
    if (ix.clouds > 1 && m > 0) {
      scratch.want(qcloud, size_t(m));
      scratch.commit(ix.device);
      k_query_clouds<<<unsigned(ceil_div(m, 256)), 256, 0, s>>>(m, ix.clouds, d_query_offsets,
                                                                 qcloud.p);
      NH_LAUNCH_CHECK();
    }
    KnnArgs a;
    a.query = d_query;
    a.qrows = qrows;
    a.qstride = qstride;
    a.m = m;
    a.k = k;
    a.epsilon = epsilon;
    a.flags = option_flags;
    a.max_radius = max_radius;
    a.out_idx = d_indices;
    a.out_d2 = d_dists2;
    a.touched = d_touched;
    a.qcloud = qcloud.p;
    run_knn(ix, a, s);
  });
}

int nabo_hip_batch_roots(const nabo_hip_index* index, int32_t* roots) {
  return guarded([&] {
    if (!index || !roots) arg_error("null argument");
    const Index& ix = index->ix;
    if (!ix.d_roots) {
      roots[0] = 0;
      return;
    }
    DeviceScope scope(ix.device);
    hipStream_t s = thread_stream(ix.device);
    NH_CHECK(hipMemcpyAsync(roots, ix.d_roots, sizeof(int32_t) * ix.clouds, hipMemcpyDeviceToHost,
                            s));
    NH_CHECK(hipStreamSynchronize(s));
  });
}

int nabo_hip_get_info(const nabo_hip_index* index, nabo_hip_info* info) {
  return guarded([&] {
    if (!index || !info) arg_error("null argument");
    const Index& ix = index->ix;
    std::memset(info, 0, sizeof(*info));
    info->n = ix.n;
    info->rows = ix.rows;
    info->dim = ix.dim;
    info->search_type = ix.search_type;
    info->bucket_size = ix.bucket_size;
    info->dim_bits = ix.dim_bits;
    info->depth = ix.depth;
    info->node_count = ix.node_count;
    info->device_bytes = ix.device_bytes;
    info->creation_flags = ix.creation_flags;
    info->device = ix.device;
    info->clouds = ix.clouds;
  });
}

int nabo_hip_get_bounds(const nabo_hip_index* index, float* min_bound, float* max_bound) {
  return guarded([&] {
    if (!index || !min_bound || !max_bound) arg_error("null argument");
    const Index& ix = index->ix;
    std::copy(ix.min_bound.begin(), ix.min_bound.end(), min_bound);
    std::copy(ix.max_bound.begin(), ix.max_bound.end(), max_bound);
  });
}

int nabo_hip_get_tree(const nabo_hip_index* index, uint32_t* nodes, int32_t* bucket_indices) {
  return guarded([&] {
    if (!index || !nodes || !bucket_indices) arg_error("null argument");
    const Index& ix = index->ix;
    if (ix.search_type != NABO_HIP_SEARCH_KDTREE) arg_error("index is not a kd-tree");
    DeviceScope scope(ix.device);
    hipStream_t s = thread_stream(ix.device);
    NH_CHECK(hipMemcpyAsync(nodes, ix.d_nodes, sizeof(uint2) * size_t(ix.node_count),
                            hipMemcpyDeviceToHost, s));
    std::vector<float> pts(size_t(ix.n) * ix.pts_stride);
    NH_CHECK(hipMemcpyAsync(pts.data(), ix.d_pts, sizeof(float) * pts.size(), hipMemcpyDeviceToHost,
                            s));
    NH_CHECK(hipStreamSynchronize(s));
    for (int64_t p = 0; p < ix.n; ++p) {
      int32_t v;
      std::memcpy(&v, &pts[size_t(p * ix.pts_stride + ix.dim)], 4);
      bucket_indices[p] = v;
    }
  });
}

int nabo_hip_serialize(const nabo_hip_index* index, void* buf, uint64_t capacity,
                       uint64_t* bytes) {
  return guarded([&] {
    if (!index || !bytes) arg_error("null argument");
    const Index& ix = index->ix;
    const SerialHeader h = serial_layout(ix);
    *bytes = h.total;
    if (!buf) return;
    if (capacity < h.total)
      arg_error("serialization buffer holds " + std::to_string(capacity) + " bytes, " +
                std::to_string(h.total) + " are needed");
    char* out = static_cast<char*>(buf);
    // Zeroed gaps, so that equal indices serialize to equal bytes.
    std::memset(out, 0, h.off[0]);
    for (int a = 0; a < 9; ++a) {
      const uint64_t end = a + 1 < 9 ? h.off[a + 1] : h.total;
      std::memset(out + h.off[a] + h.len[a], 0, end - h.off[a] - h.len[a]);
    }
    std::memcpy(out, &h, sizeof(h));
    DeviceScope scope(ix.device);
    hipStream_t s = thread_stream(ix.device);
    const void* dev[6] = {ix.d_nodes, ix.d_links, ix.d_pts, ix.d_roots, ix.d_cloud, ix.d_cloud_lp};
    for (int a = 0; a < 6; ++a)
      if (h.len[a])
        NH_CHECK(hipMemcpyAsync(out + h.off[a], dev[a], h.len[a], hipMemcpyDeviceToHost, s));
    if (h.len[kArOffsets]) std::memcpy(out + h.off[kArOffsets], ix.cloud_offsets.data(), h.len[kArOffsets]);
    if (h.len[kArMin]) std::memcpy(out + h.off[kArMin], ix.min_bound.data(), h.len[kArMin]);
    if (h.len[kArMax]) std::memcpy(out + h.off[kArMax], ix.max_bound.data(), h.len[kArMax]);
    NH_CHECK(hipStreamSynchronize(s));
  });
}

int nabo_hip_deserialize(const void* buf, uint64_t bytes, int32_t device, nabo_hip_index** out) {
  return guarded([&] {
    if (!buf || !out) arg_error("null argument");
    *out = nullptr;
    SerialHeader h;
    if (bytes < sizeof(h)) arg_error("serialized index is truncated");
    std::memcpy(&h, buf, sizeof(h));
    if (std::memcmp(h.magic, kSerialMagic, 8) != 0 || h.version != kSerialVersion ||
        h.header_bytes != sizeof(SerialHeader))
      arg_error("buffer is not a serialized index of this library version");
    if (h.total > bytes) arg_error("serialized index is truncated");
    for (int a = 0; a < 9; ++a)
      if (h.off[a] > h.total || h.len[a] > h.total - h.off[a])
        arg_error("serialized index has an array outside the buffer");
    const bool kd = h.search_type == NABO_HIP_SEARCH_KDTREE;
    if (!kd && h.search_type != NABO_HIP_SEARCH_BRUTE_FORCE) arg_error("unknown search type");
    if (h.n <= 0 || h.dim <= 0 || h.clouds < 1 ||
        (h.len[kArOffsets] != 0 && h.len[kArOffsets] != sizeof(int64_t) * uint64_t(h.clouds + 1)))
      arg_error("serialized index header is inconsistent");
    if (kd && (h.node_count < 1 ||
               h.len[kArNodes] != sizeof(uint2) * uint64_t(h.node_count + 1) ||
               h.len[kArLinks] != sizeof(uint4) * uint64_t(h.node_count) ||
               h.len[kArPts] != sizeof(float) * uint64_t(h.n + kPtsPad) * uint64_t(h.pts_stride) ||
               (h.clouds > 1) != (h.len[kArRoots] != 0)))
      arg_error("serialized kd-tree arrays do not match its header");
    if (!kd && h.len[kArCloud] != sizeof(float) * uint64_t(h.n) * uint64_t(h.dim))
      arg_error("serialized brute-force arrays do not match its header");
    if (h.params.heap != NABO_HIP_HEAP_LINEAR && h.params.heap != NABO_HIP_HEAP_STL)
      arg_error("serialized index has an unknown heap type");
    int dev = device;
    if (dev < 0) NH_CHECK(hipGetDevice(&dev));
    DeviceScope scope(dev);
    prepare_device(dev);
    hipStream_t s = thread_stream(dev);
    auto* hnd = new nabo_hip_index();
    Index& ix = hnd->ix;
    try {
      ix.device = dev;
      ix.search_type = h.search_type;
      ix.rows = h.rows;
      ix.dim = h.dim;
      ix.clouds = h.clouds;
      ix.n = h.n;
      ix.creation_flags = h.creation_flags;
      ix.bucket_size = h.bucket_size;
      ix.dim_bits = h.dim_bits;
      ix.dim_mask = h.dim_mask;
      ix.params = h.params;
      ix.depth = h.depth;
      ix.pts_stride = h.pts_stride;
      ix.bf_stride = h.bf_stride;
      ix.lp_kp = h.lp_kp;
      ix.node_count = h.node_count;
      ix.lp_n_pad = h.lp_n_pad;
      ix.device_bytes = h.device_bytes;
      const char* in = static_cast<const char*>(buf);
      void** dev_ptr[6] = {reinterpret_cast<void**>(&ix.d_nodes), reinterpret_cast<void**>(&ix.d_links),
                           reinterpret_cast<void**>(&ix.d_pts), reinterpret_cast<void**>(&ix.d_roots),
                           reinterpret_cast<void**>(&ix.d_cloud),
                           reinterpret_cast<void**>(&ix.d_cloud_lp)};
      for (int a = 0; a < 6; ++a) {
        if (!h.len[a]) continue;
        NH_CHECK(hipMalloc(dev_ptr[a], h.len[a]));
        NH_CHECK(hipMemcpyAsync(*dev_ptr[a], in + h.off[a], h.len[a], hipMemcpyHostToDevice, s));
      }
      ix.cloud_offsets.resize(h.len[kArOffsets] / sizeof(int64_t));
      if (h.len[kArOffsets])
        std::memcpy(ix.cloud_offsets.data(), in + h.off[kArOffsets], h.len[kArOffsets]);
      ix.min_bound.resize(h.len[kArMin] / sizeof(float));
      ix.max_bound.resize(h.len[kArMax] / sizeof(float));
      if (h.len[kArMin]) std::memcpy(ix.min_bound.data(), in + h.off[kArMin], h.len[kArMin]);
      if (h.len[kArMax]) std::memcpy(ix.max_bound.data(), in + h.off[kArMax], h.len[kArMax]);
      NH_CHECK(hipStreamSynchronize(s));
      if (kd) kd_make_boxes(ix, s);
    } catch (...) {
      delete hnd;
      throw;
    }
    *out = hnd;
  });
}

}  // extern "C"
