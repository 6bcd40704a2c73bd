// Served kNN (nabo_hip_server_*): resident trees under an HBM budget with
// LRU eviction, and a worker thread that admits requests continuously onto
// a set of stream slots.
//
// The server is a client of the C ABI: trees are nabo_hip_create indices,
// searches are nabo_hip_knn_device launches, so a served result is the
// result of nabo_hip_knn on the same cloud alone.
//
// Host copies stay off the worker: the submitting thread stages a request's
// queries into a pinned buffer (from a pool), the device writes its results
// into pinned buffers, and the waiting thread copies them out.
//
// Worker loop, one pass:
//   1. retire finished slots (event query): mark their requests done, wake
//      the waiters;
//   2. scan the pending requests in arrival order; a request whose tree is
//      resident takes a free slot, together with the pending requests on the
//      same cloud with the same k / epsilon / flags / radius (up to
//      max_merge_queries): copy-in, search and copy-out are enqueued on the
//      slot's stream; a request whose tree is not resident is handed to the
//      builder thread and stays pending, so it does not hold back the
//      requests behind it.
// Builder thread: evicts least recently used trees that no slot is using
// until the budget holds (with a host cache, an evicted tree is first
// serialized into pinned host memory), then builds on its own stream, or
// re-admits the tree from its host copy with one copy, concurrently with
// the searches. Slots run concurrently on the device; a request that arrives
// while others run is admitted as soon as a slot and its tree are ready.

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "internal.h"

using nabo_hip::throw_hip;

namespace {

using nabo_hip::Error;

double now_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// A serialized tree (nabo_hip_serialize) in pinned host memory.
struct HostCopy {
  void* p = nullptr;
  uint64_t bytes = 0;
  uint64_t last_use = 0;
  HostCopy() = default;
  HostCopy(const HostCopy&) = delete;
  HostCopy& operator=(const HostCopy&) = delete;
  ~HostCopy() {
    if (p) (void)hipHostFree(p);
  }
};

struct Cloud {
  std::vector<float> data;  // rows x n, column stride `stride`
  int64_t n = 0;
  int32_t rows = 0;
  int64_t stride = 0;
  int32_t dim = 0;
  nabo_hip_params params{};
  nabo_hip_index* index = nullptr;  // set once resident
  int64_t bytes = 0;
  uint64_t last_use = 0;
  int in_use = 0;          // slots searching it
  bool building = false;   // queued for, or in, the builder thread
  int build_status = NABO_HIP_OK;
  std::string build_error;
  std::unique_ptr<HostCopy> host;  // kept while within the host cache budget
};

struct Request {
  int64_t id = 0;
  int32_t cloud = 0;
  const float* query = nullptr;
  int32_t qrows = 0;
  int64_t qstride = 0;
  int64_t m = 0;
  int32_t k = 0;
  float eps = 0.f;
  uint32_t flags = 0;
  float radius = 0.f;
  int32_t* out_idx = nullptr;
  float* out_d2 = nullptr;
  float* h_q = nullptr;     // pinned, m x dim (compacted queries)
  int32_t* h_i = nullptr;   // pinned, m x k
  float* h_d = nullptr;     // pinned, m x k
  size_t cap_q = 0, cap_i = 0, cap_d = 0;
  double t_submit = 0.0;
  double latency = 0.0;
  int status = NABO_HIP_OK;
  std::string error;
  bool done = false;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t event = nullptr;
  bool busy = false;
  int32_t cloud = -1;
  std::vector<Request*> reqs;
  int64_t cap_q = 0, cap_r = 0;  // floats of the query buffer, entries of the result buffers
  float* d_q = nullptr;
  int32_t* d_i = nullptr;
  float* d_d = nullptr;
};

// Pinned host buffers by power-of-two size class, reused across requests
// (hipHostMalloc is far slower than a pool hit).
class PinnedPool {
 public:
  void* get(size_t bytes, size_t* cap) {
    size_t c = 256;
    while (c < bytes) c <<= 1;
    {
      std::lock_guard<std::mutex> lock(mu_);
      auto& v = free_[c];
      if (!v.empty()) {
        void* p = v.back();
        v.pop_back();
        *cap = c;
        return p;
      }
    }
    void* p = nullptr;
    NH_CHECK(hipHostMalloc(&p, c));
    *cap = c;
    return p;
  }
  void put(void* p, size_t cap) {
    if (!p) return;
    std::lock_guard<std::mutex> lock(mu_);
    free_[cap].push_back(p);
  }
  ~PinnedPool() {
    for (auto& kv : free_)
      for (void* p : kv.second) (void)hipHostFree(p);
  }

 private:
  std::mutex mu_;
  std::unordered_map<size_t, std::vector<void*>> free_;
};

}  // namespace

struct nabo_hip_server {
  nabo_hip_server_config cfg{};
  int device = 0;
  std::mutex mu;
  std::condition_variable wake;       // worker
  std::condition_variable finished;   // waiters
  std::deque<Request*> pending;
  std::unordered_map<int64_t, Request*> requests;
  std::vector<std::unique_ptr<Cloud>> clouds;  // stable addresses
  std::vector<Slot> slots;
  std::thread worker;
  std::thread builder;
  std::condition_variable build_wake;
  std::deque<int32_t> build_queue;
  PinnedPool pinned;
  bool stop = false;
  int64_t next_id = 1;
  uint64_t clock = 0;
  int64_t resident = 0;
  int64_t host_cached = 0;  // bytes of serialized trees in host memory
  nabo_hip_server_stats stats{};

  void run();
  void run_builder();
  bool retire();
  bool admit();
  void ensure_capacity(Slot& s, int64_t q_floats, int64_t results);
  void release_slot_buffers(Slot& s);
  int64_t slot_bytes() const {
    int64_t b = 0;
    for (const auto& s : slots) b += s.cap_q * 4 + s.cap_r * 8;
    return b;
  }
  void release(Request* r) {
    pinned.put(r->h_q, r->cap_q);
    pinned.put(r->h_i, r->cap_i);
    pinned.put(r->h_d, r->cap_d);
    r->h_q = nullptr;
    r->h_i = nullptr;
    r->h_d = nullptr;
  }
};

namespace {
template <typename F>
int server_guard(F&& f) {
  (void)hipGetLastError();  // an error left unread by an earlier call is not this call's
  try {
    f();
    nabo_hip::set_last_error("");
    return NABO_HIP_OK;
  } catch (const Error& e) {
    nabo_hip::set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    nabo_hip::set_last_error(e.what());
    return NABO_HIP_ERR_ARGUMENT;
  }
}

void check(int rc) {
  if (rc != NABO_HIP_OK) throw Error(rc, nabo_hip_last_error());
}
}  // namespace

void nabo_hip_server::ensure_capacity(Slot& s, int64_t q_floats, int64_t results) {
  if (q_floats > s.cap_q) {
    if (s.d_q) (void)hipFree(s.d_q);
    s.d_q = nullptr;
    NH_CHECK(hipMalloc(reinterpret_cast<void**>(&s.d_q), sizeof(float) * q_floats));
    s.cap_q = q_floats;
  }
  if (results > s.cap_r) {
    if (s.d_i) (void)hipFree(s.d_i);
    if (s.d_d) (void)hipFree(s.d_d);
    s.d_i = nullptr;
    s.d_d = nullptr;
    NH_CHECK(hipMalloc(reinterpret_cast<void**>(&s.d_i), sizeof(int32_t) * results));
    NH_CHECK(hipMalloc(reinterpret_cast<void**>(&s.d_d), sizeof(float) * results));
    s.cap_r = results;
  }
}

void nabo_hip_server::release_slot_buffers(Slot& s) {
  if (s.d_q) (void)hipFree(s.d_q);
  if (s.d_i) (void)hipFree(s.d_i);
  if (s.d_d) (void)hipFree(s.d_d);
  s = Slot{s.stream, s.event};
}

// Builder thread: for each queued cloud, evicts least recently used trees
// that no slot uses until the budget holds (the budget is a target: a tree
// larger than it is still built when nothing is left to evict), then builds
// the tree on this thread's stream, or deserializes its host copy. Host
// copies are kept up to host_cache_bytes, least recently used dropped first;
// a copy stays valid for good since a registered cloud never changes.
void nabo_hip_server::run_builder() {
  (void)hipSetDevice(device);
  std::unique_lock<std::mutex> lock(mu);
  for (;;) {
    build_wake.wait(lock, [&] { return stop || !build_queue.empty(); });
    if (build_queue.empty()) {
      if (stop) break;
      continue;
    }
    const int32_t id = build_queue.front();
    build_queue.pop_front();
    Cloud& c = *clouds[id];
    // Leaf-ordered points (16 B each) plus 24 B per node; a bucket of b
    // points makes about 4n/b nodes at most in practice.
    const int64_t bucket = std::max<int64_t>(1, int64_t(c.params.bucket_size));
    const int64_t estimate = c.n * int64_t(sizeof(float)) * 4 + 24 * (4 * c.n / bucket + 2);
    // Trees with pending requests are not evicted: a tree built for waiting
    // requests stays until they have run, so builds cannot undo each other.
    std::vector<char> wanted(clouds.size(), 0);
    for (const Request* r : pending) wanted[size_t(r->cloud)] = 1;
    std::vector<std::pair<Cloud*, nabo_hip_index*>> victims;
    while (cfg.hbm_budget_bytes > 0 && resident + slot_bytes() + estimate > cfg.hbm_budget_bytes) {
      Cloud* victim = nullptr;
      for (size_t i = 0; i < clouds.size(); ++i) {
        Cloud* o = clouds[i].get();
        if (o != &c && o->index && !o->in_use && !wanted[i] &&
            (!victim || o->last_use < victim->last_use))
          victim = o;
      }
      if (!victim) break;
      victims.emplace_back(victim, victim->index);
      victim->index = nullptr;
      resident -= victim->bytes;
      victim->bytes = 0;
      ++stats.evictions;
    }
    lock.unlock();
    // Evicted trees: keep a serialized copy in pinned host memory when the
    // host cache allows (only this thread touches the copies).
    std::vector<std::unique_ptr<HostCopy>> copies(victims.size());
    for (size_t v = 0; v < victims.size(); ++v) {
      Cloud* vc = victims[v].first;
      if (cfg.host_cache_bytes > 0 && !vc->host) {
        uint64_t bytes = 0;
        if (nabo_hip_serialize(victims[v].second, nullptr, 0, &bytes) == NABO_HIP_OK &&
            int64_t(bytes) <= cfg.host_cache_bytes) {
          auto hc = std::make_unique<HostCopy>();
          if (hipHostMalloc(&hc->p, bytes, hipHostMallocDefault) == hipSuccess &&
              nabo_hip_serialize(victims[v].second, hc->p, bytes, &hc->bytes) == NABO_HIP_OK)
            copies[v] = std::move(hc);
        }
      }
      nabo_hip_destroy(victims[v].second);
    }
    // Re-admission from the host copy when there is one, else a build.
    nabo_hip_index* index = nullptr;
    const bool readmit = c.host != nullptr;
    int rc;
    if (readmit) {
      rc = nabo_hip_deserialize(c.host->p, c.host->bytes, device, &index);
    } else {
      rc = nabo_hip_create(c.data.data(), c.n, c.rows, c.stride, c.dim, NABO_HIP_SEARCH_KDTREE, 0,
                           &c.params, &index);
    }
    std::string err = rc == NABO_HIP_OK ? "" : nabo_hip_last_error();
    nabo_hip_info info{};
    if (rc == NABO_HIP_OK) nabo_hip_get_info(index, &info);
    std::vector<std::unique_ptr<HostCopy>> dropped;
    lock.lock();
    for (size_t v = 0; v < victims.size(); ++v) {
      if (!copies[v]) continue;
      host_cached += int64_t(copies[v]->bytes);
      copies[v]->last_use = ++clock;
      victims[v].first->host = std::move(copies[v]);
    }
    if (c.host) c.host->last_use = ++clock;
    // Host cache over budget: drop the least recently used copies.
    while (host_cached > cfg.host_cache_bytes) {
      Cloud* oldest = nullptr;
      for (auto& o : clouds)
        if (o->host && (!oldest || o->host->last_use < oldest->host->last_use)) oldest = o.get();
      if (!oldest) break;
      host_cached -= int64_t(oldest->host->bytes);
      dropped.push_back(std::move(oldest->host));
    }
    c.building = false;
    c.build_status = rc;
    c.build_error = err;
    if (rc == NABO_HIP_OK) {
      c.index = index;
      c.bytes = info.device_bytes;
      resident += c.bytes;
      if (readmit)
        ++stats.readmits;
      else
        ++stats.builds;
      stats.peak_bytes = std::max(stats.peak_bytes, resident + slot_bytes());
    }
    wake.notify_all();
    lock.unlock();
    dropped.clear();  // hipHostFree outside the lock
    lock.lock();
  }
}

bool nabo_hip_server::retire() {
  bool any = false;
  for (auto& s : slots) {
    if (!s.busy) continue;
    const hipError_t q = hipEventQuery(s.event);
    (void)hipGetLastError();  // q carries the result; keep the launch checks clean
    if (q == hipErrorNotReady) continue;
    const double t = now_ms();
    for (Request* r : s.reqs) {
      if (q != hipSuccess) {
        r->status = NABO_HIP_ERR_HIP;
        r->error = hipGetErrorString(q);
      }
      r->latency = t - r->t_submit;
      r->done = true;
      ++stats.completed;
    }
    clouds[s.cloud]->in_use--;
    s.reqs.clear();
    s.busy = false;
    any = true;
  }
  if (any) finished.notify_all();
  return any;
}

bool nabo_hip_server::admit() {
  bool any = false;
  auto fail = [&](Request* r, int code, const std::string& msg) {
    r->status = code;
    r->error = msg;
    r->done = true;
    r->latency = now_ms() - r->t_submit;
    ++stats.completed;
    finished.notify_all();
  };
  for (auto it = pending.begin(); it != pending.end();) {
    Request* head = *it;
    Cloud& c = *clouds[head->cloud];
    if (!c.index) {
      if (c.build_status != NABO_HIP_OK && !c.building) {  // a failed build fails its requests
        fail(head, c.build_status, c.build_error);
        it = pending.erase(it);
        any = true;
        continue;
      }
      if (!c.building) {
        c.building = true;
        build_queue.push_back(head->cloud);
        build_wake.notify_one();
      }
      ++it;
      continue;
    }
    Slot* s = nullptr;
    for (auto& sl : slots)
      if (!sl.busy) {
        s = &sl;
        break;
      }
    if (!s) break;
    // The group: head and the pending requests it can share a launch with.
    // Erasing from the middle of a deque invalidates every iterator, so the
    // scan of `pending` restarts from its beginning afterwards.
    std::vector<Request*> group;
    int64_t m = 0;
    for (auto jt = it; jt != pending.end();) {
      Request* r = *jt;
      const bool same = r->cloud == head->cloud && r->k == head->k && r->eps == head->eps &&
                        r->flags == head->flags && r->radius == head->radius;
      if (same && (group.empty() || m + r->m <= cfg.max_merge_queries)) {
        group.push_back(r);
        m += r->m;
        jt = pending.erase(jt);
      } else {
        ++jt;
      }
    }
    it = pending.begin();
    try {
      const int32_t dim = c.dim;
      const int32_t kk = head->k;
      ensure_capacity(*s, std::max<int64_t>(1, m) * dim, std::max<int64_t>(1, m) * kk);
      int64_t off = 0;
      for (Request* r : group) {
        NH_CHECK(hipMemcpyAsync(s->d_q + off * dim, r->h_q, sizeof(float) * r->m * dim,
                                hipMemcpyHostToDevice, s->stream));
        off += r->m;
      }
      check(nabo_hip_knn_device(c.index, s->d_q, dim, dim, m, kk, head->eps, head->flags,
                                head->radius, nullptr, s->d_i, s->d_d, nullptr, s->stream));
      off = 0;
      for (Request* r : group) {
        NH_CHECK(hipMemcpyAsync(r->h_i, s->d_i + off * kk, sizeof(int32_t) * r->m * kk,
                                hipMemcpyDeviceToHost, s->stream));
        NH_CHECK(hipMemcpyAsync(r->h_d, s->d_d + off * kk, sizeof(float) * r->m * kk,
                                hipMemcpyDeviceToHost, s->stream));
        off += r->m;
      }
      NH_CHECK(hipEventRecord(s->event, s->stream));
    } catch (const Error& e) {
      for (Request* r : group) fail(r, e.code, e.what());
      any = true;
      continue;
    }
    c.in_use++;
    c.last_use = ++clock;
    s->cloud = head->cloud;
    s->reqs = std::move(group);
    s->busy = true;
    ++stats.launches;
    stats.merged_requests += int64_t(s->reqs.size()) - 1;
    stats.peak_bytes = std::max(stats.peak_bytes, resident + slot_bytes());
    any = true;
  }
  return any;
}

void nabo_hip_server::run() {
  (void)hipSetDevice(device);
  std::unique_lock<std::mutex> lock(mu);
  for (;;) {
    const bool busy_slots = std::any_of(slots.begin(), slots.end(), [](const Slot& s) {
      return s.busy;
    });
    if (stop && pending.empty() && !busy_slots) break;
    (void)busy_slots;
    const bool progressed = retire() | admit();
    if (progressed) continue;
    if (busy_slots) {
      // work in flight: poll the slot events at a short interval
      lock.unlock();
      std::this_thread::sleep_for(std::chrono::microseconds(20));
      lock.lock();
    } else if (std::any_of(pending.begin(), pending.end(),
                           [&](Request* r) { return clouds[r->cloud]->index != nullptr; })) {
      continue;
    } else {
      // nothing runnable: wait for a submit, a finished build or shutdown
      wake.wait_for(lock, std::chrono::milliseconds(50));
    }
  }
}

extern "C" {

void nabo_hip_server_config_default(nabo_hip_server_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->hbm_budget_bytes = 0;
  c->slots = 4;
  c->max_merge_queries = int64_t(1) << 22;
  c->device = -1;
}

int nabo_hip_server_create(const nabo_hip_server_config* config, nabo_hip_server** out) {
  return server_guard([&] {
    if (!out) throw Error(NABO_HIP_ERR_ARGUMENT, "output handle pointer is null");
    *out = nullptr;
    auto* s = new nabo_hip_server();
    nabo_hip_server_config_default(&s->cfg);
    if (config) s->cfg = *config;
    if (s->cfg.slots < 1 || s->cfg.slots > 64) {
      delete s;
      throw Error(NABO_HIP_ERR_ARGUMENT, "slots must be in [1, 64]");
    }
    if (s->cfg.max_merge_queries < 1) s->cfg.max_merge_queries = 1;
    int dev = s->cfg.device;
    if (dev < 0) NH_CHECK(hipGetDevice(&dev));
    s->device = dev;
    NH_CHECK(hipSetDevice(dev));
    s->slots.resize(size_t(s->cfg.slots));
    for (auto& sl : s->slots) {
      NH_CHECK(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking));
      NH_CHECK(hipEventCreateWithFlags(&sl.event, hipEventDisableTiming));
    }
    s->worker = std::thread([s] { s->run(); });
    s->builder = std::thread([s] { s->run_builder(); });
    *out = s;
  });
}

int nabo_hip_server_destroy(nabo_hip_server* s) {
  return server_guard([&] {
    if (!s) return;
    {
      std::lock_guard<std::mutex> lock(s->mu);
      s->stop = true;
    }
    s->wake.notify_all();
    if (s->worker.joinable()) s->worker.join();
    s->build_wake.notify_all();
    if (s->builder.joinable()) s->builder.join();
    (void)hipSetDevice(s->device);
    for (auto& sl : s->slots) {
      s->release_slot_buffers(sl);
      (void)hipEventDestroy(sl.event);
      (void)hipStreamDestroy(sl.stream);
    }
    for (auto& c : s->clouds)
      if (c->index) nabo_hip_destroy(c->index);
    for (auto& kv : s->requests) {
      s->release(kv.second);
      delete kv.second;
    }
    delete s;
  });
}

int nabo_hip_server_add_cloud(nabo_hip_server* s, const float* cloud, int64_t n, int32_t rows,
                              int64_t stride, int32_t dim, const nabo_hip_params* params,
                              int32_t* cloud_id) {
  return server_guard([&] {
    if (!s || !cloud || !cloud_id) throw Error(NABO_HIP_ERR_ARGUMENT, "null argument");
    if (n <= 0 || rows <= 0 || stride < rows || dim <= 0)
      throw Error(NABO_HIP_ERR_ARGUMENT, "invalid cloud shape");
    auto c = std::make_unique<Cloud>();
    c->n = n;
    c->rows = rows;
    c->stride = rows;
    c->dim = std::min(dim, rows);
    nabo_hip_params_default(&c->params);
    if (params) c->params = *params;
    c->data.resize(size_t(n) * rows);
    for (int64_t i = 0; i < n; ++i)
      std::memcpy(c->data.data() + i * rows, cloud + i * stride, sizeof(float) * rows);
    std::lock_guard<std::mutex> lock(s->mu);
    *cloud_id = int32_t(s->clouds.size());
    s->clouds.push_back(std::move(c));
  });
}

int nabo_hip_server_submit(nabo_hip_server* s, int32_t cloud_id, const float* query,
                           int32_t qrows, int64_t qstride, int64_t m, int32_t k, float epsilon,
                           uint32_t option_flags, float max_radius, int32_t* indices,
                           float* dists2, int64_t* request_id) {
  return server_guard([&] {
    if (!s || !request_id) throw Error(NABO_HIP_ERR_ARGUMENT, "null argument");
    auto* r = new Request();
    int32_t dim = 0;
    {
      std::lock_guard<std::mutex> lock(s->mu);
      if (cloud_id >= 0 && cloud_id < int32_t(s->clouds.size())) dim = s->clouds[cloud_id]->dim;
    }
    // Stage the compacted queries and take the result buffers here, in the
    // caller's thread.
    if (m > 0 && k > 0 && query && dim > 0 && qrows >= dim && qstride >= qrows) {
      try {
        r->h_q = static_cast<float*>(s->pinned.get(sizeof(float) * m * dim, &r->cap_q));
        r->h_i = static_cast<int32_t*>(s->pinned.get(sizeof(int32_t) * m * k, &r->cap_i));
        r->h_d = static_cast<float*>(s->pinned.get(sizeof(float) * m * k, &r->cap_d));
      } catch (...) {
        s->release(r);
        delete r;
        throw;
      }
      if (qstride == dim)
        nabo_hip::host_copy(r->h_q, query, sizeof(float) * size_t(m) * size_t(dim));
      else
        for (int64_t i = 0; i < m; ++i)
          std::memcpy(r->h_q + i * dim, query + i * qstride, sizeof(float) * dim);
    }
    {
      std::lock_guard<std::mutex> lock(s->mu);
      if (cloud_id < 0 || cloud_id >= int32_t(s->clouds.size())) {
        s->release(r);
        delete r;
        throw Error(NABO_HIP_ERR_ARGUMENT, "unknown cloud id");
      }
      const Cloud& c = *s->clouds[cloud_id];
      const bool self = option_flags & NABO_HIP_ALLOW_SELF_MATCH;
      std::string err;
      if (self && k > c.n)
        err = "Requesting more points (" + std::to_string(k) + ") than available in cloud (" +
              std::to_string(c.n) + ")";
      else if (!self && k > c.n - 1)
        err = "Requesting more points (" + std::to_string(k) +
              ") than available in cloud minus 1 (" + std::to_string(c.n - 1) +
              ") (as self match is forbidden)";
      else if (qrows < c.dim)
        err = "Query has less dimensions (" + std::to_string(qrows) +
              ") than requested for cloud (" + std::to_string(c.dim) + ")";
      else if (option_flags > 3u)
        err = "OR-ed value of option flags (" + std::to_string(option_flags) +
              ") is larger than maximal valid value (3)";
      else if (k < 1 || m < 0 || (m > 0 && (!query || !indices || !dists2 || qstride < qrows)))
        err = "invalid request arguments";
      if (!err.empty()) {
        s->release(r);
        delete r;
        throw Error(NABO_HIP_ERR_ARGUMENT, err);
      }
      r->id = s->next_id++;
      r->cloud = cloud_id;
      r->query = query;
      r->qrows = qrows;
      r->qstride = qstride;
      r->m = m;
      r->k = k;
      r->eps = epsilon;
      r->flags = option_flags;
      r->radius = max_radius;
      r->out_idx = indices;
      r->out_d2 = dists2;
      r->t_submit = now_ms();
      s->requests[r->id] = r;
      ++s->stats.submitted;
      if (m == 0) {
        r->done = true;
        ++s->stats.completed;
      } else {
        s->pending.push_back(r);
      }
      *request_id = r->id;
    }
    s->wake.notify_one();
  });
}

int nabo_hip_server_wait(nabo_hip_server* s, int64_t request_id, double* latency_ms) {
  int status = NABO_HIP_OK;
  const int rc = server_guard([&] {
    if (!s) throw Error(NABO_HIP_ERR_ARGUMENT, "null server");
    std::unique_lock<std::mutex> lock(s->mu);
    auto it = s->requests.find(request_id);
    if (it == s->requests.end()) throw Error(NABO_HIP_ERR_ARGUMENT, "unknown request id");
    Request* r = it->second;
    if (!r->done && r->m > 0 && r->out_idx && r->out_d2) {
      // the output pages take their first-touch faults while the device works
      const size_t bytes = sizeof(float) * size_t(r->m) * size_t(r->k);
      int32_t* oi = r->out_idx;
      float* od = r->out_d2;
      lock.unlock();
      nabo_hip::host_touch(oi, bytes);
      nabo_hip::host_touch(od, bytes);
      lock.lock();
    }
    s->finished.wait(lock, [&] { return r->done; });
    if (latency_ms) *latency_ms = r->latency;
    status = r->status;
    const std::string err = r->error;
    s->requests.erase(request_id);  // by key: `it` may predate a rehash
    lock.unlock();
    if (status == NABO_HIP_OK && r->m > 0) {
      nabo_hip::host_copy(r->out_idx, r->h_i, sizeof(int32_t) * r->m * r->k);
      nabo_hip::host_copy(r->out_d2, r->h_d, sizeof(float) * r->m * r->k);
    }
    s->release(r);
    delete r;
    if (status != NABO_HIP_OK) throw Error(status, err);
  });
  return rc;
}

int nabo_hip_server_get_stats(const nabo_hip_server* s, nabo_hip_server_stats* out) {
  return server_guard([&] {
    if (!s || !out) throw Error(NABO_HIP_ERR_ARGUMENT, "null argument");
    auto* ms = const_cast<nabo_hip_server*>(s);
    std::lock_guard<std::mutex> lock(ms->mu);
    *out = s->stats;
    int64_t trees = 0;
    for (const auto& c : s->clouds) trees += c->index != nullptr;
    out->resident_trees = trees;
    out->resident_bytes = s->resident;
    out->host_cached_bytes = s->host_cached;
  });
}

}  // extern "C"
