// Served kNN (nabo_hip_server_*): resident trees under an HBM budget with
// LRU eviction, and a worker thread that admits requests continuously onto
// a set of stream slots.
//
// The server is a client of the C ABI: trees are nabo_hip_create indices,
// searches are nabo_hip_knn_device launches, so a served result is the
// result of nabo_hip_knn on the same cloud alone.
//
// Worker loop, one pass:
//   1. retire finished slots (event query): copy the pinned outputs into the
//      callers' buffers, wake the waiters;
//   2. while a slot is free and requests are pending (arrival order): make
//      the request's tree resident (evicting least recently used trees that
//      no slot is using until the budget holds), merge the pending requests
//      on the same cloud with the same k / epsilon / flags / radius behind
//      it (up to max_merge_queries), and enqueue copy-in, search and
//      copy-out on the slot's stream.
// Slots run concurrently on the device; a request that arrives while others
// run is admitted as soon as a slot and memory allow.

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
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

struct Cloud {
  std::vector<float> data;  // rows x n, column stride `stride`
  int64_t n = 0;
  int32_t rows = 0;
  int64_t stride = 0;
  int32_t dim = 0;
  nabo_hip_params params{};
  nabo_hip_index* index = nullptr;
  int64_t bytes = 0;
  uint64_t last_use = 0;
  int in_use = 0;
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
  int64_t cap_q = 0, cap_r = 0;  // floats of query staging, entries of result staging
  float* d_q = nullptr;
  int32_t* d_i = nullptr;
  float* d_d = nullptr;
  float* h_q = nullptr;
  int32_t* h_i = nullptr;
  float* h_d = nullptr;
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
  std::vector<Cloud> clouds;
  std::vector<Slot> slots;
  std::thread worker;
  bool stop = false;
  int64_t next_id = 1;
  uint64_t clock = 0;
  int64_t resident = 0;
  nabo_hip_server_stats stats{};

  void run();
  bool retire();
  bool admit();
  void make_resident(Cloud& c, int32_t self);
  void ensure_capacity(Slot& s, int64_t q_floats, int64_t results);
  void release_slot_buffers(Slot& s);
  int64_t slot_bytes() const {
    int64_t b = 0;
    for (const auto& s : slots) b += s.cap_q * 4 + s.cap_r * 8;
    return b;
  }
};

namespace {
template <typename F>
int server_guard(F&& f) {
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
    if (s.h_q) (void)hipHostFree(s.h_q);
    s.d_q = nullptr;
    s.h_q = nullptr;
    NH_CHECK(hipMalloc(reinterpret_cast<void**>(&s.d_q), sizeof(float) * q_floats));
    NH_CHECK(hipHostMalloc(reinterpret_cast<void**>(&s.h_q), sizeof(float) * q_floats));
    s.cap_q = q_floats;
  }
  if (results > s.cap_r) {
    if (s.d_i) (void)hipFree(s.d_i);
    if (s.d_d) (void)hipFree(s.d_d);
    if (s.h_i) (void)hipHostFree(s.h_i);
    if (s.h_d) (void)hipHostFree(s.h_d);
    s.d_i = nullptr;
    s.d_d = nullptr;
    s.h_i = nullptr;
    s.h_d = nullptr;
    NH_CHECK(hipMalloc(reinterpret_cast<void**>(&s.d_i), sizeof(int32_t) * results));
    NH_CHECK(hipMalloc(reinterpret_cast<void**>(&s.d_d), sizeof(float) * results));
    NH_CHECK(hipHostMalloc(reinterpret_cast<void**>(&s.h_i), sizeof(int32_t) * results));
    NH_CHECK(hipHostMalloc(reinterpret_cast<void**>(&s.h_d), sizeof(float) * results));
    s.cap_r = results;
  }
}

void nabo_hip_server::release_slot_buffers(Slot& s) {
  if (s.d_q) (void)hipFree(s.d_q);
  if (s.d_i) (void)hipFree(s.d_i);
  if (s.d_d) (void)hipFree(s.d_d);
  if (s.h_q) (void)hipHostFree(s.h_q);
  if (s.h_i) (void)hipHostFree(s.h_i);
  if (s.h_d) (void)hipHostFree(s.h_d);
  s = Slot{s.stream, s.event};
}

// Builds the tree of `c` if needed, first evicting least recently used trees
// that no slot uses until the budget holds (the budget is a target: a tree
// larger than it is still built when nothing is left to evict).
void nabo_hip_server::make_resident(Cloud& c, int32_t self) {
  if (c.index) return;
  const int64_t estimate = c.n * int64_t(sizeof(float)) * 4 + c.n * 8;
  const int64_t budget = cfg.hbm_budget_bytes;
  while (budget > 0 && resident + slot_bytes() + estimate > budget) {
    int32_t victim = -1;
    for (int32_t i = 0; i < int32_t(clouds.size()); ++i) {
      const Cloud& o = clouds[i];
      if (i == self || !o.index || o.in_use) continue;
      if (victim < 0 || o.last_use < clouds[victim].last_use) victim = i;
    }
    if (victim < 0) break;
    check(nabo_hip_destroy(clouds[victim].index));
    clouds[victim].index = nullptr;
    resident -= clouds[victim].bytes;
    clouds[victim].bytes = 0;
    ++stats.evictions;
  }
  check(nabo_hip_create(c.data.data(), c.n, c.rows, c.stride, c.dim, NABO_HIP_SEARCH_KDTREE, 0,
                        &c.params, &c.index));
  nabo_hip_info info;
  check(nabo_hip_get_info(c.index, &info));
  c.bytes = info.device_bytes;
  resident += c.bytes;
  ++stats.builds;
  stats.peak_bytes = std::max(stats.peak_bytes, resident + slot_bytes());
}

bool nabo_hip_server::retire() {
  bool any = false;
  for (auto& s : slots) {
    if (!s.busy) continue;
    const hipError_t q = hipEventQuery(s.event);
    if (q == hipErrorNotReady) continue;
    int64_t off = 0;
    const double t = now_ms();
    for (Request* r : s.reqs) {
      if (q != hipSuccess) {
        r->status = NABO_HIP_ERR_HIP;
        r->error = hipGetErrorString(q);
      } else {
        std::memcpy(r->out_idx, s.h_i + off * r->k, sizeof(int32_t) * r->m * r->k);
        std::memcpy(r->out_d2, s.h_d + off * r->k, sizeof(float) * r->m * r->k);
      }
      off += r->m;
      r->latency = t - r->t_submit;
      r->done = true;
      ++stats.completed;
    }
    clouds[s.cloud].in_use--;
    s.reqs.clear();
    s.busy = false;
    any = true;
  }
  if (any) finished.notify_all();
  return any;
}

bool nabo_hip_server::admit() {
  bool any = false;
  for (auto& s : slots) {
    if (s.busy || pending.empty()) continue;
    Request* head = pending.front();
    Cloud& c = clouds[head->cloud];
    std::vector<Request*> group;
    int64_t m = 0;
    for (auto it = pending.begin(); it != pending.end();) {
      Request* r = *it;
      const bool same = r->cloud == head->cloud && r->k == head->k && r->eps == head->eps &&
                        r->flags == head->flags && r->radius == head->radius;
      if (same && (group.empty() || m + r->m <= cfg.max_merge_queries)) {
        group.push_back(r);
        m += r->m;
        it = pending.erase(it);
      } else {
        ++it;
      }
    }
    try {
      make_resident(c, head->cloud);
      const int32_t dim = c.dim;
      ensure_capacity(s, std::max<int64_t>(1, m) * dim, std::max<int64_t>(1, m) * head->k);
      int64_t off = 0;
      for (Request* r : group) {  // compact each query set to dim floats per point
        for (int64_t i = 0; i < r->m; ++i)
          std::memcpy(s.h_q + (off + i) * dim, r->query + i * r->qstride, sizeof(float) * dim);
        off += r->m;
      }
      NH_CHECK(hipMemcpyAsync(s.d_q, s.h_q, sizeof(float) * m * dim, hipMemcpyHostToDevice,
                              s.stream));
      check(nabo_hip_knn_device(c.index, s.d_q, dim, dim, m, head->k, head->eps, head->flags,
                                head->radius, nullptr, s.d_i, s.d_d, nullptr, s.stream));
      NH_CHECK(hipMemcpyAsync(s.h_i, s.d_i, sizeof(int32_t) * m * head->k,
                              hipMemcpyDeviceToHost, s.stream));
      NH_CHECK(hipMemcpyAsync(s.h_d, s.d_d, sizeof(float) * m * head->k, hipMemcpyDeviceToHost,
                              s.stream));
      NH_CHECK(hipEventRecord(s.event, s.stream));
    } catch (const Error& e) {
      for (Request* r : group) {
        r->status = e.code;
        r->error = e.what();
        r->done = true;
        r->latency = now_ms() - r->t_submit;
        ++stats.completed;
      }
      finished.notify_all();
      continue;
    }
    c.in_use++;
    c.last_use = ++clock;
    s.cloud = head->cloud;
    s.reqs = std::move(group);
    s.busy = true;
    ++stats.launches;
    stats.merged_requests += int64_t(s.reqs.size()) - 1;
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
    const bool progressed = retire() | admit();
    if (progressed) continue;
    if (busy_slots) {
      // work in flight: poll the slot events at a short interval
      lock.unlock();
      std::this_thread::sleep_for(std::chrono::microseconds(20));
      lock.lock();
    } else {
      wake.wait(lock, [&] { return stop || !pending.empty(); });
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
    (void)hipSetDevice(s->device);
    for (auto& sl : s->slots) {
      s->release_slot_buffers(sl);
      (void)hipEventDestroy(sl.event);
      (void)hipStreamDestroy(sl.stream);
    }
    for (auto& c : s->clouds)
      if (c.index) nabo_hip_destroy(c.index);
    for (auto& kv : s->requests) delete kv.second;
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
    Cloud c;
    c.n = n;
    c.rows = rows;
    c.stride = rows;
    c.dim = std::min(dim, rows);
    nabo_hip_params_default(&c.params);
    if (params) c.params = *params;
    c.data.resize(size_t(n) * rows);
    for (int64_t i = 0; i < n; ++i)
      std::memcpy(c.data.data() + i * rows, cloud + i * stride, sizeof(float) * rows);
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
    {
      std::lock_guard<std::mutex> lock(s->mu);
      if (cloud_id < 0 || cloud_id >= int32_t(s->clouds.size())) {
        delete r;
        throw Error(NABO_HIP_ERR_ARGUMENT, "unknown cloud id");
      }
      const Cloud& c = s->clouds[cloud_id];
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
    s->finished.wait(lock, [&] { return r->done; });
    if (latency_ms) *latency_ms = r->latency;
    status = r->status;
    const std::string err = r->error;
    s->requests.erase(it);
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
    for (const auto& c : s->clouds) trees += c.index != nullptr;
    out->resident_trees = trees;
    out->resident_bytes = s->resident;
  });
}

}  // extern "C"
