// Batched kd-tree build and search recorded as one HIP graph
// (nabo_hip_batch_plan_*).
//
// nabo_hip_batch_create builds many trees in one pass, but its level loop
// reads each level's split count on the host, and a knn call allocates and
// sizes its temporaries per call. A plan does both once: at creation it
// allocates the trees' arrays at their largest size (2n nodes), the build's
// and the search's scratch, and records the whole launch sequence (the
// build with a fixed number of levels, the per-query cloud ids, the query
// order and the search) as kernel nodes of one graph, which it
// instantiates. Each launch then rebuilds every tree from the current
// contents of the point buffer and searches the current query buffer with
// one graph launch: no host round trip, no allocation. The graph holds
// kernel nodes only (no memory, memset or event nodes), and every pointer
// it holds is to memory the plan owns or the caller gave, alive until the
// plan is destroyed.
//
// The level budget: a tree needs one level per split depth, unknown before
// the data is seen. With levels == 0 the plan builds the data once at
// creation (eagerly) and takes its depth plus kLevelMargin as the budget,
// recording as many levels of the build's level loop as that build ran plus
// kLevelMargin (the deeper levels are finished in LDS); a launch whose data
// needs more loop levels, or is deeper than the budget, finishes no tree and
// runs no search (both check a completeness word the build writes), which
// nabo_hip_batch_plan_status reports.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

#include "internal.h"
#include "launch.h"

namespace nabo_hip {
namespace {

constexpr int kLevelMargin = 4;
constexpr int kMaxPlanLevels = 33;  // trees of at most 32 split levels (LDS search stacks)

template <class T>
void dev_alloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) return;
  const hipError_t e = hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * count);
  if (e != hipSuccess) {
    *p = nullptr;
    throw Error(NABO_HIP_ERR_MEMORY, "device allocation of " + std::to_string(sizeof(T) * count) +
                                         " bytes failed: " + hipGetErrorString(e));
  }
}

}  // namespace
}  // namespace nabo_hip

struct nabo_hip_batch_plan {
  nabo_hip::Index ix;  // the plan's trees (arrays owned, freed with it)
  int levels = 0;
  int64_t m = 0;
  void* build_scratch = nullptr;
  void* search_scratch = nullptr;
  int32_t* qcloud = nullptr;
  int64_t* d_query_offsets = nullptr;
  int32_t* ok = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t last = nullptr;  // stream of the last launch
  bool launched = false;
  std::vector<std::shared_ptr<void>> arguments;  // the graph's kernel argument values

  ~nabo_hip_batch_plan() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (void* p : {build_scratch, search_scratch, static_cast<void*>(qcloud),
                    static_cast<void*>(d_query_offsets), static_cast<void*>(ok)})
      if (p) (void)hipFree(p);
  }
};

using namespace nabo_hip;

extern "C" {

int nabo_hip_batch_plan_create(const float* d_points, const int64_t* offsets, int32_t clouds,
                               int32_t rows, int64_t stride, int32_t dim,
                               const nabo_hip_params* params, const float* d_query,
                               int64_t qstride, const int64_t* query_offsets, int32_t k,
                               float epsilon, uint32_t option_flags, float max_radius,
                               int32_t* d_indices, float* d_dists2, int32_t levels,
                               nabo_hip_batch_plan** out) {
  return run_guarded([&] {
    if (!out) throw Error(NABO_HIP_ERR_ARGUMENT, "output handle pointer is null");
    *out = nullptr;
    if (clouds < 1 || !offsets || !query_offsets)
      throw Error(NABO_HIP_ERR_ARGUMENT, "a batch needs at least one cloud and its offsets");
    nabo_hip_params p;
    nabo_hip_params_default(&p);
    if (params) p = *params;
    std::vector<int64_t> off(offsets, offsets + clouds + 1);
    std::vector<int64_t> qoff(query_offsets, query_offsets + clouds + 1);
    if (off[0] != 0 || qoff[0] != 0)
      throw Error(NABO_HIP_ERR_ARGUMENT, "cloud and query offsets must start at 0");
    for (int32_t c = 0; c < clouds; ++c) {
      if (off[c + 1] < off[c] || qoff[c + 1] < qoff[c])
        throw Error(NABO_HIP_ERR_ARGUMENT, "offsets must not decrease");
      check_create(off[c + 1] - off[c], rows, stride, dim, NABO_HIP_SEARCH_KDTREE, p);
    }
    const int64_t n = off[clouds], m = qoff[clouds];
    if (n >= kKdPointsLimit)
      throw Error(NABO_HIP_ERR_UNSUPPORTED, "batches of 2^29 points or more are not supported");
    const int used_dim = std::min(dim, rows);
    if (used_dim != 2 && used_dim != 3)
      throw Error(NABO_HIP_ERR_UNSUPPORTED, "batch plans serve 2-D and 3-D clouds");
    if (p.heap != NABO_HIP_HEAP_LINEAR)
      throw Error(NABO_HIP_ERR_UNSUPPORTED,
                  "batch plans return KDTREE_HIP (IndexHeapBruteForceVector) results");
    if (levels < 0 || levels > kMaxPlanLevels)
      throw Error(NABO_HIP_ERR_ARGUMENT, "levels must be 0 (learned) or 1 to " +
                                             std::to_string(kMaxPlanLevels));
    if (!d_points || (m > 0 && (!d_query || !d_indices || !d_dists2)))
      throw Error(NABO_HIP_ERR_ARGUMENT, "point, query or output pointer is null");
    int dev = p.device;
    if (dev < 0) NH_CHECK(hipGetDevice(&dev));
    DeviceScope scope(dev);

    auto* plan = new nabo_hip_batch_plan();
    try {
      Index& ix = plan->ix;
      ix.search_type = NABO_HIP_SEARCH_KDTREE;
      ix.device = dev;
      ix.rows = rows;
      ix.dim = used_dim;
      ix.n = n;
      ix.params = p;
      ix.bucket_size = p.bucket_size;
      ix.dim_bits = storage_bits(uint32_t(ix.dim));
      ix.dim_mask = (1u << ix.dim_bits) - 1u;
      ix.pts_stride = 4;
      ix.clouds = clouds;
      ix.cloud_offsets = off;
      check_knn(ix, rows, qstride, m, k, option_flags);

      // The level budget: learned from one eager build of the data.
      int loop_levels = levels;
      if (levels == 0) {
        nabo_hip_index* probe = nullptr;
        hipStream_t s = nullptr;
        NH_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        const int rc = nabo_hip_batch_create(d_points, off.data(), clouds, rows, stride, dim, 0,
                                             &p, s, &probe);
        nabo_hip_info info{};
        int probe_loop = 0;
        if (rc == NABO_HIP_OK) {
          nabo_hip_get_info(probe, &info);
          probe_loop = probe->ix.build_levels;
        }
        nabo_hip_destroy(probe);
        (void)hipStreamDestroy(s);
        if (rc != NABO_HIP_OK) throw Error(rc, nabo_hip_last_error());
        levels = std::min(kMaxPlanLevels, info.depth + 1 + kLevelMargin);
        loop_levels = std::min(levels, probe_loop + kLevelMargin);
        if (info.depth + 1 > kMaxPlanLevels)
          throw Error(NABO_HIP_ERR_UNSUPPORTED,
                      "batch plans serve trees of at most " + std::to_string(kMaxPlanLevels - 1) +
                          " levels; this batch's deepest tree has " +
                          std::to_string(info.depth));
      }
      plan->levels = levels;
      plan->m = m;
      ix.depth = levels - 1;  // the deepest tree the recorded levels can build
      ix.node_count = 2 * n;  // capacity (the search reads only reachable nodes)
      ix.min_bound.assign(size_t(clouds) * ix.dim, 0.f);
      ix.max_bound.assign(size_t(clouds) * ix.dim, 0.f);

      dev_alloc(&ix.d_pts, size_t(n + kPtsPad) * ix.pts_stride);
      dev_alloc(&ix.d_nodes, size_t(2 * n + 1));
      dev_alloc(&ix.d_links, size_t(2 * n));
      dev_alloc(&ix.d_roots, size_t(clouds));
      dev_alloc(&ix.d_boxes, size_t(6) * clouds);
      ix.device_bytes = int64_t(sizeof(float)) * (n + kPtsPad) * ix.pts_stride +
                        int64_t(sizeof(uint2)) * (2 * n + 1) + int64_t(sizeof(uint4)) * 2 * n;
      dev_alloc(&plan->ok, 1);
      dev_alloc(&plan->qcloud, size_t(std::max<int64_t>(m, 1)));
      dev_alloc(&plan->d_query_offsets, size_t(clouds) + 1);
      NH_CHECK(hipMemcpy(plan->d_query_offsets, qoff.data(), sizeof(int64_t) * qoff.size(),
                         hipMemcpyHostToDevice));

      const KdBuildShape sh{n, ix.dim, ix.pts_stride, clouds, ix.bucket_size, ix.dim_bits,
                            ix.dim_mask};
      const size_t build_bytes = kd_build_record_bytes(sh);
      NH_CHECK(hipMalloc(&plan->build_scratch, build_bytes));

      KnnArgs a;
      a.query = d_query;
      a.qrows = rows;
      a.qstride = qstride;
      a.m = m;
      a.k = k;
      a.epsilon = epsilon;
      a.flags = option_flags;
      a.max_radius = max_radius;
      a.out_idx = d_indices;
      a.out_d2 = d_dists2;
      a.qcloud = plan->qcloud;
      a.ok = plan->ok;
      size_t search_bytes = 0;
      a.measure = &search_bytes;
      {
        Launcher sizing(static_cast<hipStream_t>(nullptr));
        kd_knn_launch(ix, a, sizing);
      }
      a.measure = nullptr;
      NH_CHECK(hipMalloc(&plan->search_scratch, std::max<size_t>(search_bytes, 256)));
      ExternalScratch ext;
      ext.base = plan->search_scratch;
      ext.bytes = std::max<size_t>(search_bytes, 256);
      a.scratch = &ext;

      NH_CHECK(hipGraphCreate(&plan->graph, 0));
      Launcher L(plan->graph);
      KdBuildOut bo;
      bo.pts = ix.d_pts;
      bo.nodes = ix.d_nodes;
      bo.links = ix.d_links;
      bo.roots = ix.d_roots;
      bo.boxes = ix.d_boxes;
      bo.ok = plan->ok;
      kd_build_record(sh, d_points, stride, off.data(), plan->build_scratch, build_bytes, bo, L,
                      loop_levels, levels - 1);
      if (m > 0 && k > 0) {
        L(k_query_clouds, dim3(unsigned(ceil_div(m, 256))), dim3(256), 0, m, clouds,
          static_cast<const int64_t*>(plan->d_query_offsets), plan->qcloud);
        kd_knn_launch(ix, a, L);
      }
      NH_CHECK(hipGraphInstantiate(&plan->exec, plan->graph, nullptr, nullptr, 0));
      plan->arguments = L.take_arguments();
    } catch (...) {
      delete plan;
      throw;
    }
    *out = plan;
  });
}

int nabo_hip_batch_plan_launch(nabo_hip_batch_plan* plan, void* hip_stream) {
  return run_guarded([&] {
    if (!plan) throw Error(NABO_HIP_ERR_ARGUMENT, "plan is null");
    DeviceScope scope(plan->ix.device);
    hipStream_t s = static_cast<hipStream_t>(hip_stream);
    NH_CHECK(hipGraphLaunch(plan->exec, s));
    plan->last = s;
    plan->launched = true;
  });
}

int nabo_hip_batch_plan_status(nabo_hip_batch_plan* plan, int32_t* complete, int32_t* levels) {
  return run_guarded([&] {
    if (!plan) throw Error(NABO_HIP_ERR_ARGUMENT, "plan is null");
    DeviceScope scope(plan->ix.device);
    if (levels) *levels = plan->levels;
    if (!plan->launched) {
      if (complete) *complete = 0;
      return;
    }
    NH_CHECK(hipStreamSynchronize(plan->last));
    int32_t ok = 0;
    NH_CHECK(hipMemcpy(&ok, plan->ok, sizeof(ok), hipMemcpyDeviceToHost));
    if (complete) *complete = ok;
  });
}

int nabo_hip_batch_plan_destroy(nabo_hip_batch_plan* plan) {
  return run_guarded([&] {
    if (!plan) return;
    DeviceScope scope(plan->ix.device);
    if (plan->launched) (void)hipStreamSynchronize(plan->last);
    delete plan;
  });
}

}  // extern "C"
