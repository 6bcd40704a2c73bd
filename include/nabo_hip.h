/* nabo_hip.h - C ABI of the MI355X (gfx950) kNN library.
 *
 * Plain C types only: pointers, sizes and status codes. Host C++
 * (include/nabo/nabo.h), Python (libnabo_amd) and any FFI bind to these.
 *
 * Layout conventions are libnabo's (nabo/nabo.h:260-272): a cloud or query
 * set is a column-major `rows x count` float matrix, one point per column,
 * consecutive points `stride` floats apart (stride >= rows); results are
 * column-major `k x m` (indices int32, dists2 float): the k results of query
 * i are contiguous at [i*k, i*k+k). Missing results are (-1, +inf)
 * (nabo/nabo.h:216-226, nabo/index_heap.h:299).
 *
 * Every function returns NABO_HIP_OK or an error code; the message of the
 * last error on the calling thread is nabo_hip_last_error().
 *
 * Device memory: besides each index's own arrays, builds and knn calls take
 * their temporaries from a per-device pool of scratch blocks that the
 * library keeps between calls (at most eight blocks of at most 1 GiB per
 * device, until the process ends). A kd-tree knn call with the IndexHeapSTL
 * heap (NABO_HIP_HEAP_STL) and k <= 96 also takes 8 * (L + k) bytes per
 * resident search lane for its admission logs, L = nabo_hip_params.stl_log
 * or by default min(256, max(64, 8k)) (about 0.2 GB for k = 16 to 96 on
 * MI355X).
 */
#ifndef NABO_HIP_H
#define NABO_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NABO_HIP_ABI_VERSION 1

/* Status codes. */
enum {
  NABO_HIP_OK = 0,
  NABO_HIP_ERR_ARGUMENT = 1,   /* the reference throws Nabo::runtime_error */
  NABO_HIP_ERR_HIP = 2,        /* HIP runtime failure */
  NABO_HIP_ERR_MEMORY = 3,     /* device allocation failed */
  NABO_HIP_ERR_UNSUPPORTED = 4 /* valid for the reference, not offered here */
};

/* Search types: the reference's SearchType values 0..5 (nabo/nabo.h:290-299)
 * keep their numbers; the HIP types are appended. */
enum {
  NABO_HIP_SEARCH_KDTREE = 6,     /* HIP kd-tree, KDTREE_LINEAR_HEAP semantics */
  NABO_HIP_SEARCH_BRUTE_FORCE = 7 /* HIP brute force, BRUTE_FORCE semantics */
};

/* Creation flags (nabo/nabo.h:302-305) and search flags (nabo/nabo.h:308-312). */
enum { NABO_HIP_TOUCH_STATISTICS = 1 };
enum { NABO_HIP_ALLOW_SELF_MATCH = 1, NABO_HIP_SORT_RESULTS = 2 };

/* Brute-force distance precisions (nabo_hip_params.bf_precision). */
enum {
  NABO_HIP_BF_FP32 = 0, /* exact FP32 sum of squared differences */
  NABO_HIP_BF_BF16 = 1, /* MFMA bf16 shortlist of k+m, exact FP32 rescoring */
  NABO_HIP_BF_FP16 = 2  /* MFMA fp16 shortlist of k+m, exact FP32 rescoring */
};

/* k-best containers (nabo_hip_params.heap): which of the reference's index
 * heaps the results follow. LINEAR is IndexHeapBruteForceVector
 * (nabo/index_heap.h:263-369; KDTREE_LINEAR_HEAP): ascending output, a new
 * entry goes after equal distances. STL is IndexHeapSTL
 * (nabo/index_heap.h:51-165; KDTREE_TREE_HEAP and BRUTE_FORCE): binary-heap
 * order unless SORT_RESULTS (then std::sort_heap order), the same entry
 * dropped among equal distances as std::pop_heap drops. Both keep the same
 * distances. */
enum { NABO_HIP_HEAP_LINEAR = 0, NABO_HIP_HEAP_STL = 1 };

typedef struct nabo_hip_index nabo_hip_index;

/* Creation parameters. The reference's Parameters map carries only
 * "bucketSize" (nabo/kdtree_cpu.cpp:230); the other fields are this
 * library's own keys. Initialise with nabo_hip_params_default(). */
typedef struct nabo_hip_params {
  uint32_t bucket_size;  /* "bucketSize", default 8, must be >= 2 */
  int32_t query_sort;    /* 1: reorder queries by home leaf before searching */
  int32_t bf_precision;  /* NABO_HIP_BF_* */
  int32_t bf_extra;      /* m: shortlist size is k + m for bf16/fp16 */
  int32_t device;        /* HIP device ordinal, -1 = current */
  int32_t heap;          /* NABO_HIP_HEAP_*; STL brute force needs bf_precision FP32 */
  int32_t wave_queries;  /* kd-tree search: queries a wavefront claims at a time, 64 x 1, 2,
                            4 or 8 (0: 64) */
  int32_t stack_place;   /* kd-tree search: NABO_HIP_STACK_* */
  int32_t stl_log;       /* kd-tree IndexHeapSTL search, k <= 96: admission-log entries per
                            lane, 16 to 4096 (0: 8k clamped to 64..256) */
  int32_t reserved[7];
} nabo_hip_params;

/* Where the kd-tree search keeps its per-lane stacks (nabo_hip_params.
 * stack_place): AUTO in LDS when the tree is at most 32 levels deep (else in
 * HBM), HBM always in device memory (more lanes resident, slower stack). */
enum { NABO_HIP_STACK_AUTO = 0, NABO_HIP_STACK_HBM = 1 };

/* Message of the last failure on this thread ("" if none). */
const char* nabo_hip_last_error(void);

/* Fills p with the defaults. */
void nabo_hip_params_default(nabo_hip_params* p);

/* Creates a search over a host cloud (rows x n, column stride `stride`).
 * Replaces NearestNeighbourSearch::create (nabo/nabo.cpp:134-155) and the
 * kd-tree / brute-force constructors (nabo/kdtree_cpu.cpp:227-272,
 * nabo/brute_force_cpu.cpp:45-60). dim is clamped to rows (nabo/nabo.cpp:71).
 * The tree is built on the GPU; the cloud is copied to HBM, so the host
 * buffer need not outlive the call. Kd-trees take fewer than 2^29 points,
 * brute force fewer than 2^30, and a knn call fewer than 2^31 queries
 * (NABO_HIP_ERR_UNSUPPORTED otherwise). */
int nabo_hip_create(const float* cloud, int64_t n, int32_t rows, int64_t stride, int32_t dim,
                    int32_t search_type, uint32_t creation_flags, const nabo_hip_params* params,
                    nabo_hip_index** out);

/* Same, for a cloud already resident in device memory (not retained). */
int nabo_hip_create_device(const float* d_cloud, int64_t n, int32_t rows, int64_t stride,
                           int32_t dim, int32_t search_type, uint32_t creation_flags,
                           const nabo_hip_params* params, void* hip_stream,
                           nabo_hip_index** out);

/* Releases an index (the reference's `delete nns`). NULL is accepted. */
int nabo_hip_destroy(nabo_hip_index* index);

/* kNN for host buffers. Replaces both virtual knn overloads
 * (nabo/nabo.h:338 and :351; nabo/kdtree_cpu.cpp:274-336,
 * nabo/brute_force_cpu.cpp:63-103): max_radii == NULL selects the scalar
 * max_radius form, otherwise one radius per query. query is qrows x m with
 * column stride qstride; indices/dists2 are k x m. *touched (nullable)
 * receives the TOUCH_STATISTICS count (the return value of the reference's
 * knn, nabo/kdtree_cpu.cpp:303), 0 when the index was created without it.
 * Argument checks follow checkSizesKnn (nabo/nabo.cpp:102-131). */
int nabo_hip_knn(const nabo_hip_index* index, const float* query, int32_t qrows, int64_t qstride,
                 int64_t m, int32_t k, float epsilon, uint32_t option_flags, float max_radius,
                 const float* max_radii, int32_t* indices, float* dists2, uint64_t* touched);

/* kNN for device buffers, enqueued on hip_stream (NULL = default stream);
 * returns once enqueued. d_touched (nullable) is a device uint64 that
 * receives the touched count. Stream-ordered scratch only: concurrent calls
 * on one index from different streams are safe (knn is const and
 * reentrant, nabo/nabo.h:179). A stream being captured into a HIP graph is
 * refused (NABO_HIP_ERR_UNSUPPORTED, nothing launched): the temporaries come
 * from a pool that reuses its blocks between calls. */
int nabo_hip_knn_device(const nabo_hip_index* index, const float* d_query, int32_t qrows,
                        int64_t qstride, int64_t m, int32_t k, float epsilon,
                        uint32_t option_flags, float max_radius, const float* d_max_radii,
                        int32_t* d_indices, float* d_dists2, uint64_t* d_touched,
                        void* hip_stream);

/* nabo_hip_knn_device that allocates nothing. No counterpart in the
 * reference: it is for callers that may not allocate device memory (tight
 * serving loops, harnesses that forbid hipMalloc). The call's temporaries
 * come from a static 64 MiB device array that each GPU the library runs on
 * holds until the process ends, instead of the scratch pool. Same arguments,
 * semantics and results as nabo_hip_knn_device, for kd-tree indices; refused
 * with NABO_HIP_ERR_UNSUPPORTED (before any search is launched) for
 * brute-force indices, IndexHeapSTL searches with k <= 96 (their admission
 * logs scale with the grid), trees deeper than 32 levels (HBM stacks) and
 * calls whose temporaries exceed the array (about 24 bytes per query with
 * query ordering: up to about 2.6 million queries). All calls on one GPU
 * share the array: calls that overlap on a GPU race and must be serialized
 * by the caller. Unlike nabo_hip_knn_device, it may be captured in a HIP
 * graph (stream capture): the captured work is kernel launches only, on the
 * index's buffers, the caller's and the static array, which outlives every
 * graph; replays share the array with every other no-allocation call on that
 * GPU and must not overlap them. */
int nabo_hip_knn_device_noalloc(const nabo_hip_index* index, const float* d_query,
                                int32_t qrows, int64_t qstride, int64_t m, int32_t k,
                                float epsilon, uint32_t option_flags, float max_radius,
                                const float* d_max_radii, int32_t* d_indices, float* d_dists2,
                                uint64_t* d_touched, void* hip_stream);

/* Batched kd-trees: `clouds` independent clouds concatenated in one device
 * buffer (cloud c = columns [offsets[c], offsets[c+1]), host offsets array
 * of clouds + 1 entries), all built in one pass on the GPU. Tree c is node
 * for node the tree nabo_hip_create builds for cloud c alone. No counterpart
 * in the reference, which builds one tree per NearestNeighbourSearch object
 * (nabo/kdtree_cpu.cpp:227-272); for callers with many small clouds. */
int nabo_hip_batch_create(const float* d_points, const int64_t* offsets, int32_t clouds,
                          int32_t rows, int64_t stride, int32_t dim, uint32_t creation_flags,
                          const nabo_hip_params* params, void* hip_stream,
                          nabo_hip_index** out);

/* kNN on batched trees, one search launch for every cloud: d_query holds m
 * queries, query set c = columns [d_query_offsets[c], d_query_offsets[c+1])
 * (device array of clouds + 1 entries, last = m) searched in tree c.
 * Outputs are k x m; indices are local to each cloud. Semantics per set are
 * those of nabo_hip_knn_device on cloud c's own index. */
int nabo_hip_batch_knn(const nabo_hip_index* index, const float* d_query, int32_t qrows,
                       int64_t qstride, int64_t m, const int64_t* d_query_offsets, int32_t k,
                       float epsilon, uint32_t option_flags, float max_radius,
                       int32_t* d_indices, float* d_dists2, uint64_t* d_touched,
                       void* hip_stream);

/* Batched build + search recorded once as a HIP graph. A plan records the
 * batched kd-tree build of `clouds` clouds (as nabo_hip_batch_create) and
 * the search of their query sets (as nabo_hip_batch_knn: query set c =
 * columns [query_offsets[c], query_offsets[c+1]) of d_query, host offsets)
 * as the kernel nodes of one graph, with every array allocated at creation.
 * Each nabo_hip_batch_plan_launch then rebuilds all trees from the current
 * contents of d_points and searches the current d_query into d_indices /
 * d_dists2 with one graph launch: no host round trip, no allocation. Change
 * the buffers' contents (same sizes) between launches to process new data;
 * they must stay valid until the plan is destroyed. No counterpart in the
 * reference, which builds one tree per NearestNeighbourSearch object
 * (nabo/kdtree_cpu.cpp:227-272); for callers that rebuild and search many
 * clouds per frame. The graph launches a fixed number of build levels:
 * `levels` (1 to 33), or 0 to build the data once at creation and take its
 * depth + 1 plus a margin of 4. A launch whose data needs more levels
 * builds and searches nothing (results stay as they were) and
 * nabo_hip_batch_plan_status reports complete = 0. Results equal
 * nabo_hip_batch_knn's; KDTREE_HIP (linear heap) semantics, 2-D and 3-D
 * clouds, trees of at most 32 split levels, no touched count. Launches of
 * one plan share its memory: do not overlap them (consecutive launches on
 * one stream are ordered). */
typedef struct nabo_hip_batch_plan nabo_hip_batch_plan;
int nabo_hip_batch_plan_create(const float* d_points, const int64_t* offsets, int32_t clouds,
                               int32_t rows, int64_t stride, int32_t dim,
                               const nabo_hip_params* params, const float* d_query,
                               int64_t qstride, const int64_t* query_offsets, int32_t k,
                               float epsilon, uint32_t option_flags, float max_radius,
                               int32_t* d_indices, float* d_dists2, int32_t levels,
                               nabo_hip_batch_plan** out);
/* Enqueues one launch of the plan's graph on hip_stream. */
int nabo_hip_batch_plan_launch(nabo_hip_batch_plan* plan, void* hip_stream);
/* Waits for the plan's last launch. *complete (nullable): 1 if its build
 * finished within the recorded levels (results valid), else 0 (also before
 * any launch); *levels (nullable): the recorded level count. */
int nabo_hip_batch_plan_status(nabo_hip_batch_plan* plan, int32_t* complete, int32_t* levels);
/* Waits for the plan's last launch and frees it. NULL is accepted. */
int nabo_hip_batch_plan_destroy(nabo_hip_batch_plan* plan);

/* Root node of each batched tree (clouds entries; {0} for a single tree). */
int nabo_hip_batch_roots(const nabo_hip_index* index, int32_t* roots);

/* Read-only properties of an index. */
typedef struct nabo_hip_info {
  int64_t n;            /* points */
  int32_t rows;         /* cloud rows */
  int32_t dim;          /* searched dimensions */
  int32_t search_type;  /* NABO_HIP_SEARCH_* */
  uint32_t bucket_size;
  uint32_t dim_bits;    /* bits of the node word holding the cut dimension */
  int32_t depth;        /* max split nodes on a root-leaf path (kd-tree) */
  int64_t node_count;   /* kd-tree nodes */
  int64_t device_bytes; /* HBM held by the index */
  uint32_t creation_flags;
  int32_t device;
  int32_t clouds;       /* trees held (1 unless batched) */
} nabo_hip_info;

int nabo_hip_get_info(const nabo_hip_index* index, nabo_hip_info* info);

/* Copies the bounding box (minBound/maxBound of nabo/nabo.h:276-279, with the
 * reference's initial values when the tree is a single bucket) to host;
 * clouds x dim values each for batched trees. */
int nabo_hip_get_bounds(const nabo_hip_index* index, float* min_bound, float* max_bound);

/* Copies the kd-tree to host: nodes as 2 x node_count uint32 (the node word
 * `dim | (child or count) << dim_bits` and the cut value bits or the bucket
 * start, nabo/nabo_private.h:138-153) and the n point indices of the buckets
 * in leaf order (BucketEntry::index, nabo/nabo_private.h:158-168). */
int nabo_hip_get_tree(const nabo_hip_index* index, uint32_t* nodes, int32_t* bucket_indices);

/* Serialized copy of a built index (kd-tree, batched kd-trees or brute
 * force) in caller-owned host memory, e.g. to keep an evicted tree in host
 * RAM and re-admit it later with one copy instead of a rebuild. No
 * counterpart in the reference, which has no tree (de)serialization
 * (nabo/nabo_private.h:158-168 keeps the tree in process memory only).
 * nabo_hip_serialize: with buf null, sets *bytes to the size needed; else
 * writes the index into buf, which must hold *bytes (capacity). Pinned host
 * memory (hipHostMalloc) makes the copies faster. */
int nabo_hip_serialize(const nabo_hip_index* index, void* buf, uint64_t capacity,
                       uint64_t* bytes);
/* Recreates an index from nabo_hip_serialize output on a device (-1: the
 * current one) without a build; searches give the same results as on the
 * serialized index. Rejects buffers that are truncated or not produced by
 * this library version. */
int nabo_hip_deserialize(const void* buf, uint64_t bytes, int32_t device, nabo_hip_index** out);

/* Kernel timing for measurement tools. While enabled, each launch of the
 * search kernels is bracketed by HIP events recorded on its own stream.
 * collect() waits for the recorded events, returns the summed kernel time in
 * milliseconds and the number of launches, and clears the record. */
int nabo_hip_kernel_timing(int32_t enable);
int nabo_hip_kernel_timing_collect(double* milliseconds, int64_t* launches);

/* ---- Served kNN -----------------------------------------------------------
 * A request server that keeps kd-trees resident in HBM. No counterpart in
 * the reference (a library without a serving layer); for callers that issue
 * many kNN requests against a pool of clouds. Clouds are registered once;
 * a tree is built on the first request that needs it and stays resident
 * until the HBM budget forces it out (least recently used first, never one
 * in use). A worker thread admits requests in arrival order as stream slots
 * and memory free up, so requests that arrive while others run overlap on
 * the device; pending requests on the same cloud with the same k, epsilon,
 * flags and radius are merged into one search launch. Each result equals
 * nabo_hip_knn on that cloud alone. */
typedef struct nabo_hip_server nabo_hip_server;

typedef struct nabo_hip_server_config {
  int64_t hbm_budget_bytes;   /* resident trees + request buffers (0: no limit) */
  int32_t slots;              /* concurrent streams (default 4) */
  int64_t max_merge_queries;  /* cap on queries merged into one launch */
  int32_t device;             /* -1: current */
  int64_t host_cache_bytes;   /* evicted trees kept serialized in pinned host
                                 memory up to this size (0: none); a cached
                                 tree is re-admitted by a copy, not a build */
  int32_t reserved[6];
} nabo_hip_server_config;

typedef struct nabo_hip_server_stats {
  int64_t submitted, completed, launches, merged_requests;
  int64_t builds, evictions, resident_trees, resident_bytes, peak_bytes;
  int64_t readmits;           /* trees re-admitted from the host cache */
  int64_t host_cached_bytes;  /* serialized trees held in host memory */
} nabo_hip_server_stats;

void nabo_hip_server_config_default(nabo_hip_server_config* c);
int nabo_hip_server_create(const nabo_hip_server_config* config, nabo_hip_server** out);
/* Stops the worker after the requests already submitted and frees all. */
int nabo_hip_server_destroy(nabo_hip_server* server);
/* Registers a host cloud (copied); the tree is built when first needed. */
int nabo_hip_server_add_cloud(nabo_hip_server* server, const float* cloud, int64_t n,
                              int32_t rows, int64_t stride, int32_t dim,
                              const nabo_hip_params* params, int32_t* cloud_id);
/* Queues a request; query/indices/dists2 are host buffers that must stay
 * valid until nabo_hip_server_wait returns for it. Arguments are checked
 * here as nabo_hip_knn checks them. */
int nabo_hip_server_submit(nabo_hip_server* server, int32_t cloud_id, const float* query,
                           int32_t qrows, int64_t qstride, int64_t m, int32_t k, float epsilon,
                           uint32_t option_flags, float max_radius, int32_t* indices,
                           float* dists2, int64_t* request_id);
/* Blocks until the request is done; returns its status and, if latency_ms
 * is not null, the time from submit to completion. */
int nabo_hip_server_wait(nabo_hip_server* server, int64_t request_id, double* latency_ms);
int nabo_hip_server_get_stats(const nabo_hip_server* server, nabo_hip_server_stats* stats);

/* Number of HIP devices visible (0 when none). */
int nabo_hip_device_count(int32_t* count);

#ifdef __cplusplus
}
#endif

#endif /* NABO_HIP_H */
