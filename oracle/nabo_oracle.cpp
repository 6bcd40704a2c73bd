// Oracle: a CPU restatement of libnabo's kd-tree, k-best heaps and brute force.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker that the HIP library is
// compared against. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it; the product path (libnabo_amd) never does.
//
// What is restated (citations are paths inside the libnabo source tree):
//   * NearestNeighbourSearch ctor bounds init  nabo/nabo.cpp:68-80
//       (maxBound starts at FLT_MIN, the smallest *positive* float)
//   * checkSizesKnn                            nabo/nabo.cpp:102-131
//   * getStorageBitCount / argMax / getBounds  nabo/kdtree_cpu.cpp:60-107
//   * buildNodes (sliding midpoint, two-pass Hoare partition, leftCount rule)
//                                              nabo/kdtree_cpu.cpp:109-225
//   * kd-tree ctor (single bucket, node guard) nabo/kdtree_cpu.cpp:227-272
//   * knn / onePointKnn / recurseKnn           nabo/kdtree_cpu.cpp:274-447
//   * Node / BucketEntry encodings             nabo/nabo_private.h:117-177
//   * IndexHeapBruteForceVector (linear heap)  nabo/index_heap.h:263-369
//   * IndexHeapSTL (binary heap, libstdc++)    nabo/index_heap.h:51-165
//   * BruteForceSearch::knn                    nabo/brute_force_cpu.cpp:63-103
//
// Parity pinning (what checks this file, see tests/test_oracle.py):
//   - known answers from python/test.py and examples/usage.cpp,
//   - brute force in FP32 and FP64 (exact kNN sets, tie band),
//   - scipy.spatial.cKDTree fixtures under tests/golden/,
//   - tree invariants (cut planes, leaf sizes, pre-order layout),
//   - hand-worked tiny trees, and planted mutants (ORACLE_MUTANT) that the
//     checks must reject.
// Brute-force dists2 use the kd-tree's sequential sum-of-squares; the
// reference's Eigen squaredNorm summation order at high dimension is
// parity-unpinned (Eigen is not available to run).
//
// Arithmetic: built with -O3 -ffp-contract=off and no -march, as the
// reference's Release flags (no FMA, every product and sum rounded to float).

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef ORACLE_MUTANT
#define ORACLE_MUTANT 0
#endif

namespace {

thread_local std::string g_err;

int fail(const std::string& msg) {
  g_err = msg;
  return 1;
}

// Number of bits needed to hold v (nabo/kdtree_cpu.cpp:60-70).
uint32_t bits_for(uint32_t v) {
  uint32_t b = 0;
  while (v) {
    ++b;
    v >>= 1;
  }
  return b;
}

struct Tree {
  int rows = 0;        // floats per cloud column
  int dim = 0;         // dimensions searched (<= rows)
  int64_t n = 0;       // points
  unsigned bucket = 8; // bucketSize parameter
  uint32_t dim_bits = 0;
  uint32_t dim_mask = 0;
  std::vector<float> cloud;          // private copy, column-major rows x n
  std::vector<uint32_t> node_word;   // dim | (child or count) << dim_bits
  std::vector<uint32_t> node_aux;    // cut value bits or bucket start
  std::vector<int32_t> bucket_index; // point index of each bucket entry
  std::vector<float> min_bound, max_bound;
  int depth = 0;                     // max split nodes on a root-leaf path
  std::vector<int32_t> build;        // the BuildPoints permutation

  float coord(int64_t idx, int d) const { return cloud[size_t(idx) * rows + d]; }
  uint32_t encode(uint32_t d, uint32_t child) const { return d | (child << dim_bits); }
};

float as_float(uint32_t u) {
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
uint32_t as_bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}

// Index of the first strictly largest extent; 0 when all extents are <= 0
// (nabo/kdtree_cpu.cpp:76-90).
unsigned widest_axis(const std::vector<float>& lo, const std::vector<float>& hi) {
  float best = 0.f;
  unsigned axis = 0;
  for (size_t d = 0; d < lo.size(); ++d) {
    const float extent = hi[d] - lo[d];
#if ORACLE_MUTANT == 6
    if (extent >= best && extent > 0.f) {
#else
    if (extent > best) {
#endif
      best = extent;
      axis = unsigned(d);
    }
  }
  return axis;
}

// Recursive sliding-midpoint build over build[first, first+count)
// (nabo/kdtree_cpu.cpp:109-225). Returns the pre-order id of the new node.
uint32_t build_nodes(Tree& t, int64_t first, int64_t count, std::vector<float> lo,
                     std::vector<float> hi, int depth) {
  const uint32_t pos = uint32_t(t.node_word.size());
  if (count <= int64_t(t.bucket)) {
    const uint32_t start = uint32_t(t.bucket_index.size());
    for (int64_t i = 0; i < count; ++i) t.bucket_index.push_back(t.build[first + i]);
    t.node_word.push_back(t.encode(uint32_t(t.dim), uint32_t(count)));
    t.node_aux.push_back(start);
    t.depth = std::max(t.depth, depth);
    return pos;
  }
  int32_t* pts = t.build.data() + first;
  const unsigned cd = widest_axis(lo, hi);
  const float ideal = (hi[cd] + lo[cd]) / 2;

  float pmin = std::numeric_limits<float>::max();
  float pmax = std::numeric_limits<float>::lowest();
  for (int64_t i = 0; i < count; ++i) {
    const float v = t.coord(pts[i], cd);
    pmin = std::min(v, pmin);
    pmax = std::max(v, pmax);
  }
  float cut = ideal;
  if (ideal < pmin)
    cut = pmin;
  else if (ideal > pmax)
    cut = pmax;

  // First pass: [0, br1) < cut <= [br1, count).
  int64_t l = 0, r = count - 1;
  for (;;) {
    while (l < count && t.coord(pts[l], cd) < cut) ++l;
    while (r >= 0 && t.coord(pts[r], cd) >= cut) --r;
    if (l > r) break;
    std::swap(pts[l], pts[r]);
    ++l;
    --r;
  }
  const int64_t br1 = l;
  // Second pass on [br1, count): [br1, br2) == cut < [br2, count).
  r = count - 1;
  for (;;) {
    while (l < count && t.coord(pts[l], cd) <= cut) ++l;
    while (r >= br1 && t.coord(pts[r], cd) > cut) --r;
    if (l > r) break;
    std::swap(pts[l], pts[r]);
    ++l;
    --r;
  }
  const int64_t br2 = l;

  int64_t left;
#if ORACLE_MUTANT == 5
  left = count / 2;
#else
  if (ideal < pmin)
    left = 1;
  else if (ideal > pmax)
    left = count - 1;
  else if (br1 > count / 2)
    left = br1;
  else if (br2 < count / 2)
    left = br2;
  else
    left = count / 2;
#endif

  std::vector<float> left_hi(hi);
  left_hi[cd] = cut;
  std::vector<float> right_lo(lo);
  right_lo[cd] = cut;

  t.node_word.push_back(0);
  t.node_aux.push_back(as_bits(cut));
  build_nodes(t, first, left, lo, left_hi, depth + 1);
  const uint32_t right = build_nodes(t, first + left, count - left, right_lo, hi, depth + 1);
  t.node_word[pos] = t.encode(cd, right);
  return pos;
}

// Sorted-array k-best (IndexHeapBruteForceVector, nabo/index_heap.h:263-369):
// ascending values, a new entry lands after entries of equal value.
struct LinearHeap {
  std::vector<int32_t> idx;
  std::vector<float> val;
  explicit LinearHeap(int k) : idx(size_t(k), -1), val(size_t(k), INFINITY) {}
  void reset() {
    std::fill(idx.begin(), idx.end(), -1);
    std::fill(val.begin(), val.end(), INFINITY);
  }
  float head() const { return val.back(); }
  void replace_head(int32_t index, float value) {
    size_t i = val.size() - 1;
    for (; i > 0; --i) {
      if (val[i - 1] > value) {
        val[i] = val[i - 1];
        idx[i] = idx[i - 1];
      } else {
        break;
      }
    }
    val[i] = value;
    idx[i] = index;
  }
  void sort() {}
  void get(int32_t* oi, float* ov) const {
    for (size_t i = 0; i < val.size(); ++i) {
      oi[i] = idx[i];
      ov[i] = val[i];
    }
  }
};

// Binary max-heap on std::push_heap / std::pop_heap (IndexHeapSTL,
// nabo/index_heap.h:51-165). It starts with one sentinel (-1, +inf) entry.
struct StlHeap {
  struct Entry {
    int32_t index;
    float value;
    bool operator<(const Entry& o) const { return value < o.value; }
  };
  std::vector<Entry> data;
  size_t k;
  explicit StlHeap(int kk) : k(size_t(kk)) {
    data.reserve(k);
    reset();
  }
  void reset() {
    data.clear();
    data.push_back(Entry{-1, INFINITY});
  }
  float head() const { return data.front().value; }
  void replace_head(int32_t index, float value) {
    if (data.size() == k) {
      std::pop_heap(data.begin(), data.end());
      data.back() = Entry{index, value};
    } else {
      data.push_back(Entry{index, value});
    }
    std::push_heap(data.begin(), data.end());
  }
  void sort() { std::sort_heap(data.begin(), data.end()); }
  void get(int32_t* oi, float* ov) const {
    size_t i = 0;
    for (; i < data.size(); ++i) {
      oi[i] = data[i].index;
      ov[i] = data[i].value;
    }
    for (; i < k; ++i) {
      oi[i] = -1;
      ov[i] = INFINITY;
    }
  }
};

// Per-query state for the recursive descent (nabo/kdtree_cpu.cpp:367-447).
template <typename Heap>
struct Searcher {
  const Tree& t;
  Heap heap;
  std::vector<float> off;
  float max_err2 = 1.f, max_r2 = INFINITY;
  bool allow_self = false;
  uint64_t touched = 0;
  uint32_t visits = 0;

  Searcher(const Tree& tree, int k) : t(tree), heap(k), off(size_t(tree.dim), 0.f) {}

  void recurse(const float* q, uint32_t n, float rd) {
    ++visits;
    const uint32_t word = t.node_word[n];
    const uint32_t cd = word & t.dim_mask;
    if (cd == uint32_t(t.dim)) {
      const uint32_t count = word >> t.dim_bits;
      const uint32_t start = t.node_aux[n];
      for (uint32_t i = 0; i < count; ++i) {
        const int32_t index = t.bucket_index[start + i];
        const float* p = &t.cloud[size_t(index) * t.rows];
        float dist = 0.f;
        for (int d = 0; d < t.dim; ++d) {
          const float diff = q[d] - p[d];
          dist += diff * diff;
        }
#if ORACLE_MUTANT == 3
        const bool self_ok = allow_self || dist >= 0.f;
#else
        const bool self_ok = allow_self || dist > FLT_EPSILON;
#endif
#if ORACLE_MUTANT == 4
        const bool in_radius = dist < max_r2;
#else
        const bool in_radius = dist <= max_r2;
#endif
        if (in_radius && dist < heap.head() && self_ok) heap.replace_head(index, dist);
      }
      touched += count;
      return;
    }
    const uint32_t right = word >> t.dim_bits;
    const float cut = as_float(t.node_aux[n]);
    const float old_off = off[cd];
    const float new_off = q[cd] - cut;
    const uint32_t near_child = new_off > 0 ? right : n + 1;
    const uint32_t far_child = new_off > 0 ? n + 1 : right;
    recurse(q, near_child, rd);
#if ORACLE_MUTANT == 2
    rd += old_off * old_off + new_off * new_off;
#else
    rd += -old_off * old_off + new_off * new_off;
#endif
    if (rd <= max_r2 && rd * max_err2 < heap.head()) {
      off[cd] = new_off;
      recurse(q, far_child, rd);
#if ORACLE_MUTANT != 1
      off[cd] = old_off;
#endif
    }
  }
};

int check_knn_args(int64_t n, int dim, int qrows, int k, unsigned flags) {
  const bool allow_self = flags & 1u;
  if (allow_self) {
    if (k > n)
      return fail("Requesting more points (" + std::to_string(k) + ") than available in cloud (" +
                  std::to_string(n) + ")");
  } else if (k > n - 1) {
    return fail("Requesting more points (" + std::to_string(k) +
                ") than available in cloud minus 1 (" + std::to_string(n - 1) +
                ") (as self match is forbidden)");
  }
  if (qrows < dim)
    return fail("Query has less dimensions (" + std::to_string(qrows) +
                ") than requested for cloud (" + std::to_string(dim) + ")");
  if (flags > 3u)
    return fail("OR-ed value of option flags (" + std::to_string(flags) +
                ") is larger than maximal valid value (3)");
  if (k < 1) return fail("k must be at least 1");
  return 0;
}

template <typename Heap>
void kd_knn_all(const Tree& t, const float* q, int qrows, int64_t m, int k, float eps,
                unsigned flags, float max_radius, const float* radii, int32_t* oi, float* od,
                uint64_t* touched_pq, uint32_t* visits_pq, int threads) {
  const bool allow_self = flags & 1u;
  const bool sort_results = flags & 2u;
  const float max_err2 = (1 + eps) * (1 + eps);
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
#pragma omp parallel
  {
    Searcher<Heap> s(t, k);
    s.allow_self = allow_self;
    s.max_err2 = max_err2;
#pragma omp for schedule(guided, 32)
    for (int64_t i = 0; i < m; ++i) {
      const float r = radii ? radii[i] : max_radius;
#if ORACLE_MUTANT == 7
      s.max_r2 = radii ? r : r * r;  // planted: a per-query radius left unsquared
#else
      s.max_r2 = r * r;  // kdtree_cpu.cpp:331-332 (per-query), :280 (scalar)
#endif
      std::fill(s.off.begin(), s.off.end(), 0.f);
      s.heap.reset();
      s.touched = 0;
      s.visits = 0;
      s.recurse(q + size_t(i) * qrows, 0, 0.f);
      if (sort_results) s.heap.sort();
      s.heap.get(oi + size_t(i) * k, od + size_t(i) * k);
      if (touched_pq) touched_pq[i] = s.touched;
      if (visits_pq) visits_pq[i] = s.visits;
    }
  }
}

}  // namespace

extern "C" {

const char* oracle_last_error() { return g_err.c_str(); }

int oracle_mutant() { return ORACLE_MUTANT; }

// Builds the kd-tree of a column-major rows x n cloud (kdtree_cpu.cpp:227-272).
void* oracle_kd_build(const float* cloud, int rows, int64_t n, int dim, unsigned bucket) {
  if (dim <= 0) {
    fail("Your space must have at least one dimension");
    return nullptr;
  }
  if (n == 0) {
    fail("Cloud has no points");
    return nullptr;
  }
  if (rows == 0) {
    fail("Cloud has 0 dimensions");
    return nullptr;
  }
  Tree* t = new Tree();
  t->rows = rows;
  t->dim = std::min(dim, rows);
  t->n = n;
  t->bucket = bucket;
  t->dim_bits = bits_for(uint32_t(t->dim));
  t->dim_mask = (1u << t->dim_bits) - 1u;
  t->cloud.assign(cloud, cloud + size_t(rows) * size_t(n));
  t->min_bound.assign(size_t(t->dim), std::numeric_limits<float>::max());
  t->max_bound.assign(size_t(t->dim), std::numeric_limits<float>::min());
  if (bucket < 2) {
    fail("Requested bucket size " + std::to_string(bucket) + ", but must be larger than 2");
    delete t;
    return nullptr;
  }
  if (n <= int64_t(bucket)) {
    for (int64_t i = 0; i < n; ++i) t->bucket_index.push_back(int32_t(i));
    t->node_word.push_back(t->encode(uint32_t(t->dim), uint32_t(n)));
    t->node_aux.push_back(0);
    return t;
  }
  const uint64_t max_nodes = (1ull << (32 - t->dim_bits)) - 1;
  const uint64_t est_nodes = uint64_t(n) / (bucket / 2);
  if (est_nodes > max_nodes) {
    fail("Cloud has a risk to have more nodes (" + std::to_string(est_nodes) +
         ") than the kd-tree allows (" + std::to_string(max_nodes) + ")");
    delete t;
    return nullptr;
  }
  t->build.resize(size_t(n));
  for (int64_t i = 0; i < n; ++i) {
    t->build[size_t(i)] = int32_t(i);
    for (int d = 0; d < t->dim; ++d) {
      const float v = t->coord(i, d);
      t->min_bound[d] = std::min(t->min_bound[d], v);
      t->max_bound[d] = std::max(t->max_bound[d], v);
    }
  }
  build_nodes(*t, 0, n, t->min_bound, t->max_bound, 0);
  t->build.clear();
  t->build.shrink_to_fit();
  return t;
}

void oracle_kd_free(void* h) { delete static_cast<Tree*>(h); }
int64_t oracle_kd_node_count(void* h) { return int64_t(static_cast<Tree*>(h)->node_word.size()); }
int oracle_kd_depth(void* h) { return static_cast<Tree*>(h)->depth; }
int oracle_kd_dim_bits(void* h) { return int(static_cast<Tree*>(h)->dim_bits); }

// nodes: 2 words per node (word, aux); bucket: n point indices in leaf order.
void oracle_kd_export(void* h, uint32_t* nodes, int32_t* bucket, float* min_b, float* max_b) {
  const Tree* t = static_cast<Tree*>(h);
  for (size_t i = 0; i < t->node_word.size(); ++i) {
    nodes[2 * i] = t->node_word[i];
    nodes[2 * i + 1] = t->node_aux[i];
  }
  std::copy(t->bucket_index.begin(), t->bucket_index.end(), bucket);
  std::copy(t->min_bound.begin(), t->min_bound.end(), min_b);
  std::copy(t->max_bound.begin(), t->max_bound.end(), max_b);
}

// heap_kind: 0 = linear heap (KDTREE_LINEAR_HEAP), 1 = STL heap (KDTREE_TREE_HEAP).
// radii (nullable) gives one radius per query (kdtree_cpu.cpp:306-336).
int oracle_kd_knn(void* h, const float* q, int qrows, int64_t m, int k, float eps, unsigned flags,
                  float max_radius, const float* radii, int heap_kind, int32_t* oi, float* od,
                  uint64_t* touched_pq, uint32_t* visits_pq, int threads) {
  const Tree* t = static_cast<Tree*>(h);
  if (check_knn_args(t->n, t->dim, qrows, k, flags)) return 1;
  if (heap_kind == 0)
    kd_knn_all<LinearHeap>(*t, q, qrows, m, k, eps, flags, max_radius, radii, oi, od,
                           touched_pq, visits_pq, threads);
  else
    kd_knn_all<StlHeap>(*t, q, qrows, m, k, eps, flags, max_radius, radii, oi, od, touched_pq,
                        visits_pq, threads);
  return 0;
}

// Brute force (brute_force_cpu.cpp:63-103): points in index order, STL heap,
// epsilon ignored. threads > 1 splits queries (the reference is serial;
// results do not depend on it).
int oracle_bf_knn(const float* cloud, int rows, int64_t n, int dim_req, const float* q, int qrows,
                  int64_t m, int k, unsigned flags, float max_radius, const float* radii,
                  int32_t* oi, float* od, int threads) {
  if (dim_req <= 0) return fail("Your space must have at least one dimension");
  const int dim = std::min(dim_req, rows);
  if (check_knn_args(n, dim, qrows, k, flags)) return 1;
  const bool allow_self = flags & 1u;
  const bool sort_results = flags & 2u;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
#pragma omp parallel
  {
    StlHeap heap(k);
#pragma omp for schedule(guided, 32)
    for (int64_t c = 0; c < m; ++c) {
      const float r = radii ? radii[c] : max_radius;
      const float r2 = r * r;
      const float* qc = q + size_t(c) * qrows;
      heap.reset();
      for (int64_t i = 0; i < n; ++i) {
        const float* p = cloud + size_t(i) * rows;
        float dist = 0.f;
        for (int d = 0; d < dim; ++d) {
          const float diff = p[d] - qc[d];
          dist += diff * diff;
        }
        if (dist <= r2 && dist < heap.head() && (allow_self || dist > FLT_EPSILON))
          heap.replace_head(int32_t(i), dist);
      }
      if (sort_results) heap.sort();
      heap.get(oi + size_t(c) * k, od + size_t(c) * k);
    }
  }
  return 0;
}

// Exact squared distances in double for given (query, point) pairs: used by
// the tie-band recheck. pairs: m x k point indices (-1 skipped -> +inf).
void oracle_dist2_f64(const float* cloud, int rows, int dim, const float* q, int qrows, int64_t m,
                      int k, const int32_t* idx, double* out) {
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < m; ++c) {
    for (int j = 0; j < k; ++j) {
      const int32_t i = idx[size_t(c) * k + j];
      if (i < 0) {
        out[size_t(c) * k + j] = INFINITY;
        continue;
      }
      double s = 0.0;
      for (int d = 0; d < dim; ++d) {
        const double diff = double(cloud[size_t(i) * rows + d]) - double(q[size_t(c) * qrows + d]);
        s += diff * diff;
      }
      out[size_t(c) * k + j] = s;
    }
  }
}

// FP64 brute-force exact k-th distance per query (ascending), for the tie band.
void oracle_bf_kth_f64(const float* cloud, int rows, int64_t n, int dim, const float* q, int qrows,
                       int64_t m, int k, unsigned flags, double* out_sorted) {
  const bool allow_self = flags & 1u;
#pragma omp parallel
  {
    std::vector<double> best;
#pragma omp for schedule(guided, 8)
    for (int64_t c = 0; c < m; ++c) {
      best.assign(size_t(k), INFINITY);
      for (int64_t i = 0; i < n; ++i) {
        double s = 0.0;
        for (int d = 0; d < dim; ++d) {
          const double diff = double(cloud[size_t(i) * rows + d]) - double(q[size_t(c) * qrows + d]);
          s += diff * diff;
        }
        if (!allow_self && s <= double(FLT_EPSILON)) continue;
        if (s < best.back()) {
          size_t j = best.size() - 1;
          while (j > 0 && best[j - 1] > s) {
            best[j] = best[j - 1];
            --j;
          }
          best[j] = s;
        }
      }
      std::copy(best.begin(), best.end(), out_sorted + size_t(c) * k);
    }
  }
}

// Wall-clock timer for the CPU baseline (tests/helpers.h:128-146 style).
double oracle_now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int oracle_max_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
