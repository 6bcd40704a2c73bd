// kd-tree kNN search on the GPU: the depth-first descent of
// KDTreeUnbalancedPtInLeavesImplicitBoundsStackOpt::recurseKnn
// (nabo/kdtree_cpu.cpp:367-447) with the sorted-array k-best of
// IndexHeapBruteForceVector (nabo/index_heap.h:263-369).
//
// One query per lane. The recursion becomes a per-lane explicit stack with
// one 8-byte entry per split ancestor of the current node:
//   pending  {far child, its rd}: pushed when the lane descends into the
//            near child; popped after the near subtree, when the reference
//            tests `rd <= maxRadius2 && rd*maxError2 < head` (:415-417 /
//            :433-435) against the k-best as it is then;
//   restore  {cut dim | flag, previous off[cd]^2}: replaces the pending entry
//            while the far subtree is visited and puts off[cd]^2 back after
//            it (:419-424 / :437-442).
// The far child's rd is formed from the same operands, in the same order,
// as the reference, so the visit order, the pruning decisions, ties and
// dists2 are bit-identical, for epsilon > 0 as well. The stack is
// slot-major ([slot][lane]): in LDS when the tree is at most 32 levels deep,
// else in HBM.
//
// The k-best array is K = next supported size >= k entries in VGPRs; the
// first K - k entries hold -inf and never move, so the head (largest kept
// distance) is always entry K - 1 and no register is indexed dynamically.
// KDTREE_TREE_HEAP instantiates the same kernel with K = 0 (IndexHeapSTL in
// LDS) or K = -1 (in the output rows, when LDS is short).
//
// Queries are first put in a spatial order (Hilbert order in the cloud's box
// for one 2-D / 3-D tree, else the order of the leaf their greedy descent
// reaches), so the lanes of a wave walk neighbouring paths; results are
// written to each query's own row.

#include <atomic>
#include <cstring>
#include <type_traits>

#include <cfloat>

#include <optional>

#include "device_common.h"
#include "internal.h"
#include "launch.h"
#include "scan.h"

namespace nabo_hip {
namespace {

constexpr int kBlock = 256;
// Search workgroups are one wave: the LDS stack of a wave is depth x 512 B,
// and one-wave workgroups let the CU fill its LDS to the last wave's worth.
constexpr int kSearchBlock = 64;
constexpr uint32_t kRestore = 0x80000000u;
constexpr int kMaxGenericDim = 64;
constexpr int kXcds = 8;       // MI355X: 8 XCDs, workgroups dispatched round-robin
constexpr int kLeafBatch = 8;  // bucket entries loaded together
static_assert(kLeafBatch <= kPtsPad, "a leaf batch may read kLeafBatch - 1 entries past a bucket");
constexpr int kLdsStackSlots = 32;  // deepest tree whose stacks fit in LDS
constexpr int kPopBatch = 4;
constexpr int kMaxCachedDevices = 64;        // stack entries read per round trip

// Coordinates / offsets of one query in VGPRs, indexed by a runtime cut
// dimension through selects (DIM > 0), a local array (DIM == 0, up to
// kMaxGenericDim) or memory (DIM == -1, any dimension).
template <int DIM>
struct Vec {
  float v[DIM];
  __device__ __forceinline__ float get(uint32_t i) const {
    if constexpr (DIM == 3) {
      // Selects of values, not of elements: a select of two elements is a
      // select of their addresses, which puts the array in scratch memory.
      const float v0 = v[0], v1 = v[1], v2 = v[2];
      const float lo = i == 1u ? v1 : v0;
      return i == 2u ? v2 : lo;
    } else if constexpr (DIM == 2) {
      const float v0 = v[0], v1 = v[1];
      return i == 1u ? v1 : v0;
    } else {
      uint32_t r = 0;
#pragma unroll
      for (int d = 0; d < DIM; ++d) r |= __float_as_uint(v[d]) & (0u - uint32_t(i == uint32_t(d)));
      return __uint_as_float(r);
    }
  }
  __device__ __forceinline__ float at(int d) const { return v[d]; }
  __device__ __forceinline__ void set(uint32_t i, float x) {
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const uint32_t m = 0u - uint32_t(i == uint32_t(d));
      v[d] = __uint_as_float((__float_as_uint(x) & m) | (__float_as_uint(v[d]) & ~m));
    }
  }
};
template <>
struct Vec<0> {
  float v[kMaxGenericDim];
  __device__ __forceinline__ float get(uint32_t i) const { return v[i]; }
  __device__ __forceinline__ float at(int d) const { return v[d]; }
  __device__ __forceinline__ void set(uint32_t i, float x) { v[i] = x; }
};
// Any dimension (above kMaxGenericDim): the values live in memory, element
// i at p[i * stride] (the query's own row, or this lane's slots of a
// slot-major array in HBM).
template <>
struct Vec<-1> {
  float* p;
  int64_t stride;
  __device__ __forceinline__ float get(uint32_t i) const { return p[int64_t(i) * stride]; }
  __device__ __forceinline__ float at(int d) const { return p[int64_t(d) * stride]; }
  __device__ __forceinline__ void set(uint32_t i, float x) { p[int64_t(i) * stride] = x; }
};

struct SearchParams {
  const uint2* nodes;
  const uint4* links;
  const float* pts;
  int pts_stride;
  uint32_t dim_bits;
  uint32_t dim_mask;
  int dim;
  const float* query;
  int64_t qstride;
  int64_t m;
  const int32_t* order;
  const int32_t* roots;   // per-cloud root node (batched trees), or null
  const int32_t* qcloud;  // per-query cloud id when roots is set
  int k;
  float max_err2;
  float max_r2;
  const float* radii;
  int32_t* out_idx;
  float* out_d2;
  unsigned long long* touched;
  uint32_t* work;  // kXcds chunk counters (zeroed before the launch)
  uint2* stack;
  int64_t stack_lanes;
  int allow_self;
  int sort_results;  // IndexHeapSTL only: sort_heap before the output
  uint2* log;          // LOG kernels: admission logs, [entry][grid lane]
  uint2* log_heap;     // LOG kernels: heaps of lanes whose log filled, [entry][grid lane]
  uint32_t log_cap;    // entries per lane
  uint32_t log_lanes;  // lanes of the grid
  float* offs;         // DIM == -1: per-lane offsets, [dimension][grid lane]
  int64_t offs_lanes;
  int rounds;          // sub-chunks of kWave queries per work claim
  int hbm_stack;       // host side: stacks in HBM even when LDS would hold them
  int log_entries;     // host side: admission-log entries per lane (0: by k)
  const int32_t* ok;   // nullable: the search returns at once while *ok == 0
                       // (a tree whose recorded build did not finish)
};

// Sorted insertion of (value, index): entries greater than value move up by
// one, value lands after the entries equal to it (index_heap.h:322-334).
// With hd ascending, the new entry j is hd[j-1] if that is greater than
// value, else min(hd[j], value): the median of (hd[j-1], hd[j], value).
// Indices follow with two selects per entry; both compares are taken on
// the old distances, so the update has no branch.
// Each entry's own select is taken inside the descending loop (entry j
// still holds its old value there), so that no K-entry temporary is live
// beside the k-best: K = 128 then fits its registers without spills, K = 64
// two waves per SIMD, K = 32 three, and smaller K need no register moves
// after the update (for K = 96 the temporary form measured faster).
// min(hd[0], value) is med3(-inf, hd[0], value): fminf would also quiet
// signalling NaNs, which neither operand can be, at two extra instructions.
template <int K>
__device__ __forceinline__ void heap_insert(float (&hd)[K], int32_t (&hi)[K], float value,
                                            int32_t index) {
  if constexpr (K != 96) {
#pragma unroll
    for (int j = K - 1; j > 0; --j) {
      const int32_t own = hd[j] > value ? index : hi[j];
      hi[j] = hd[j - 1] > value ? hi[j - 1] : own;
      hd[j] = __builtin_amdgcn_fmed3f(hd[j - 1], hd[j], value);
    }
    hi[0] = hd[0] > value ? index : hi[0];
    hd[0] = __builtin_amdgcn_fmed3f(-INFINITY, hd[0], value);
  } else {
    int32_t moved[K];
#pragma unroll
    for (int j = 0; j < K; ++j) moved[j] = hd[j] > value ? index : hi[j];
#pragma unroll
    for (int j = K - 1; j > 0; --j) {
      hi[j] = hd[j - 1] > value ? hi[j - 1] : moved[j];
      hd[j] = __builtin_amdgcn_fmed3f(hd[j - 1], hd[j], value);
    }
    hi[0] = moved[0];
    hd[0] = fminf(hd[0], value);
  }
}

// The k-best of a query: for K > 0, IndexHeapBruteForceVector as K sorted
// entries in VGPRs (hd, hi); for K = -1, IndexHeapSTL as a heap in the
// query's output row, updated at each admission (k above 96); for K = -2,
// IndexHeapBruteForceVector as a sorted array in the output row (k above
// 128). All admit the same distances (each drops one largest entry for the
// new one, so their multisets of values, and their heads, are always
// equal): the descent and its pruning are the same; only the entry kept
// among equal distances and the output order differ between the two heaps.
template <int K>
constexpr int kArr = K > 0 ? K : 1;
template <int K>
using HeapOf = typename std::conditional<K == -2, RowSortedHeap, StlHeap<RowHeapStore>>::type;

template <int K>
__device__ __forceinline__ void best_insert(float (&hd)[kArr<K>], int32_t (&hi)[kArr<K>],
                                            HeapOf<K>& heap, float value, int32_t index) {
  if constexpr (K <= 0)
    heap.replace_head(index, value);
  else
    heap_insert<K>(hd, hi, value, index);
}

// IndexHeapSTL from a log (LOG kernels, KDTREE_TREE_HEAP): the sorted k-best
// in VGPRs decides admissions and pruning as for the linear heap, and every
// admission is also appended to the lane's log in HBM. When the query ends
// its stack is empty, and the log is replayed into a binary heap held in the
// stack's LDS slots. The replay runs for all lanes of the wave together,
// where heap moves made at each admission would run for the few lanes that
// admit, and the heap takes no LDS beyond the stacks. A lane whose log fills
// before the query ends moves its heap to HBM (log_heap) and replays the log
// there, then and at the end. Both heaps take the same admissions in the
// same order, so the result is the one of IndexHeapSTL updated on the fly.
struct NoLog {
  __device__ __forceinline__ void add(float, int32_t) {}
};
struct AdmitLog {
  uint2* base;      // entry j of this lane at base[j * stride]
  uint32_t stride;  // lanes of the grid
  uint32_t cap;
  uint32_t n;
  bool spilled;                      // the log filled: the heap is in HBM
  StlHeap<StridedHeapStore> spill;   // that heap (store and k set up front)
  template <class Heap>
  __device__ void replay(Heap& heap) {
    for (uint32_t j = 0; j < n; ++j) {
      const uint2 e = base[size_t(j) * stride];
      heap.replace_head(int32_t(e.x), __uint_as_float(e.y));
    }
    n = 0;
  }
  __device__ __forceinline__ void add(float value, int32_t index) {
    if (n == cap) {
      if (!spilled) {
        spilled = true;
        spill.reset(spill.st, spill.k);
      }
      replay(spill);
    }
    base[size_t(n) * stride] = make_uint2(uint32_t(index), __float_as_uint(value));
    ++n;
  }
};

// Admission tests of a bucket entry (kdtree_cpu.cpp:392-394) and of a far
// child (:415-417 / :433-435). With epsilon = 0 (EXACT) the far-child test
// `rd <= r2 && rd * 1 < head` and the entry test `d <= r2 && d < head` are
// both `x < min(head, r2n)`, r2n being the next float above r2 (+inf for an
// infinite radius, -inf for a NaN one, which admits nothing as in the
// reference): one compare instead of two (and a multiply).
template <bool EXACT>
struct Admit {
  float r2, r2n, err2;
  __device__ __forceinline__ Admit(float r2_, float err2_) : r2(r2_), err2(err2_) {
    // r2 = r * r is +0 or more: the next float is the next bit pattern
    // (FLT_MAX goes to +inf).
    r2n = r2_ != r2_ ? -INFINITY
                     : (r2_ == INFINITY ? INFINITY
                                        : __uint_as_float(__float_as_uint(fmaxf(r2_, 0.f)) + 1u));
  }
  // min(head, r2n) as med3(-inf, head, r2n): no NaN-quieting instructions
  // (head is never NaN; a NaN r2n was mapped to -inf above).
  __device__ __forceinline__ bool node(float rd, float head) const {
    return EXACT ? rd < __builtin_amdgcn_fmed3f(-INFINITY, head, r2n)
                 : (rd <= r2 && rd * err2 < head);
  }
  __device__ __forceinline__ bool point(float d, float head) const {
    return EXACT ? d < __builtin_amdgcn_fmed3f(-INFINITY, head, r2n) : (d <= r2 && d < head);
  }
};

__device__ __forceinline__ uint2 right_word(const uint4* links, uint32_t n) {
  return reinterpret_cast<const uint2*>(links)[2 * size_t(n)];
}

template <int DIM>
__device__ __forceinline__ float point_dist(const SearchParams& a, const Vec<DIM>& q, int64_t e,
                                            int32_t* index) {
  if constexpr (DIM > 0 && DIM <= 3) {
    const float4 p = reinterpret_cast<const float4*>(a.pts)[e];
    const float c[4] = {p.x, p.y, p.z, p.w};
    float dist = 0.f;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const float diff = q.v[d] - c[d];
      dist = dist + diff * diff;
    }
    *index = __float_as_int(c[DIM]);
    return dist;
  } else {
    const float* p = a.pts + e * a.pts_stride;
    const int dim = DIM > 0 ? DIM : a.dim;
    float dist = 0.f;
    for (int d = 0; d < dim; ++d) {
      const float diff = q.at(d) - p[d];
      dist = dist + diff * diff;
    }
    *index = __float_as_int(p[dim]);
    return dist;
  }
}

// Scans one leaf bucket: loads up to kLeafBatch entries at once, then tests
// them in bucket order (kdtree_cpu.cpp:380-398).
template <int DIM, int K, bool EXACT, class Log>
__device__ __forceinline__ void scan_leaf(const SearchParams& a, const Vec<DIM>& q, uint32_t count,
                                          int64_t start, const Admit<EXACT>& adm,
                                          float (&hd)[kArr<K>], int32_t (&hi)[kArr<K>],
                                          HeapOf<K>& heap, Log& lg) {
  if constexpr (DIM > 0 && DIM <= 3) {
    const float4* pts4 = reinterpret_cast<const float4*>(a.pts) + start;
    for (uint32_t j0 = 0; j0 < count; j0 += kLeafBatch) {
      // Unconditional loads: entries past the bucket belong to the next
      // buckets or to the zeroed padding after the last one.
      float4 pv[kLeafBatch];
#pragma unroll
      for (int j = 0; j < kLeafBatch; ++j) pv[j] = pts4[j0 + j];
#pragma unroll
      for (int j = 0; j < kLeafBatch; ++j) {
        if (j0 + j < count) {
          const float c[4] = {pv[j].x, pv[j].y, pv[j].z, pv[j].w};
          float dist = 0.f;
#pragma unroll
          for (int d = 0; d < DIM; ++d) {
            const float diff = q.v[d] - c[d];
            dist = dist + diff * diff;
          }
          if (adm.point(dist, K <= 0 ? heap.head : hd[kArr<K> - 1]) &&
              (a.allow_self || dist > FLT_EPSILON)) {
            best_insert<K>(hd, hi, heap, dist, __float_as_int(c[DIM]));
            lg.add(dist, __float_as_int(c[DIM]));
          }
        }
      }
    }
  } else {
    for (uint32_t j = 0; j < count; ++j) {
      int32_t index;
      const float dist = point_dist<DIM>(a, q, start + j, &index);
      if (adm.point(dist, K <= 0 ? heap.head : hd[kArr<K> - 1]) &&
          (a.allow_self || dist > FLT_EPSILON)) {
        best_insert<K>(hd, hi, heap, dist, index);
        lg.add(dist, index);
      }
    }
  }
}

// Work mapping: the queries (in leaf order) are cut into chunks of 64, one
// per wave at a time. Workgroups are dispatched round-robin over the 8 XCDs,
// so workgroup b runs on XCD b % 8; each XCD owns one contiguous eighth of
// the chunks, which keeps the tree region it touches small, and its waves
// take chunks from it through an atomic counter; a wave whose XCD's eighth
// is exhausted takes chunks from the other eighths, so no XCD idles while
// another still has work.
// Up to K = 32 three waves per SIMD fit without spills (at most 168 VGPRs),
// K = 64 two (256 VGPRs, a few bytes of spill for DIM 3); for K = 16 the
// third wave is one more than the LDS stack of a 28-level tree lets a CU
// hold.
template <int DIM, int K, bool LDS_STACK, bool EXACT, bool LOG = false>
__global__ __launch_bounds__(kSearchBlock, (K <= 32 ? 3 : K <= 64 ? 2 : 1)) void k_kd_search(SearchParams a) {
  // Per-lane stack: in LDS when the tree is shallow enough ([slot][lane] of
  // this workgroup), else in HBM ([slot][lane] over the grid).
  if (a.ok && *a.ok == 0) return;
  extern __shared__ uint2 lds_stack[];
  char* const stack_base = reinterpret_cast<char*>(LDS_STACK ? lds_stack : a.stack);
  auto slot = [stack_base](uint32_t off) -> uint2& {
    return *reinterpret_cast<uint2*>(stack_base + off);
  };
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t lane_id = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int xcd = blockIdx.x % kXcds;
  // Work units: chunks of a.rounds sub-chunks of kWave queries (one query
  // per lane); a wave claims a chunk and searches its sub-chunks in turn.
  const int rounds = a.rounds > 0 ? a.rounds : 1;
  const int64_t subs = (a.m + kWave - 1) / kWave;
  const int64_t chunks = (subs + rounds - 1) / rounds;

  const int dim = DIM > 0 ? DIM : a.dim;
  const uint32_t leaf_tag = uint32_t(dim);
  // Slot s of this lane's stack is entry s * lanes + lane of the stack; the
  // running offset is in bytes and stays in 32 bits (8 * lanes * depth <
  // 2^32 is checked on the host).
  const uint32_t base_off =
      uint32_t(sizeof(uint2)) * (LDS_STACK ? uint32_t(threadIdx.x) : uint32_t(lane_id));
  const uint32_t sstride =
      uint32_t(sizeof(uint2)) * (LDS_STACK ? uint32_t(kSearchBlock) : uint32_t(a.stack_lanes));
  uint64_t touched = 0;

  // Output rows are written as 16-byte vectors when k = K and both output
  // arrays are 16-byte aligned (rows are then 16-byte aligned too).
  const bool vec_out =
      K >= 4 && a.k == K &&
      ((reinterpret_cast<uintptr_t>(a.out_idx) | reinterpret_cast<uintptr_t>(a.out_d2)) & 15) == 0;

  // The claim of the next chunk is issued before the current chunk's
  // results are stored, so its round trip overlaps the stores.
  int owner = xcd;   // eighth being drained
  int visited = 0;   // eighths found exhausted
  uint32_t claimed = 0;
  if (lane == 0) claimed = atomicAdd(&a.work[owner], 1u);
  int64_t c = -1;
  int r = rounds;  // next sub-chunk of chunk c (rounds: claim a chunk)
  for (;;) {
    if (r == rounds) {
      c = -1;
      uint32_t got = __shfl(claimed, 0, kWave);
      while (visited < kXcds) {
        const int64_t begin = chunks * owner / kXcds;
        const int64_t end = chunks * (owner + 1) / kXcds;
        if (begin + int64_t(got) < end) {
          c = begin + got;
          break;
        }
        owner = (owner + 1) % kXcds;
        if (++visited < kXcds) {
          got = 0;
          if (lane == 0) got = atomicAdd(&a.work[owner], 1u);
          got = __shfl(got, 0, kWave);
        }
      }
      if (c < 0) break;
      r = 0;
    }
    // (wave-uniform) the last sub-chunk of this chunk: the next chunk is
    // claimed before this one's results are stored
    const int64_t sub = c * rounds + r;
    ++r;
    const bool last = r == rounds || sub + 1 >= subs;
    if (last) r = rounds;
    const int64_t i = sub * kWave + lane;
    if (i >= a.m) {
      // Only the last sub-chunk has such lanes, and its lane 0 is never one.
      continue;
    }
    const int64_t qi = a.order ? a.order[i] : i;
    // off2[d] = off[d] * off[d] (rounded as the reference rounds it): the
    // reference only ever uses off[cd] squared, and -(a * a) is (-a) * a.
    Vec<DIM> q, off2;
    const float* qp = a.query + qi * a.qstride;
    if constexpr (DIM == -1) {
      q.p = const_cast<float*>(qp);
      q.stride = 1;
      off2.p = a.offs + lane_id;
      off2.stride = a.offs_lanes;
      for (int d = 0; d < dim; ++d) off2.set(uint32_t(d), 0.f);
    } else {
      for (int d = 0; d < dim; ++d) {
        q.v[d] = qp[d];
        off2.v[d] = 0.f;
      }
    }
    float hd[kArr<K>];
    int32_t hi[kArr<K>];
    HeapOf<K> heap;
    typename std::conditional<LOG, AdmitLog, NoLog>::type lg;
    if constexpr (LOG) {
      static_assert(K > 0, "logged IndexHeapSTL runs beside the sorted k-best");
      static_assert(kSearchBlock == kWave,
                    "the heap's LDS slots (kWave apart) are the stack's (kSearchBlock apart)");
      lg.base = a.log + lane_id;
      lg.stride = a.log_lanes;
      lg.cap = a.log_cap;
      lg.n = 0;
      lg.spilled = false;
      lg.spill.st = StridedHeapStore{a.log_heap + lane_id, a.log_lanes};
      lg.spill.k = a.k;
    }
    if constexpr (K < 0) {
      heap.reset(RowHeapStore{a.out_idx + qi * a.k, a.out_d2 + qi * a.k}, a.k);
    } else {
#pragma unroll
      for (int j = 0; j < K; ++j) {
        hd[j] = j < K - a.k ? -INFINITY : INFINITY;
        hi[j] = -1;
      }
    }
    float r2 = a.max_r2;
    if (a.radii) {
      const float r = a.radii[qi];
      r2 = r * r;
    }
    const Admit<EXACT> adm(r2, a.max_err2);

    float rd = 0.f;
    uint32_t top = base_off;  // offset of the next free slot

    // The split at internal node `node` (word w): pushes the far child when
    // it passes the test and returns the near child; go_right tells which.
    // The far child is tested again after the near subtree, against a head
    // that can only have shrunk: a test failed now fails then.
    auto split = [&](uint32_t node, uint2 w, bool& go_right) -> uint32_t {
      const uint32_t cd = w.x & a.dim_mask;
      const float new_off = q.get(cd) - __uint_as_float(w.y);
      const float rd_far = rd + (-off2.get(cd) + new_off * new_off);
      const uint32_t right = w.x >> a.dim_bits;
      go_right = new_off > 0;
      if (adm.node(rd_far, K <= 0 ? heap.head : hd[kArr<K> - 1])) {
        slot(top) = make_uint2(go_right ? node + 1 : right, __float_as_uint(rd_far));
        top += sstride;
      }
      return go_right ? right : node + 1;
    };
    // Node n's word and both its children's words (left child = n + 1).
    uint32_t n = a.roots ? uint32_t(a.roots[a.qcloud[qi]]) : 0u;
    uint2 nd = a.nodes[n], lf = (a.nodes + n)[1], rt = right_word(a.links, n);
    for (;;) {
      // Descend to a leaf, two levels per round trip to memory. The wave
      // leaves this loop only when every lane stands on a leaf, so the leaf
      // scan below runs once for all lanes instead of once per node step of
      // the slowest lane.
      while ((nd.x & a.dim_mask) != leaf_tag) {
        bool go_right;
        n = split(n, nd, go_right);
        nd = go_right ? rt : lf;
        if ((nd.x & a.dim_mask) == leaf_tag) break;
        n = split(n, nd, go_right);
        nd = a.nodes[n];
        lf = (a.nodes + n)[1];
        rt = right_word(a.links, n);
      }
      const uint32_t count = nd.x >> a.dim_bits;
      scan_leaf<DIM, K, EXACT>(a, q, count, int64_t(nd.y), adm, hd, hi, heap, lg);
      touched += count;

      // Pop: kPopBatch entries per round trip to the stack. Restore entries
      // put their offset back, pending entries that fail the test are
      // dropped, and the first one that passes is entered.
      bool descend = false;
      while (!descend && top != base_off) {
        uint2 e[kPopBatch];
#pragma unroll
        for (int j = 0; j < kPopBatch; ++j) {
          const bool has = top - base_off >= uint32_t(j + 1) * sstride;
          e[j] = slot(has ? top - uint32_t(j + 1) * sstride : top - sstride);
        }
#pragma unroll
        for (int j = 0; j < kPopBatch; ++j) {
          if (!descend && top != base_off) {
            top -= sstride;
            if (e[j].x & kRestore) {
              off2.set(e[j].x & ~kRestore, __uint_as_float(e[j].y));
            } else {
              const float rdf = __uint_as_float(e[j].y);
              if (adm.node(rdf, K <= 0 ? heap.head : hd[kArr<K> - 1])) {
                descend = true;
                n = e[j].x;
                rd = rdf;
              }
            }
          }
        }
      }
      if (!descend) break;
      {
        // Enter far child n: its split comes from the links, with its own
        // word and its children's words in the same round trip.
        const uint4 lk = a.links[n];
        nd = a.nodes[n];
        lf = (a.nodes + n)[1];
        rt = make_uint2(lk.x, lk.y);
        const uint32_t pcd = lk.z;
        const float new_off = q.get(pcd) - __uint_as_float(lk.w);
        slot(top) = make_uint2(kRestore | pcd, __float_as_uint(off2.get(pcd)));
        top += sstride;
        off2.set(pcd, new_off * new_off);
      }
    }

    if (last && lane == 0) claimed = atomicAdd(&a.work[owner], 1u);
    int32_t* oi = a.out_idx + qi * a.k;
    float* od = a.out_d2 + qi * a.k;
    if constexpr (LOG) {
      if (lg.spilled) {
        lg.replay(lg.spill);
        if (a.sort_results) lg.spill.sort();
        lg.spill.write(oi, od);
      } else {
        // the stack is empty: its first k LDS slots of this lane hold the heap
        StlHeap<LdsHeapStore> heap_lds;
        heap_lds.reset(LdsHeapStore{lds_stack + threadIdx.x}, a.k);
        lg.replay(heap_lds);
        if (a.sort_results) heap_lds.sort();
        heap_lds.write(oi, od);
      }
    } else if constexpr (K <= 0) {
      if (a.sort_results) heap.sort();
      heap.write(oi, od);
    } else if (vec_out) {
#pragma unroll
      for (int j = 0; j < K / 4; ++j) {
        reinterpret_cast<int4*>(oi)[j] =
            make_int4(hi[4 * j], hi[4 * j + 1], hi[4 * j + 2], hi[4 * j + 3]);
        reinterpret_cast<float4*>(od)[j] =
            make_float4(hd[4 * j], hd[4 * j + 1], hd[4 * j + 2], hd[4 * j + 3]);
      }
    } else {
#pragma unroll
      for (int j = 0; j < K; ++j) {
        if (j >= K - a.k) {
          oi[j - (K - a.k)] = hi[j];
          od[j - (K - a.k)] = hd[j];
        }
      }
    }
  }

  if (a.touched) {
    uint32_t lo = uint32_t(touched), hi32 = uint32_t(touched >> 32);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const uint32_t olo = (uint32_t)__shfl_xor((int)lo, o, kWave);
      const uint32_t ohi = (uint32_t)__shfl_xor((int)hi32, o, kWave);
      const uint64_t t = ((uint64_t(hi32) << 32) | lo) + ((uint64_t(ohi) << 32) | olo);
      lo = uint32_t(t);
      hi32 = uint32_t(t >> 32);
    }
    const uint64_t sum = (uint64_t(hi32) << 32) | lo;
    if (lane == 0 && sum) atomicAdd(a.touched, (unsigned long long)sum);
  }
}

// Greedy descent to the leaf a query falls into; its bucket start orders the
// queries. For 2-D and 3-D the query sits in registers and each round trip
// reads a node with both children (two levels per round trip).
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_locate(const uint2* __restrict__ nodes,
                                                   const uint4* __restrict__ links,
                                                   uint32_t dim_bits, uint32_t dim_mask, int dim_rt,
                                                   const float* __restrict__ query,
                                                   int64_t qstride, int64_t m, uint32_t shift,
                                                   const int32_t* __restrict__ roots,
                                                   const int32_t* __restrict__ qcloud,
                                                   uint32_t* __restrict__ keys) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int dim = DIM > 0 ? DIM : dim_rt;
  const uint32_t leaf_tag = uint32_t(dim);
  const float* qp = query + i * qstride;
  uint32_t n = roots ? uint32_t(roots[qcloud[i]]) : 0u;
  uint2 nd;
  if constexpr (DIM == 2 || DIM == 3) {
    Vec<DIM> q;
#pragma unroll
    for (int d = 0; d < DIM; ++d) q.v[d] = qp[d];
    nd = nodes[n];
    uint2 lf = (nodes + n)[1], rt = right_word(links, n);
    while ((nd.x & dim_mask) != leaf_tag) {
      bool go_right = q.get(nd.x & dim_mask) - __uint_as_float(nd.y) > 0;
      n = go_right ? (nd.x >> dim_bits) : n + 1;
      nd = go_right ? rt : lf;
      if ((nd.x & dim_mask) == leaf_tag) break;
      go_right = q.get(nd.x & dim_mask) - __uint_as_float(nd.y) > 0;
      n = go_right ? (nd.x >> dim_bits) : n + 1;
      nd = nodes[n];
      lf = (nodes + n)[1];
      rt = right_word(links, n);
    }
  } else {
    for (;;) {
      nd = nodes[n];
      const uint32_t cd = nd.x & dim_mask;
      if (cd == leaf_tag) break;
      const float new_off = qp[cd] - __uint_as_float(nd.y);
      n = new_off > 0 ? (nd.x >> dim_bits) : n + 1;
    }
  }
  keys[i] = nd.y >> shift;
}

// Hilbert-curve index of a 2-D / 3-D query in the cloud's bounding box with
// axis_bits bits per axis (coordinates outside the box clamp to it, NaN to
// its low corner). Consecutive indices are neighbouring cells (Morton order
// jumps), so the 64 queries of a wave cover a more compact region: 2 % less
// search time than Morton keys on the 1M LiDAR workload. Grid cell to
// index by Skilling's transpose form (J. Skilling, "Programming the Hilbert
// curve", AIP Conf. Proc. 707, 2004): undo the excess rotations from the
// top bit down, Gray-code across the axes, then interleave the bits.
struct CurveBox {
  int axis_bits;
  float lo[3], sc[3];
};

// Bit b of the low 10 (spread3) or 16 (spread2) bits of v moved to bit 3b
// (2b): the interleave of a Morton / Hilbert key with constant masks.
__device__ __forceinline__ uint32_t spread3(uint32_t v) {
  v &= 0x3ffu;
  v = (v ^ (v << 16)) & 0xff0000ffu;
  v = (v ^ (v << 8)) & 0x0300f00fu;
  v = (v ^ (v << 4)) & 0x030c30c3u;
  v = (v ^ (v << 2)) & 0x09249249u;
  return v;
}
__device__ __forceinline__ uint32_t spread2(uint32_t v) {
  v &= 0xffffu;
  v = (v ^ (v << 8)) & 0x00ff00ffu;
  v = (v ^ (v << 4)) & 0x0f0f0f0fu;
  v = (v ^ (v << 2)) & 0x33333333u;
  v = (v ^ (v << 1)) & 0x55555555u;
  return v;
}
template <int DIM>
__device__ __forceinline__ uint32_t hilbert_code(const float* qp, const CurveBox& box) {
  static_assert(DIM == 2 || DIM == 3, "keys of 2-D and 3-D queries");
  const float top = float((1u << box.axis_bits) - 1u);
  uint32_t x[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d)
    x[d] = uint32_t(fminf(fmaxf((qp[d] - box.lo[d]) * box.sc[d], 0.f), top));
  const uint32_t msb = 1u << (box.axis_bits - 1);
  for (uint32_t q = msb; q > 1; q >>= 1) {
    const uint32_t low = q - 1;
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
      if (x[i] & q) {
        x[0] ^= low;
      } else {
        const uint32_t t = (x[0] ^ x[i]) & low;
        x[0] ^= t;
        x[i] ^= t;
      }
    }
  }
#pragma unroll
  for (int i = 1; i < DIM; ++i) x[i] ^= x[i - 1];
  // flip = the xor of q - 1 over the bits q > 1 of the last coordinate:
  // bit j is the parity of its bits above j (a suffix parity, shifted)
  uint32_t y = x[DIM - 1];
  y ^= y >> 1;
  y ^= y >> 2;
  y ^= y >> 4;
  y ^= y >> 8;
  y ^= y >> 16;
  const uint32_t flip = y >> 1;
  // bit b of coordinate i goes to bit b * DIM + (DIM - 1 - i)
  uint32_t code = 0;
#pragma unroll
  for (int i = 0; i < DIM; ++i) {
    const uint32_t xi = x[i] ^ flip;
    code |= (DIM == 3 ? spread3(xi) : spread2(xi)) << (DIM - 1 - i);
  }
  return code;
}

// Zeroes the search's chunk counters (when the queries are not sorted).
__global__ void k_zero_counters(uint32_t* counters, int64_t n) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e < n) counters[e] = 0;
}

// Query ordering: a two-level bucket sort of the keys with workgroup-local
// (LDS) counters only; global atomics from every query serialize at the
// memory side. The order inside a bucket is arbitrary: results do not
// depend on the order in which queries are searched.
//   1. k_sort_hist: per tile of kSortTile queries, the count of each coarse
//      bucket (the key's top coarse_bits), written bucket-major
//      counts[bucket * tiles + tile]; for Hilbert keys it also computes and
//      stores the keys, and its first workgroup zeroes the search's chunk
//      counters (one launch instead of three);
//   2. exclusive scan of counts: where each (bucket, tile) pair starts;
//   3. k_sort_scatter: each tile places its queries in their coarse buckets;
//   4. k_sort_fine: one workgroup per coarse bucket places its queries by
//      the key's low fine_bits.
constexpr int kSortTile = 4096;
constexpr int kHistBlock = 1024;  // sort kernels' workgroups (hist, scatter: 4 queries per thread)
constexpr int kFineBlock = 512;   // fine pass: one workgroup per coarse bucket (~490 queries for 1M)
constexpr int kCoarseBits = 11;
constexpr int kFineBitsMax = 10;

// MDIM = 2 / 3: keys are the queries' Hilbert indices (batched trees: with
// the cloud id above, boxes per cloud), computed here and stored; MDIM = 0:
// keys were computed before.
template <int MDIM>
__global__ __launch_bounds__(kHistBlock) void k_sort_hist(uint32_t* __restrict__ keys,
                                                      const float* __restrict__ query,
                                                      int64_t qstride, CurveBox box,
                                                      const int32_t* __restrict__ qcloud,
                                                      const float* __restrict__ boxes,
                                                      uint32_t curve_bits, int64_t m,
                                                      int fine_bits, int coarse_bins,
                                                      int64_t tiles, uint32_t* __restrict__ counts,
                                                      uint32_t* __restrict__ counters) {
  __shared__ uint32_t h[1 << kCoarseBits];
  const int64_t tile = blockIdx.x;
  if (tile == 0 && threadIdx.x < kXcds) counters[threadIdx.x] = 0;
  for (int b = threadIdx.x; b < coarse_bins; b += kHistBlock) h[b] = 0;
  __syncthreads();
  const int64_t end = min(m, (tile + 1) * kSortTile);
  for (int64_t q = tile * kSortTile + threadIdx.x; q < end; q += kHistBlock) {
    uint32_t key;
    if constexpr (MDIM > 0) {
      if (qcloud) {
        // batched trees: the cloud id above the index in that cloud's box
        const int32_t c = qcloud[q];
        const float top = float((1u << box.axis_bits) - 1u);
        CurveBox cb{box.axis_bits, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
#pragma unroll
        for (int d = 0; d < MDIM; ++d) {
          cb.lo[d] = boxes[int64_t(c) * 6 + d];
          cb.sc[d] = boxes[int64_t(c) * 6 + 3 + d] * top;
        }
        key = (uint32_t(c) << curve_bits) | hilbert_code<MDIM>(query + q * qstride, cb);
      } else {
        key = hilbert_code<MDIM>(query + q * qstride, box);
      }
      keys[q] = key;
    } else {
      key = keys[q];
    }
    atomicAdd(&h[key >> fine_bits], 1u);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < coarse_bins; b += kHistBlock) counts[int64_t(b) * tiles + tile] = h[b];
}

__global__ __launch_bounds__(kHistBlock) void k_sort_scatter(const uint32_t* __restrict__ keys,
                                                         int64_t m, int fine_bits,
                                                         int coarse_bins, int64_t tiles,
                                                         const uint32_t* __restrict__ starts,
                                                         uint2* __restrict__ fpairs) {
  __shared__ uint32_t cur[1 << kCoarseBits];
  const int64_t tile = blockIdx.x;
  for (int b = threadIdx.x; b < coarse_bins; b += kHistBlock)
    cur[b] = starts[int64_t(b) * tiles + tile];
  __syncthreads();
  const uint32_t fine_mask = (1u << fine_bits) - 1u;
  const int64_t end = min(m, (tile + 1) * kSortTile);
  for (int64_t q = tile * kSortTile + threadIdx.x; q < end; q += kHistBlock) {
    const uint32_t key = keys[q];
    const uint32_t pos = atomicAdd(&cur[key >> fine_bits], 1u);
    fpairs[pos] = make_uint2(key & fine_mask, uint32_t(q));  // {fine key, query}
  }
}

__global__ __launch_bounds__(kFineBlock) void k_sort_fine(const uint2* __restrict__ fpairs,
                                                      const uint32_t* __restrict__ starts,
                                                      int64_t tiles, int64_t m, int fine_bits,
                                                      int coarse_bins, int32_t* __restrict__ order) {
  __shared__ uint32_t h[1 << kFineBitsMax];
  __shared__ uint32_t wave_sum[kFineBlock / kWave];
  const int bin = blockIdx.x;
  const uint32_t begin = starts[int64_t(bin) * tiles];
  const uint32_t end = bin + 1 < coarse_bins ? starts[int64_t(bin + 1) * tiles] : uint32_t(m);
  if (end - begin <= 1 || fine_bits == 0) {
    for (uint32_t e = begin + threadIdx.x; e < end; e += kFineBlock) order[e] = int32_t(fpairs[e].y);
    return;
  }
  const int bins = 1 << fine_bits;
  for (int b = threadIdx.x; b < bins; b += kFineBlock) h[b] = 0;
  __syncthreads();
  for (uint32_t e = begin + threadIdx.x; e < end; e += kFineBlock) atomicAdd(&h[fpairs[e].x], 1u);
  __syncthreads();
  // Exclusive scan of the bins: each thread owns bins [t * per, (t + 1) * per).
  const int per = (bins + kFineBlock - 1) / kFineBlock;
  const int b0 = threadIdx.x * per;
  uint32_t sum = 0;
  for (int b = b0; b < min(bins, b0 + per); ++b) sum += h[b];
  // inclusive scan of the partial sums: within each wave by shuffles, then
  // across the block's waves through LDS (one barrier)
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  uint32_t incl = sum;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const uint32_t up = __shfl_up(incl, o, kWave);
    if (lane >= o) incl += up;
  }
  if (lane == kWave - 1) wave_sum[wave] = incl;
  __syncthreads();
  for (int w = 0; w < wave; ++w) incl += wave_sum[w];
  uint32_t run = begin + incl - sum;
  for (int b = b0; b < min(bins, b0 + per); ++b) {
    const uint32_t c = h[b];
    h[b] = run;
    run += c;
  }
  __syncthreads();
  for (uint32_t e = begin + threadIdx.x; e < end; e += kFineBlock) {
    const uint2 f = fpairs[e];
    order[atomicAdd(&h[f.x], 1u)] = int32_t(f.y);
  }
}

// Persistent grid: as many workgroups as the device holds at once (from
// the occupancy of this instantiation), a multiple of the XCD count.
// The occupancy is cached per device and LDS size (tree depth, heap size).
constexpr int kLdsSlotsMax = 65536 / (kSearchBlock * sizeof(uint2));  // 64 KiB per workgroup

template <int DIM, int K, bool LDS, bool EXACT, bool LOG>
unsigned resident_grid(size_t lds, int cus, int dev) {
  static std::atomic<int> cache[kMaxCachedDevices][kLdsSlotsMax + 1];
  const size_t slot = std::min<size_t>(lds / (kSearchBlock * sizeof(uint2)), kLdsSlotsMax);
  int per_cu = dev >= 0 && dev < kMaxCachedDevices ? cache[dev][slot].load() : 0;
  if (per_cu < 1) {
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu,
                                                     k_kd_search<DIM, K, LDS, EXACT, LOG>,
                                                     kSearchBlock, lds) != hipSuccess ||
        per_cu < 1)
      per_cu = 1;
    if (dev >= 0 && dev < kMaxCachedDevices) cache[dev][slot].store(per_cu);
  }
  return unsigned(std::max(kXcds, (cus * per_cu) / kXcds * kXcds));
}

// LDS of one search workgroup: the stacks (depth slots per lane, when the
// tree is at most kLdsStackSlots deep); for LOG at least k slots per lane,
// which hold the IndexHeapSTL heaps once the stacks are empty.
template <bool LOG>
size_t search_lds(int depth, int k) {
  const size_t stack = depth <= kLdsStackSlots ? size_t(depth) : 0;
  const size_t slots = LOG ? std::max(stack, size_t(k)) : stack;
  return slots * kSearchBlock * sizeof(uint2);
}

// Admission log entries per lane (LOG kernels); a full log is replayed and
// emptied, so this bounds memory, not the number of admissions.
uint32_t search_log_cap(int k, int entries) {
  return uint32_t(entries > 0 ? entries : std::min(256, std::max(64, 8 * k)));
}

template <int DIM, int K, bool EXACT, bool LOG>
void launch_search_as(SearchParams p, int cus, int dev, int depth, DevBuf<uint2>& stack,
                      const ExternalScratch* ext, Launcher& L) {
  const size_t lds = search_lds<LOG>(depth, p.k);
  const bool lds_stack = depth <= kLdsStackSlots && !p.hbm_stack;
  // Caller-owned scratch holds the call's fixed-size temporaries only: the
  // admission logs and HBM stacks scale with the resident grid.
  if (ext && LOG)
    throw Error(NABO_HIP_ERR_UNSUPPORTED,
                "IndexHeapSTL searches with k <= 96 keep per-lane admission logs, which a "
                "caller-owned scratch does not hold");
  if (ext && !lds_stack)
    throw Error(NABO_HIP_ERR_UNSUPPORTED,
                "trees deeper than " + std::to_string(kLdsStackSlots) +
                    " levels keep their search stacks in HBM, which a caller-owned scratch "
                    "does not hold");
  const unsigned grid = lds_stack ? resident_grid<DIM, K, true, EXACT, LOG>(lds, cus, dev)
                                  : resident_grid<DIM, K, false, EXACT, LOG>(lds, cus, dev);
  const int64_t lanes = int64_t(grid) * kSearchBlock;
  if (ext && DIM == -1)
    throw Error(NABO_HIP_ERR_UNSUPPORTED,
                "searches above " + std::to_string(kMaxGenericDim) +
                    " dimensions keep per-lane offsets in HBM, which a caller-owned scratch "
                    "does not hold");
  // the LOG kernels' admission logs and spilled heaps; DIM == -1: offsets
  Scratch scratch(L.stream());
  Span<uint2> log, log_heap;
  Span<float> offs;
  if constexpr (LOG) {
    p.log_cap = search_log_cap(p.k, p.log_entries);
    p.log_lanes = uint32_t(lanes);
    scratch.want(log, size_t(p.log_cap) * size_t(lanes));
    scratch.want(log_heap, size_t(p.k) * size_t(lanes));
  }
  if constexpr (DIM == -1) scratch.want(offs, size_t(p.dim) * size_t(lanes));
  if constexpr (LOG || DIM == -1) {
    scratch.commit(dev);
    p.log = log.p;
    p.log_heap = log_heap.p;
    p.offs = offs.p;
    p.offs_lanes = lanes;
  }
  // measurement events bracket eager launches only
  std::optional<KernelTimer> timer;
  if (lds_stack) {
    if (!L.recording()) timer.emplace(L.stream());
    L(k_kd_search<DIM, K, true, EXACT, LOG>, dim3(grid), dim3(kSearchBlock), lds, p);
  } else {
    if (uint64_t(lanes) * uint64_t(depth + 1) * sizeof(uint2) >= (uint64_t(1) << 32))
      throw Error(NABO_HIP_ERR_UNSUPPORTED, "kd-tree too deep for the search stack");
    stack.alloc(size_t(lanes) * size_t(depth), L.stream());
    p.stack = stack.p;
    p.stack_lanes = lanes;
    if (!L.recording()) timer.emplace(L.stream());
    L(k_kd_search<DIM, K, false, EXACT, LOG>, dim3(grid), dim3(kSearchBlock), lds, p);
  }
}

template <int DIM, int K, bool LOG = false>
void launch_search(SearchParams p, int cus, int dev, int depth, DevBuf<uint2>& stack,
                   const ExternalScratch* ext, Launcher& L) {
  // epsilon = 0 (max_err2 == 1, so the multiply is exact) takes the
  // one-compare admission tests.
  if (p.max_err2 == 1.f)
    launch_search_as<DIM, K, true, LOG>(p, cus, dev, depth, stack, ext, L);
  else
    launch_search_as<DIM, K, false, LOG>(p, cus, dev, depth, stack, ext, L);
}

// Sorted k-best of K >= k entries (LOG: with the logged IndexHeapSTL).
template <int DIM, bool LOG>
void dispatch_sorted(const SearchParams& p, int cus, int dev, int depth, DevBuf<uint2>& stack,
                     const ExternalScratch* ext, Launcher& L) {
  const int k = p.k;
  if (k <= 1)
    launch_search<DIM, 1, LOG>(p, cus, dev, depth, stack, ext, L);
  else if (k <= 2)
    launch_search<DIM, 2, LOG>(p, cus, dev, depth, stack, ext, L);
  else if (k <= 4)
    launch_search<DIM, 4, LOG>(p, cus, dev, depth, stack, ext, L);
  else if (k <= 8)
    launch_search<DIM, 8, LOG>(p, cus, dev, depth, stack, ext, L);
  else if (k <= 16)
    launch_search<DIM, 16, LOG>(p, cus, dev, depth, stack, ext, L);
  else if (k <= 32)
    launch_search<DIM, 32, LOG>(p, cus, dev, depth, stack, ext, L);
  else if (k <= 64)
    launch_search<DIM, 64, LOG>(p, cus, dev, depth, stack, ext, L);
  else if (k <= 96)
    launch_search<DIM, 96, LOG>(p, cus, dev, depth, stack, ext, L);
  else if constexpr (LOG)
    throw Error(NABO_HIP_ERR_UNSUPPORTED, "the logged IndexHeapSTL search runs up to k = 96");
  else if (k <= kMaxK)
    launch_search<DIM, 128, LOG>(p, cus, dev, depth, stack, ext, L);
  else
    launch_search<DIM, -2, LOG>(p, cus, dev, depth, stack, ext, L);
}

template <int DIM>
void dispatch_k(const SearchParams& p, bool stl, int cus, int dev, int depth,
                DevBuf<uint2>& stack, const ExternalScratch* ext, Launcher& L) {
  // IndexHeapSTL: the logged form up to k = 96 (its heaps in the stacks'
  // LDS slots), else the heap in the output rows.
  static_assert(std::max(kLdsStackSlots, 96) * kSearchBlock * sizeof(uint2) <= 65536,
                "96-entry heaps fit in 64 KiB of LDS");
  if (stl && p.k <= 96)
    dispatch_sorted<DIM, true>(p, cus, dev, depth, stack, ext, L);
  else if (stl)
    launch_search<DIM, -1>(p, cus, dev, depth, stack, ext, L);
  else
    dispatch_sorted<DIM, false>(p, cus, dev, depth, stack, ext, L);
}

uint32_t bits_needed(uint64_t v) {
  uint32_t b = 0;
  while (v) {
    ++b;
    v >>= 1;
  }
  return b;
}

}  // namespace

// Cloud id of each query of a batched search (binary search in the offsets).
__global__ void k_query_clouds(int64_t m, int32_t clouds, const int64_t* __restrict__ offsets,
                               int32_t* __restrict__ qcloud) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
  int lo = 0, hi = clouds - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (offsets[mid] <= i)
      lo = mid;
    else
      hi = mid - 1;
  }
  qcloud[i] = lo;
}

void kd_knn(const Index& ix, const KnnArgs& a, hipStream_t s) {
  Launcher L(s);
  kd_knn_launch(ix, a, L);
}

void kd_knn_launch(const Index& ix, const KnnArgs& a, Launcher& L) {
  if (a.m == 0) return;
  if (a.scratch) {
    // refused before anything is launched (launch_search_as checks again)
    if (ix.params.heap == NABO_HIP_HEAP_STL && a.k <= 96)
      throw Error(NABO_HIP_ERR_UNSUPPORTED,
                  "IndexHeapSTL searches with k <= 96 keep per-lane admission logs, which a "
                  "caller-owned scratch does not hold");
    if (ix.depth > kLdsStackSlots || ix.params.stack_place == NABO_HIP_STACK_HBM)
      throw Error(NABO_HIP_ERR_UNSUPPORTED,
                  "trees deeper than " + std::to_string(kLdsStackSlots) +
                      " levels, and indices made with HBM stacks, keep their search stacks in "
                      "HBM, which a caller-owned scratch does not hold");
  }

  SearchParams p{};
  p.nodes = ix.d_nodes;
  p.links = ix.d_links;
  p.pts = ix.d_pts;
  p.pts_stride = ix.pts_stride;
  p.dim_bits = ix.dim_bits;
  p.dim_mask = ix.dim_mask;
  p.dim = ix.dim;
  p.query = a.query;
  p.qstride = a.qstride;
  p.m = a.m;
  p.k = a.k;
  p.max_err2 = (1 + a.epsilon) * (1 + a.epsilon);
  p.max_r2 = a.max_radius * a.max_radius;
  p.radii = a.radii;
  p.out_idx = a.out_idx;
  p.out_d2 = a.out_d2;
  p.touched = reinterpret_cast<unsigned long long*>(a.touched);
  p.allow_self = (a.flags & NABO_HIP_ALLOW_SELF_MATCH) ? 1 : 0;
  p.sort_results = (a.flags & NABO_HIP_SORT_RESULTS) ? 1 : 0;
  p.ok = a.ok;
  p.rounds = ix.params.wave_queries > 0 ? ix.params.wave_queries / kWave : 1;
  p.hbm_stack = ix.params.stack_place == NABO_HIP_STACK_HBM ? 1 : 0;
  p.log_entries = ix.params.stl_log;
  const bool stl = ix.params.heap == NABO_HIP_HEAP_STL;
  // Per-query cloud ids select the tree (batched trees, and single trees
  // searched in batched form, whose roots and boxes are in device memory).
  const bool batched = ix.clouds > 1 || (a.qcloud && ix.d_roots);
  if (batched) {
    if (!a.qcloud || !ix.d_roots)
      throw Error(NABO_HIP_ERR_ARGUMENT, "batched trees need per-query cloud ids");
    p.roots = ix.d_roots;
    p.qcloud = a.qcloud;
  }

  // Query order: keys, then the two-level bucket sort. 2-D / 3-D trees:
  // key = Hilbert index of the query in its cloud's box, with up to a few
  // cells per query (at most 2^21 cells); batched trees put the cloud id
  // above it (the index then gets the bits left of 21). Lanes of a wave then
  // walk neighbouring paths. Other dimensions: key = bucket start of the
  // leaf the greedy descent reaches, cut to 16 bits (bucket starts are
  // global, so this also groups the queries of batched trees by cloud).
  // Scratch of the call: chunk counters, then the sort's arrays, sized
  // before anything is launched.
  const bool sort = ix.params.query_sort && a.m > 1 && ix.node_count > 1;
  const int cloud_bits = ix.clouds > 1 ? int(bits_needed(uint64_t(ix.clouds - 1))) : 0;
  const int max_bits = std::min(ix.dim == 3 ? 7 : 10, (21 - cloud_bits) / std::max(ix.dim, 1));
  const bool curve = (ix.dim == 2 || ix.dim == 3) && max_bits >= 2;
  int key_bits = 0, axis_bits = 0;
  uint32_t locate_shift = 0;
  if (sort && curve) {
    axis_bits = std::max(1, std::min(max_bits, (int(bits_needed(uint64_t(a.m))) + 2) / ix.dim));
    key_bits = cloud_bits + axis_bits * ix.dim;
  } else if (sort) {
    const uint32_t full_bits = std::max(1u, bits_needed(uint64_t(ix.n)));
    locate_shift = full_bits > 16 ? full_bits - 16 : 0;
    key_bits = int(full_bits - locate_shift);
  }
  // Keys have at most kCoarseBits + kFineBitsMax bits (Hilbert: 7 x 3 or
  // 10 x 2; leaf keys: 16), so the fine pass orders the rest exactly.
  static_assert(kCoarseBits + kFineBitsMax >= 21, "sort keys fit the two passes");
  const int coarse_bits = std::min(kCoarseBits, key_bits);
  const int fine_bits = key_bits - coarse_bits;
  const int coarse_bins = 1 << coarse_bits;
  const int64_t tiles = ceil_div(a.m, kSortTile);
  const int64_t ncounts = int64_t(coarse_bins) * tiles;

  Scratch scratch(L.stream(), a.scratch ? a.scratch->base : nullptr,
                  a.scratch ? a.scratch->bytes : 0);
  Span<uint32_t> counters, keys, counts, starts, partials;
  Span<uint2> fpairs;
  Span<int32_t> order;
  scratch.want(counters, kXcds);
  if (sort) {
    scratch.want(keys, size_t(a.m));
    scratch.want(counts, size_t(ncounts));
    scratch.want(starts, size_t(ncounts));
    scratch.want(fpairs, size_t(a.m));
    scratch.want(order, size_t(a.m));
    scratch.want(partials, kScanGrid);
  }
  if (a.measure) {
    *a.measure = scratch.bytes();
    return;
  }
  int dev = 0;
  NH_CHECK(hipGetDevice(&dev));
  scratch.commit(dev);

  p.work = counters.p;
  if (!sort) L(k_zero_counters, dim3(1), dim3(kBlock), 0, counters.p, int64_t(kXcds));
  if (sort) {
    const unsigned g = unsigned(ceil_div(a.m, kBlock));
    CurveBox box{axis_bits, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
    if (curve) {
      // Cell scale of the cloud's box (one tree, host bounds); batched
      // trees read theirs from d_boxes (a cloud of at most bucketSize
      // points keeps the reference's initial bounds, max < min: scale 0).
      const float top = float((1u << axis_bits) - 1u);
      if (!batched) {
        for (int d = 0; d < ix.dim; ++d) {
          box.lo[d] = ix.min_bound[size_t(d)];
          const float ext = ix.max_bound[size_t(d)] - box.lo[d];
          box.sc[d] = ext > 0.f ? top / ext : 0.f;
        }
      }
      if (batched && !ix.d_boxes) throw Error(NABO_HIP_ERR_HIP, "batched tree without boxes");
      const float* d_boxes = batched ? ix.d_boxes : nullptr;
      const uint32_t curve_bits = uint32_t(axis_bits * ix.dim);
      L(ix.dim == 3 ? k_sort_hist<3> : k_sort_hist<2>, dim3(unsigned(tiles)), dim3(kHistBlock), 0,
        keys.p, a.query, a.qstride, box, p.qcloud, d_boxes, curve_bits, a.m, fine_bits,
        coarse_bins, tiles, counts.p, counters.p);
    } else {
      L(ix.dim == 3 ? k_locate<3> : ix.dim == 2 ? k_locate<2> : k_locate<0>, dim3(g), dim3(kBlock),
        0, static_cast<const uint2*>(ix.d_nodes), static_cast<const uint4*>(ix.d_links),
        ix.dim_bits, ix.dim_mask, ix.dim, a.query, a.qstride, a.m, locate_shift, p.roots,
        p.qcloud, keys.p);
      L(k_sort_hist<0>, dim3(unsigned(tiles)), dim3(kHistBlock), 0, keys.p, a.query, a.qstride,
        box, static_cast<const int32_t*>(nullptr), static_cast<const float*>(nullptr),
        uint32_t(0), a.m, fine_bits, coarse_bins, tiles, counts.p, counters.p);
    }
    exclusive_scan<uint32_t>(L, counts.p, starts.p, ScanCount{nullptr, ncounts, nullptr},
                             partials.p, static_cast<uint32_t*>(nullptr));
    L(k_sort_scatter, dim3(unsigned(tiles)), dim3(kHistBlock), 0,
      static_cast<const uint32_t*>(keys.p), a.m, fine_bits, coarse_bins, tiles,
      static_cast<const uint32_t*>(starts.p), fpairs.p);
    L(k_sort_fine, dim3(unsigned(coarse_bins)), dim3(kFineBlock), 0,
      static_cast<const uint2*>(fpairs.p), static_cast<const uint32_t*>(starts.p), tiles, a.m,
      fine_bits, coarse_bins, order.p);
    p.order = order.p;
  }

  int cus = 256;
  NH_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int depth = std::max(1, ix.depth);
  DevBuf<uint2> stack;
  if (ix.dim == 3)
    dispatch_k<3>(p, stl, cus, dev, depth, stack, a.scratch, L);
  else if (ix.dim == 2)
    dispatch_k<2>(p, stl, cus, dev, depth, stack, a.scratch, L);
  else if (ix.dim <= kMaxGenericDim)
    dispatch_k<0>(p, stl, cus, dev, depth, stack, a.scratch, L);
  else
    dispatch_k<-1>(p, stl, cus, dev, depth, stack, a.scratch, L);
}

}  // namespace nabo_hip
