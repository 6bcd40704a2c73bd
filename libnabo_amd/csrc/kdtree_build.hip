// kd-tree build on the GPU: the unbalanced points-in-leaves tree of
// KDTreeUnbalancedPtInLeavesImplicitBoundsStackOpt (nabo/kdtree_cpu.cpp:109-272),
// node for node and bucket entry for bucket entry.
//
// The recursive build is run breadth-first, one level of segments at a time:
//   1. classify each segment: leaf if count <= bucketSize, else cut dimension
//      = first widest cell extent, ideal cut = cell midpoint
//      (nabo/kdtree_cpu.cpp:117-135, argMax :76-90);
//   2. min/max of the points on the cut dimension (getBounds :93-107), cut
//      value clamped to them (:141-147);
//   3. the two Hoare passes (:149-177). A pass swaps the i-th out-of-place
//      element counted from the left with the i-th out-of-place element
//      counted from the right, so its result is a pure function of ranks:
//      a global exclusive scan of the out-of-place flags gives every rank and
//      the swap pairs follow without any sequential scan;
//   4. leftCount rule (:180-190) and the children's clipped cells (:207-212).
// The points travel as records (coordinates, then the point's index within
// its cloud), swapped in place by the passes level after level: each level
// reads the cut coordinate contiguously, a pass touches only the points it
// moves, and the final order is the leaf-ordered point array the search
// reads (the reference's bucket vector order).
//
// A level's segment count and first BFS id live in device memory
// (BuildLevel, two alternating slots), and every kernel reads them there:
// the launches of a level are the same whatever its size. The eager build
// reads the number of splits of each level from a polled pinned mailbox and
// stops at the first level without one; a build recorded in a HIP graph
// launches a fixed number of levels, and kernels of levels past the last
// return at once; a word in device memory tells whether the levels
// sufficed, and when they did not, nothing is built from the partial
// records (the finalize kernels, and the search that follows in a graph,
// return at once).
//
// Segments of at most a few thousand points leave the level loop: one
// workgroup finishes such a segment's whole subtree in LDS (k_local_finish),
// with the same steps in the same arithmetic (min/max, cut clamp, counts,
// both Hoare passes by ranks from a workgroup scan, leftCount rule, clipped
// cells), so the tree is the same; its nodes take ids from a range of their
// own (2n + 2 * the segment's first point onwards), after the level loop's.
//
// Node numbering: the reference emits nodes in pre-order with the left
// child right after its parent (:215-223). With a full binary tree, the
// nodes before node v in pre-order are its level(v) ancestors and the
// complete subtrees left of it, 2 L - R nodes, where L counts the leaves
// whose bucket starts before v's first point and R the right-child edges on
// v's path; trees of batched clouds are stored one after the other, so
//   pre(v) = 2 * leaves_before(start(v)) + level(v) - R(v) - cloud(v),
// with leaves_before an exclusive scan of leaf-start flags over the points.

#include <cstring>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <mutex>
#include <vector>

#include "device_common.h"
#include "internal.h"
#include "launch.h"
#include "scan.h"

namespace nabo_hip {
namespace {

constexpr int kBlock = 256;
constexpr int kGridStride = 2048;  // workgroups of launches over segments / nodes

struct Seg {  // one level of segments
  int32_t* start;
  int32_t* count;
  int32_t* rc;  // right-child edges on the segment's path
  float* lo;    // S x dim
  float* hi;    // S x dim
};

struct SegStat {
  uint32_t* cd;
  float* ideal;
  float* cut;
  uint32_t* minkey;
  uint32_t* maxkey;
  uint32_t* lt;
  uint32_t* eq;
  uint32_t* split;
  uint32_t* childbase;
  int32_t* left;
};

struct Bfs {  // all nodes in BFS order
  uint32_t* cd;     // cut dimension, or dim for a leaf
  uint32_t* aux;    // cut value bits, or bucket start
  uint32_t* count;  // leaf point count
  int32_t* left;
  int32_t* right;
  int32_t* start;  // first point of the node's range
  int32_t* level;
  int32_t* rc;      // right-child edges on the path
  uint32_t* pre;    // pre-order id
};

// A segment finished in LDS (k_local_finish): its points, level, right-child
// edges, node id and cell.
struct LocalSeg {
  int32_t start, count, level, rc, node;
  float lo[3], hi[3];
};
// Build-wide counters in device memory, zeroed at the build's start.
struct BuildCounters {
  uint32_t local_count;  // LocalSeg entries
  uint32_t claim;        // entries taken by k_local_finish workgroups
  uint32_t local_nodes;  // nodes made by k_local_finish (the segments' own nodes excluded)
  int32_t max_level;     // deepest level holding a node
};
constexpr int kLocalT = 2048;     // points of a segment finished in LDS, at most
constexpr int kLocalSegs = 228;   // splitting segments of one LDS level, at most
constexpr int kLocalBlock = 512;
constexpr int kLocalPerCu = 2;    // k_local_finish workgroups resident per CU (LDS)
constexpr int kLocalItems = kLocalT / kLocalBlock;  // consecutive points per thread

// Largest segment finished in LDS: 2-D and 3-D clouds (16-byte records);
// a level of a segment of T points has at most T / (bucket + 1) splitting
// segments, which must fit the LDS tables.
inline int local_points(int dim, uint32_t bucket) {
  if (dim < 1 || dim > 3) return 0;
  return int(std::min<int64_t>(kLocalT, int64_t(kLocalSegs) * (int64_t(bucket) + 1)));
}

constexpr int kTile = 2048;              // points per workgroup in the tiled passes
constexpr int kItems = kTile / kBlock;   // consecutive points per thread

template <int OP>
__device__ __forceinline__ uint32_t red(uint32_t a, uint32_t b) {
  return OP == 0 ? min(a, b) : (OP == 1 ? max(a, b) : a + b);
}
template <int OP>
__device__ __forceinline__ uint32_t red_identity() {
  return OP == 0 ? 0xffffffffu : 0u;
}
template <int OP>
__device__ __forceinline__ void red_atomic(uint32_t* p, uint32_t v) {
  if (OP == 0)
    atomicMin(p, v);
  else if (OP == 1)
    atomicMax(p, v);
  else
    atomicAdd(p, v);
}

// Per-tile partial results of the two segments a tile may share with its
// neighbours (those of its first and last items): slot 2t and 2t + 1 of
// tile t. A tile made of one segment writes it to both slots (the second
// with identity values), so the partials of a segment stay contiguous.
struct TileParts {
  int32_t* seg;
  uint32_t* a;
  uint32_t* b;
};

// Reduction by segment of two quantities over one tile, without contention:
// thread t holds items [kItems t, kItems (t + 1)) of the tile in position
// order, seg < 0 meaning no item; equal segments are contiguous. A segment
// that lies inside the tile gets one atomic per quantity (no other tile
// touches it); the tile's first and last segments go to its partials, or to
// atomics as well when parts is null. Targets: ta[seg * stride],
// tb[seg * stride]. Same-address atomics from many workgroups serialize at
// the memory side, which made the top levels of the build slow.
template <int OPA, int OPB>
__device__ void tile_reduce(const int32_t (&seg)[kItems], const uint32_t (&va)[kItems],
                            const uint32_t (&vb)[kItems], uint32_t* ta, uint32_t* tb,
                            int64_t stride, const TileParts* parts, int64_t tile) {
  __shared__ int32_t s_first[kBlock], s_last[kBlock], s_flag[kBlock];
  __shared__ uint32_t s_a[kBlock], s_b[kBlock];
  const int t = threadIdx.x;
  // Runs inside the thread: the first one (head), interior ones (complete,
  // emitted here) and the last one (tail).
  int32_t cs = seg[0];
  uint32_t ca = va[0], cb = vb[0];
  int32_t hs = seg[0];
  uint32_t ha = 0, hb = 0;
  bool single = true;
  for (int j = 1; j < kItems; ++j) {
    if (seg[j] == cs) {
      ca = red<OPA>(ca, va[j]);
      cb = red<OPB>(cb, vb[j]);
      continue;
    }
    if (single) {
      ha = ca;
      hb = cb;
      single = false;
    } else if (cs >= 0) {
      red_atomic<OPA>(ta + cs * stride, ca);
      red_atomic<OPB>(tb + cs * stride, cb);
    }
    cs = seg[j];
    ca = va[j];
    cb = vb[j];
  }
  s_first[t] = seg[0];
  s_last[t] = cs;
  __syncthreads();
  const int32_t prev_last = t > 0 ? s_last[t - 1] : -1;
  const int32_t next_first = t + 1 < kBlock ? s_first[t + 1] : -1;
  const int32_t tile_first = s_first[0], tile_last = s_last[kBlock - 1];
  // Inclusive segmented scan of the tails over the threads: a tail
  // continues the previous thread's tail when the thread is one run of
  // that segment.
  int32_t flag = !(single && cs >= 0 && cs == prev_last);
  uint32_t sa = ca, sb = cb;
  for (int d = 1; d < kBlock; d <<= 1) {
    s_flag[t] = flag;
    s_a[t] = sa;
    s_b[t] = sb;
    __syncthreads();
    if (t >= d && !flag) {
      const int32_t of = s_flag[t - d];
      sa = red<OPA>(s_a[t - d], sa);
      sb = red<OPB>(s_b[t - d], sb);
      flag = of;
    }
    __syncthreads();
  }
  // Scanned value of the previous thread's tail, for the head run.
  s_a[t] = sa;
  s_b[t] = sb;
  __syncthreads();
  auto emit = [&](int32_t sgm, uint32_t a, uint32_t b) {
    if (sgm < 0) return;
    if (parts && (sgm == tile_first || sgm == tile_last)) {
      const int slot = sgm == tile_first ? 0 : 1;
      parts->seg[2 * tile + slot] = sgm;
      parts->a[2 * tile + slot] = a;
      parts->b[2 * tile + slot] = b;
    } else {
      red_atomic<OPA>(ta + sgm * stride, a);
      red_atomic<OPB>(tb + sgm * stride, b);
    }
  };
  if (!single) {
    uint32_t a = ha, b = hb;
    if (hs >= 0 && hs == prev_last) {
      a = red<OPA>(s_a[t - 1], a);
      b = red<OPB>(s_b[t - 1], b);
    }
    emit(hs, a, b);
  }
  if (cs != next_first) emit(cs, sa, sb);
  if (parts) {
    if (t == 0 && tile_first < 0) parts->seg[2 * tile] = -1;
    if (t == kBlock - 1 && (tile_last < 0 || tile_last == tile_first)) {
      parts->seg[2 * tile + 1] = tile_last < 0 ? -1 : tile_last;
      parts->a[2 * tile + 1] = red_identity<OPA>();
      parts->b[2 * tile + 1] = red_identity<OPB>();
    }
  }
}

// Second stage: the partials of `groups` reductions ([group][2 * tiles]),
// reduced the same way with atomics only (targets ta / tb + group * gstride).
template <int OPA, int OPB>
__global__ __launch_bounds__(kBlock) void k_tile_partials(int64_t nparts, TileParts parts,
                                                          uint32_t* ta, uint32_t* tb,
                                                          int64_t stride, int64_t gstride,
                                                          const int32_t* gate) {
  if (gate && *gate == 0) return;
  const int64_t g = blockIdx.y;
  const int64_t base = int64_t(blockIdx.x) * kTile + int64_t(threadIdx.x) * kItems;
  int32_t sg[kItems];
  uint32_t a[kItems], b[kItems];
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const int64_t e = base + j;
    const bool ok = e < nparts;
    sg[j] = ok ? parts.seg[g * nparts + e] : -1;
    a[j] = ok ? parts.a[g * nparts + e] : red_identity<OPA>();
    b[j] = ok ? parts.b[g * nparts + e] : red_identity<OPB>();
  }
  tile_reduce<OPA, OPB>(sg, a, b, ta + g * gstride, tb + g * gstride, stride, nullptr, 0);
}

// Per-cloud bounds (kdtree_cpu.cpp:256-267): keys[c*2*dim + d] holds the min
// key of dimension d, keys[c*2*dim + dim + d] the max key. Partials per
// dimension: parts [d][2 * tiles].
__global__ __launch_bounds__(kBlock) void k_bounds(int64_t n, const float* __restrict__ cloud,
                                                   int64_t stride, int dim,
                                                   const int32_t* __restrict__ cid,
                                                   uint32_t* __restrict__ keys, TileParts parts,
                                                   int64_t nparts) {
  const int64_t tile = blockIdx.x;
  const int64_t base = tile * kTile + int64_t(threadIdx.x) * kItems;
  int32_t sg[kItems];
#pragma unroll
  for (int j = 0; j < kItems; ++j) sg[j] = base + j < n ? cid[base + j] : -1;
  for (int d = 0; d < dim; ++d) {
    uint32_t kv[kItems];
#pragma unroll
    for (int j = 0; j < kItems; ++j)
      kv[j] = base + j < n ? float_key(cloud[(base + j) * stride + d]) : 0u;
    const TileParts pd{parts.seg + d * nparts, parts.a + d * nparts, parts.b + d * nparts};
    tile_reduce<0, 1>(sg, kv, kv, keys + d, keys + dim + d, 2 * dim, &pd, tile);
    __syncthreads();
  }
}

__global__ void k_init_keys(int32_t clouds, int dim, uint32_t* keys) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= int64_t(clouds) * dim) return;
  const int64_t c = e / dim;
  const int d = int(e % dim);
  keys[c * 2 * dim + d] = float_key(FLT_MAX);        // minBound starts at max() (nabo.cpp:73)
  keys[c * 2 * dim + dim + d] = float_key(FLT_MIN);  // maxBound starts at min() > 0 (nabo.cpp:74)
}

// Level 0: one root segment per cloud, its cell = the cloud's bounds; the
// first level's state (clouds segments, BFS ids from 0).
__global__ void k_roots(int32_t clouds, int dim, const int64_t* __restrict__ offsets,
                        const uint32_t* __restrict__ keys, Seg seg, float* __restrict__ bounds,
                        BuildLevel* level0) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e == 0) {
    level0->S = clouds;
    level0->base = 0;
  }
  if (e >= int64_t(clouds) * dim) return;
  const int64_t c = e / dim;
  const int d = int(e % dim);
  const float lo = key_float(keys[c * 2 * dim + d]);
  const float hi = key_float(keys[c * 2 * dim + dim + d]);
  seg.lo[c * dim + d] = lo;
  seg.hi[c * dim + d] = hi;
  bounds[c * 2 * dim + d] = lo;
  bounds[c * 2 * dim + dim + d] = hi;
  if (d == 0) {
    seg.start[c] = int32_t(offsets[c]);
    seg.count[c] = int32_t(offsets[c + 1] - offsets[c]);
    seg.rc[c] = 0;
  }
}

// Cloud id of every point (binary search in the offsets), its record
// (coordinates, then its index within its cloud as bits), every point in its
// cloud's root segment.
__global__ void k_init_points(int64_t n, int32_t clouds, const int64_t* __restrict__ offsets,
                              const float* __restrict__ cloud, int64_t stride, int dim, int ps,
                              int32_t* __restrict__ cid, float* __restrict__ recs,
                              int32_t* __restrict__ pseg) {
  const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int lo = 0, hi = clouds - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (offsets[mid] <= p)
      lo = mid;
    else
      hi = mid - 1;
  }
  cid[p] = lo;
  pseg[p] = lo;
  float* r = recs + p * ps;
  const float* src = cloud + p * stride;
  for (int d = 0; d < dim; ++d) r[d] = src[d];
  r[dim] = __int_as_float(int32_t(p - offsets[lo]));
  for (int d = dim + 1; d < ps; ++d) r[d] = 0.f;
}

__global__ void k_seg_classify(const BuildLevel* __restrict__ cur, int32_t level, Seg seg,
                               SegStat st, Bfs bfs, uint32_t bucket, int dim, int32_t local_t,
                               LocalSeg* __restrict__ local, BuildCounters* __restrict__ ctr) {
  const int32_t S = cur->S;
  const int64_t base = cur->base;
  if (blockIdx.x == 0 && threadIdx.x == 0 && S > 0) atomicMax(&ctr->max_level, level);
  for (int32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < S; s += gridDim.x * blockDim.x) {
    const int32_t cnt = seg.count[s];
    const int64_t node = base + s;
    bfs.start[node] = seg.start[s];
    bfs.level[node] = level;
    bfs.rc[node] = seg.rc[s];
    if (cnt <= int32_t(bucket)) {
      st.split[s] = 0;
      bfs.cd[node] = uint32_t(dim);
      bfs.aux[node] = uint32_t(seg.start[s]);
      bfs.count[node] = uint32_t(cnt);
      continue;
    }
    if (cnt <= local_t) {
      // finished in LDS after the level loop (its node words are written there)
      st.split[s] = 0;
      LocalSeg& e = local[atomicAdd(&ctr->local_count, 1u)];
      e.start = seg.start[s];
      e.count = cnt;
      e.level = level;
      e.rc = seg.rc[s];
      e.node = int32_t(node);
      for (int d = 0; d < dim; ++d) {
        e.lo[d] = seg.lo[int64_t(s) * dim + d];
        e.hi[d] = seg.hi[int64_t(s) * dim + d];
      }
      continue;
    }
    st.split[s] = 1;
    const float* lo = seg.lo + int64_t(s) * dim;
    const float* hi = seg.hi + int64_t(s) * dim;
    float best = 0.f;
    uint32_t axis = 0;
    for (int d = 0; d < dim; ++d) {
      const float extent = hi[d] - lo[d];
      if (extent > best) {
        best = extent;
        axis = uint32_t(d);
      }
    }
    st.cd[s] = axis;
    st.ideal[s] = (hi[axis] + lo[axis]) * 0.5f;  // exact: same value as /2
    st.minkey[s] = 0xffffffffu;
    st.maxkey[s] = 0u;
    st.lt[s] = 0u;
    st.eq[s] = 0u;
  }
}

// Number of splits of a level, posted to the host (eager builds): the
// count, a system-scope fence, then the sequence number the host waits for.
struct Mailbox {
  int32_t total;
  int32_t extra;
  uint32_t seq;
};

// The next level's state: twice the splits of this one, BFS ids after it.
__global__ void k_level_next(const BuildLevel* __restrict__ cur, BuildLevel* __restrict__ nxt,
                             const uint32_t* __restrict__ splits, Mailbox* box, uint32_t seq) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int32_t total = cur->S > 0 ? int32_t(*splits) : 0;
  nxt->S = 2 * total;
  nxt->base = cur->base + cur->S;
  if (box) {
    __hip_atomic_store(&box->total, total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __threadfence_system();
    __hip_atomic_store(&box->seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// The point kernels below return at once when the level splits nothing
// (*gate, the next level's segment count, is 0).

// Reads the cut coordinate of every point still inside a splitting segment
// (one point per thread: each point's loads depend on its segment's).
__global__ void k_pt_gather(int64_t n, const int32_t* __restrict__ gate,
                            int32_t* __restrict__ pseg, const float* __restrict__ recs, int ps,
                            SegStat st, float* __restrict__ xs) {
  if (*gate == 0) return;
  const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int32_t s = pseg[p];
  if (s < 0) return;
  if (st.split[s] == 0) {
    pseg[p] = -1;  // its segment became a leaf
    return;
  }
  xs[p] = recs[p * ps + st.cd[s]];
}

// Min / max of the gathered cut coordinates per segment.
__global__ __launch_bounds__(kBlock) void k_pt_minmax(int64_t n, const int32_t* __restrict__ gate,
                                                      const int32_t* __restrict__ pseg,
                                                      const float* __restrict__ xs, SegStat st,
                                                      TileParts parts) {
  if (*gate == 0) return;
  const int64_t tile = blockIdx.x;
  const int64_t base = tile * kTile + int64_t(threadIdx.x) * kItems;
  int32_t sg[kItems];
  uint32_t kv[kItems];
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const int64_t p = base + j;
    const int s = p < n ? pseg[p] : -1;
    sg[j] = s;
    kv[j] = s >= 0 ? float_key(xs[p]) : 0u;
  }
  tile_reduce<0, 1>(sg, kv, kv, st.minkey, st.maxkey, 1, &parts, tile);
}

__global__ void k_seg_cut(const BuildLevel* __restrict__ cur, SegStat st) {
  const int32_t S = cur->S;
  for (int32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < S; s += gridDim.x * blockDim.x) {
    if (st.split[s] == 0) continue;
    const float pmin = key_float(st.minkey[s]);
    const float pmax = key_float(st.maxkey[s]);
    const float ideal = st.ideal[s];
    float cut;
    if (ideal < pmin)
      cut = pmin;
    else if (ideal > pmax)
      cut = pmax;
    else
      cut = ideal;
    st.cut[s] = cut;
  }
}

__global__ __launch_bounds__(kBlock) void k_pt_count(int64_t n, const int32_t* __restrict__ gate,
                                                     const int32_t* __restrict__ pseg,
                                                     const float* __restrict__ xs, SegStat st,
                                                     TileParts parts) {
  if (*gate == 0) return;
  const int64_t tile = blockIdx.x;
  const int64_t base = tile * kTile + int64_t(threadIdx.x) * kItems;
  int32_t sg[kItems];
  uint32_t lt[kItems], eq[kItems];
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const int64_t p = base + j;
    const int s = p < n ? pseg[p] : -1;
    sg[j] = s;
    lt[j] = 0u;
    eq[j] = 0u;
    if (s >= 0) {
      const float x = xs[p];
      const float cut = st.cut[s];
      lt[j] = x < cut ? 1u : 0u;
      eq[j] = x == cut ? 1u : 0u;
    }
  }
  tile_reduce<2, 2>(sg, lt, eq, st.lt, st.eq, 1, &parts, tile);
}

// Out-of-place flags of one Hoare pass: low word = belongs right but sits
// left, high word = belongs left but sits right.
__device__ __forceinline__ uint64_t pt_flag(int64_t p, int pass, const int32_t* __restrict__ pseg,
                                            const float* __restrict__ xs, const Seg& seg,
                                            const SegStat& st) {
  const int s = pseg[p];
  uint64_t f = 0;
  if (s >= 0) {
    const int32_t pos = int32_t(p) - seg.start[s];
    const float x = xs[p];
    const float cut = st.cut[s];
    const int32_t br1 = int32_t(st.lt[s]);
    bool fl, fr;
    if (pass == 1) {
      fl = pos < br1 && !(x < cut);
      fr = pos >= br1 && x < cut;
    } else {
      // After pass 1 every x in [br1, count) is >= cut, so "> cut" is
      // "!= cut" and "<= cut" is "== cut"; written with == the two counts
      // match by construction.
      const int32_t br2 = br1 + int32_t(st.eq[s]);
      fl = pos >= br1 && pos < br2 && !(x == cut);
      fr = pos >= br2 && x == cut;
    }
    f = uint64_t(fl) | (uint64_t(fr) << 32);
  }
  return f;
}

// The flags of every point, laid out as the scan splits them (workgroup b
// of kScanGrid writes chunk b of scan_chunk(n)), with each workgroup's sum
// in partials[b]: the scan's first pass, without reading the flags again.
constexpr int kFlagThreads = 1024;  // per workgroup: more points of a chunk in flight
__global__ __launch_bounds__(kFlagThreads) void k_pt_flags(
    int64_t n, const int32_t* __restrict__ gate, int pass, const int32_t* __restrict__ pseg,
    const float* __restrict__ xs, Seg seg, SegStat st, uint64_t* __restrict__ flags,
    uint64_t* __restrict__ partials) {
  if (*gate == 0) return;
  __shared__ uint64_t wave_sum[kFlagThreads / kWave];
  const int64_t chunk = scan_chunk(n);
  const int64_t begin = int64_t(blockIdx.x) * chunk;
  const int64_t end = begin + chunk < n ? begin + chunk : n;
  uint64_t sum = 0;
  for (int64_t p = begin + threadIdx.x; p < end; p += kFlagThreads) {
    const uint64_t f = pt_flag(p, pass, pseg, xs, seg, st);
    flags[p] = f;
    sum += f;
  }
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) sum += __shfl_xor(sum, o, kWave);
  if ((threadIdx.x & (kWave - 1)) == 0) wave_sum[threadIdx.x / kWave] = sum;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t total = 0;
#pragma unroll
    for (int w = 0; w < kFlagThreads / kWave; ++w) total += wave_sum[w];
    partials[blockIdx.x] = total;
  }
}

__device__ __forceinline__ void pass_ranks(int64_t p, int s, Seg seg,
                                           const uint64_t* __restrict__ flags,
                                           const uint64_t* __restrict__ excl, uint64_t f,
                                           int32_t* start_out, uint32_t* rank_l,
                                           uint32_t* rank_r) {
  const int32_t start = seg.start[s];
  const int32_t last = start + seg.count[s] - 1;
  const uint64_t base = excl[start];
  const uint64_t mine = excl[p];
  *start_out = start;
  *rank_l = uint32_t(mine) - uint32_t(base);
  if (f >> 32) {
    const uint64_t end = excl[last] + flags[last];
    const uint32_t total_r = uint32_t(end >> 32) - uint32_t(base >> 32);
    const uint32_t from_left = uint32_t(mine >> 32) - uint32_t(base >> 32);
    *rank_r = total_r - 1u - from_left;
  } else {
    *rank_r = 0;
  }
}

// Slot of every right-misplaced point, by rank: the i-th left-misplaced
// point of a segment (from the left) swaps with the i-th right-misplaced
// one counted from the right.
__global__ void k_pt_slots(int64_t n, const int32_t* __restrict__ gate,
                           const int32_t* __restrict__ pseg, Seg seg,
                           const uint64_t* __restrict__ flags, const uint64_t* __restrict__ excl,
                           int32_t* __restrict__ slot_r) {
  if (*gate == 0) return;
  const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint64_t f = flags[p];
  if ((f >> 32) == 0) return;
  const int s = pseg[p];
  int32_t start;
  uint32_t rl, rr;
  pass_ranks(p, s, seg, flags, excl, f, &start, &rl, &rr);
  slot_r[start + rr] = int32_t(p);
}

// The pass's swaps, in place: the thread of each left-misplaced point swaps
// its record and cut coordinate with its partner's (ps floats per record;
// 4 for 2-D and 3-D clouds, one 16-byte access). Pairs are disjoint, so no
// two threads touch one point.
__global__ void k_pt_swap(int64_t n, const int32_t* __restrict__ gate,
                          const int32_t* __restrict__ pseg, Seg seg,
                          const uint64_t* __restrict__ flags, const uint64_t* __restrict__ excl,
                          const int32_t* __restrict__ slot_r, float* __restrict__ recs,
                          float* __restrict__ xs, int ps) {
  if (*gate == 0) return;
  const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint64_t f = flags[p];
  if (uint32_t(f) == 0) return;
  const int s = pseg[p];
  const int32_t start = seg.start[s];
  const uint32_t rl = uint32_t(excl[p]) - uint32_t(excl[start]);
  const int64_t q = slot_r[start + rl];
  if (ps == 4) {
    float4* r4 = reinterpret_cast<float4*>(recs);
    const float4 a = r4[p], b = r4[q];
    r4[p] = b;
    r4[q] = a;
  } else {
    for (int d = 0; d < ps; ++d) {
      const float a = recs[p * ps + d], b = recs[q * ps + d];
      recs[p * ps + d] = b;
      recs[q * ps + d] = a;
    }
  }
  const float xa = xs[p], xb = xs[q];
  xs[p] = xb;
  xs[q] = xa;
}

__global__ void k_seg_children(const BuildLevel* __restrict__ cur, Seg seg, SegStat st, Bfs bfs,
                               Seg next, int dim) {
  const int32_t S = cur->S;
  const int64_t node_base = cur->base;
  const int64_t next_base = node_base + S;
  for (int32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < S; s += gridDim.x * blockDim.x) {
    if (st.split[s] == 0) continue;
    const int32_t cnt = seg.count[s];
    const int32_t start = seg.start[s];
    const float ideal = st.ideal[s];
    const float pmin = key_float(st.minkey[s]);
    const float pmax = key_float(st.maxkey[s]);
    const int32_t br1 = int32_t(st.lt[s]);
    const int32_t br2 = br1 + int32_t(st.eq[s]);
    int32_t left;
    if (ideal < pmin)
      left = 1;
    else if (ideal > pmax)
      left = cnt - 1;
    else if (br1 > cnt / 2)
      left = br1;
    else if (br2 < cnt / 2)
      left = br2;
    else
      left = cnt / 2;
    left = min(max(left, 1), cnt - 1);  // no-op for finite coordinates
    st.left[s] = left;
    const uint32_t cd = st.cd[s];
    const float cut = st.cut[s];
    const int32_t c = int32_t(2 * st.childbase[s]);
    next.start[c] = start;
    next.count[c] = left;
    next.start[c + 1] = start + left;
    next.count[c + 1] = cnt - left;
    next.rc[c] = seg.rc[s];
    next.rc[c + 1] = seg.rc[s] + 1;
    for (int d = 0; d < dim; ++d) {
      const float lo = seg.lo[int64_t(s) * dim + d];
      const float hi = seg.hi[int64_t(s) * dim + d];
      next.lo[int64_t(c) * dim + d] = lo;
      next.hi[int64_t(c) * dim + d] = (uint32_t(d) == cd) ? cut : hi;
      next.lo[int64_t(c + 1) * dim + d] = (uint32_t(d) == cd) ? cut : lo;
      next.hi[int64_t(c + 1) * dim + d] = hi;
    }
    const int64_t node = node_base + s;
    bfs.cd[node] = cd;
    bfs.aux[node] = __float_as_uint(cut);
    bfs.left[node] = int32_t(next_base + c);
    bfs.right[node] = int32_t(next_base + c + 1);
  }
}

__global__ void k_pt_reseg(int64_t n, const int32_t* __restrict__ gate,
                           int32_t* __restrict__ pseg, Seg seg, SegStat st) {
  if (*gate == 0) return;
  const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int s = pseg[p];
  if (s < 0) return;
  const int32_t pos = int32_t(p) - seg.start[s];
  pseg[p] = int32_t(2 * st.childbase[s]) + (pos < st.left[s] ? 0 : 1);
}

// Exclusive prefix sum over the workgroup's kLocalBlock threads (wave
// shuffles, then the waves' totals through LDS). One barrier: the caller
// does not write wave_tot again before every thread has passed another.
__device__ __forceinline__ uint32_t block_excl_sum(uint32_t v, uint32_t* wave_tot) {
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  uint32_t incl = v;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const uint32_t up = __shfl_up(incl, o, kWave);
    if (lane >= o) incl += up;
  }
  if (lane == kWave - 1) wave_tot[w] = incl;
  __syncthreads();
  uint32_t add = 0;
  for (int i = 0; i < w; ++i) add += wave_tot[i];
  return add + incl - v;
}

__device__ __forceinline__ float rec_coord(const float4& r, uint32_t cd) {
  const float lo = cd == 1u ? r.y : r.x;
  return cd == 2u ? r.z : lo;
}

// The segments that split at one level of an LDS subtree.
template <int DIM>
struct LocalTable {
  int32_t start[kLocalSegs], count[kLocalSegs], rc[kLocalSegs], node[kLocalSegs];
  float lo[kLocalSegs * DIM], hi[kLocalSegs * DIM];
  uint32_t cd[kLocalSegs], mn[kLocalSegs], mx[kLocalSegs], lt[kLocalSegs], eq[kLocalSegs];
  float ideal[kLocalSegs];

  // Entry k: position in the subtree's points, count, right-child count,
  // node id and cell; its cut dimension (the first widest extent) and ideal
  // cut (the cell's midpoint); min / max and counts cleared.
  __device__ __forceinline__ void put(int k, int32_t st, int32_t cnt, int32_t r, int32_t nd,
                                      const float* l, const float* h) {
    start[k] = st;
    count[k] = cnt;
    rc[k] = r;
    node[k] = nd;
    float best = 0.f;
    uint32_t axis = 0;
    float mid = (h[0] + l[0]) * 0.5f;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      lo[k * DIM + d] = l[d];
      hi[k * DIM + d] = h[d];
      const float extent = h[d] - l[d];
      if (extent > best) {
        best = extent;
        axis = uint32_t(d);
        mid = (h[d] + l[d]) * 0.5f;
      }
    }
    cd[k] = axis;
    ideal[k] = mid;
    mn[k] = 0xffffffffu;
    mx[k] = 0u;
    lt[k] = 0u;
    eq[k] = 0u;
  }
};

// Finishes segments of at most kLocalT points in LDS, one workgroup per
// segment (taken from the LocalSeg list through a counter): level by level,
// every splitting segment of the subtree at once, the steps of the level
// loop above on LDS copies of the records (thread t owns points
// [kLocalItems t, kLocalItems (t + 1)) and keeps their segments in
// registers; per-segment min/max and counts by runs of equal segments, one
// LDS atomic per run; each Hoare pass's ranks from a workgroup scan of the
// packed out-of-place flags). A level reads its segments from one table and
// writes the next level's into the other, with their cut dimensions and
// ideal cuts, so a level takes ten workgroup barriers. The subtree's nodes
// take ids r0 + 2 * start + j (j < 2 * count - 2, in creation order); the
// segment's own node keeps its level-loop id.
template <int DIM>
__global__ __launch_bounds__(kLocalBlock) void k_local_finish(
    const LocalSeg* __restrict__ list, BuildCounters* __restrict__ ctr, float4* __restrict__ recs,
    Bfs bfs, int64_t r0, uint32_t bucket) {
  __shared__ float4 rec[kLocalT];
  __shared__ int16_t slot_r[kLocalT];
  __shared__ uint32_t excl[kLocalT + 1];
  __shared__ LocalTable<DIM> tab[2];
  __shared__ int32_t c_left[kLocalSegs];
  __shared__ int16_t c_child[2 * kLocalSegs];
  __shared__ uint32_t wave_tot[2][kLocalBlock / kWave];  // one per Hoare pass
  __shared__ int32_t s_entry, s_next[2], s_nodes, s_lmax;
  const int t = threadIdx.x;
  for (;;) {
    if (t == 0) s_entry = int32_t(atomicAdd(&ctr->claim, 1u));
    __syncthreads();
    const int32_t e = s_entry;
    if (e >= int32_t(__hip_atomic_load(&ctr->local_count, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT)))
      return;
    const int32_t C = list[e].count;
    const int64_t S0 = list[e].start;
    int32_t level = list[e].level;
    for (int p = t; p < C; p += kLocalBlock) rec[p] = recs[S0 + p];
    // segment of each of this thread's points in the level's table; -1:
    // in a leaf, or past the segment
    int g[kLocalItems];
#pragma unroll
    for (int j = 0; j < kLocalItems; ++j) g[j] = t * kLocalItems + j < C ? 0 : -1;
    if (t == 0) {
      float lo[DIM], hi[DIM];
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        lo[d] = list[e].lo[d];
        hi[d] = list[e].hi[d];
      }
      tab[0].put(0, 0, C, list[e].rc, list[e].node, lo, hi);
      s_next[0] = 0;
      s_next[1] = 0;
      s_nodes = 0;
      s_lmax = level;
    }
    __syncthreads();
    int ns = 1, cur = 0;
    while (ns > 0) {
      LocalTable<DIM>& ct = tab[cur];
      LocalTable<DIM>& nt = tab[cur ^ 1];
      // This thread's points for the level, in registers: cut dimension and
      // coordinate along it. The loads of all items are issued before the
      // first LDS atomic, which would order them.
      uint32_t cdv[kLocalItems];
      float x[kLocalItems];
#pragma unroll
      for (int j = 0; j < kLocalItems; ++j) cdv[j] = g[j] >= 0 ? ct.cd[g[j]] : 0u;
#pragma unroll
      for (int j = 0; j < kLocalItems; ++j) x[j] = rec_coord(rec[t * kLocalItems + j], cdv[j]);
      // min / max of the cut coordinate
      {
        int run = -1;
        uint32_t mn = 0xffffffffu, mx = 0u;
#pragma unroll
        for (int j = 0; j < kLocalItems; ++j) {
          if (g[j] != run) {
            if (run >= 0) {
              atomicMin(&ct.mn[run], mn);
              atomicMax(&ct.mx[run], mx);
            }
            run = g[j];
            mn = 0xffffffffu;
            mx = 0u;
          }
          if (g[j] >= 0) {
            const uint32_t key = float_key(x[j]);
            mn = min(mn, key);
            mx = max(mx, key);
          }
        }
        if (run >= 0) {
          atomicMin(&ct.mn[run], mn);
          atomicMax(&ct.mx[run], mx);
        }
      }
      __syncthreads();
      // every thread has read this table's count: the counter of the table
      // the next level fills can be cleared
      if (t == 0) s_next[cur] = 0;
      // the cut (the ideal cut clamped to the points' range), points below /
      // equal to it
      float cut[kLocalItems];
      int32_t sst[kLocalItems], scnt[kLocalItems];
#pragma unroll
      for (int j = 0; j < kLocalItems; ++j) {
        cut[j] = 0.f;
        sst[j] = 0;
        scnt[j] = 0;
        if (g[j] >= 0) {
          const float pmin = key_float(ct.mn[g[j]]), pmax = key_float(ct.mx[g[j]]);
          const float ideal = ct.ideal[g[j]];
          cut[j] = ideal < pmin ? pmin : (ideal > pmax ? pmax : ideal);
          sst[j] = ct.start[g[j]];
          scnt[j] = ct.count[g[j]];
        }
      }
      {
        int run = -1;
        uint32_t lt = 0u, eq = 0u;
#pragma unroll
        for (int j = 0; j < kLocalItems; ++j) {
          if (g[j] != run) {
            if (run >= 0) {
              atomicAdd(&ct.lt[run], lt);
              atomicAdd(&ct.eq[run], eq);
            }
            run = g[j];
            lt = 0u;
            eq = 0u;
          }
          if (g[j] >= 0) {
            lt += x[j] < cut[j] ? 1u : 0u;
            eq += x[j] == cut[j] ? 1u : 0u;
          }
        }
        if (run >= 0) {
          atomicAdd(&ct.lt[run], lt);
          atomicAdd(&ct.eq[run], eq);
        }
      }
      __syncthreads();
      int32_t br1[kLocalItems], br2[kLocalItems];
#pragma unroll
      for (int j = 0; j < kLocalItems; ++j) {
        br1[j] = g[j] >= 0 ? int32_t(ct.lt[g[j]]) : 0;
        br2[j] = br1[j] + (g[j] >= 0 ? int32_t(ct.eq[g[j]]) : 0);
      }
      // the two Hoare passes: flags (low half: belongs right but sits left,
      // high half: belongs left but sits right), ranks, pairs, swaps
      for (int pass = 1; pass <= 2; ++pass) {
        if (pass == 2) {
          // the first pass's swaps moved records between threads
#pragma unroll
          for (int j = 0; j < kLocalItems; ++j)
            x[j] = rec_coord(rec[t * kLocalItems + j], cdv[j]);
        }
        uint32_t f[kLocalItems];
        uint32_t sum = 0;
#pragma unroll
        for (int j = 0; j < kLocalItems; ++j) {
          f[j] = 0u;
          if (g[j] >= 0) {
            const int32_t pos = t * kLocalItems + j - sst[j];
            bool fl, fr;
            if (pass == 1) {
              fl = pos < br1[j] && !(x[j] < cut[j]);
              fr = pos >= br1[j] && x[j] < cut[j];
            } else {
              fl = pos >= br1[j] && pos < br2[j] && !(x[j] == cut[j]);
              fr = pos >= br2[j] && x[j] == cut[j];
            }
            f[j] = uint32_t(fl) | (uint32_t(fr) << 16);
          }
          sum += f[j];
        }
        uint32_t ex[kLocalItems];
        uint32_t run = block_excl_sum(sum, wave_tot[pass - 1]);
#pragma unroll
        for (int j = 0; j < kLocalItems; ++j) {
          ex[j] = run;
          excl[t * kLocalItems + j] = run;
          run += f[j];
        }
        if (t == kLocalBlock - 1) excl[kLocalT] = run;
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kLocalItems; ++j) {
          if ((f[j] >> 16) == 0u) continue;
          const uint32_t base = excl[sst[j]] >> 16;
          const uint32_t total_r = (excl[sst[j] + scnt[j]] >> 16) - base;
          const uint32_t from_left = (ex[j] >> 16) - base;
          slot_r[sst[j] + int32_t(total_r - 1u - from_left)] = int16_t(t * kLocalItems + j);
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kLocalItems; ++j) {
          if ((f[j] & 0xffffu) == 0u) continue;
          const int p = t * kLocalItems + j;
          const int q = slot_r[sst[j] + int32_t((ex[j] & 0xffffu) - (excl[sst[j]] & 0xffffu))];
          const float4 a = rec[p];
          rec[p] = rec[q];
          rec[q] = a;
        }
        // the second pass's swaps are read only after the barrier that
        // ends the level
        if (pass == 1) __syncthreads();
      }
      // leftCount rule, node words, children (the next level's table)
      for (int s = t; s < ns; s += kLocalBlock) {
        const int32_t cnt = ct.count[s];
        const float ideal = ct.ideal[s];
        const float pmin = key_float(ct.mn[s]), pmax = key_float(ct.mx[s]);
        const int32_t b1 = int32_t(ct.lt[s]);
        const int32_t b2 = b1 + int32_t(ct.eq[s]);
        int32_t left;
        if (ideal < pmin)
          left = 1;
        else if (ideal > pmax)
          left = cnt - 1;
        else if (b1 > cnt / 2)
          left = b1;
        else if (b2 < cnt / 2)
          left = b2;
        else
          left = cnt / 2;
        left = min(max(left, 1), cnt - 1);
        c_left[s] = left;
        const uint32_t cd = ct.cd[s];
        const float cut_s = ideal < pmin ? pmin : (ideal > pmax ? pmax : ideal);
        const int64_t id0 = r0 + 2 * S0 + atomicAdd(&s_nodes, 2);
        const int32_t node = ct.node[s];
        bfs.cd[node] = cd;
        bfs.aux[node] = __float_as_uint(cut_s);
        bfs.left[node] = int32_t(id0);
        bfs.right[node] = int32_t(id0 + 1);
        for (int side = 0; side < 2; ++side) {
          const int32_t cs = ct.start[s] + (side ? left : 0);
          const int32_t cc = side ? cnt - left : left;
          const int32_t rc = ct.rc[s] + side;
          const int64_t id = id0 + side;
          bfs.start[id] = int32_t(S0 + cs);
          bfs.level[id] = level + 1;
          bfs.rc[id] = rc;
          if (cc <= int32_t(bucket)) {
            bfs.cd[id] = uint32_t(DIM);
            bfs.aux[id] = uint32_t(S0 + cs);
            bfs.count[id] = uint32_t(cc);
            c_child[2 * s + side] = -1;
            atomicMax(&s_lmax, level + 1);
          } else {
            float lo[DIM], hi[DIM];
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
              lo[d] = (side && uint32_t(d) == cd) ? cut_s : ct.lo[s * DIM + d];
              hi[d] = (!side && uint32_t(d) == cd) ? cut_s : ct.hi[s * DIM + d];
            }
            const int k = atomicAdd(&s_next[cur ^ 1], 1);
            nt.put(k, cs, cc, rc, int32_t(id), lo, hi);
            c_child[2 * s + side] = int16_t(k);
          }
        }
      }
      __syncthreads();
#pragma unroll
      for (int j = 0; j < kLocalItems; ++j) {
        const int p = t * kLocalItems + j;
        if (g[j] >= 0) g[j] = c_child[2 * g[j] + (p - sst[j] < c_left[g[j]] ? 0 : 1)];
      }
      ns = s_next[cur ^ 1];
      cur ^= 1;
      ++level;
    }
    for (int p = t; p < C; p += kLocalBlock) recs[S0 + p] = rec[p];
    if (t == 0) {
      atomicAdd(&ctr->local_nodes, uint32_t(s_nodes));
      atomicMax(&ctr->max_level, s_lmax);
    }
    __syncthreads();
  }
}

// Node count and depth after the LDS subtrees: tot->base = all nodes,
// tot->S = the deepest level; posted to the host for eager builds.
__global__ void k_build_totals(const BuildLevel* __restrict__ fin,
                               const BuildCounters* __restrict__ ctr, BuildLevel* __restrict__ tot,
                               Mailbox* box, uint32_t seq) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int64_t total = fin->base + int64_t(ctr->local_nodes);
  tot->base = total;
  tot->S = ctr->max_level;
  if (box) {
    __hip_atomic_store(&box->total, int32_t(total), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&box->extra, ctr->max_level, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __threadfence_system();
    __hip_atomic_store(&box->seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// The finalize kernels below return at once when *ok is 0 (a recorded
// build whose levels did not finish the tree: its records are partial);
// eager builds pass null.
__global__ void k_fill_u32(uint32_t* __restrict__ p, int64_t count, uint32_t v,
                           const int32_t* __restrict__ ok) {
  if (ok && *ok == 0) return;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
       i += int64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}

// Node ids: the level loop's [0, fin->base), then the LDS subtrees' range
// [r0, r0 + local) whose unused ids have level -1. node_id maps a position
// in that sequence to an id, or -1.
struct NodeIds {
  int64_t r0, local;
  __device__ __forceinline__ int64_t count(const BuildLevel* fin) const { return fin->base + local; }
  __device__ __forceinline__ int64_t id(const BuildLevel* fin, const Bfs& bfs, int64_t v) const {
    if (v < fin->base) return v;
    const int64_t u = r0 + (v - fin->base);
    return bfs.level[u] >= 0 ? u : -1;
  }
};

__global__ void k_mark_leaves(const BuildLevel* __restrict__ fin, Bfs bfs, int dim, NodeIds ids,
                              uint32_t* __restrict__ leaf_start, const int32_t* __restrict__ ok) {
  if (ok && *ok == 0) return;
  const int64_t count = ids.count(fin);
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
       i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t v = ids.id(fin, bfs, i);
    if (v >= 0 && bfs.cd[v] == uint32_t(dim)) leaf_start[bfs.start[v]] = 1u;
  }
}

__global__ void k_preorder(const BuildLevel* __restrict__ fin, Bfs bfs, NodeIds ids,
                           const uint32_t* __restrict__ leaves_before,
                           const int32_t* __restrict__ cid, const int32_t* __restrict__ ok) {
  if (ok && *ok == 0) return;
  const int64_t count = ids.count(fin);
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
       i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t v = ids.id(fin, bfs, i);
    if (v < 0) continue;
    const int32_t s = bfs.start[v];
    bfs.pre[v] = 2u * leaves_before[s] + uint32_t(bfs.level[v]) - uint32_t(bfs.rc[v]) -
                 uint32_t(cid[s]);
  }
}

__global__ void k_encode(const BuildLevel* __restrict__ fin, Bfs bfs, NodeIds ids,
                         uint32_t dim_bits, int dim, uint2* __restrict__ nodes,
                         const int32_t* __restrict__ ok) {
  if (ok && *ok == 0) return;
  const int64_t count = ids.count(fin);
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
       i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t v = ids.id(fin, bfs, i);
    if (v < 0) continue;
    const uint32_t cd = bfs.cd[v];
    uint2 out;
    if (cd == uint32_t(dim)) {
      out.x = uint32_t(dim) | (bfs.count[v] << dim_bits);
      out.y = bfs.aux[v];
    } else {
      out.x = cd | (bfs.pre[bfs.right[v]] << dim_bits);
      out.y = bfs.aux[v];
    }
    nodes[bfs.pre[v]] = out;
  }
}

// Search-side links, by pre-order node p: links[p].xy is the node word of
// p's right child (the left child is node p + 1), so one round trip reads a
// node and both its children; links[c].zw is {cut dim, cut bits} of the
// split of c's parent, so a far child is entered without reading its parent
// again. nodes[count] is a padding leaf, so nodes[p + 1] exists for every p.
// tot->base: the node count.
__global__ void k_links_self(const BuildLevel* __restrict__ tot, const uint2* __restrict__ nodes,
                             uint32_t dim_bits, uint32_t dim_mask, int dim,
                             uint4* __restrict__ links, uint2* __restrict__ all_nodes,
                             const int32_t* __restrict__ ok) {
  if (ok && *ok == 0) return;
  const int64_t count = tot->base;
  for (int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; p < count;
       p += int64_t(gridDim.x) * blockDim.x) {
    const uint2 nd = nodes[p];
    uint4 l = make_uint4(0u, 0u, 0u, 0u);
    if ((nd.x & dim_mask) != uint32_t(dim)) {
      const uint2 r = nodes[nd.x >> dim_bits];
      l.x = r.x;
      l.y = r.y;
    }
    links[p] = l;
    if (p == 0) all_nodes[count] = make_uint2(uint32_t(dim), 0u);
  }
}

__global__ void k_links_parent(const BuildLevel* __restrict__ tot,
                               const uint2* __restrict__ nodes, uint32_t dim_bits,
                               uint32_t dim_mask, int dim, uint4* __restrict__ links,
                               const int32_t* __restrict__ ok) {
  if (ok && *ok == 0) return;
  const int64_t count = tot->base;
  for (int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; p < count;
       p += int64_t(gridDim.x) * blockDim.x) {
    const uint2 nd = nodes[p];
    const uint32_t cd = nd.x & dim_mask;
    if (cd == uint32_t(dim)) continue;
    const uint2 split = make_uint2(cd, nd.y);
    reinterpret_cast<uint2*>(links + p + 1)[1] = split;
    reinterpret_cast<uint2*>(links + (nd.x >> dim_bits))[1] = split;
  }
}

// Zeroed records after the last point (a leaf scan may read kPtsPad entries
// past its bucket).
__global__ void k_pad_points(int64_t n, int ps, float* __restrict__ recs) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e < int64_t(kPtsPad) * ps) recs[n * ps + e] = 0.f;
}

// Root of cloud c: the nodes of the trees before it, 2 leaves - 1 each.
__global__ void k_roots_out(int32_t clouds, const int64_t* __restrict__ offsets,
                            const uint32_t* __restrict__ leaves_before,
                            int32_t* __restrict__ roots, const int32_t* __restrict__ ok) {
  if (ok && *ok == 0) return;
  const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < clouds) roots[c] = int32_t(2u * leaves_before[offsets[c]] - uint32_t(c));
}

// Per-cloud query-order boxes of 2-D / 3-D batched trees, {min bound[3],
// 1 / extent[3]}: a cloud of at most bucketSize points keeps the
// reference's initial bounds (kdtree_cpu.cpp:236-243), whose extent is
// negative, so its scale is 0.
__global__ void k_boxes(int32_t clouds, int dim, const int64_t* __restrict__ offsets,
                        uint32_t bucket, const float* __restrict__ bounds,
                        float* __restrict__ boxes) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= int64_t(clouds) * 3) return;
  const int64_t c = e / 3;
  const int d = int(e % 3);
  float lo = 0.f, sc = 0.f;
  if (d < dim) {
    const bool one_bucket = offsets[c + 1] - offsets[c] <= int64_t(bucket);
    lo = one_bucket ? FLT_MAX : bounds[c * 2 * dim + d];
    const float hi = one_bucket ? FLT_MIN : bounds[c * 2 * dim + dim + d];
    const float ext = hi - lo;
    sc = ext > 0.f ? 1.f / ext : 0.f;
  }
  boxes[c * 6 + d] = lo;
  boxes[c * 6 + 3 + d] = sc;
}

// 1 when the recorded levels finished the tree (the level after the last
// one has no segments) and its deepest level is at most max_level (the
// depth the search's stacks are sized for), else 0.
__global__ void k_build_ok(const BuildLevel* __restrict__ fin, const BuildLevel* __restrict__ tot,
                           int32_t max_level, int32_t* __restrict__ ok) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *ok = (fin->S == 0 && tot->S <= max_level) ? 1 : 0;
}

inline unsigned grid_for(int64_t n) { return unsigned(std::max<int64_t>(1, ceil_div(n, kBlock))); }
inline unsigned tiles_for(int64_t n) { return unsigned(std::max<int64_t>(1, ceil_div(n, kTile))); }

// Pinned host mailboxes for the per-level split counts, pooled: the host
// polls one instead of a device-to-host copy and a stream synchronization
// (that round trip took ~45 us per level, 28 levels for 1M points).
struct HostMailbox {
  Mailbox* box;
  uint32_t seq;
};
std::mutex g_mailbox_mu;
std::vector<HostMailbox> g_mailboxes;

class LevelMailbox {
 public:
  LevelMailbox() {
    {
      std::lock_guard<std::mutex> lock(g_mailbox_mu);
      if (!g_mailboxes.empty()) {
        m_ = g_mailboxes.back();
        g_mailboxes.pop_back();
      }
    }
    if (!m_.box) {
      // coherent (fine-grained): the kernel's system-scope stores reach the
      // host without waiting for the kernel's end
      NH_CHECK(hipHostMalloc(reinterpret_cast<void**>(&m_.box), sizeof(Mailbox),
                             hipHostMallocCoherent));
      m_.box->total = 0;
      m_.box->extra = 0;
      m_.box->seq = 0;
      m_.seq = 0;
    }
  }
  ~LevelMailbox() {
    std::lock_guard<std::mutex> lock(g_mailbox_mu);
    g_mailboxes.push_back(m_);
  }
  LevelMailbox(const LevelMailbox&) = delete;
  LevelMailbox& operator=(const LevelMailbox&) = delete;

  Mailbox* box() const { return m_.box; }
  uint32_t next() { return ++m_.seq; }
  // The second value posted with the last `seq` waited for.
  int32_t extra() const { return __atomic_load_n(&m_.box->extra, __ATOMIC_RELAXED); }
  // Waits until the kernel has posted `seq`; after ~20 ms of polling it
  // falls back to synchronizing the stream, which reports a failed launch.
  int32_t wait(uint32_t seq, hipStream_t s) {
    const auto t0 = std::chrono::steady_clock::now();
    for (uint32_t spins = 0;; ++spins) {
      if (__atomic_load_n(&m_.box->seq, __ATOMIC_ACQUIRE) == seq)
        return __atomic_load_n(&m_.box->total, __ATOMIC_RELAXED);
      if ((spins & 1023u) == 1023u &&
          std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(20)) {
        NH_CHECK(hipStreamSynchronize(s));
        if (__atomic_load_n(&m_.box->seq, __ATOMIC_ACQUIRE) == seq)
          return __atomic_load_n(&m_.box->total, __ATOMIC_RELAXED);
        throw Error(NABO_HIP_ERR_HIP, "kd-tree build: level count not posted");
      }
    }
  }

 private:
  HostMailbox m_{nullptr, 0};
};

// Scratch arrays of one build, carved from one block.
struct BuildScratch {
  Span<int32_t> seg_start[2], seg_count[2], seg_rc[2];
  Span<float> seg_lo[2], seg_hi[2];
  Span<uint32_t> st_cd, st_minkey, st_maxkey, st_lt, st_eq, st_split, st_childbase;
  Span<float> st_ideal, st_cut;
  Span<int32_t> st_left;
  Span<uint32_t> b_cd, b_aux, b_count, b_pre;
  Span<int32_t> b_left, b_right, b_start, b_level, b_rc;
  Span<int32_t> pseg, slot_r, cid;
  Span<float> xs;
  Span<uint64_t> flags, excl, partials64;
  Span<uint32_t> keys, leaf_start, leaves_before, partials32, splits;
  Span<float> bounds;
  Span<int64_t> d_offsets;
  Span<int32_t> part_seg;
  Span<uint32_t> part_a, part_b;
  Span<BuildLevel> levels;  // two alternating level states, then the totals
  Span<LocalSeg> local;     // segments finished in LDS
  Span<BuildCounters> counters;
  int32_t local_t = 0;      // largest segment finished in LDS (0: none)

  void want(Scratch& sc, int64_t n, int dim, int ps, int32_t clouds, uint32_t bucket) {
    const int64_t seg_cap = std::max<int64_t>(n, clouds);
    local_t = local_points(dim, bucket);
    // node ids: the level loop's below 2n, the LDS subtrees' in [2n, 4n)
    const int64_t max_nodes = local_t > 0 ? 4 * n : 2 * n;
    for (int b = 0; b < 2; ++b) {
      sc.want(seg_start[b], size_t(seg_cap));
      sc.want(seg_count[b], size_t(seg_cap));
      sc.want(seg_rc[b], size_t(seg_cap));
      sc.want(seg_lo[b], size_t(seg_cap) * dim);
      sc.want(seg_hi[b], size_t(seg_cap) * dim);
    }
    for (Span<uint32_t>* a : {&st_cd, &st_minkey, &st_maxkey, &st_lt, &st_eq, &st_split,
                              &st_childbase})
      sc.want(*a, size_t(seg_cap));
    sc.want(st_ideal, size_t(seg_cap));
    sc.want(st_cut, size_t(seg_cap));
    sc.want(st_left, size_t(seg_cap));
    for (Span<uint32_t>* a : {&b_cd, &b_aux, &b_count, &b_pre}) sc.want(*a, size_t(max_nodes));
    for (Span<int32_t>* a : {&b_left, &b_right, &b_start, &b_level, &b_rc})
      sc.want(*a, size_t(max_nodes));
    for (Span<int32_t>* a : {&pseg, &slot_r, &cid}) sc.want(*a, size_t(n));
    sc.want(xs, size_t(n));
    sc.want(flags, size_t(n));
    sc.want(excl, size_t(n));
    sc.want(partials64, kScanGrid);
    sc.want(keys, size_t(2) * clouds * dim);
    sc.want(leaf_start, size_t(n) + 1);
    sc.want(leaves_before, size_t(n) + 1);
    sc.want(partials32, kScanGrid);
    sc.want(splits, 1);
    sc.want(bounds, size_t(2) * clouds * dim);
    sc.want(d_offsets, size_t(clouds) + 1);
    const int64_t nparts = 2 * int64_t(tiles_for(n));
    sc.want(part_seg, size_t(nparts) * dim);
    sc.want(part_a, size_t(nparts) * dim);
    sc.want(part_b, size_t(nparts) * dim);
    sc.want(levels, 3);
    sc.want(local, local_t > 0 ? size_t(n / (int64_t(bucket) + 1) + 1) : 1);
    sc.want(counters, 1);
  }
};

// The build's launch sequence. The offsets must already be in
// scratch d_offsets. Eager (levels == 0): levels are launched until one
// splits nothing, reading each level's split count from the mailbox;
// *levels_out and *nodes_out receive the tree's depth + 1 and node count,
// *loop_out the levels the loop ran, and the node arrays are allocated once
// the count is known (out.nodes / out.links null on entry). Recorded
// (levels > 0): exactly `levels` levels of the loop, node arrays
// preallocated for 2n nodes; out.ok receives 0 if the tree needed more
// levels or is deeper than max_depth, in which case the node arrays are left
// unwritten.
void kd_build_sequence(const KdBuildShape& sh, const float* d_cloud, int64_t stride,
                       BuildScratch& sc, KdBuildOut& out, Launcher& L, int levels,
                       int max_depth, int* levels_out, int* loop_out, int64_t* nodes_out) {
  const int64_t n = sh.n;
  const int dim = sh.dim;
  const int ps = sh.pts_stride;
  const int32_t clouds = sh.clouds;
  const bool eager = levels == 0;
  Seg seg[2];
  for (int b = 0; b < 2; ++b)
    seg[b] = Seg{sc.seg_start[b].p, sc.seg_count[b].p, sc.seg_rc[b].p, sc.seg_lo[b].p,
                 sc.seg_hi[b].p};
  SegStat st{sc.st_cd.p, sc.st_ideal.p, sc.st_cut.p, sc.st_minkey.p, sc.st_maxkey.p, sc.st_lt.p,
             sc.st_eq.p, sc.st_split.p, sc.st_childbase.p, sc.st_left.p};
  Bfs bfs{sc.b_cd.p, sc.b_aux.p, sc.b_count.p, sc.b_left.p, sc.b_right.p,
          sc.b_start.p, sc.b_level.p, sc.b_rc.p, sc.b_pre.p};
  const int64_t nparts = 2 * int64_t(tiles_for(n));
  const TileParts parts{sc.part_seg.p, sc.part_a.p, sc.part_b.p};
  const unsigned gpart = unsigned(ceil_div(nparts, kTile));
  float* recs = out.pts;
  BuildLevel* lvl = sc.levels.p;
  BuildLevel* tot = lvl + 2;
  const int64_t r0 = 2 * n;  // first id of the LDS subtrees' nodes
  const NodeIds ids{r0, sc.local_t > 0 ? 2 * n : 0};
  BuildCounters* ctr = sc.counters.p;
  L(k_fill_u32, dim3(1), dim3(64), 0, reinterpret_cast<uint32_t*>(ctr),
    int64_t(sizeof(BuildCounters) / sizeof(uint32_t)), 0u, static_cast<const int32_t*>(nullptr));
  if (sc.local_t > 0)
    L(k_fill_u32, dim3(unsigned(std::min<int64_t>(kGridStride, grid_for(2 * n)))), dim3(kBlock), 0,
      reinterpret_cast<uint32_t*>(bfs.level + r0), 2 * n, 0xffffffffu,
      static_cast<const int32_t*>(nullptr));

  // Bounds pass (kdtree_cpu.cpp:256-267) and the root segments.
  const int64_t cd_elems = int64_t(clouds) * dim;
  L(k_init_points, dim3(grid_for(n)), dim3(kBlock), 0, n, clouds, sc.d_offsets.p, d_cloud, stride,
    dim, ps, sc.cid.p, recs, sc.pseg.p);
  L(k_init_keys, dim3(grid_for(cd_elems)), dim3(kBlock), 0, clouds, dim, sc.keys.p);
  L(k_bounds, dim3(tiles_for(n)), dim3(kBlock), 0, n, d_cloud, stride, dim, sc.cid.p, sc.keys.p,
    parts, nparts);
  L(k_tile_partials<0, 1>, dim3(gpart, unsigned(dim)), dim3(kBlock), 0, nparts, parts, sc.keys.p,
    sc.keys.p + dim, int64_t(2 * dim), int64_t(1), nullptr);
  L(k_roots, dim3(grid_for(cd_elems)), dim3(kBlock), 0, clouds, dim, sc.d_offsets.p, sc.keys.p,
    seg[0], sc.bounds.p, lvl);

  const unsigned gs = unsigned(std::min<int64_t>(kGridStride, grid_for(std::max<int64_t>(n, clouds))));
  const unsigned gp = grid_for(n), gt = tiles_for(n);
  LevelMailbox mailbox;
  int32_t S = clouds;  // eager: the host's copy of the level's segment count
  int64_t node_base = 0;
  int level = 0;
  for (;; ++level) {
    if (!eager && level == levels) break;
    BuildLevel* cur = lvl + (level & 1);
    BuildLevel* nxt = lvl + ((level + 1) & 1);
    const int32_t* gate = &nxt->S;
    const Seg sg = seg[level & 1], nx = seg[(level + 1) & 1];
    if (eager && node_base + S > 2 * n)
      throw Error(NABO_HIP_ERR_HIP, "kd-tree build overflow");
    L(k_seg_classify, dim3(gs), dim3(kBlock), 0, static_cast<const BuildLevel*>(cur), level, sg,
      st, bfs, sh.bucket, dim, sc.local_t, sc.local.p, ctr);
    exclusive_scan<uint32_t>(L, st.split, st.childbase, ScanCount{&cur->S, 0, nullptr},
                             sc.partials32.p, sc.splits.p);
    const uint32_t seq = eager ? mailbox.next() : 0u;
    L(k_level_next, dim3(1), dim3(64), 0, static_cast<const BuildLevel*>(cur), nxt,
      static_cast<const uint32_t*>(sc.splits.p), eager ? mailbox.box() : nullptr, seq);
    int32_t T = 1;
    if (eager) {
      T = mailbox.wait(seq, L.stream());
      node_base += S;
      S = 2 * T;
      if (T == 0) {
        ++level;
        break;
      }
    }
    L(k_pt_gather, dim3(gp), dim3(kBlock), 0, n, gate, sc.pseg.p,
      static_cast<const float*>(recs), ps, st, sc.xs.p);
    L(k_pt_minmax, dim3(gt), dim3(kBlock), 0, n, gate, static_cast<const int32_t*>(sc.pseg.p),
      static_cast<const float*>(sc.xs.p), st, parts);
    L(k_tile_partials<0, 1>, dim3(gpart), dim3(kBlock), 0, nparts, parts, st.minkey, st.maxkey,
      int64_t(1), int64_t(0), gate);
    L(k_seg_cut, dim3(gs), dim3(kBlock), 0, static_cast<const BuildLevel*>(cur), st);
    L(k_pt_count, dim3(gt), dim3(kBlock), 0, n, gate, static_cast<const int32_t*>(sc.pseg.p),
      static_cast<const float*>(sc.xs.p), st, parts);
    L(k_tile_partials<2, 2>, dim3(gpart), dim3(kBlock), 0, nparts, parts, st.lt, st.eq,
      int64_t(1), int64_t(0), gate);
    for (int pass = 1; pass <= 2; ++pass) {
      L(k_pt_flags, dim3(kScanGrid), dim3(kFlagThreads), 0, n, gate, pass,
        static_cast<const int32_t*>(sc.pseg.p), static_cast<const float*>(sc.xs.p), sg, st,
        sc.flags.p, sc.partials64.p);
      exclusive_scan_from_partials<uint64_t>(L, sc.flags.p, sc.excl.p,
                                             ScanCount{nullptr, n, gate}, sc.partials64.p,
                                             static_cast<uint64_t*>(nullptr));
      L(k_pt_slots, dim3(gp), dim3(kBlock), 0, n, gate, static_cast<const int32_t*>(sc.pseg.p),
        sg, static_cast<const uint64_t*>(sc.flags.p), static_cast<const uint64_t*>(sc.excl.p),
        sc.slot_r.p);
      L(k_pt_swap, dim3(gp), dim3(kBlock), 0, n, gate, static_cast<const int32_t*>(sc.pseg.p),
        sg, static_cast<const uint64_t*>(sc.flags.p), static_cast<const uint64_t*>(sc.excl.p),
        static_cast<const int32_t*>(sc.slot_r.p), recs, sc.xs.p, ps);
    }
    L(k_seg_children, dim3(gs), dim3(kBlock), 0, static_cast<const BuildLevel*>(cur), sg, st,
      bfs, nx, dim);
    L(k_pt_reseg, dim3(gp), dim3(kBlock), 0, n, gate, sc.pseg.p, sg, st);
  }
  const int levels_run = level;
  BuildLevel* fin = lvl + (levels_run & 1);

  // The segments left to LDS, then the node count and depth.
  if (sc.local_t > 0) {
    int dev = 0, cus = 256;
    NH_CHECK(hipGetDevice(&dev));
    NH_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const unsigned g = unsigned(std::min<int64_t>(int64_t(sc.local.count),
                                                  int64_t(std::max(1, cus)) * kLocalPerCu));
    if (dim == 3)
      L(k_local_finish<3>, dim3(g), dim3(kLocalBlock), 0, static_cast<const LocalSeg*>(sc.local.p),
        ctr, reinterpret_cast<float4*>(recs), bfs, r0, sh.bucket);
    else if (dim == 2)
      L(k_local_finish<2>, dim3(g), dim3(kLocalBlock), 0, static_cast<const LocalSeg*>(sc.local.p),
        ctr, reinterpret_cast<float4*>(recs), bfs, r0, sh.bucket);
    else
      L(k_local_finish<1>, dim3(g), dim3(kLocalBlock), 0, static_cast<const LocalSeg*>(sc.local.p),
        ctr, reinterpret_cast<float4*>(recs), bfs, r0, sh.bucket);
  }
  {
    const uint32_t seq = eager ? mailbox.next() : 0u;
    L(k_build_totals, dim3(1), dim3(64), 0, static_cast<const BuildLevel*>(fin),
      static_cast<const BuildCounters*>(ctr), tot, eager ? mailbox.box() : nullptr, seq);
    if (eager) {
      node_base = mailbox.wait(seq, L.stream());
      if (levels_out) *levels_out = mailbox.extra() + 1;
    }
  }
  if (!eager && levels_out) *levels_out = levels;
  if (loop_out) *loop_out = levels_run;
  if (nodes_out) *nodes_out = eager ? node_base : -1;

  // Pre-order ids from the leaf-start scan, then the node words and links.
  if (eager) {
    NH_CHECK(hipMalloc(reinterpret_cast<void**>(&out.nodes), sizeof(uint2) * (node_base + 1)));
    NH_CHECK(hipMalloc(reinterpret_cast<void**>(&out.links), sizeof(uint4) * node_base));
  }
  const int32_t* ok = out.ok;
  if (out.ok)
    L(k_build_ok, dim3(1), dim3(64), 0, static_cast<const BuildLevel*>(fin),
      static_cast<const BuildLevel*>(tot), int32_t(max_depth), out.ok);
  const unsigned gn = unsigned(std::min<int64_t>(kGridStride, grid_for(2 * n)));
  L(k_fill_u32, dim3(gn), dim3(kBlock), 0, sc.leaf_start.p, n + 1, 0u, ok);
  L(k_mark_leaves, dim3(gn), dim3(kBlock), 0, static_cast<const BuildLevel*>(fin), bfs, dim, ids,
    sc.leaf_start.p, ok);
  exclusive_scan<uint32_t>(L, sc.leaf_start.p, sc.leaves_before.p, ScanCount{nullptr, n + 1, ok},
                           sc.partials32.p, static_cast<uint32_t*>(nullptr));
  L(k_preorder, dim3(gn), dim3(kBlock), 0, static_cast<const BuildLevel*>(fin), bfs, ids,
    static_cast<const uint32_t*>(sc.leaves_before.p), static_cast<const int32_t*>(sc.cid.p), ok);
  L(k_encode, dim3(gn), dim3(kBlock), 0, static_cast<const BuildLevel*>(fin), bfs, ids,
    sh.dim_bits, dim, out.nodes, ok);
  L(k_links_self, dim3(gn), dim3(kBlock), 0, static_cast<const BuildLevel*>(tot),
    static_cast<const uint2*>(out.nodes), sh.dim_bits, sh.dim_mask, dim, out.links, out.nodes, ok);
  L(k_links_parent, dim3(gn), dim3(kBlock), 0, static_cast<const BuildLevel*>(tot),
    static_cast<const uint2*>(out.nodes), sh.dim_bits, sh.dim_mask, dim, out.links, ok);
  L(k_pad_points, dim3(grid_for(int64_t(kPtsPad) * ps)), dim3(kBlock), 0, n, ps, out.pts);
  if (out.roots)
    L(k_roots_out, dim3(grid_for(clouds)), dim3(kBlock), 0, clouds,
      static_cast<const int64_t*>(sc.d_offsets.p),
      static_cast<const uint32_t*>(sc.leaves_before.p), out.roots, ok);
  if (out.boxes)
    L(k_boxes, dim3(grid_for(int64_t(clouds) * 3)), dim3(kBlock), 0, clouds, dim,
      static_cast<const int64_t*>(sc.d_offsets.p), sh.bucket,
      static_cast<const float*>(sc.bounds.p), out.boxes);
}

}  // namespace

size_t kd_build_record_bytes(const KdBuildShape& sh) {
  Scratch sc(nullptr);
  BuildScratch b;
  b.want(sc, sh.n, sh.dim, sh.pts_stride, sh.clouds, sh.bucket);
  return sc.bytes();
}

void kd_build_record(const KdBuildShape& sh, const float* d_cloud, int64_t stride,
                     const int64_t* h_offsets, void* scratch, size_t scratch_bytes,
                     KdBuildOut& out, Launcher& L, int levels, int max_depth) {
  if (levels < 1) throw Error(NABO_HIP_ERR_ARGUMENT, "a recorded build needs at least one level");
  Scratch sc(nullptr, scratch, scratch_bytes);
  BuildScratch b;
  b.want(sc, sh.n, sh.dim, sh.pts_stride, sh.clouds, sh.bucket);
  sc.commit(0);
  NH_CHECK(hipMemcpy(b.d_offsets.p, h_offsets, sizeof(int64_t) * (size_t(sh.clouds) + 1),
                     hipMemcpyHostToDevice));
  kd_build_sequence(sh, d_cloud, stride, b, out, L, levels, max_depth, nullptr, nullptr, nullptr);
}

void kd_build(Index& ix, const float* d_cloud, int64_t stride,
              const std::vector<int64_t>& offsets, hipStream_t s) {
  const int64_t n = ix.n;
  const int dim = ix.dim;
  const int32_t clouds = int32_t(offsets.size() - 1);
  ix.clouds = clouds;
  ix.cloud_offsets = offsets;
  ix.pts_stride = int((dim + 1 + 3) / 4 * 4);
  const KdBuildShape sh{n, dim, ix.pts_stride, clouds, ix.bucket_size, ix.dim_bits, ix.dim_mask};

  NH_CHECK(hipMalloc(reinterpret_cast<void**>(&ix.d_pts),
                     sizeof(float) * size_t(n + kPtsPad) * ix.pts_stride));
  Scratch scratch(s);
  BuildScratch sc;
  sc.want(scratch, n, dim, ix.pts_stride, clouds, ix.bucket_size);
  int dev = 0;
  NH_CHECK(hipGetDevice(&dev));
  scratch.commit(dev);
  NH_CHECK(hipMemcpyAsync(sc.d_offsets.p, offsets.data(), sizeof(int64_t) * offsets.size(),
                          hipMemcpyHostToDevice, s));
  KdBuildOut out{};
  out.pts = ix.d_pts;
  if (clouds > 1) {
    NH_CHECK(hipMalloc(reinterpret_cast<void**>(&ix.d_roots), sizeof(int32_t) * clouds));
    out.roots = ix.d_roots;
    if (dim == 2 || dim == 3) {
      NH_CHECK(hipMalloc(reinterpret_cast<void**>(&ix.d_boxes), sizeof(float) * 6 * clouds));
      out.boxes = ix.d_boxes;
    }
  }
  Launcher L(s);
  int levels = 0, loop = 0;
  int64_t count = 0;
  kd_build_sequence(sh, d_cloud, stride, sc, out, L, 0, 0, &levels, &loop, &count);
  ix.d_nodes = out.nodes;
  ix.d_links = out.links;
  std::vector<float> hb(size_t(2) * clouds * dim);
  NH_CHECK(hipMemcpyAsync(hb.data(), sc.bounds.p, sizeof(float) * hb.size(), hipMemcpyDeviceToHost,
                          s));
  NH_CHECK(hipStreamSynchronize(s));
  // A cloud of at most bucketSize points is one bucket and keeps the initial
  // bounds (kdtree_cpu.cpp:236-243).
  ix.min_bound.assign(size_t(clouds) * dim, FLT_MAX);
  ix.max_bound.assign(size_t(clouds) * dim, FLT_MIN);
  for (int32_t c = 0; c < clouds; ++c) {
    if (offsets[c + 1] - offsets[c] <= int64_t(ix.bucket_size)) continue;
    for (int d = 0; d < dim; ++d) {
      ix.min_bound[size_t(c) * dim + d] = hb[size_t(c) * 2 * dim + d];
      ix.max_bound[size_t(c) * dim + d] = hb[size_t(c) * 2 * dim + dim + d];
    }
  }
  ix.node_count = count;
  ix.depth = levels - 1;
  ix.build_levels = loop;
}

void kd_make_boxes(Index& ix, hipStream_t s) {
  if (ix.clouds <= 1 || (ix.dim != 2 && ix.dim != 3) || ix.d_boxes) return;
  std::vector<float> hb(size_t(ix.clouds) * 6, 0.f);
  for (int32_t c = 0; c < ix.clouds; ++c)
    for (int d = 0; d < ix.dim; ++d) {
      const float lo = ix.min_bound[size_t(c) * ix.dim + d];
      const float ext = ix.max_bound[size_t(c) * ix.dim + d] - lo;
      hb[size_t(c) * 6 + d] = lo;
      hb[size_t(c) * 6 + 3 + d] = ext > 0.f ? 1.f / ext : 0.f;
    }
  NH_CHECK(hipMalloc(reinterpret_cast<void**>(&ix.d_boxes), sizeof(float) * hb.size()));
  NH_CHECK(hipMemcpyAsync(ix.d_boxes, hb.data(), sizeof(float) * hb.size(),
                          hipMemcpyHostToDevice, s));
  NH_CHECK(hipStreamSynchronize(s));
}

}  // namespace nabo_hip
