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
// Nodes are recorded in BFS order; subtree sizes (bottom-up) then give the
// reference's pre-order numbering (left child = parent + 1, :215-223).
// Bucket starts are the positions of the leaf segments in the final
// permutation, which is the reference's bucket vector order.

#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <mutex>
#include <vector>

#include "device_common.h"
#include "internal.h"

namespace nabo_hip {
namespace {

constexpr int kBlock = 256;

struct Seg {  // one level of segments
  int32_t* start;
  int32_t* count;
  float* lo;  // S x dim
  float* hi;  // S x dim
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
  uint32_t* size;   // nodes in the subtree
  uint32_t* pre;    // pre-order id
};

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
                                                          int64_t stride, int64_t gstride) {
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

// Level 0: one root segment per cloud, its cell = the cloud's bounds.
__global__ void k_roots(int32_t clouds, int dim, const int64_t* __restrict__ offsets,
                        const uint32_t* __restrict__ keys, Seg seg, float* __restrict__ bounds) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
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
  }
}

// Cloud id of every point (binary search in the offsets), identity order,
// every point in its cloud's root segment.
__global__ void k_init_points(int64_t n, int32_t clouds, const int64_t* __restrict__ offsets,
                              int32_t* __restrict__ cid, int32_t* __restrict__ perm,
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
  perm[p] = int32_t(p);
  pseg[p] = lo;
}

__global__ void k_seg_classify(int32_t S, Seg seg, SegStat st, Bfs bfs, int64_t node_base,
                               uint32_t bucket, int dim) {
  const int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  const int32_t cnt = seg.count[s];
  const int64_t node = node_base + s;
  if (cnt <= int32_t(bucket)) {
    st.split[s] = 0;
    bfs.cd[node] = uint32_t(dim);
    bfs.aux[node] = uint32_t(seg.start[s]);
    bfs.count[node] = uint32_t(cnt);
    return;
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

// Number of segments that split at this level, posted to the host: the
// count, a system-scope fence, then the sequence number the host waits for.
struct Mailbox {
  int32_t total;
  uint32_t seq;
};
__global__ void k_split_total(int32_t S, const uint32_t* split, const uint32_t* childbase,
                              Mailbox* box, uint32_t seq) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const int32_t total = int32_t(childbase[S - 1] + split[S - 1]);
    __hip_atomic_store(&box->total, total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __threadfence_system();
    __hip_atomic_store(&box->seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// Gathers the cut coordinate of every point still inside a splitting segment
// and reduces its min/max per segment.
__global__ __launch_bounds__(kBlock) void k_pt_gather(int64_t n, int32_t* __restrict__ pseg,
                                                      const int32_t* __restrict__ perm,
                                                      const float* __restrict__ cloud,
                                                      int64_t stride, SegStat st,
                                                      float* __restrict__ xs, TileParts parts) {
  const int64_t tile = blockIdx.x;
  const int64_t base = tile * kTile + int64_t(threadIdx.x) * kItems;
  int32_t sg[kItems];
  uint32_t kv[kItems];
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const int64_t p = base + j;
    int s = p < n ? pseg[p] : -1;
    kv[j] = 0u;
    if (s >= 0 && st.split[s] == 0) {
      pseg[p] = -1;  // its segment became a leaf
      s = -1;
    }
    if (s >= 0) {
      const float x = cloud[int64_t(perm[p]) * stride + st.cd[s]];
      xs[p] = x;
      kv[j] = float_key(x);
    }
    sg[j] = s;
  }
  tile_reduce<0, 1>(sg, kv, kv, st.minkey, st.maxkey, 1, &parts, tile);
}

__global__ void k_seg_cut(int32_t S, SegStat st) {
  const int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S || st.split[s] == 0) return;
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

__global__ __launch_bounds__(kBlock) void k_pt_count(int64_t n, const int32_t* __restrict__ pseg,
                                                     const float* __restrict__ xs, SegStat st,
                                                     TileParts parts) {
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
__global__ void k_pt_flags(int64_t n, int pass, const int32_t* __restrict__ pseg,
                           const float* __restrict__ xs, Seg seg, SegStat st,
                           uint64_t* __restrict__ flags) {
  const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= n) return;
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
  flags[p] = f;
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

__global__ void k_pt_slots(int64_t n, const int32_t* __restrict__ pseg, Seg seg,
                           const uint64_t* __restrict__ flags, const uint64_t* __restrict__ excl,
                           int32_t* __restrict__ slot_l, int32_t* __restrict__ slot_r) {
  const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint64_t f = flags[p];
  if (f == 0) return;
  const int s = pseg[p];
  int32_t start;
  uint32_t rl, rr;
  pass_ranks(p, s, seg, flags, excl, f, &start, &rl, &rr);
  if (uint32_t(f))
    slot_l[start + rl] = int32_t(p);
  else
    slot_r[start + rr] = int32_t(p);
}

__global__ void k_pt_move(int64_t n, const int32_t* __restrict__ pseg, Seg seg,
                          const uint64_t* __restrict__ flags, const uint64_t* __restrict__ excl,
                          const int32_t* __restrict__ slot_l, const int32_t* __restrict__ slot_r,
                          const int32_t* __restrict__ perm, const float* __restrict__ xs,
                          int32_t* __restrict__ perm_out, float* __restrict__ xs_out) {
  const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint64_t f = flags[p];
  int64_t dst = p;
  if (f) {
    const int s = pseg[p];
    int32_t start;
    uint32_t rl, rr;
    pass_ranks(p, s, seg, flags, excl, f, &start, &rl, &rr);
    dst = uint32_t(f) ? slot_r[start + rl] : slot_l[start + rr];
  }
  perm_out[dst] = perm[p];
  xs_out[dst] = xs[p];
}

__global__ void k_seg_children(int32_t S, Seg seg, SegStat st, Bfs bfs, int64_t node_base,
                               int64_t next_base, Seg next, int dim) {
  const int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S || st.split[s] == 0) return;
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

__global__ void k_pt_reseg(int64_t n, int32_t* __restrict__ pseg, Seg seg, SegStat st) {
  const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int s = pseg[p];
  if (s < 0) return;
  const int32_t pos = int32_t(p) - seg.start[s];
  pseg[p] = int32_t(2 * st.childbase[s]) + (pos < st.left[s] ? 0 : 1);
}

__global__ void k_subtree_size(int64_t base, int32_t S, Bfs bfs, int dim) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= S) return;
  const int64_t node = base + i;
  if (bfs.cd[node] == uint32_t(dim))
    bfs.size[node] = 1;
  else
    bfs.size[node] = 1 + bfs.size[bfs.left[node]] + bfs.size[bfs.right[node]];
}

// Pre-order ids: a root takes its tree's base (trees are stored one after
// the other), then left = parent + 1 and right = parent + 1 + |left subtree|.
__global__ void k_preorder(int64_t base, int32_t S, Bfs bfs, int dim,
                           const uint32_t* __restrict__ root_base) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= S) return;
  const int64_t node = base + i;
  if (root_base) bfs.pre[node] = root_base[node];
  if (bfs.cd[node] == uint32_t(dim)) return;
  const uint32_t pre = bfs.pre[node];
  const int32_t l = bfs.left[node];
  bfs.pre[l] = pre + 1;
  bfs.pre[bfs.right[node]] = pre + 1 + bfs.size[l];
}

__global__ void k_encode(int64_t count, Bfs bfs, uint32_t dim_bits, int dim, uint2* nodes) {
  const int64_t node = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (node >= count) return;
  const uint32_t cd = bfs.cd[node];
  uint2 out;
  if (cd == uint32_t(dim)) {
    out.x = uint32_t(dim) | (bfs.count[node] << dim_bits);
    out.y = bfs.aux[node];
  } else {
    out.x = cd | (bfs.pre[bfs.right[node]] << dim_bits);
    out.y = bfs.aux[node];
  }
  nodes[bfs.pre[node]] = out;
}

// Leaf-ordered copy of the points: pts_stride floats per entry, coordinates
// then the point's index within its own cloud (as bits) in slot dim.
__global__ void k_pack_points(int64_t n, const int32_t* __restrict__ perm,
                              const int32_t* __restrict__ cid, const int64_t* __restrict__ offsets,
                              const float* __restrict__ cloud, int64_t stride, int dim,
                              int pts_stride, float* __restrict__ pts) {
  const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= n + kPtsPad) return;
  if (p >= n) {
    for (int d = 0; d < pts_stride; ++d) pts[p * pts_stride + d] = 0.f;
    return;
  }
  const int32_t idx = perm[p];
  float* dst = pts + p * pts_stride;
  const float* src = cloud + int64_t(idx) * stride;
  for (int d = 0; d < dim; ++d) dst[d] = src[d];
  dst[dim] = __int_as_float(int32_t(idx - offsets[cid[idx]]));
  for (int d = dim + 1; d < pts_stride; ++d) dst[d] = 0.f;
}

// Search-side links, by pre-order node p: links[p].xy is the node word of
// p's right child (the left child is node p + 1), so one round trip reads a
// node and both its children; links[c].zw is {cut dim, cut bits} of the
// split of c's parent, so a far child is entered without reading its parent
// again. nodes[count] is a padding leaf, so nodes[p + 1] exists for every p.
__global__ void k_links_self(int64_t count, const uint2* __restrict__ nodes, uint32_t dim_bits,
                             uint32_t dim_mask, int dim, uint4* __restrict__ links,
                             uint2* __restrict__ pad) {
  const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= count) return;
  const uint2 nd = nodes[p];
  uint4 l = make_uint4(0u, 0u, 0u, 0u);
  if ((nd.x & dim_mask) != uint32_t(dim)) {
    const uint2 r = nodes[nd.x >> dim_bits];
    l.x = r.x;
    l.y = r.y;
  }
  links[p] = l;
  if (p == 0) *pad = make_uint2(uint32_t(dim), 0u);
}

__global__ void k_links_parent(int64_t count, const uint2* __restrict__ nodes, uint32_t dim_bits,
                               uint32_t dim_mask, int dim, uint4* __restrict__ links) {
  const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= count) return;
  const uint2 nd = nodes[p];
  const uint32_t cd = nd.x & dim_mask;
  if (cd == uint32_t(dim)) return;
  const uint2 split = make_uint2(cd, nd.y);
  reinterpret_cast<uint2*>(links + p + 1)[1] = split;
  reinterpret_cast<uint2*>(links + (nd.x >> dim_bits))[1] = split;
}

__global__ void k_roots_out(int32_t clouds, const uint32_t* __restrict__ root_base,
                            int32_t* __restrict__ roots) {
  const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < clouds) roots[c] = int32_t(root_base[c]);
}

inline unsigned grid_for(int64_t n) { return unsigned(std::max<int64_t>(1, ceil_div(n, kBlock))); }

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

inline unsigned tiles_for(int64_t n) { return unsigned(std::max<int64_t>(1, ceil_div(n, kTile))); }

}  // namespace

void kd_build(Index& ix, const float* d_cloud, int64_t stride,
              const std::vector<int64_t>& offsets, hipStream_t s) {
  const int64_t n = ix.n;
  const int dim = ix.dim;
  const int32_t clouds = int32_t(offsets.size() - 1);
  ix.clouds = clouds;
  ix.cloud_offsets = offsets;
  ix.pts_stride = int((dim + 1 + 3) / 4 * 4);

  NH_CHECK(hipMalloc(reinterpret_cast<void**>(&ix.d_pts),
                     sizeof(float) * size_t(n + kPtsPad) * ix.pts_stride));
  float* pts = ix.d_pts;

  const int64_t max_nodes = 2 * n;
  // Level buffers (double-buffered), per-segment stats, BFS node records,
  // per-point arrays: one scratch block for the whole build.
  Scratch scratch(s);
  Span<int32_t> seg_start[2], seg_count[2];
  Span<float> seg_lo[2], seg_hi[2];
  const int64_t seg_cap = std::max<int64_t>(n, clouds);
  for (int b = 0; b < 2; ++b) {
    scratch.want(seg_start[b], size_t(seg_cap));
    scratch.want(seg_count[b], size_t(seg_cap));
    scratch.want(seg_lo[b], size_t(seg_cap) * dim);
    scratch.want(seg_hi[b], size_t(seg_cap) * dim);
  }
  Span<uint32_t> st_cd, st_minkey, st_maxkey, st_lt, st_eq, st_split, st_childbase;
  Span<float> st_ideal, st_cut;
  Span<int32_t> st_left;
  for (Span<uint32_t>* a : {&st_cd, &st_minkey, &st_maxkey, &st_lt, &st_eq, &st_split,
                            &st_childbase})
    scratch.want(*a, size_t(seg_cap));
  scratch.want(st_ideal, size_t(seg_cap));
  scratch.want(st_cut, size_t(seg_cap));
  scratch.want(st_left, size_t(seg_cap));
  Span<uint32_t> b_cd, b_aux, b_count, b_size, b_pre;
  Span<int32_t> b_left, b_right;
  for (Span<uint32_t>* a : {&b_cd, &b_aux, &b_count, &b_size, &b_pre})
    scratch.want(*a, size_t(max_nodes));
  scratch.want(b_left, size_t(max_nodes));
  scratch.want(b_right, size_t(max_nodes));

  Span<int32_t> perm, perm2, pseg, slot_l, slot_r, cid;
  for (Span<int32_t>* a : {&perm, &perm2, &pseg, &slot_l, &slot_r, &cid})
    scratch.want(*a, size_t(n));
  Span<float> xs, xs2;
  scratch.want(xs, size_t(n));
  scratch.want(xs2, size_t(n));
  Span<uint64_t> flags, excl;
  scratch.want(flags, size_t(n));
  scratch.want(excl, size_t(n));
  Span<uint32_t> keys, root_base;
  scratch.want(keys, size_t(2) * clouds * dim);
  scratch.want(root_base, size_t(clouds));
  Span<float> bounds;
  scratch.want(bounds, size_t(2) * clouds * dim);
  Span<int64_t> d_offsets;
  scratch.want(d_offsets, offsets.size());

  // Scan scratch, sized for the largest scan (n uint64).
  size_t scan_bytes64 = 0, scan_bytes32 = 0;
  NH_CHECK(rocprim::exclusive_scan(nullptr, scan_bytes64, flags.p, excl.p, uint64_t(0), size_t(n),
                                   rocprim::plus<uint64_t>(), s));
  NH_CHECK(rocprim::exclusive_scan(nullptr, scan_bytes32, st_split.p, st_childbase.p, uint32_t(0),
                                   size_t(seg_cap), rocprim::plus<uint32_t>(), s));
  Span<uint8_t> scan_tmp;
  scratch.want(scan_tmp, std::max(scan_bytes64, scan_bytes32) + 16);
  // Tile partials of the reductions: 2 per tile, dim groups for the bounds.
  const int64_t nparts = 2 * int64_t(tiles_for(n));
  Span<int32_t> part_seg;
  Span<uint32_t> part_a, part_b;
  scratch.want(part_seg, size_t(nparts) * dim);
  scratch.want(part_a, size_t(nparts) * dim);
  scratch.want(part_b, size_t(nparts) * dim);
  int dev = 0;
  NH_CHECK(hipGetDevice(&dev));
  scratch.commit(dev);
  NH_CHECK(hipMemcpyAsync(d_offsets.p, offsets.data(), sizeof(int64_t) * offsets.size(),
                          hipMemcpyHostToDevice, s));
  const TileParts parts{part_seg.p, part_a.p, part_b.p};
  const unsigned gpart = unsigned(ceil_div(nparts, kTile));

  SegStat st{st_cd.p, st_ideal.p, st_cut.p, st_minkey.p, st_maxkey.p, st_lt.p, st_eq.p,
             st_split.p, st_childbase.p, st_left.p};
  Bfs bfs{b_cd.p, b_aux.p, b_count.p, b_left.p, b_right.p, b_size.p, b_pre.p};
  auto level = [&](int b) {
    return Seg{seg_start[b].p, seg_count[b].p, seg_lo[b].p, seg_hi[b].p};
  };

  // Bounds pass (kdtree_cpu.cpp:256-267) and the root segments.
  const int64_t cd_elems = int64_t(clouds) * dim;
  k_init_points<<<grid_for(n), kBlock, 0, s>>>(n, clouds, d_offsets.p, cid.p, perm.p, pseg.p);
  k_init_keys<<<grid_for(cd_elems), kBlock, 0, s>>>(clouds, dim, keys.p);
  k_bounds<<<tiles_for(n), kBlock, 0, s>>>(n, d_cloud, stride, dim, cid.p, keys.p, parts, nparts);
  k_tile_partials<0, 1><<<dim3(gpart, unsigned(dim)), kBlock, 0, s>>>(nparts, parts, keys.p,
                                                                      keys.p + dim, 2 * dim, 1);
  k_roots<<<grid_for(cd_elems), kBlock, 0, s>>>(clouds, dim, d_offsets.p, keys.p, level(0),
                                                 bounds.p);
  NH_LAUNCH_CHECK();
  std::vector<float> hb(size_t(2) * clouds * dim);
  NH_CHECK(hipMemcpyAsync(hb.data(), bounds.p, sizeof(float) * hb.size(), hipMemcpyDeviceToHost, s));

  std::vector<int64_t> level_base;
  std::vector<int32_t> level_size;
  int32_t S = clouds;
  int64_t node_base = 0;
  int cur = 0;
  LevelMailbox mailbox;
  while (S > 0) {
    if (node_base + S > max_nodes || int64_t(level_base.size()) > n)
      throw Error(NABO_HIP_ERR_HIP, "kd-tree build overflow");
    level_base.push_back(node_base);
    level_size.push_back(S);
    Seg sg = level(cur);
    Seg nx = level(cur ^ 1);
    const unsigned gs = grid_for(S);
    k_seg_classify<<<gs, kBlock, 0, s>>>(S, sg, st, bfs, node_base, ix.bucket_size, dim);
    size_t bytes = scan_tmp.count;
    NH_CHECK(rocprim::exclusive_scan(scan_tmp.p, bytes, st.split, st.childbase, uint32_t(0),
                                     size_t(S), rocprim::plus<uint32_t>(), s));
    const uint32_t seq = mailbox.next();
    k_split_total<<<1, 64, 0, s>>>(S, st.split, st.childbase, mailbox.box(), seq);
    NH_LAUNCH_CHECK();
    const int32_t T = mailbox.wait(seq, s);
    if (T > 0) {
      const unsigned gp = grid_for(n);
      const unsigned gt = tiles_for(n);
      k_pt_gather<<<gt, kBlock, 0, s>>>(n, pseg.p, perm.p, d_cloud, stride, st, xs.p, parts);
      k_tile_partials<0, 1><<<gpart, kBlock, 0, s>>>(nparts, parts, st.minkey, st.maxkey, 1, 0);
      k_seg_cut<<<gs, kBlock, 0, s>>>(S, st);
      k_pt_count<<<gt, kBlock, 0, s>>>(n, pseg.p, xs.p, st, parts);
      k_tile_partials<2, 2><<<gpart, kBlock, 0, s>>>(nparts, parts, st.lt, st.eq, 1, 0);
      for (int pass = 1; pass <= 2; ++pass) {
        k_pt_flags<<<gp, kBlock, 0, s>>>(n, pass, pseg.p, xs.p, sg, st, flags.p);
        bytes = scan_tmp.count;
        NH_CHECK(rocprim::exclusive_scan(scan_tmp.p, bytes, flags.p, excl.p, uint64_t(0),
                                         size_t(n), rocprim::plus<uint64_t>(), s));
        k_pt_slots<<<gp, kBlock, 0, s>>>(n, pseg.p, sg, flags.p, excl.p, slot_l.p, slot_r.p);
        k_pt_move<<<gp, kBlock, 0, s>>>(n, pseg.p, sg, flags.p, excl.p, slot_l.p, slot_r.p,
                                        perm.p, xs.p, perm2.p, xs2.p);
        std::swap(perm.p, perm2.p);
        std::swap(xs.p, xs2.p);
      }
      k_seg_children<<<gs, kBlock, 0, s>>>(S, sg, st, bfs, node_base, node_base + S, nx, dim);
      k_pt_reseg<<<gp, kBlock, 0, s>>>(n, pseg.p, sg, st);
      NH_LAUNCH_CHECK();
    }
    node_base += S;
    S = 2 * T;
    cur ^= 1;
  }
  const int64_t count = node_base;
  const int levels = int(level_base.size());

  for (int L = levels - 1; L >= 0; --L)
    k_subtree_size<<<grid_for(level_size[L]), kBlock, 0, s>>>(level_base[L], level_size[L], bfs,
                                                              dim);
  // Trees one after the other: the root of cloud c (BFS id c) starts at the
  // sum of the sizes of the trees before it.
  size_t bytes = scan_tmp.count;
  NH_CHECK(rocprim::exclusive_scan(scan_tmp.p, bytes, bfs.size, root_base.p, uint32_t(0),
                                   size_t(clouds), rocprim::plus<uint32_t>(), s));
  for (int L = 0; L < levels; ++L)
    k_preorder<<<grid_for(level_size[L]), kBlock, 0, s>>>(level_base[L], level_size[L], bfs, dim,
                                                          L == 0 ? root_base.p : nullptr);
  NH_CHECK(hipMalloc(reinterpret_cast<void**>(&ix.d_nodes), sizeof(uint2) * (count + 1)));
  NH_CHECK(hipMalloc(reinterpret_cast<void**>(&ix.d_links), sizeof(uint4) * count));
  k_encode<<<grid_for(count), kBlock, 0, s>>>(count, bfs, ix.dim_bits, dim, ix.d_nodes);
  k_links_self<<<grid_for(count), kBlock, 0, s>>>(count, ix.d_nodes, ix.dim_bits, ix.dim_mask,
                                                  dim, ix.d_links, ix.d_nodes + count);
  k_links_parent<<<grid_for(count), kBlock, 0, s>>>(count, ix.d_nodes, ix.dim_bits, ix.dim_mask,
                                                    dim, ix.d_links);
  k_pack_points<<<grid_for(n + kPtsPad), kBlock, 0, s>>>(n, perm.p, cid.p, d_offsets.p, d_cloud, stride,
                                                dim, ix.pts_stride, pts);
  if (clouds > 1) {
    NH_CHECK(hipMalloc(reinterpret_cast<void**>(&ix.d_roots), sizeof(int32_t) * clouds));
    k_roots_out<<<grid_for(clouds), kBlock, 0, s>>>(clouds, root_base.p, ix.d_roots);
  }
  NH_LAUNCH_CHECK();
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
  kd_make_boxes(ix, s);
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
