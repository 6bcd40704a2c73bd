// Brute-force kNN on the GPU (BruteForceSearch, nabo/brute_force_cpu.cpp:63-103).
//
// FP32 path: every (query, point) squared distance is the sequential
// sum of squared differences in point index order of dimensions, each
// product and sum rounded on its own. Points are visited in index order with
// the reference's admission rule (dist <= maxRadius2, dist < head, and
// dist > FLT_EPSILON unless self match is allowed), so among equal distances
// the lowest index is kept. Results are always in ascending distance (then
// index) order.
//
// Work split: a block holds 256 queries (one per lane, query in VGPRs) and one
// contiguous chunk of the points, staged through LDS in tiles; each lane
// keeps a sorted k-best of its chunk; a merge kernel then combines the chunks
// of a query in chunk order, which preserves the lowest-index tie rule.
//
// IndexHeapSTL form (the reference's own heap, brute_force_cpu.cpp:79): the
// same chunk pass also logs each chunk's admissions; one lane per query then
// replays its logs in chunk order through a binary heap in LDS (StlHeap)
// with the reference's admission test, and queries whose logs overflowed
// are redone by one lane scanning the whole cloud in index order. Results
// are bit-identical to the reference's, including the order and which of
// equal distances are kept.

#include <atomic>
#include <cfloat>
#include <type_traits>

#include "device_common.h"
#include "internal.h"

namespace nabo_hip {
namespace {

constexpr int kBlock = 256;
constexpr int kTile = 64;
constexpr int kBfUnroll = 4;  // STL form: points whose distances are summed together
static_assert(kTile % kBfUnroll == 0, "a tile holds whole groups of points");

template <int K>
__device__ __forceinline__ void bf_insert(float (&hd)[K], int32_t (&hi)[K], float value,
                                          int32_t index) {
#pragma unroll
  for (int j = K - 1; j > 0; --j) {
    const bool shift = hd[j - 1] > value;
    const bool place = !shift && hd[j] > value;
    hd[j] = shift ? hd[j - 1] : (place ? value : hd[j]);
    hi[j] = shift ? hi[j - 1] : (place ? index : hi[j]);
  }
  const bool place0 = hd[0] > value;
  hd[0] = place0 ? value : hd[0];
  hi[0] = place0 ? index : hi[0];
}

// Sequential sum of squared differences over DIMP coordinates, where those
// from dim on are zero in both the point and the query: each adds +0 to a
// sum that is +0 or more, which leaves it unchanged, so the result is the
// sum over the first dim coordinates, bit for bit, without a branch per
// coordinate.
template <int DIMP>
__device__ __forceinline__ float padded_dist(const float* pp, const float (&q)[DIMP]) {
  float dist = 0.f;
#pragma unroll
  for (int d = 0; d < DIMP; ++d) {
    const float diff = pp[d] - q[d];
    dist = dist + diff * diff;
  }
  return dist;
}

// Tiles of kTile points in LDS, DIMP floats per point, zero from dim on
// (the zeros are written once; later tiles only overwrite the first dim
// floats of each row). A tile is read from the packed cloud (dim floats per
// point, src 16-byte aligned) with 16-byte loads issued together (the first
// round before the barrier, so they are in flight while other waves finish
// the previous tile); the barrier keeps the previous tile intact until
// every thread is done with it.
template <int DIMP, int BLOCK>
__device__ __forceinline__ void zero_tile(float4* tile4) {
  for (int e = threadIdx.x; e < kTile * DIMP / 4; e += BLOCK) tile4[e] = make_float4(0.f, 0.f, 0.f, 0.f);
}
template <int DIMP, int BLOCK>
__device__ __forceinline__ void load_round(float4 (&v)[(kTile * DIMP / 4 + BLOCK - 1) / BLOCK <= 12
                                                          ? (kTile * DIMP / 4 + BLOCK - 1) / BLOCK
                                                          : 8],
                                           const float* src, int floats, int r0) {
  constexpr int kRound = sizeof(v) / sizeof(v[0]);
#pragma unroll
  for (int j = 0; j < kRound; ++j) {
    const int e = threadIdx.x + (r0 + j) * BLOCK;
    if (4 * e + 4 <= floats) {
      v[j] = reinterpret_cast<const float4*>(src)[e];
    } else {
      v[j] = make_float4(0.f, 0.f, 0.f, 0.f);
      if (4 * e < floats) v[j].x = src[4 * e];
      if (4 * e + 1 < floats) v[j].y = src[4 * e + 1];
      if (4 * e + 2 < floats) v[j].z = src[4 * e + 2];
    }
  }
}
template <int DIMP, int BLOCK, int N>
__device__ __forceinline__ void scatter_round(float* tile, const float4 (&v)[N], int floats,
                                              int dim, int r0) {
#pragma unroll
  for (int j = 0; j < N; ++j) {
    const int e = threadIdx.x + (r0 + j) * BLOCK;
    if (4 * e < floats) {
      int pt = (4 * e) / dim, d = 4 * e - pt * dim;
      const float c[4] = {v[j].x, v[j].y, v[j].z, v[j].w};
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        if (4 * e + i < floats) tile[pt * DIMP + d] = c[i];
        if (++d == dim) {
          d = 0;
          ++pt;
        }
      }
    }
  }
}
template <int DIMP, int BLOCK>
__device__ __forceinline__ void load_tile(float4* tile4, const float* src, int floats, int dim) {
  constexpr int kVec = (kTile * DIMP / 4 + BLOCK - 1) / BLOCK;  // float4 loads per thread
  constexpr int kRound = kVec <= 12 ? kVec : 8;                 // issued together
  float* tile = reinterpret_cast<float*>(tile4);
  float4 v[kRound];
  load_round<DIMP, BLOCK>(v, src, floats, 0);  // in flight while others finish the last tile
  __syncthreads();
  scatter_round<DIMP, BLOCK>(tile, v, floats, dim, 0);
#pragma unroll
  for (int r0 = kRound; r0 < kVec; r0 += kRound) {
    load_round<DIMP, BLOCK>(v, src, floats, r0);
    scatter_round<DIMP, BLOCK>(tile, v, floats, dim, r0);
  }
}

struct BfParams {
  const float* cloud;
  int64_t cstride;
  int dim;
  int64_t n;
  const float* query;
  int64_t qstride;
  int64_t m;
  int k;
  float max_r2;
  const float* radii;
  int allow_self;
  int64_t chunk;
  int chunks;
  float* part_d;   // chunks x m x k
  int32_t* part_i;
  // IndexHeapSTL in one pass: per (chunk, query) admission logs, slot-major
  // log[(chunk * log_cap + j) * m + query], their lengths log_n[chunk * m +
  // query] (may exceed log_cap: overflow), and the queries to redo by a full
  // scan (redo[query] != 0).
  uint2* log;
  uint32_t* log_n;
  int log_cap;
  int32_t* redo;
};

// DIMP: compile-time bound on dim (query held in VGPRs). LOG: also append
// every admission of the chunk's own k-best (index, dist) to the query's
// log, instead of writing the chunk's k-best.
template <int DIMP, int K, bool LOG>
__global__ __launch_bounds__(kBlock) void k_bf_chunk(BfParams a) {
  __shared__ float4 tile4[kTile * DIMP / 4];
  const float* tile = reinterpret_cast<const float*>(tile4);
  const int64_t qi = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  const int chunk_id = blockIdx.y;
  const int64_t p0 = int64_t(chunk_id) * a.chunk;
  const int64_t p1 = min(a.n, p0 + a.chunk);
  const int dim = a.dim;
  const bool valid = qi < a.m;

  float q[DIMP];
#pragma unroll
  for (int d = 0; d < DIMP; ++d) q[d] = (valid && d < dim) ? a.query[qi * a.qstride + d] : 0.f;
  float r2 = a.max_r2;
  if (valid && a.radii) {
    const float r = a.radii[qi];
    r2 = r * r;
  }
  float hd[K];
  int32_t hi[K];
#pragma unroll
  for (int j = 0; j < K; ++j) {
    hd[j] = j < K - a.k ? -INFINITY : INFINITY;
    hi[j] = -1;
  }

  uint32_t nlog = 0;
  zero_tile<DIMP, kBlock>(tile4);
  for (int64_t t0 = p0; t0 < p1; t0 += kTile) {
    const int tn = int(min<int64_t>(kTile, p1 - t0));
    // p0 is a multiple of kTile (the host rounds the chunk size), so the
    // tile's first float is 16-byte aligned in the packed cloud
    load_tile<DIMP, kBlock>(tile4, a.cloud + t0 * dim, tn * dim, dim);
    __syncthreads();
    if (valid) {
      for (int pt = 0; pt < tn; ++pt) {
        const float dist = padded_dist<DIMP>(tile + pt * DIMP, q);
        if (dist <= r2 && dist < hd[K - 1] && (a.allow_self || dist > FLT_EPSILON)) {
          if constexpr (LOG) {
            if (nlog < uint32_t(a.log_cap))
              a.log[(int64_t(chunk_id) * a.log_cap + nlog) * a.m + qi] =
                  make_uint2(uint32_t(t0 + pt), __float_as_uint(dist));
            ++nlog;
          }
          bf_insert<K>(hd, hi, dist, int32_t(t0 + pt));
        }
      }
    }
  }
  if (!valid) return;
  if constexpr (LOG) {
    a.log_n[int64_t(chunk_id) * a.m + qi] = nlog;
    return;
  }
  const int64_t o = (int64_t(chunk_id) * a.m + qi) * a.k;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    if (j >= K - a.k) {
      a.part_d[o + j - (K - a.k)] = hd[j];
      a.part_i[o + j - (K - a.k)] = hi[j];
    }
  }
}

template <int K>
__global__ __launch_bounds__(kBlock) void k_bf_merge(BfParams a, int32_t* __restrict__ out_idx,
                                                     float* __restrict__ out_d2) {
  const int64_t qi = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (qi >= a.m) return;
  float hd[K];
  int32_t hi[K];
#pragma unroll
  for (int j = 0; j < K; ++j) {
    hd[j] = j < K - a.k ? -INFINITY : INFINITY;
    hi[j] = -1;
  }
  for (int c = 0; c < a.chunks; ++c) {
    const int64_t o = (int64_t(c) * a.m + qi) * a.k;
    for (int j = 0; j < a.k; ++j) {
      const float v = a.part_d[o + j];
      if (v < hd[K - 1]) bf_insert<K>(hd, hi, v, a.part_i[o + j]);
    }
  }
#pragma unroll
  for (int j = 0; j < K; ++j) {
    if (j >= K - a.k) {
      out_idx[qi * a.k + j - (K - a.k)] = hi[j];
      out_d2[qi * a.k + j - (K - a.k)] = hd[j];
    }
  }
}

constexpr int kStlBlock = 64;  // one wave per workgroup: m / 64 waves spread over the CUs
constexpr size_t kStlLogBytes = size_t(768) << 20;  // admission logs per batch of queries

template <int DIMP>
__global__ __launch_bounds__(kStlBlock) void k_bf_stl(BfParams a, int sort_results,
                                                      int32_t* __restrict__ out_idx,
                                                      float* __restrict__ out_d2) {
  // A tile is kTile points, DIMP floats each, zero from dim on. It is read
  // from the packed cloud (dim floats per point) with 16-byte loads.
  __shared__ float4 tile4[kTile * DIMP / 4];
  float* tile = reinterpret_cast<float*>(tile4);
  extern __shared__ uint2 heaps[];  // k slots x kStlBlock lanes
  const int64_t qi = int64_t(blockIdx.x) * kStlBlock + threadIdx.x;
  const int dim = a.dim;
  // with redo set, only the queries it marks (the rest are done)
  const bool valid = qi < a.m && (!a.redo || a.redo[qi] != 0);
  if (!__syncthreads_or(valid)) return;

  float q[DIMP];
#pragma unroll
  for (int d = 0; d < DIMP; ++d) q[d] = (valid && d < dim) ? a.query[qi * a.qstride + d] : 0.f;
  float r2 = a.max_r2;
  if (valid && a.radii) {
    const float r = a.radii[qi];
    r2 = r * r;
  }
  StlHeap<LdsHeapStore> heap;
  if (valid) heap.reset(LdsHeapStore{heaps + threadIdx.x}, a.k);
  zero_tile<DIMP, kStlBlock>(tile4);

  for (int64_t t0 = 0; t0 < a.n; t0 += kTile) {
    const int tn = int(min<int64_t>(kTile, a.n - t0));
    // t0 is a multiple of kTile: 16-byte aligned in the packed cloud
    load_tile<DIMP, kStlBlock>(tile4, a.cloud + t0 * dim, tn * dim, dim);
    __syncthreads();
    if (valid) {
      // kBfUnroll independent sums at a time (one wave per SIMD here, so the
      // wave's own instructions must overlap); rows past tn are read but not
      // admitted. Admissions stay in index order.
      constexpr int kUnroll = DIMP <= 36 ? kBfUnroll : (DIMP <= 64 ? 2 : 1);  // within VGPRs
      for (int pt = 0; pt < tn; pt += kUnroll) {
        float dist[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) dist[u] = padded_dist<DIMP>(tile + (pt + u) * DIMP, q);
#pragma unroll
        for (int u = 0; u < kUnroll; ++u)
          if (pt + u < tn && dist[u] <= r2 && dist[u] < heap.head &&
              (a.allow_self || dist[u] > FLT_EPSILON))
            heap.replace_head(int32_t(t0 + pt + u), dist[u]);
      }
    }
  }
  if (!valid) return;
  if (sort_results) heap.sort();
  heap.write(out_idx + qi * a.k, out_d2 + qi * a.k);
}

// Replays a query's chunk logs in chunk order through an IndexHeapSTL with
// the global admission test. A chunk's own k-best starts empty, so its head
// is never below the global head at the same point: every point the
// reference admits is in its chunk's log, in order, and the replay admits
// exactly those. A query with an overflowed log is left to a full scan.
__global__ __launch_bounds__(kStlBlock) void k_bf_replay(BfParams a, int sort_results,
                                                         int32_t* __restrict__ out_idx,
                                                         float* __restrict__ out_d2) {
  extern __shared__ uint2 heaps[];  // k slots x kStlBlock lanes
  const int64_t qi = int64_t(blockIdx.x) * kStlBlock + threadIdx.x;
  if (qi >= a.m) return;
  StlHeap<LdsHeapStore> heap;
  heap.reset(LdsHeapStore{heaps + threadIdx.x}, a.k);
  for (int c = 0; c < a.chunks; ++c) {
    const uint32_t n = a.log_n[int64_t(c) * a.m + qi];
    if (n > uint32_t(a.log_cap)) {
      a.redo[qi] = 1;
      return;
    }
    const uint2* lg = a.log + int64_t(c) * a.log_cap * a.m + qi;
    for (uint32_t j = 0; j < n; ++j) {
      const uint2 e = lg[int64_t(j) * a.m];
      const float d = __uint_as_float(e.y);
      if (d < heap.head) heap.replace_head(int32_t(e.x), d);
    }
  }
  a.redo[qi] = 0;
  if (sort_results) heap.sort();
  heap.write(out_idx + qi * a.k, out_d2 + qi * a.k);
}

template <int DIMP>
void launch_bf_stl(const BfParams& p, int sort_results, hipStream_t s, int32_t* oi, float* od) {
  if (p.cstride != p.dim) throw Error(NABO_HIP_ERR_ARGUMENT, "brute-force cloud is not packed");
  const size_t heap_lds = size_t(p.k) * kStlBlock * sizeof(uint2);
  static_assert(kTile * 128 * sizeof(float) + 64 * kStlBlock * sizeof(uint2) <= 65536,
                "tile and heaps fit the 64 KiB of LDS a workgroup may hold");
  k_bf_stl<DIMP><<<unsigned(ceil_div(p.m, kStlBlock)), kStlBlock, heap_lds, s>>>(p, sort_results,
                                                                                 oi, od);
}

// LOG: the chunk pass of the IndexHeapSTL form (logs only, no merge).
template <int DIMP, int K, bool LOG>
void launch_bf(const BfParams& p, hipStream_t s, int32_t* oi, float* od) {
  dim3 grid(unsigned(ceil_div(p.m, kBlock)), unsigned(p.chunks));
  k_bf_chunk<DIMP, K, LOG><<<grid, kBlock, 0, s>>>(p);
  if constexpr (!LOG) k_bf_merge<K><<<unsigned(ceil_div(p.m, kBlock)), kBlock, 0, s>>>(p, oi, od);
}

template <int DIMP, bool LOG = false>
void bf_dispatch_k(const BfParams& p, hipStream_t s, int32_t* oi, float* od) {
  if (p.k <= 1)
    launch_bf<DIMP, 1, LOG>(p, s, oi, od);
  else if (p.k <= 4)
    launch_bf<DIMP, 4, LOG>(p, s, oi, od);
  else if (p.k <= 8)
    launch_bf<DIMP, 8, LOG>(p, s, oi, od);
  else if (p.k <= 16)
    launch_bf<DIMP, 16, LOG>(p, s, oi, od);
  else if (p.k <= 32)
    launch_bf<DIMP, 32, LOG>(p, s, oi, od);
  else
    launch_bf<DIMP, 64, LOG>(p, s, oi, od);
}

// IndexHeapSTL brute force: chunk pass with logs, replay, then a full scan
// of the queries whose logs overflowed.
template <int DIMP>
void launch_bf_stl_logged(const BfParams& p, int sort_results, hipStream_t s, int32_t* oi,
                          float* od) {
  bf_dispatch_k<DIMP, true>(p, s, oi, od);
  const size_t heap_lds = size_t(p.k) * kStlBlock * sizeof(uint2);
  k_bf_replay<<<unsigned(ceil_div(p.m, kStlBlock)), kStlBlock, heap_lds, s>>>(p, sort_results,
                                                                                oi, od);
  launch_bf_stl<DIMP>(p, sort_results, s, oi, od);
}

__global__ void k_copy_cloud(int64_t n, int dim, const float* __restrict__ src, int64_t stride,
                             float* __restrict__ dst) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n * dim) return;
  const int64_t p = e / dim;
  const int d = int(e % dim);
  dst[e] = src[p * stride + d];
}


// ---------------------------------------------------------------------------
// Mixed-precision shortlist (bf16 or fp16 MFMA, fp32 accumulation).
//
// Ranking value of (query q, point p): |p|^2 - 2 q.p, which orders points as
// the squared distance |q|^2 + |p|^2 - 2 q.p does for a fixed query. It is
// one GEMM: point row [p~_0..p~_{d-1}, hi(|p|^2), lo(|p|^2), 0...] times query
// column [-2q~_0..-2q~_{d-1}, 1, 1, 0...] (x~ = x rounded to 16 bits; the
// norm is split in two 16-bit parts so that it keeps ~16 significant bits;
// -2q is exact). K is padded to a multiple of 16 for
// v_mfma_f32_32x32x16_{bf16,f16}.
//
// MFMA operand/result layout (gfx950, checked by a probe): A[i][k] in lane
// i + 32*(k/8), element k%8; B[k][j] in lane j + 32*(k/8); D[i][j] in lane
// j + 32*((i/4)%2), register (i%4) + 4*(i/8). With points as rows and
// queries as columns, lane l holds 16 ranking values of query l%32, for the
// rows of half l/32 of each 32-point tile.
//
// Each lane keeps the L = k + m smallest ranking values of its half of a
// chunk of points (candidates buffered in LDS, k_bf_shortlist); a second
// kernel takes the L smallest of a query's lists
// (the shortlist), recomputes their squared distances exactly in fp32 with
// the sequential sum, applies the reference's admission rule and keeps the
// k best by (distance, index).

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kMfmaTile = 32;  // points (rows) and queries (columns) per MFMA
constexpr int kShortBuf = 16;  // candidates buffered per lane (LDS) before insertion
static_assert(kShortBuf >= 8, "8 registers can bring 8 candidates to an emptied buffer");

template <bool F16>
struct Lp;
template <>
struct Lp<false> {
  typedef __bf16 T;
  typedef bf16x8 V;
  static __device__ __forceinline__ f32x16 mfma(V a, V b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
  }
};
template <>
struct Lp<true> {
  typedef _Float16 T;
  typedef f16x8 V;
  static __device__ __forceinline__ f32x16 mfma(V a, V b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
  }
};

// Point rows: coordinates, then the two halves of |p|^2 (fp32 sequential
// sum); padding rows get an infinite norm so that they rank last. Stored in
// the MFMA's A-operand order, so that each K-step of a 32-row tile is one
// contiguous 1 KiB load (lane l reads 16 bytes at 16 l): column c of row p
// sits at ((p / 32) * KS + c / 16) * 512 + (p % 32 + 32 * (c % 16 / 8)) * 8
// + c % 8, KS = kp / 16.
__device__ __forceinline__ int64_t lp_offset(int64_t p, int c, int kp) {
  const int64_t tile = p / kMfmaTile;
  const int row = int(p % kMfmaTile);
  const int ks = c / 16;
  const int lane = row + 32 * ((c % 16) / 8);
  return (tile * (kp / 16) + ks) * 512 + lane * 8 + c % 8;
}

template <bool F16>
__global__ void k_lp_points(int64_t n, int64_t n_pad, int dim, int kp,
                            const float* __restrict__ cloud, int64_t stride,
                            typename Lp<F16>::T* __restrict__ out) {
  typedef typename Lp<F16>::T T;
  const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= n_pad) return;
  if (p >= n) {
    for (int c = 0; c < kp; ++c) out[lp_offset(p, c, kp)] = T(c == dim ? INFINITY : 0.f);
    return;
  }
  const float* src = cloud + p * stride;
  float norm = 0.f;
  for (int d = 0; d < dim; ++d) {
    const float x = src[d];
    out[lp_offset(p, d, kp)] = T(x);
    norm = norm + x * x;
  }
  const T hi = T(norm);
  out[lp_offset(p, dim, kp)] = hi;
  out[lp_offset(p, dim + 1, kp)] = T(norm - float(hi));
  for (int c = dim + 2; c < kp; ++c) out[lp_offset(p, c, kp)] = T(0.f);
}

// Query columns: -2 q, then 1, 1 for the norm halves. Also resets the
// query's shared threshold (order_key of +inf).
template <bool F16>
__global__ void k_lp_queries(int64_t m, int64_t m_pad, int dim, int kp,
                             const float* __restrict__ query, int64_t qstride,
                             typename Lp<F16>::T* __restrict__ out, int32_t* __restrict__ thr) {
  typedef typename Lp<F16>::T T;
  const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= m_pad) return;
  T* row = out + q * kp;
  for (int c = 0; c < kp; ++c) {
    float v = 0.f;
    if (q < m) v = c < dim ? -2.f * query[q * qstride + c] : (c < dim + 2 ? 1.f : 0.f);
    row[c] = T(v);
  }
  thr[q] = 0x7F800000;
}

// Signed-integer keys in float order (both zeros map to 0), for atomicMin
// on floats; order_value inverts them. The key after k is that of the next
// float up.
__device__ __forceinline__ int32_t order_key(float f) {
  const int32_t b = __float_as_int(f);
  return f == 0.f ? 0 : (b >= 0 ? b : b ^ 0x7FFFFFFF);
}
__device__ __forceinline__ float order_value(int32_t k) {
  return __int_as_float(k >= 0 ? k : k ^ 0x7FFFFFFF);
}
constexpr int32_t kKeyInf = 0x7F800000;

// Sorted insertion of a ranking value (ties keep the earlier entry first),
// in place from the top: entry j takes entry j-1 when that is greater than
// value, else value when entry j is, else stays. The index selects are bit
// blends so that they stay branch-free without temporaries.
__device__ __forceinline__ int32_t blend(uint32_t mask, int32_t a, int32_t b) {
  return int32_t((uint32_t(a) & mask) | (uint32_t(b) & ~mask));
}
template <int L>
__device__ __forceinline__ void list_insert(float (&ld)[L], int32_t (&li)[L], float value,
                                            int32_t index) {
  uint32_t cur = 0u - uint32_t(ld[L - 1] > value);
#pragma unroll
  for (int j = L - 1; j > 0; --j) {
    const uint32_t prev = 0u - uint32_t(ld[j - 1] > value);
    li[j] = blend(prev, li[j - 1], blend(cur, index, li[j]));
    ld[j] = __builtin_amdgcn_fmed3f(ld[j - 1], ld[j], value);
    cur = prev;
  }
  li[0] = blend(cur, index, li[0]);
  ld[0] = fminf(ld[0], value);
}

struct ShortlistParams {
  const void* points;   // n_pad x kp 16-bit
  const void* queries;  // m_pad x kp 16-bit
  int kp;
  int64_t n;
  int64_t m_pad;
  int64_t chunk;        // points per chunk (multiple of 32)
  int l;                // k + m
  float* part_v;        // chunks x m_pad x 2 x LCAP
  int32_t* part_i;
  int32_t* thr;         // m_pad: lowest list head of each query so far (order_key)
};

// grid: (m_pad / 128, chunks); 4 waves, each 32 queries.
//
// Candidate upkeep: each of a tile's 16 values per lane that is below the
// query's threshold (the lower of the heads of its two lists, read at the
// tile's start) is appended to the lane's buffer in LDS ([slot][thread];
// the test is one compare per value and a branch that skips the append for
// the whole wave when no lane has that value as a candidate, which late in
// a chunk is most values). Before each half tile (8 values), if some lane's
// buffer could not take 8 more, every lane moves its buffer into its sorted
// list (each entry still below the head is inserted), so the O(L)
// insertions run for all lanes at once, about once per kShortBuf - 7
// candidates of the busiest lane, instead of once per candidate of any
// lane. Buffered values may have been passed by the head since; they are
// tested again at the move, so each list ends as the L smallest values of
// its rows that were under the threshold, as with one insertion per value
// in the same order. The next tile's point operands are loaded while a
// tile's values are tested.
template <bool F16, int LCAP, int KS>
__device__ __forceinline__ void shortlist_scan(const ShortlistParams& a, float* __restrict__ bv,
                                               int32_t* __restrict__ bi) {
  typedef typename Lp<F16>::T T;
  typedef typename Lp<F16>::V V;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int half = lane >> 5;
  const int64_t q = int64_t(blockIdx.x) * 128 + wave * 32 + (lane & 31);
  const int64_t p0 = int64_t(blockIdx.y) * a.chunk;
  const int64_t p1 = min(p0 + a.chunk, (a.n + kMfmaTile - 1) / kMfmaTile * kMfmaTile);
  const T* qrow = static_cast<const T*>(a.queries) + q * a.kp + 8 * half;
  V b[KS];
#pragma unroll
  for (int ks = 0; ks < KS; ++ks) b[ks] = *reinterpret_cast<const V*>(qrow + 16 * ks);

  float ld[LCAP];
  int32_t li[LCAP];
#pragma unroll
  for (int j = 0; j < LCAP; ++j) {
    ld[j] = j < LCAP - a.l ? -INFINITY : INFINITY;
    li[j] = -1;
  }
  int nb = 0;  // buffered candidates of this lane
  // The candidate threshold. A value v enters this lane's buffer when
  // v < own head (the list's own rule) and v <= h for h the head of any
  // other list of the query: the partner lane's (lane ^ 32: the other rows
  // of the tile) and those published by the query's other chunks (the
  // lowest in thr). Such a list holds l values at or below h, so a value
  // above h is not among the query's l smallest: the lists' merge is the
  // same as without the threshold, ties included. Heads only fall, so a
  // stale h is still valid: the threshold is refreshed at each flush only
  // (measured: refreshing it every 64 to 1024 tiles as well costs more than
  // it saves).
  int32_t* gq = a.thr + q;
  float head = INFINITY;
  auto refresh = [&](int32_t published) {
    const float own = ld[LCAP - 1];
    const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(own), __float_as_uint(own),
                                                     false, false);
    const int32_t other = min(min(published, order_key(own)),
                              min(order_key(__uint_as_float(sw[0])),
                                  order_key(__uint_as_float(sw[1]))));
    const float th = other >= kKeyInf ? INFINITY : order_value(other + 1);
    head = __builtin_amdgcn_fmed3f(own, th, -INFINITY);  // the lower (no NaN quieting)
  };
  auto flush = [&]() {
    const int32_t published = atomicMin(gq, order_key(ld[LCAP - 1]));
#pragma unroll 1
    for (int e = 0; e < kShortBuf; ++e) {
      if (!__builtin_amdgcn_ballot_w64(e < nb)) break;
      if (e < nb) {
        const float v = bv[e * 256 + tid];
        if (v < ld[LCAP - 1]) list_insert<LCAP>(ld, li, v, bi[e * 256 + tid]);
      }
    }
    nb = 0;
    atomicMin(gq, order_key(ld[LCAP - 1]));
    refresh(published);
  };
  // tile t's K-step ks: 512 elements at (t / 32 * KS + ks) * 512, lane's 8 at 8 * lane
  const T* pbase = static_cast<const T*>(a.points) + lane * 8;
  auto load = [&](V (&dst)[KS], int64_t t) {
    const T* tp = pbase + (t / kMfmaTile) * (KS * 512);
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) dst[ks] = *reinterpret_cast<const V*>(tp + ks * 512);
  };
  // One tile: the MFMA, then the candidate test. The threshold is the lower
  // of this lane's head and its partner's (lane ^ 32: the same query, the
  // other rows): the partner's list already holds l values at or below its
  // head, so a value at or above it is not among the query's l smallest.
  auto tile = [&](const V (&op)[KS], int64_t t) {
    f32x16 acc = Lp<F16>::mfma(op[0], b[0], f32x16{});
#pragma unroll
    for (int ks = 1; ks < KS; ++ks) acc = Lp<F16>::mfma(op[ks], b[ks], acc);
    const int32_t base = int32_t(t) + 4 * half;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      // room for the next 8 registers' candidates in every lane's buffer
      if (r % 8 == 0 && __builtin_amdgcn_ballot_w64(nb > kShortBuf - 8)) flush();
      const bool mine = acc[r] < head;
      if (__builtin_amdgcn_ballot_w64(mine)) {
        // keeps the wave-wide test a branch of its own (not folded into the
        // lane mask of the append, which would then be issued for every value)
        asm volatile("");
        if (mine) {
          bv[nb * 256 + tid] = acc[r];
          bi[nb * 256 + tid] = base + (r & 3) + 8 * (r >> 2);
          ++nb;
        }
      }
    }
  };
  // Two operand sets in turn: one tile's rows load while the other's are
  // tested (the last tile reloads its own rows: in bounds).
  V pa[KS], pb[KS];
  if (p0 < p1) load(pa, p0);
  for (int64_t t = p0; t < p1; t += 2 * kMfmaTile) {
    const int64_t t1 = t + kMfmaTile;
    load(pb, t1 < p1 ? t1 : t);
    tile(pa, t);
    if (t1 >= p1) break;
    load(pa, t1 + kMfmaTile < p1 ? t1 + kMfmaTile : t1);
    tile(pb, t1);
  }
  flush();
  float* ov = a.part_v + ((int64_t(blockIdx.y) * a.m_pad + q) * 2 + half) * a.l;
  int32_t* oi = a.part_i + ((int64_t(blockIdx.y) * a.m_pad + q) * 2 + half) * a.l;
#pragma unroll
  for (int j = 0; j < LCAP; ++j) {
    if (j >= LCAP - a.l) {
      ov[j - (LCAP - a.l)] = ld[j];
      oi[j - (LCAP - a.l)] = li[j];
    }
  }
}

template <bool F16, int LCAP>
__global__ __launch_bounds__(256) void k_bf_shortlist(ShortlistParams a) {
  __shared__ float bv[kShortBuf * 256];
  __shared__ int32_t bi[kShortBuf * 256];
  switch (a.kp / 16) {
    case 1: shortlist_scan<F16, LCAP, 1>(a, bv, bi); break;
    case 2: shortlist_scan<F16, LCAP, 2>(a, bv, bi); break;
    case 3: shortlist_scan<F16, LCAP, 3>(a, bv, bi); break;
    default: shortlist_scan<F16, LCAP, 4>(a, bv, bi); break;
  }
}

// Shortlist of each query: the l smallest ranking values over its
// 2 x chunks lists, taken in list order (a value equal to the list's head
// does not enter, so among equal values the earlier list's entry is kept).
// One thread per query; each list is read four entries at a time (lists are
// ascending: the first entry not below the head ends the list). The
// shortlist's indices are written over the query's first list, which only
// this thread reads.
template <int LCAP>
__global__ __launch_bounds__(64) void k_bf_lp_merge(ShortlistParams s, int chunks, int64_t m) {
  const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= m) return;
  float ld[LCAP];
  int32_t li[LCAP];
#pragma unroll
  for (int j = 0; j < LCAP; ++j) {
    ld[j] = j < LCAP - s.l ? -INFINITY : INFINITY;
    li[j] = -1;
  }
  for (int c = 0; c < chunks * 2; ++c) {
    const int64_t o = ((int64_t(c >> 1) * s.m_pad + q) * 2 + (c & 1)) * s.l;
    bool done = false;
    for (int j = 0; j < s.l && !done; j += 4) {
      float v[4];
      int32_t ix[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const bool in = j + u < s.l;
        v[u] = in ? s.part_v[o + j + u] : INFINITY;
        ix[u] = in ? s.part_i[o + j + u] : -1;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (done || !(v[u] < ld[LCAP - 1])) {
          done = true;
        } else {
          list_insert<LCAP>(ld, li, v[u], ix[u]);
        }
      }
    }
  }
  int32_t* out = s.part_i + q * 2 * s.l;
#pragma unroll
  for (int j = 0; j < LCAP; ++j)
    if (j >= LCAP - s.l) out[j - (LCAP - s.l)] = li[j];
}

// Exact rescoring of the shortlists: G lanes per query (G >= l), lane j
// takes the shortlist's j-th index, computes its squared distance in fp32
// with the sequential sum (the exact path's arithmetic), and applies the
// reference's admission rule (within the radius, finite, not a self match
// unless allowed). The k best by (distance, index) are the admitted
// candidates of rank < k, the rank being the number of admitted candidates
// of the query before it in that order (indices in a shortlist are
// distinct, so ranks are too); output positions past the admitted count
// get (-1, inf). Equal to inserting the candidates one by one into a
// sorted k-best.
template <int G>
__global__ __launch_bounds__(256) void k_bf_rescore(ShortlistParams s, BfParams a,
                                                    int32_t* __restrict__ out_idx,
                                                    float* __restrict__ out_d2) {
  const int64_t gid = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t q = gid / G;
  const int j = int(gid % G);
  const bool live = q < a.m;
  int32_t idx = -1;
  if (live && j < s.l) idx = s.part_i[q * 2 * s.l + j];
  float dist = INFINITY;
  bool admit = false;
  if (live && idx >= 0 && idx < a.n) {
    const float* pp = a.cloud + int64_t(idx) * a.cstride;
    const float* qp = a.query + q * a.qstride;
    float sum = 0.f;
    int d = 0;
    for (; d + 8 <= a.dim; d += 8) {
      float pv[8], qv[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        pv[u] = pp[d + u];
        qv[u] = qp[d + u];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const float diff = pv[u] - qv[u];
        sum = sum + diff * diff;
      }
    }
    for (; d < a.dim; ++d) {
      const float diff = pp[d] - qp[d];
      sum = sum + diff * diff;
    }
    float r2 = a.max_r2;
    if (a.radii) {
      const float r = a.radii[q];
      r2 = r * r;
    }
    dist = sum;
    admit = sum <= r2 && sum < INFINITY && (a.allow_self || sum > FLT_EPSILON);
  }
  const float key_d = admit ? dist : INFINITY;
  const int32_t key_i = admit ? idx : INT32_MAX;
  int rank = 0;
  int admitted = admit ? 1 : 0;
#pragma unroll 4
  for (int o = 1; o < G; ++o) {
    const int src = (j + o) % G;
    const float od = __shfl(key_d, src, G);
    const int32_t oi = __shfl(key_i, src, G);
    const int oadmit = __shfl(admit ? 1 : 0, src, G);
    admitted += oadmit;
    rank += (oadmit && (od < key_d || (od == key_d && oi < key_i))) ? 1 : 0;
  }
  if (!live) return;
  if (admit && rank < a.k) {
    out_idx[q * a.k + rank] = idx;
    out_d2[q * a.k + rank] = dist;
  }
  if (j >= admitted && j < a.k) {
    out_idx[q * a.k + j] = -1;
    out_d2[q * a.k + j] = INFINITY;
  }
}

template <bool F16, int LCAP>
void launch_shortlist(const ShortlistParams& s, const BfParams& p, hipStream_t st,
                      int32_t* oi, float* od) {
  k_bf_shortlist<F16, LCAP><<<dim3(unsigned(s.m_pad / 128), unsigned(p.chunks)), 256, 0, st>>>(s);
  k_bf_lp_merge<LCAP><<<unsigned(ceil_div(p.m, 64)), 64, 0, st>>>(s, p.chunks, p.m);
  constexpr int G = LCAP <= 16 ? 16 : (LCAP <= 32 ? 32 : 64);
  k_bf_rescore<G><<<unsigned(ceil_div(p.m * G, 256)), 256, 0, st>>>(s, p, oi, od);
}

template <bool F16>
void shortlist_l(const ShortlistParams& s, const BfParams& p, hipStream_t st, int32_t* oi,
                 float* od) {
  if (s.l <= 16)
    launch_shortlist<F16, 16>(s, p, st, oi, od);
  else if (s.l <= 24)
    launch_shortlist<F16, 24>(s, p, st, oi, od);
  else if (s.l <= 32)
    launch_shortlist<F16, 32>(s, p, st, oi, od);
  else if (s.l <= 40)
    launch_shortlist<F16, 40>(s, p, st, oi, od);
  else if (s.l <= 48)
    launch_shortlist<F16, 48>(s, p, st, oi, od);
  else
    launch_shortlist<F16, 64>(s, p, st, oi, od);
}

// Workgroups of k_bf_shortlist one CU holds at once (registers grow with
// the list length; the LDS buffer is fixed), cached per instantiation.
template <bool F16, int LCAP>
int shortlist_per_cu_of() {
  static std::atomic<int> cache{0};
  int v = cache.load();
  if (v < 1) {
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, k_bf_shortlist<F16, LCAP>, 256, 0) !=
            hipSuccess ||
        v < 1)
      v = 1;
    cache.store(v);
  }
  return v;
}

template <bool F16>
int shortlist_per_cu(int l) {
  if (l <= 16) return shortlist_per_cu_of<F16, 16>();
  if (l <= 24) return shortlist_per_cu_of<F16, 24>();
  if (l <= 32) return shortlist_per_cu_of<F16, 32>();
  if (l <= 40) return shortlist_per_cu_of<F16, 40>();
  if (l <= 48) return shortlist_per_cu_of<F16, 48>();
  return shortlist_per_cu_of<F16, 64>();
}

// Chunks of points for the shortlist: at least enough workgroups to fill
// the device once (each chunk fills its lists from empty, about
// l * (1 + ln(chunk / l)) candidates, so fewer, longer chunks are cheaper
// per point), then, up to twice that, the count whose last round of
// workgroups is fullest (a grid of 1.15 device-fulls takes as long as 2).
int64_t shortlist_chunks(int64_t qb, int64_t cap, int64_t max_chunks) {
  auto fill = [&](int64_t c) {
    const int64_t blocks = qb * c;
    return double(blocks) / double(ceil_div(blocks, cap) * cap);
  };
  const int64_t c0 = std::min(max_chunks, std::max<int64_t>(1, ceil_div(cap, qb)));
  int64_t best = c0;
  for (int64_t c = c0 + 1; c <= std::min(max_chunks, 2 * c0); ++c)
    if (fill(c) > fill(best) + 0.05) best = c;
  return best;
}

// Brute force beyond the tiled kernels' sizes (k above 64, dim above 128,
// or a shortlist longer than 64): one query per thread scans the cloud in
// index order (brute_force_cpu.cpp:81-98), its k-best in its own output
// row, either sorted (BRUTE_FORCE_HIP: ascending, a new entry after equal
// ones, so ties keep the lower index; the shortlist precisions take this
// exact path too) or as IndexHeapSTL (BRUTE_FORCE: heap order, sort_heap
// with SORT_RESULTS). Distances are the sequential sum the tiled kernels
// compute. No tiling: it is the unbounded form, not the fast one.
template <bool STL>
__global__ __launch_bounds__(kBlock) void k_bf_rows(BfParams a, int sort, int32_t* out_idx,
                                                    float* out_d2) {
  const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= a.m) return;
  const float* qp = a.query + q * a.qstride;
  float r2 = a.max_r2;
  if (a.radii) {
    const float r = a.radii[q];
    r2 = r * r;
  }
  const RowHeapStore row{out_idx + q * a.k, out_d2 + q * a.k};
  typename std::conditional<STL, StlHeap<RowHeapStore>, RowSortedHeap>::type heap;
  heap.reset(row, a.k);
  for (int64_t i = 0; i < a.n; ++i) {
    const float* pp = a.cloud + i * a.cstride;
    float dist = 0.f;
    for (int d = 0; d < a.dim; ++d) {
      const float diff = pp[d] - qp[d];
      dist = dist + diff * diff;
    }
    if (dist <= r2 && dist < heap.head && (a.allow_self || dist > FLT_EPSILON))
      heap.replace_head(int32_t(i), dist);
  }
  if (sort) heap.sort();
  heap.write(row.idx, row.val);
}

}  // namespace

constexpr int kBfMaxK = 64;
constexpr int kBfMaxDim = 128;

constexpr int kLpMaxDim = 62;  // dim + 2 norm columns within 4 MFMA K-steps

void bf_prepare(Index& ix, const float* d_cloud, int64_t stride, hipStream_t s) {
  const int prec = ix.params.bf_precision;
  if (prec != NABO_HIP_BF_FP32 && prec != NABO_HIP_BF_BF16 && prec != NABO_HIP_BF_FP16)
    throw Error(NABO_HIP_ERR_ARGUMENT, "unknown brute-force precision");
  if (ix.params.heap == NABO_HIP_HEAP_STL && prec != NABO_HIP_BF_FP32)
    throw Error(NABO_HIP_ERR_UNSUPPORTED,
                "the STL-heap brute force (the reference's order) is exact FP32 only");
  if (prec != NABO_HIP_BF_FP32) {
    if (ix.dim > kLpMaxDim)
      throw Error(NABO_HIP_ERR_UNSUPPORTED, "the bf16/fp16 shortlist supports dim <= 62");
    if (ix.params.bf_extra < 0) throw Error(NABO_HIP_ERR_ARGUMENT, "bfExtra must be >= 0");
    ix.lp_kp = int(ceil_div(ix.dim + 2, 16) * 16);
    ix.lp_n_pad = ceil_div(ix.n, kMfmaTile) * kMfmaTile;
    NH_CHECK(hipMalloc(reinterpret_cast<void**>(&ix.d_cloud_lp),
                       sizeof(uint16_t) * size_t(ix.lp_n_pad) * ix.lp_kp));
    const unsigned g = unsigned(ceil_div(ix.lp_n_pad, kBlock));
    if (prec == NABO_HIP_BF_FP16)
      k_lp_points<true><<<g, kBlock, 0, s>>>(ix.n, ix.lp_n_pad, ix.dim, ix.lp_kp, d_cloud, stride,
                                              reinterpret_cast<_Float16*>(ix.d_cloud_lp));
    else
      k_lp_points<false><<<g, kBlock, 0, s>>>(ix.n, ix.lp_n_pad, ix.dim, ix.lp_kp, d_cloud, stride,
                                               reinterpret_cast<__bf16*>(ix.d_cloud_lp));
    NH_LAUNCH_CHECK();
  }
  ix.bf_stride = ix.dim;
  NH_CHECK(hipMalloc(reinterpret_cast<void**>(&ix.d_cloud), sizeof(float) * size_t(ix.n) * ix.dim));
  const int64_t total = ix.n * ix.dim;
  k_copy_cloud<<<unsigned(ceil_div(total, kBlock)), kBlock, 0, s>>>(ix.n, ix.dim, d_cloud, stride,
                                                                     ix.d_cloud);
  NH_LAUNCH_CHECK();
  // Bounds as BruteForceSearch computes them (rowwise min/max, brute_force_cpu.cpp:49-50).
  std::vector<float> host(static_cast<size_t>(total));
  NH_CHECK(hipMemcpyAsync(host.data(), ix.d_cloud, sizeof(float) * total, hipMemcpyDeviceToHost, s));
  NH_CHECK(hipStreamSynchronize(s));
  ix.min_bound.assign(size_t(ix.dim), FLT_MAX);
  ix.max_bound.assign(size_t(ix.dim), -FLT_MAX);
  for (int64_t p = 0; p < ix.n; ++p)
    for (int d = 0; d < ix.dim; ++d) {
      const float v = host[size_t(p * ix.dim + d)];
      ix.min_bound[d] = std::min(ix.min_bound[d], v);
      ix.max_bound[d] = std::max(ix.max_bound[d], v);
    }
}

void bf_knn(const Index& ix, const KnnArgs& a, hipStream_t s) {
  if (a.m == 0) return;
  BfParams p{};
  p.cloud = ix.d_cloud;
  p.cstride = ix.bf_stride;
  p.dim = ix.dim;
  p.n = ix.n;
  p.query = a.query;
  p.qstride = a.qstride;
  p.m = a.m;
  p.k = a.k;
  p.max_r2 = a.max_radius * a.max_radius;
  p.radii = a.radii;
  p.allow_self = (a.flags & NABO_HIP_ALLOW_SELF_MATCH) ? 1 : 0;
  if (p.cstride != p.dim) throw Error(NABO_HIP_ERR_ARGUMENT, "brute-force cloud is not packed");
  const bool lp = ix.params.bf_precision != NABO_HIP_BF_FP32;
  if (a.k > kBfMaxK || ix.dim > kBfMaxDim || (lp && a.k + ix.params.bf_extra > 64)) {
    const int sort = (a.flags & NABO_HIP_SORT_RESULTS) ? 1 : 0;
    const unsigned g = unsigned(ceil_div(a.m, kBlock));
    KernelTimer timer(s);
    if (ix.params.heap == NABO_HIP_HEAP_STL)
      k_bf_rows<true><<<g, kBlock, 0, s>>>(p, sort, a.out_idx, a.out_d2);
    else
      k_bf_rows<false><<<g, kBlock, 0, s>>>(p, sort, a.out_idx, a.out_d2);
    NH_LAUNCH_CHECK();
    return;
  }
  const int64_t qblocks = ceil_div(a.m, kBlock);
  int64_t chunks = std::max<int64_t>(1, ceil_div(2048, qblocks));
  chunks = std::min<int64_t>(chunks, std::max<int64_t>(1, ceil_div(ix.n, 1024)));
  chunks = std::min<int64_t>(chunks, 65535);
  p.chunk = ceil_div(ceil_div(ix.n, chunks), kTile) * kTile;  // tiles start 16-byte aligned
  p.chunks = int(ceil_div(ix.n, p.chunk));
  if (ix.params.bf_precision != NABO_HIP_BF_FP32) {
    const int l = a.k + ix.params.bf_extra;
    const bool f16 = ix.params.bf_precision == NABO_HIP_BF_FP16;
    ShortlistParams sp{};
    sp.kp = ix.lp_kp;
    sp.n = ix.n;
    sp.m_pad = ceil_div(a.m, 128) * 128;
    sp.l = l;
    sp.points = ix.d_cloud_lp;
    DevBuf<uint16_t> qlp(size_t(sp.m_pad) * sp.kp, s);
    DevBuf<int32_t> thr(size_t(sp.m_pad), s);
    const unsigned gq = unsigned(ceil_div(sp.m_pad, kBlock));
    if (f16)
      k_lp_queries<true><<<gq, kBlock, 0, s>>>(a.m, sp.m_pad, ix.dim, sp.kp, a.query, a.qstride,
                                                reinterpret_cast<_Float16*>(qlp.p), thr.p);
    else
      k_lp_queries<false><<<gq, kBlock, 0, s>>>(a.m, sp.m_pad, ix.dim, sp.kp, a.query, a.qstride,
                                                 reinterpret_cast<__bf16*>(qlp.p), thr.p);
    sp.queries = qlp.p;
    sp.thr = thr.p;
    const int64_t qb = sp.m_pad / 128;
    int dev = 0, cus = 256;
    NH_CHECK(hipGetDevice(&dev));
    NH_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int64_t cap = int64_t(cus) * (f16 ? shortlist_per_cu<true>(l) : shortlist_per_cu<false>(l));
    const int64_t chunks_lp =
        shortlist_chunks(qb, cap, std::min<int64_t>(65535, std::max<int64_t>(1, ix.lp_n_pad / 1024)));
    sp.chunk = ceil_div(ceil_div(ix.lp_n_pad, chunks_lp), kMfmaTile) * kMfmaTile;
    p.chunks = int(ceil_div(ix.lp_n_pad, sp.chunk));
    DevBuf<float> pv(size_t(p.chunks) * sp.m_pad * 2 * l, s);
    DevBuf<int32_t> pidx(size_t(p.chunks) * sp.m_pad * 2 * l, s);
    sp.part_v = pv.p;
    sp.part_i = pidx.p;
    KernelTimer timer(s);
    if (f16)
      shortlist_l<true>(sp, p, s, a.out_idx, a.out_d2);
    else
      shortlist_l<false>(sp, p, s, a.out_idx, a.out_d2);
    NH_LAUNCH_CHECK();
    return;
  }
  if (ix.params.heap == NABO_HIP_HEAP_STL) {
    const int sort = (a.flags & NABO_HIP_SORT_RESULTS) ? 1 : 0;
    // Chunk logs of up to `cap` admissions per (chunk, query); a chunk's
    // own k-best admits about k (1 + ln(chunk / k)) points, so a log
    // overflows only for adversarial orders (those queries are redone by a
    // full scan). Queries go in batches that keep the logs within
    // kStlLogBytes.
    const int cap = std::min(1024, std::max(64, 16 * a.k));
    // At least kBlock queries per batch within kStlLogBytes: fewer, longer
    // chunks when the logs of kBlock queries would not fit.
    const size_t per_chunk = size_t(cap) * sizeof(uint2) + sizeof(uint32_t);
    const int64_t max_chunks =
        std::max<int64_t>(1, int64_t(kStlLogBytes / (size_t(kBlock) * per_chunk)));
    if (p.chunks > max_chunks) {
      p.chunk = ceil_div(ceil_div(ix.n, max_chunks), kTile) * kTile;
      p.chunks = int(ceil_div(ix.n, p.chunk));
    }
    const size_t per_query = size_t(p.chunks) * per_chunk + sizeof(int32_t);
    int64_t mb = std::max<int64_t>(kBlock, int64_t(kStlLogBytes / per_query) / kBlock * kBlock);
    mb = std::min(mb, a.m);
    Scratch scratch(s);
    Span<uint2> log;
    Span<uint32_t> log_n;
    Span<int32_t> redo;
    scratch.want(log, size_t(p.chunks) * size_t(cap) * size_t(mb));
    scratch.want(log_n, size_t(p.chunks) * size_t(mb));
    scratch.want(redo, size_t(mb));
    scratch.commit(ix.device);
    KernelTimer timer(s);
    for (int64_t q0 = 0; q0 < a.m; q0 += mb) {
      BfParams b = p;
      b.query = a.query + q0 * a.qstride;
      b.m = std::min(mb, a.m - q0);
      b.radii = a.radii ? a.radii + q0 : nullptr;
      b.log = log.p;
      b.log_n = log_n.p;
      b.log_cap = cap;
      b.redo = redo.p;
      int32_t* oi = a.out_idx + q0 * a.k;
      float* od = a.out_d2 + q0 * a.k;
      if (ix.dim <= 4)
        launch_bf_stl_logged<4>(b, sort, s, oi, od);
      else if (ix.dim <= 8)
        launch_bf_stl_logged<8>(b, sort, s, oi, od);
      else if (ix.dim <= 16)
        launch_bf_stl_logged<16>(b, sort, s, oi, od);
      else if (ix.dim <= 36)
        launch_bf_stl_logged<36>(b, sort, s, oi, od);
      else if (ix.dim <= 64)
        launch_bf_stl_logged<64>(b, sort, s, oi, od);
      else
        launch_bf_stl_logged<128>(b, sort, s, oi, od);
      NH_LAUNCH_CHECK();
    }
    return;
  }
  Scratch scratch(s);
  Span<float> pd;
  Span<int32_t> pi;
  scratch.want(pd, size_t(p.chunks) * a.m * a.k);
  scratch.want(pi, size_t(p.chunks) * a.m * a.k);
  scratch.commit(ix.device);
  p.part_d = pd.p;
  p.part_i = pi.p;
  KernelTimer timer(s);
  if (ix.dim <= 4)
    bf_dispatch_k<4>(p, s, a.out_idx, a.out_d2);
  else if (ix.dim <= 8)
    bf_dispatch_k<8>(p, s, a.out_idx, a.out_d2);
  else if (ix.dim <= 16)
    bf_dispatch_k<16>(p, s, a.out_idx, a.out_d2);
  else if (ix.dim <= 36)
    bf_dispatch_k<36>(p, s, a.out_idx, a.out_d2);
  else if (ix.dim <= 64)
    bf_dispatch_k<64>(p, s, a.out_idx, a.out_d2);
  else
    bf_dispatch_k<128>(p, s, a.out_idx, a.out_d2);
  NH_LAUNCH_CHECK();
  (void)a.touched;
}

}  // namespace nabo_hip
