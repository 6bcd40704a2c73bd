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

#include <cfloat>

#include "device_common.h"
#include "internal.h"

namespace nabo_hip {
namespace {

constexpr int kBlock = 256;
constexpr int kTile = 64;

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
};

// DIMP: compile-time bound on dim (query held in VGPRs).
template <int DIMP, int K>
__global__ __launch_bounds__(kBlock) void k_bf_chunk(BfParams a) {
  __shared__ float tile[kTile * DIMP];
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

  for (int64_t t0 = p0; t0 < p1; t0 += kTile) {
    const int tn = int(min<int64_t>(kTile, p1 - t0));
    __syncthreads();
    for (int e = threadIdx.x; e < kTile * DIMP; e += kBlock) {
      const int pt = e / DIMP, d = e % DIMP;
      tile[e] = (pt < tn && d < dim) ? a.cloud[(t0 + pt) * a.cstride + d] : 0.f;
    }
    __syncthreads();
    if (valid) {
      for (int pt = 0; pt < tn; ++pt) {
        const float* pp = tile + pt * DIMP;
        float dist = 0.f;
#pragma unroll
        for (int d = 0; d < DIMP; ++d) {
          if (d < dim) {
            const float diff = pp[d] - q[d];
            dist = dist + diff * diff;
          }
        }
        if (dist <= r2 && dist < hd[K - 1] && (a.allow_self || dist > FLT_EPSILON))
          bf_insert<K>(hd, hi, dist, int32_t(t0 + pt));
      }
    }
  }
  if (!valid) return;
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

template <int DIMP, int K>
void launch_bf(const BfParams& p, hipStream_t s, int32_t* oi, float* od) {
  dim3 grid(unsigned(ceil_div(p.m, kBlock)), unsigned(p.chunks));
  k_bf_chunk<DIMP, K><<<grid, kBlock, 0, s>>>(p);
  k_bf_merge<K><<<unsigned(ceil_div(p.m, kBlock)), kBlock, 0, s>>>(p, oi, od);
}

template <int DIMP>
void bf_dispatch_k(const BfParams& p, hipStream_t s, int32_t* oi, float* od) {
  if (p.k <= 1)
    launch_bf<DIMP, 1>(p, s, oi, od);
  else if (p.k <= 4)
    launch_bf<DIMP, 4>(p, s, oi, od);
  else if (p.k <= 8)
    launch_bf<DIMP, 8>(p, s, oi, od);
  else if (p.k <= 16)
    launch_bf<DIMP, 16>(p, s, oi, od);
  else if (p.k <= 32)
    launch_bf<DIMP, 32>(p, s, oi, od);
  else
    launch_bf<DIMP, 64>(p, s, oi, od);
}

__global__ void k_copy_cloud(int64_t n, int dim, const float* __restrict__ src, int64_t stride,
                             float* __restrict__ dst) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n * dim) return;
  const int64_t p = e / dim;
  const int d = int(e % dim);
  dst[e] = src[p * stride + d];
}

}  // namespace

constexpr int kBfMaxK = 64;
constexpr int kBfMaxDim = 128;

void bf_prepare(Index& ix, const float* d_cloud, int64_t stride, hipStream_t s) {
  if (ix.dim > kBfMaxDim)
    throw Error(NABO_HIP_ERR_UNSUPPORTED, "dim = " + std::to_string(ix.dim) +
                                              " exceeds the HIP brute-force limit of " +
                                              std::to_string(kBfMaxDim));
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
  if (a.k > kBfMaxK)
    throw Error(NABO_HIP_ERR_UNSUPPORTED, "k = " + std::to_string(a.k) +
                                              " exceeds the HIP brute-force limit of " +
                                              std::to_string(kBfMaxK));
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
  const int64_t qblocks = ceil_div(a.m, kBlock);
  int64_t chunks = std::max<int64_t>(1, ceil_div(2048, qblocks));
  chunks = std::min<int64_t>(chunks, std::max<int64_t>(1, ceil_div(ix.n, 1024)));
  chunks = std::min<int64_t>(chunks, 65535);
  p.chunk = ceil_div(ix.n, chunks);
  p.chunks = int(ceil_div(ix.n, p.chunk));
  DevBuf<float> pd(size_t(p.chunks) * a.m * a.k, s);
  DevBuf<int32_t> pi(size_t(p.chunks) * a.m * a.k, s);
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
