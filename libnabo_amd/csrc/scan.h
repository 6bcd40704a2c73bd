// Exclusive prefix sums over device arrays whose length may live in device
// memory, in three launches of fixed shape:
//   k_scan_partials: workgroup b sums its chunk of the input (whole tiles
//                    of kScanTile elements; chunks cover [0, count));
//   k_scan_spine:    one workgroup turns the kScanGrid sums into chunk
//                    offsets, and writes the grand total;
//   k_scan_apply:    workgroup b scans its chunk again from its offset.
// The count is taken from the host or read from device memory (a kd-tree
// level's segment count), and a gate word read from device memory can make
// all three return at once, so the same launches serve every level of a
// build recorded in a graph.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_common.h"
#include "launch.h"

namespace nabo_hip {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;
constexpr int kScanGrid = 1024;  // workgroups of the partial and apply passes

// Element count: *dev when dev is set, else host. gate (nullable): the
// launches do nothing while *gate == 0.
struct ScanCount {
  const int32_t* dev = nullptr;
  int64_t host = 0;
  const int32_t* gate = nullptr;
};

__device__ __forceinline__ bool scan_gated(const ScanCount& c) { return c.gate && *c.gate == 0; }
__device__ __forceinline__ int64_t scan_count(const ScanCount& c) {
  return c.dev ? int64_t(*c.dev) : c.host;
}
// Elements per workgroup: whole tiles, the same split in both passes.
__device__ __forceinline__ int64_t scan_chunk(int64_t count) {
  const int64_t tiles = (count + kScanTile - 1) / kScanTile;
  return (tiles + kScanGrid - 1) / kScanGrid * kScanTile;
}

// Exclusive scan of one value per thread over a workgroup of kScanThreads;
// *total receives the sum (all threads).
template <class T>
__device__ __forceinline__ T block_exclusive(T v, T* total) {
  constexpr int kWaves = kScanThreads / kWave;
  __shared__ T wave_sum[kWaves];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  T incl = v;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const T up = __shfl_up(incl, o, kWave);
    if (lane >= o) incl += up;
  }
  if (lane == kWave - 1) wave_sum[wave] = incl;
  __syncthreads();
  T before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) {
    const T s = wave_sum[w];
    before += w < wave ? s : T(0);
    all += s;
  }
  __syncthreads();  // wave_sum is reused by the next call
  *total = all;
  return before + incl - v;
}

template <class T>
__global__ __launch_bounds__(kScanThreads) void k_scan_partials(const T* __restrict__ in,
                                                                ScanCount cnt,
                                                                T* __restrict__ partials) {
  if (scan_gated(cnt)) return;
  const int64_t count = scan_count(cnt);
  const int64_t chunk = scan_chunk(count);
  const int64_t begin = int64_t(blockIdx.x) * chunk;
  const int64_t end = begin + chunk < count ? begin + chunk : count;
  T sum = 0;
  for (int64_t i = begin + threadIdx.x; i < end; i += kScanThreads) sum += in[i];
  T total;
  block_exclusive(sum, &total);
  if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

// One workgroup of kScanGrid threads.
template <class T>
__global__ __launch_bounds__(kScanGrid) void k_scan_spine(ScanCount cnt, T* __restrict__ partials,
                                                          T* __restrict__ total_out) {
  if (scan_gated(cnt)) return;
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  constexpr int kWaves = kScanGrid / kWave;
  __shared__ T wave_sum[kWaves];
  const T v = partials[threadIdx.x];
  T incl = v;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const T up = __shfl_up(incl, o, kWave);
    if (lane >= o) incl += up;
  }
  if (lane == kWave - 1) wave_sum[wave] = incl;
  __syncthreads();
  T before = 0, all = 0;
  for (int w = 0; w < kWaves; ++w) {
    before += w < wave ? wave_sum[w] : T(0);
    all += wave_sum[w];
  }
  partials[threadIdx.x] = before + incl - v;
  if (threadIdx.x == 0 && total_out) *total_out = all;
}

// Tiles go through LDS: loaded and stored with consecutive threads on
// consecutive elements, scanned with kScanItems consecutive elements per
// thread (element e at e + e / kScanItems, so those reads spread over the
// banks).
__device__ __forceinline__ int scan_pad(int e) { return e + e / kScanItems; }

template <class T>
__global__ __launch_bounds__(kScanThreads) void k_scan_apply(const T* __restrict__ in,
                                                             T* __restrict__ out, ScanCount cnt,
                                                             const T* __restrict__ partials) {
  if (scan_gated(cnt)) return;
  __shared__ T tv[kScanTile + kScanTile / kScanItems];
  const int64_t count = scan_count(cnt);
  const int64_t chunk = scan_chunk(count);
  const int64_t begin = int64_t(blockIdx.x) * chunk;
  const int64_t end = begin + chunk < count ? begin + chunk : count;
  const int t = threadIdx.x;
  T run = partials[blockIdx.x];
  for (int64_t tile = begin; tile < end; tile += kScanTile) {
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
      const int e = j * kScanThreads + t;
      tv[scan_pad(e)] = tile + e < end ? in[tile + e] : T(0);
    }
    __syncthreads();
    T v[kScanItems];
    T sum = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
      v[j] = tv[scan_pad(t * kScanItems + j)];
      sum += v[j];
    }
    T total;
    T at = run + block_exclusive(sum, &total);
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
      tv[scan_pad(t * kScanItems + j)] = at;
      at += v[j];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
      const int e = j * kScanThreads + t;
      if (tile + e < end) out[tile + e] = tv[scan_pad(e)];
    }
    run += total;
    __syncthreads();
  }
}

// out[i] = in[0] + ... + in[i-1] for i < count; *total (nullable) = the sum
// of all count elements. partials: kScanGrid elements of scratch.
template <class T>
void exclusive_scan(Launcher& L, const T* in, T* out, ScanCount cnt, T* partials, T* total) {
  L(k_scan_partials<T>, dim3(kScanGrid), dim3(kScanThreads), 0, in, cnt, partials);
  L(k_scan_spine<T>, dim3(1), dim3(kScanGrid), 0, cnt, partials, total);
  L(k_scan_apply<T>, dim3(kScanGrid), dim3(kScanThreads), 0, in, out, cnt,
    static_cast<const T*>(partials));
}

// The same scan when the kernel that wrote `in` also wrote partials[b], the
// sum of chunk b of the split above (workgroup b of kScanGrid covering
// [b * scan_chunk(count), ...)): the spine and the apply only.
template <class T>
void exclusive_scan_from_partials(Launcher& L, const T* in, T* out, ScanCount cnt, T* partials,
                                  T* total) {
  L(k_scan_spine<T>, dim3(1), dim3(kScanGrid), 0, cnt, partials, total);
  L(k_scan_apply<T>, dim3(kScanGrid), dim3(kScanThreads), 0, in, out, cnt,
    static_cast<const T*>(partials));
}

}  // namespace nabo_hip
