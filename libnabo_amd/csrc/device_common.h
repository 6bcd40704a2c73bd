// Device helpers shared by the kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

// Every product and sum below is rounded on its own, as on the reference's
// x86-64 build without FMA: contraction into fma would change dists2 bits,
// ties and epsilon pruning.
#pragma clang fp contract(off)

namespace nabo_hip {

constexpr int kWave = 64;

// Monotone map of a float onto uint32 so that unsigned min/max order floats.
__device__ __forceinline__ uint32_t float_key(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key_float(uint32_t k) {
  const uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
  return __uint_as_float(u);
}

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, o, kWave));
  return v;
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, kWave));
  return v;
}
__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += (uint32_t)__shfl_xor((int)v, o, kWave);
  return v;
}

// Reduces (seg, value) pairs of the active lanes of a wave per distinct seg
// and applies one atomic per segment. All lanes of the wave must call it.
// OP: 0 = min, 1 = max, 2 = add.
template <int OP>
__device__ __forceinline__ void wave_segmented_atomic(bool active, int seg, uint32_t value,
                                                      uint32_t* target) {
  const uint32_t identity = OP == 0 ? 0xffffffffu : 0u;
  uint64_t pending = __ballot(active);
  const int lane = __lane_id();
  while (pending) {
    const int leader = __ffsll((long long)pending) - 1;
    const int lseg = __shfl(seg, leader, kWave);
    const bool mine = active && seg == lseg;
    const uint64_t group = __ballot(mine);
    uint32_t v = mine ? value : identity;
    if (OP == 0)
      v = wave_min_u32(v);
    else if (OP == 1)
      v = wave_max_u32(v);
    else
      v = wave_sum_u32(v);
    if (lane == leader) {
      if (OP == 0)
        atomicMin(&target[lseg], v);
      else if (OP == 1)
        atomicMax(&target[lseg], v);
      else
        atomicAdd(&target[lseg], v);
    }
    if (mine) active = false;
    pending &= ~group;
  }
}

}  // namespace nabo_hip
