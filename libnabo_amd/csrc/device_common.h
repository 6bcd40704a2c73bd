// Device helpers shared by the kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
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

// The k-best of IndexHeapSTL (nabo/index_heap.h:51-165) for one lane: a
// binary max-heap by value that starts as one sentinel (-1, +inf). Entries
// move as libstdc++'s std::pop_heap, std::push_heap and std::sort_heap move
// them (the reference's heap calls): a pop moves the hole left at the root
// down to a leaf, always to the larger child (the right one unless the left
// is strictly larger), then sifts the refill entry up; a push sifts up while
// the parent is strictly smaller. So the output order, and which entry
// leaves among equal maxima, are the reference's. The root's value
// (headValue) stays in a register. An entry is {index, value bits}; Store
// maps entry j of the lane to memory only that lane touches, so its reads
// see its own writes.
struct LdsHeapStore {  // entry j at base[j * kWave] (slot-major over a wave's lanes)
  uint2* base;
  __device__ __forceinline__ uint2 get(int j) const { return base[j * kWave]; }
  __device__ __forceinline__ void set(int j, uint2 e) { base[j * kWave] = e; }
};
struct StridedHeapStore {  // entry j at base[j * stride] (slot-major over a grid's lanes)
  uint2* base;
  uint32_t stride;
  __device__ __forceinline__ uint2 get(int j) const { return base[size_t(j) * stride]; }
  __device__ __forceinline__ void set(int j, uint2 e) { base[size_t(j) * stride] = e; }
};
struct RowHeapStore {  // entry j in the query's output row: idx[j], val[j]
  int32_t* idx;
  float* val;
  __device__ __forceinline__ uint2 get(int j) const {
    return make_uint2(uint32_t(idx[j]), __float_as_uint(val[j]));
  }
  __device__ __forceinline__ void set(int j, uint2 e) {
    idx[j] = int32_t(e.x);
    val[j] = __uint_as_float(e.y);
  }
};

// IndexHeapBruteForceVector (nabo/index_heap.h:263-369) as a sorted array
// in the query's output row, for k beyond the register-resident sizes: an
// admission shifts the larger entries up by one and lands after the equal
// ones (replaceHead, :322-334); the row is the output as it stands.
struct RowSortedHeap {
  int32_t* idx;
  float* val;
  int k;
  float head;
  __device__ void reset(RowHeapStore st, int kk) {
    idx = st.idx;
    val = st.val;
    k = kk;
    for (int j = 0; j < k; ++j) {
      idx[j] = -1;
      val[j] = INFINITY;
    }
    head = INFINITY;
  }
  __device__ void replace_head(int32_t index, float v) {
    int j = k - 1;
    for (; j > 0; --j) {
      const float prev = val[j - 1];
      if (!(prev > v)) break;
      val[j] = prev;
      idx[j] = idx[j - 1];
    }
    val[j] = v;
    idx[j] = index;
    head = val[k - 1];
  }
  __device__ void sort() {}
  __device__ void write(int32_t*, float*) const {}
};

template <class Store>
struct StlHeap {
  Store st;
  int size;
  int k;
  float head;

  static __device__ __forceinline__ float value(uint2 e) { return __uint_as_float(e.y); }
  __device__ __forceinline__ void reset(Store s, int kk) {
    st = s;
    k = kk;
    size = 1;
    head = INFINITY;
    st.set(0, make_uint2(0xffffffffu, __float_as_uint(INFINITY)));
  }
  __device__ __forceinline__ void put(int j, uint2 e) {
    st.set(j, e);
    if (j == 0) head = value(e);
  }
  // Sift-up of e from `hole` while its parent is smaller, not above `top`.
  __device__ void sift_up(int hole, int top, uint2 e) {
    while (hole > top) {
      const int parent = (hole - 1) / 2;
      const uint2 pe = st.get(parent);
      if (!(value(pe) < value(e))) break;
      put(hole, pe);
      hole = parent;
    }
    put(hole, e);
  }
  // Refills the heap [0, len) whose slot `hole` is empty with e.
  __device__ void adjust(int hole, int len, uint2 e) {
    const int top = hole;
    int child = hole;
    while (child < (len - 1) / 2) {
      child = 2 * (child + 1);
      uint2 ce = st.get(child);
      const uint2 le = st.get(child - 1);
      if (value(ce) < value(le)) {
        --child;
        ce = le;
      }
      put(hole, ce);
      hole = child;
    }
    if ((len & 1) == 0 && child == (len - 2) / 2) {
      child = 2 * (child + 1) - 1;
      put(hole, st.get(child));
      hole = child;
    }
    sift_up(hole, top, e);
  }
  // replaceHead (index_heap.h:103-117): with k entries, pop the root (its
  // slot, the last one, is overwritten at once by the new entry), else
  // append; then push the new entry.
  __device__ void replace_head(int32_t index, float v) {
    if (size == k) {
      if (size > 1) adjust(0, size - 1, st.get(size - 1));
    } else {
      ++size;
    }
    sift_up(size - 1, 0, make_uint2(uint32_t(index), __float_as_uint(v)));
  }
  // sort() when SORT_RESULTS (index_heap.h:120-124).
  __device__ void sort() {
    for (int len = size; len > 1; --len) {
      const uint2 last = st.get(len - 1);
      st.set(len - 1, st.get(0));
      adjust(0, len - 1, last);
    }
  }
  // getData (index_heap.h:130-145): the entries, then (-1, +inf) up to k.
  __device__ void write(int32_t* oi, float* od) const {
    for (int j = 0; j < k; ++j) {
      const uint2 e = j < size ? st.get(j) : make_uint2(0xffffffffu, __float_as_uint(INFINITY));
      oi[j] = int32_t(e.x);
      od[j] = value(e);
    }
  }
};

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
