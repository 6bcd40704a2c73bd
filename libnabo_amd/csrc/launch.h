// Kernel launches that either run on a stream or are recorded as kernel
// nodes of a HIP graph.
//
// A Launcher made on a stream launches each kernel at once; one made on a
// graph appends each kernel as a node depending on the previous one, so the
// graph replays the same sequence. Both go through the runtime's
// argument-array form (hipLaunchKernel / hipGraphAddKernelNode): the
// arguments are converted to the kernel's parameter types first, so each
// pointer in the array addresses a value of exactly the type the kernel
// reads. A recording keeps its argument values alive as long as the graph
// holder keeps them (take_arguments), whether or not the runtime copies them
// when a node is added.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <tuple>
#include <utility>
#include <vector>

#include "internal.h"

namespace nabo_hip {

class Launcher {
 public:
  explicit Launcher(hipStream_t s) : s_(s) {}
  explicit Launcher(hipGraph_t g) : g_(g) {}

  bool recording() const { return g_ != nullptr; }
  // The recorded launches' argument values (keep while the graph lives).
  std::vector<std::shared_ptr<void>> take_arguments() { return std::move(keep_); }
  hipStream_t stream() const { return s_; }

  template <typename... P, typename... A>
  void operator()(void (*kernel)(P...), dim3 grid, dim3 block, size_t lds, A&&... args) {
    static_assert(sizeof...(P) == sizeof...(A), "argument count differs from the kernel's");
    using Values = std::tuple<std::decay_t<P>...>;
    if (!g_) {
      Values values(std::forward<A>(args)...);
      void* ptrs[sizeof...(P) > 0 ? sizeof...(P) : 1];
      fill(ptrs, values, std::index_sequence_for<P...>{});
      NH_CHECK(hipLaunchKernel(reinterpret_cast<const void*>(kernel), grid, block, ptrs, lds, s_));
      return;
    }
    // Recorded: the argument values outlive the recording (kept with the
    // Launcher, then by whoever holds the graph), whether or not the
    // runtime copies them.
    auto values = std::make_shared<Values>(std::forward<A>(args)...);
    keep_.push_back(values);
    void* ptrs[sizeof...(P) > 0 ? sizeof...(P) : 1];
    fill(ptrs, *values, std::index_sequence_for<P...>{});
    hipKernelNodeParams p{};
    p.func = reinterpret_cast<void*>(kernel);
    p.gridDim = grid;
    p.blockDim = block;
    p.sharedMemBytes = unsigned(lds);
    p.kernelParams = ptrs;
    p.extra = nullptr;
    hipGraphNode_t node;
    NH_CHECK(hipGraphAddKernelNode(&node, g_, last_ ? &last_ : nullptr, last_ ? 1 : 0, &p));
    last_ = node;
  }

 private:
  template <typename Tuple, size_t... I>
  static void fill(void** ptrs, Tuple& t, std::index_sequence<I...>) {
    ((ptrs[I] = static_cast<void*>(&std::get<I>(t))), ...);
  }
  hipStream_t s_ = nullptr;
  hipGraph_t g_ = nullptr;
  hipGraphNode_t last_ = nullptr;
  std::vector<std::shared_ptr<void>> keep_;
};

}  // namespace nabo_hip
