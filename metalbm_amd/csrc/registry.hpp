// registry.hpp -- run-time selection of the compile-time specialised kernels.
//
// The reference picks its operators at compile time through global aliases
// (Collision_, ForcingScheme_, Equilibrium_, L: Collision.h:913-914,
// ForcingScheme.h:200, Equilibrium.h:330, Lattice.h:804).  Here every
// supported combination is instantiated once and looked up by key, so one
// library serves every configuration without recompiling.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>

#include "step_kernels.hpp"

namespace lbm {

using StepFn = hipError_t (*)(const StepArgs&, int fsx, int block,
                              int grid, hipStream_t);

// vec = 0 selects the one-cell-per-lane kernel; vec = V > 0 the kernel in
// which each lane owns V consecutive cells (aligned vector accesses).
struct KernelKey {
  int lattice, collision, src, eq, storage, compute, shifted, vec;
  int64_t code() const {
    return (((((((int64_t)lattice * 4 + collision) * 8 + src) * 4 + eq) * 4 +
              storage) * 4 + compute) * 2 + shifted) * 16 + vec;
  }
};

struct Registry {
  std::map<int64_t, StepFn> step;
  void add(const KernelKey& k, StepFn f) { step[k.code()] = f; }
  StepFn find(const KernelKey& k) const {
    auto it = step.find(k.code());
    return it == step.end() ? nullptr : it->second;
  }
};

void register_step_f64_d2q9(Registry& r);
void register_step_f64_d3q19(Registry& r);
void register_step_f64_d3q27(Registry& r);
void register_step_f32_d2q9(Registry& r);
void register_step_f32_d3q19(Registry& r);
void register_step_f32_d3q27(Registry& r);

const Registry& registry();

}  // namespace lbm
