// instantiate.hpp -- instantiates and registers the fused step kernels for
// one compute type.  Included by exactly one translation unit per compute
// type, so each type can be built with its own floating-point contraction
// mode (FP64: off, for reference-order parity; FP32: fused).
#pragma once

#include "fused_step.hpp"
#include "fused_step_vec.hpp"
#include "registry.hpp"

namespace lbm {

template <class L, class S, class C, int FS, int EQ, bool SH, int CO>
hipError_t launch_step(const StepArgs& a, int fsx, int block, int grid,
                       hipStream_t st) {
  if constexpr (CO == (int)CollisionType::BGK) {
    if (a.nt_loads)
      hipLaunchKernelGGL((k_bgk<L, S, C, FS, EQ, SH, 256, true>), dim3(grid),
                         dim3(block), 0, st, a, fsx);
    else
      hipLaunchKernelGGL((k_bgk<L, S, C, FS, EQ, SH, 256, false>), dim3(grid),
                         dim3(block), 0, st, a, fsx);
  }
  else if (a.alpha_mode == 0)
    hipLaunchKernelGGL((k_elbm<L, S, C, FS, EQ, SH, 256, 0>), dim3(grid),
                       dim3(block), 0, st, a, fsx);
  else if (a.alpha_mode == 3)
    // plain Newton from the stored alpha to the residual tolerance (MODE 4)
    hipLaunchKernelGGL((k_elbm<L, S, C, FS, EQ, SH, 256, 4>), dim3(grid),
                       dim3(block), 0, st, a, fsx);
  else {
    // Modes 1 and 2: the certified correction, then plain Newton for the
    // cells it deferred (MODE 3, striding over the list).
    hipLaunchKernelGGL((k_defer_reset<0>), dim3(1), dim3(64), 0, st, a.defer_count);
    if (a.alpha_mode == 2) {
      hipLaunchKernelGGL((k_elbm<L, S, C, FS, EQ, SH, 256, 2>), dim3(grid),
                         dim3(block), 0, st, a, fsx);
    } else if constexpr (MlpShape<L>::NIN > 0) {
      // The learned solver stages the network in LDS (dynamic shared memory).
      using M = MfmaLayout<MlpShape<L>::NIN, MlpShape<L>::H1, MlpShape<L>::H2>;
      hipLaunchKernelGGL((k_elbm<L, S, C, FS, EQ, SH, 256, 1>), dim3(grid),
                         dim3(block), M::BYTES, st, a, fsx);
    } else {
      return hipErrorInvalidValue;
    }
    // a capped grid striding over the list, or (log-form plain Newton) one
    // thread per possible entry
    const int full = (a.N + 255) / 256;
    hipLaunchKernelGGL((k_elbm<L, S, C, FS, EQ, SH, 256, 3>),
                       dim3(kLogResidual<L> ? full : std::min(full, kDeferBlocks)), dim3(256),
                       0, st, a, fsx);
  }
  return hipGetLastError();
}

template <class L, class S, class C, int FS, int EQ, bool SH, int V>
hipError_t launch_step_vec(const StepArgs& a, int fsx, int block, int grid,
                           hipStream_t st) {
  (void)block;
  hipLaunchKernelGGL((k_bgk_vec<L, S, C, FS, EQ, SH, V>), dim3(grid), dim3(64 * kVecWaves),
                     0, st, a, fsx);
  return hipGetLastError();
}

template <class L, class S, class C, bool SH, int... Vs>
void register_combo(Registry& r, int storage, int compute) {
  auto add_eq = [&](auto eqc) {
    constexpr int EQ = decltype(eqc)::value;
    auto add_fs = [&](auto fsc) {
      constexpr int FS = decltype(fsc)::value;
      r.add({(int)L::Type, (int)CollisionType::BGK, FS, EQ, storage, compute, SH, 0},
            &launch_step<L, S, C, FS, EQ, SH, (int)CollisionType::BGK>);
      r.add({(int)L::Type, (int)CollisionType::ELBM, FS, EQ, storage, compute, SH, 0},
            &launch_step<L, S, C, FS, EQ, SH, (int)CollisionType::ELBM>);
      (r.add({(int)L::Type, (int)CollisionType::BGK, FS, EQ, storage, compute, SH, Vs},
             &launch_step_vec<L, S, C, FS, EQ, SH, Vs>), ...);
    };
    add_fs(std::integral_constant<int, SRC_ZERO>{});
    add_fs(std::integral_constant<int, SRC_GUO>{});
    add_fs(std::integral_constant<int, SRC_EDM>{});
  };
  add_eq(std::integral_constant<int, (int)EquilibriumType::TruncationMa3>{});
  // There is no Exact equilibrium for D3Q19 (Equilibrium.h:36-126).
  if constexpr (L::Type != LatticeType::D3Q19)
    add_eq(std::integral_constant<int, (int)EquilibriumType::Exact>{});
}

}  // namespace lbm
