// config_rules.hpp -- host-side rules of the C ABI with no HIP dependency:
// configuration validation and the kernel-shape choice.  Included by
// capi.hip; the host sanitizer build (tests/test_sanitizers.py) compiles it
// for the CPU on its own.
#pragma once

#include <cmath>
#include <string>

#include "../../include/metalbm_amd.h"

namespace lbm {

// LBM_OK, or LBM_E_INVALID with the reason in why.  The reference takes its
// configuration from compile-time constants and has no run-time check.
inline int validate_config(const lbm_config* c, std::string& why) {
  const int lat = c->lattice;
  if (lat != LBM_D2Q9 && lat != LBM_D3Q19 && lat != LBM_D3Q27) {
    why = "lattice must be D2Q9 (3), D3Q19 (9) or D3Q27 (10)";
    return LBM_E_INVALID;
  }
  if (c->collision != LBM_BGK && c->collision != LBM_ELBM) {
    why = "collision must be BGK (1) or ELBM (2)";
    return LBM_E_INVALID;
  }
  if (c->forcing_scheme < LBM_FS_NONE || c->forcing_scheme > LBM_FS_EDM) {
    why = "forcing_scheme must be None (1), Guo (2), ShanChen (3) or EDM (4)";
    return LBM_E_INVALID;
  }
  if (c->force != LBM_FORCE_NONE && (c->force < LBM_FORCE_CONSTANT ||
                                      c->force > LBM_FORCE_TURBULENT_2D)) {
    why = "force must be None (0), Constant (4), Sinusoidal (5), Kolmogorov (6), "
          "ConstantShell (7), EnergyRemoval (8) or Turbulent2D (9)";
    return LBM_E_INVALID;
  }
  if ((c->force == LBM_FORCE_CONSTANT_SHELL || c->force == LBM_FORCE_TURBULENT_2D) &&
      (c->force_kmin < 0 || c->force_kmax < c->force_kmin)) {
    why = "spectral force shell needs 0 <= force_kmin <= force_kmax";
    return LBM_E_INVALID;
  }
  if ((c->force == LBM_FORCE_ENERGY_REMOVAL || c->force == LBM_FORCE_TURBULENT_2D) &&
      (c->removal_kmin < 0 || c->removal_kmax < c->removal_kmin)) {
    why = "energy removal shell needs 0 <= removal_kmin <= removal_kmax";
    return LBM_E_INVALID;
  }
  if (c->equilibrium != LBM_EQ_MA3 && c->equilibrium != LBM_EQ_EXACT) {
    why = "equilibrium must be TruncationMa3 (3) or Exact (1)";
    return LBM_E_INVALID;
  }
  if (c->equilibrium == LBM_EQ_EXACT && lat == LBM_D3Q19) {
    why = "there is no Exact equilibrium for D3Q19";
    return LBM_E_INVALID;
  }
  if (c->nx <= 0 || c->ny <= 0 || c->nz <= 0) {
    why = "nx, ny, nz must be positive";
    return LBM_E_INVALID;
  }
  if (lat == LBM_D2Q9 && c->nz != 1) {
    why = "nz must be 1 for D2Q9";
    return LBM_E_INVALID;
  }
  const double cells = (double)c->nx * c->ny * c->nz;
  if (cells >= 2147483648.0) {
    why = "nx*ny*nz must be below 2^31";
    return LBM_E_INVALID;
  }
  if (!(c->tau > 0.0) || !std::isfinite(c->tau)) {
    why = "tau must be positive and finite";
    return LBM_E_INVALID;
  }
  const int D = lat == LBM_D2Q9 ? 2 : 3;
  if (c->force == LBM_FORCE_SINUSOIDAL)
    for (int d = 0; d < D; ++d)
      if (c->wavelength[d] == 0.0) {
        why = "Sinusoidal force needs nonzero wavelength in every dimension";
        return LBM_E_INVALID;
      }
  if (c->force == LBM_FORCE_KOLMOGOROV && c->wavelength[0] == 0.0) {
    why = "Kolmogorov force needs a nonzero wavelength[0]";
    return LBM_E_INVALID;
  }
  if (c->storage < LBM_F64 || c->storage > LBM_F16 || c->compute < LBM_F64 ||
      c->compute > LBM_F32) {
    why = "storage must be F64/F32/F16 and compute F64/F32";
    return LBM_E_INVALID;
  }
  if (c->storage == LBM_F16 && !c->shifted) {
    why = "F16 storage requires shifted populations (f - w)";
    return LBM_E_INVALID;
  }
  if (c->compute == LBM_F64 && c->storage != LBM_F64) {
    why = "F64 compute requires F64 storage";
    return LBM_E_INVALID;
  }
  if (c->compute == LBM_F32 && c->storage == LBM_F64 && !c->shifted) {
    why = "F32 compute with F64 storage requires shifted populations";
    return LBM_E_INVALID;
  }
  if (c->kernel < LBM_KERNEL_AUTO || c->kernel > LBM_KERNEL_VECTOR) {
    why = "kernel must be AUTO (0), CELL (1) or VECTOR (2)";
    return LBM_E_INVALID;
  }
  if (c->kernel == LBM_KERNEL_VECTOR && c->collision != LBM_BGK) {
    why = "the vector kernel implements BGK only";
    return LBM_E_INVALID;
  }
  if (c->collision == LBM_ELBM && c->alpha_mode != LBM_ALPHA_NEWTON &&
      c->alpha_mode != LBM_ALPHA_LEARNED && c->alpha_mode != LBM_ALPHA_WARM_RESIDUAL &&
      c->alpha_mode != LBM_ALPHA_NEWTON_RESIDUAL) {
    why = "alpha_mode must be NEWTON (0), LEARNED (1), WARM_RESIDUAL (2) or NEWTON_RESIDUAL (3)";
    return LBM_E_INVALID;
  }
  if (c->collision == LBM_ELBM && c->alpha_mode == LBM_ALPHA_LEARNED &&
      lat == LBM_D3Q19) {
    why = "learned alpha is defined for D2Q9 and D3Q27 (tensor moment basis)";
    return LBM_E_INVALID;
  }
  if (c->collision == LBM_ELBM && c->alpha_mode != LBM_ALPHA_NEWTON &&
      !(c->residual_tol > 0.0)) {
    why = "residual_tol must be positive";
    return LBM_E_INVALID;
  }
  return LBM_OK;
}

// Workgroup size of the one-cell-per-lane kernel for a contiguous extent nk:
// one tile per row up to 256 cells, else the largest of 256/192/128 dividing
// the row, else 256.
inline int pick_block(int nk) {
  if (nk <= 256) return ((nk + 63) / 64) * 64;
  for (int b : {256, 192, 128})
    if (nk % b == 0) return b;
  return 256;
}

// Cells per lane V of the vector kernel for a validated configuration, or 0
// for one cell per lane.  instantiated(V) tells whether the vector kernel of
// this configuration exists with that V.
template <class F>
int pick_vector_width(const lbm_config* cfg, F&& instantiated) {
  const int D = cfg->lattice == LBM_D2Q9 ? 2 : 3;
  const int nk = D == 3 ? cfg->nz : cfg->ny;
  const int elem = cfg->storage == LBM_F64 ? 8 : (cfg->storage == LBM_F32 ? 4 : 2);
  // LBM_KERNEL_AUTO takes the vector kernel only where it measured faster
  // than one cell per lane (MI355X, 100-step runs): D3Q19 with FP32 storage
  // at 256^3 (+2%) and FP64 storage with FP32 compute (D3Q27 384^3, +2%).
  // One cell per lane won for D3Q19 FP64 (+12%), D3Q27 FP16 (+10%) and FP32
  // (+4%) storage, and D2Q9 128^2 FP64 (+48%): more wavefronts in flight.
  const bool vector_preferred =
      (cfg->lattice == LBM_D3Q19 && cfg->storage == LBM_F32) ||
      (cfg->storage == LBM_F64 && cfg->compute == LBM_F32);
  if (cfg->collision != LBM_BGK ||
      !(cfg->kernel == LBM_KERNEL_VECTOR ||
        (cfg->kernel == LBM_KERNEL_AUTO && vector_preferred)))
    return 0;
  // Vector width V, in this order (measured on D3Q27 384^3 rows):
  //  1. 16-byte lanes (V sizeof(S) = 16) even if the last 64-lane segment
  //     of a row is partial (FP32 V=4 split 64 + 32 lanes: +3% over V=2);
  //  2. else the widest V whose 64-lane segments tile the row (FP16 V=2,
  //     three full segments: 37.2k MLUPS, V=4 split 64 + 32: 31.8k);
  //  3. else the widest V dividing the row.
  // FP16 V=8 only for D2Q9: 27 x 8 (19 x 8) population values alone fill a
  // wavefront's registers (one wavefront per SIMD, 23.8k MLUPS).
  static const int cands[3][3] = {{2, 0, 0}, {4, 2, 0}, {8, 4, 2}};
  static const int cands_f16_3d[3] = {4, 2, 0};
  const int* cand = (cfg->storage == LBM_F16 && D == 3) ? cands_f16_3d : cands[cfg->storage];
  for (int k = 0; k < 3; ++k) {
    const int V = cand[k];
    if (V != 0 && V * elem == 16 && nk % V == 0 && instantiated(V)) return V;
  }
  for (int k = 0; k < 3; ++k) {
    const int V = cand[k];
    if (V != 0 && nk % (64 * V) == 0 && instantiated(V)) return V;
  }
  for (int k = 0; k < 3; ++k) {
    const int V = cand[k];
    if (V != 0 && nk % V == 0 && instantiated(V)) return V;
  }
  return 0;
}

}  // namespace lbm
