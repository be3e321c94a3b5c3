// step_f32_d3q27.hip -- FP32-compute instantiations for D3Q27 (shifted f - w
// storage in FP32 / FP16 / FP64, and unshifted FP32).  Built with FMA
// contraction.
#include "instantiate.hpp"

namespace lbm {

void register_step_f32_d3q27(Registry& r) {
  register_combo<Lattice<LatticeType::D3Q27>, float, float, false, 4, 2>(r, TYPE_F32, TYPE_F32);
  register_combo<Lattice<LatticeType::D3Q27>, float, float, true, 4, 2>(r, TYPE_F32, TYPE_F32);
  register_combo<Lattice<LatticeType::D3Q27>, _Float16, float, true, 8, 4, 2>(r, TYPE_F16, TYPE_F32);
  register_combo<Lattice<LatticeType::D3Q27>, double, float, true, 2>(r, TYPE_F64, TYPE_F32);
}

}  // namespace lbm
