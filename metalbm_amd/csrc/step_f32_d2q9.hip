// step_f32_d2q9.hip -- FP32-compute instantiations for D2Q9 (shifted f - w
// storage in FP32 / FP16 / FP64, and unshifted FP32).  Built with FMA
// contraction.
#include "instantiate.hpp"

namespace lbm {

void register_step_f32_d2q9(Registry& r) {
  register_combo<Lattice<LatticeType::D2Q9>, float, float, false, 4, 2>(r, TYPE_F32, TYPE_F32);
  register_combo<Lattice<LatticeType::D2Q9>, float, float, true, 4, 2>(r, TYPE_F32, TYPE_F32);
  register_combo<Lattice<LatticeType::D2Q9>, _Float16, float, true, 8, 4, 2>(r, TYPE_F16, TYPE_F32);
  register_combo<Lattice<LatticeType::D2Q9>, double, float, true, 2>(r, TYPE_F64, TYPE_F32);
}

}  // namespace lbm
