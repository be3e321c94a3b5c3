// step_f32.hip -- FP32-compute instantiations (shifted f - w storage in
// FP32 / FP16 / FP64, and unshifted FP32).  Built with FMA contraction.
#include "instantiate.hpp"

namespace lbm {

void register_step_f32(Registry& r) {
  register_all_lattices<float, float, false, 4, 2>(r, TYPE_F32, TYPE_F32);
  register_all_lattices<float, float, true, 4, 2>(r, TYPE_F32, TYPE_F32);
  register_all_lattices<_Float16, float, true, 8, 4, 2>(r, TYPE_F16, TYPE_F32);
  register_all_lattices<double, float, true, 2>(r, TYPE_F64, TYPE_F32);
}

}  // namespace lbm
