// step_f64_d2q9.hip -- FP64-compute instantiations for D2Q9.  Built with
// -ffp-contract=off: with unshifted FP64 storage these kernels evaluate every
// operator in the reference's operation order, separately rounded, for
// bitwise parity with the FP64 oracle wherever no libm call is involved.
#include "instantiate.hpp"

namespace lbm {

void register_step_f64_d2q9(Registry& r) {
  register_combo<Lattice<LatticeType::D2Q9>, double, double, false, 2>(r, TYPE_F64, TYPE_F64);
  register_combo<Lattice<LatticeType::D2Q9>, double, double, true, 2>(r, TYPE_F64, TYPE_F64);
}

}  // namespace lbm
