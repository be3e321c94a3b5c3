// lattice.hpp -- velocity sets and the reference's enum vocabulary.
//
// The enum classes carry the integer values of the reference's own enums
// (include/metaLBM/Options.h:12-41) so a configuration written against the
// reference means the same thing here.  Velocity sets are in the reference
// order (include/metaLBM/Lattice.h:188-206 D2Q9, 578-608 D3Q19, 657-699
// D3Q27): index 0 is rest, 1..faceQ have c_x = -1, faceQ+1..2 faceQ are their
// mirrors with c_x = +1.
#pragma once

#include <hip/hip_runtime.h>

namespace lbm {

enum class LatticeType : int {
  Generic = 0, D1Q3, D2Q5, D2Q9, D2Q13, D2Q17, D2Q21, D2Q37,
  D3Q15, D3Q19, D3Q27, D3Q33
};
enum class CollisionType : int { GenericSRT = 0, BGK = 1, ELBM = 2 };
enum class ForcingSchemeType : int {
  Generic = 0, None = 1, Guo = 2, ShanChen = 3, ExactDifferenceMethod = 4
};
enum class ForceType : int {
  None = 0, Generic = 1, GenericTimeIndependent = 2, GenericTimeDependent = 3,
  Constant = 4, Sinusoidal = 5, Kolmogorov = 6
};
enum class EquilibriumType : int {
  Generic = 0, Exact = 1, TruncationMa2 = 2, TruncationMa3 = 3
};

// Lattice<LatticeType>: compile-time descriptor.  celerity(i, d) and
// weight(i) are constexpr so a fully unrolled kernel folds every c_i and w_i
// into immediates.  inv_cs2 = 3 for all three sets (Lattice.h:150, 540, 619).
template <LatticeType LT>
struct Lattice;

template <>
struct Lattice<LatticeType::D2Q9> {
  static constexpr LatticeType Type = LatticeType::D2Q9;
  static constexpr int dimD = 2, dimQ = 9, faceQ = 3;
  static constexpr double inv_cs2 = 3.0;
  static constexpr double cs2 = 1.0 / 3.0;
  static constexpr int C[9][3] = {{0, 0, 0},  {-1, 1, 0}, {-1, 0, 0},
                                  {-1, -1, 0}, {1, -1, 0}, {1, 0, 0},
                                  {1, 1, 0},  {0, -1, 0}, {0, 1, 0}};
  static constexpr double W[9] = {4.0 / 9.0,  1.0 / 36.0, 1.0 / 9.0,
                                  1.0 / 36.0, 1.0 / 36.0, 1.0 / 9.0,
                                  1.0 / 36.0, 1.0 / 9.0,  1.0 / 9.0};
  __host__ __device__ static constexpr int celerity(int i, int d) { return C[i][d]; }
  __host__ __device__ static constexpr double weight(int i) { return W[i]; }
};

template <>
struct Lattice<LatticeType::D3Q19> {
  static constexpr LatticeType Type = LatticeType::D3Q19;
  static constexpr int dimD = 3, dimQ = 19, faceQ = 5;
  static constexpr double inv_cs2 = 3.0;
  static constexpr double cs2 = 1.0 / 3.0;
  static constexpr int C[19][3] = {
      {0, 0, 0},  {-1, 0, 0}, {-1, -1, 0}, {-1, 1, 0},  {-1, 0, -1},
      {-1, 0, 1}, {1, 0, 0},  {1, 1, 0},   {1, -1, 0},  {1, 0, 1},
      {1, 0, -1}, {0, -1, 0}, {0, 0, -1},  {0, -1, -1}, {0, -1, 1},
      {0, 1, 0},  {0, 0, 1},  {0, 1, 1},   {0, 1, -1}};
  static constexpr double W[19] = {
      1.0 / 3.0,  1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
      1.0 / 36.0, 1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
      1.0 / 36.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0,
      1.0 / 18.0, 1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0};
  __host__ __device__ static constexpr int celerity(int i, int d) { return C[i][d]; }
  __host__ __device__ static constexpr double weight(int i) { return W[i]; }
};

template <>
struct Lattice<LatticeType::D3Q27> {
  static constexpr LatticeType Type = LatticeType::D3Q27;
  static constexpr int dimD = 3, dimQ = 27, faceQ = 9;
  static constexpr double inv_cs2 = 3.0;
  static constexpr double cs2 = 1.0 / 3.0;
  static constexpr int C[27][3] = {
      {0, 0, 0},    {-1, 0, 0},   {-1, -1, 0}, {-1, 1, 0},  {-1, 0, -1},
      {-1, 0, 1},   {-1, -1, -1}, {-1, -1, 1}, {-1, 1, -1}, {-1, 1, 1},
      {1, 0, 0},    {1, 1, 0},    {1, -1, 0},  {1, 0, 1},   {1, 0, -1},
      {1, 1, 1},    {1, 1, -1},   {1, -1, 1},  {1, -1, -1}, {0, -1, 0},
      {0, 0, -1},   {0, -1, -1},  {0, -1, 1},  {0, 1, 0},   {0, 0, 1},
      {0, 1, 1},    {0, 1, -1}};
  static constexpr double W[27] = {
      8.0 / 27.0,  2.0 / 27.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,
      1.0 / 54.0,  1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0,
      2.0 / 27.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,
      1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0, 2.0 / 27.0,
      2.0 / 27.0,  1.0 / 54.0,  1.0 / 54.0,  2.0 / 27.0,  2.0 / 27.0,
      1.0 / 54.0,  1.0 / 54.0};
  __host__ __device__ static constexpr int celerity(int i, int d) { return C[i][d]; }
  __host__ __device__ static constexpr double weight(int i) { return W[i]; }
};

}  // namespace lbm
