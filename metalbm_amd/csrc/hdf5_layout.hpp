// hdf5_layout.hpp -- the reference's on-disk field layout, with no HDF5 or
// HIP dependency: dimensions gSD::pLength() -- the global extents with the
// last one padded to 2 (n/2 + 1) for FFTW's in-place real transforms
// (Domain.h:53-57, MathVector.h:331-344) -- and the copies between the dense
// [x][y][z] arrays of the C ABI (z fastest) and that padded layout.  Used by
// hdf5_io.cpp; the host sanitizer build compiles it on its own.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace lbm {

typedef unsigned long long hsize_t;  // HDF5's extent type

struct Grid {
  int D = 0, Q = 0, n[3] = {1, 1, 1};
  int64_t N = 0;
  hsize_t dims[3] = {1, 1, 1};  // padded, as gSD::pLength()
  size_t padded = 0;
};

inline void set_extents(Grid& g, int D, int Q, int nx, int ny, int nz) {
  g.D = D;
  g.Q = Q;
  g.n[0] = nx;
  g.n[1] = ny;
  g.n[2] = D == 3 ? nz : 1;
  g.N = (int64_t)nx * ny * g.n[2];
  g.padded = 1;
  for (int d = 0; d < g.D; ++d) {
    g.dims[d] = (hsize_t)(d == g.D - 1 ? 2 * (g.n[d] / 2 + 1) : g.n[d]);
    g.padded *= (size_t)g.dims[d];
  }
}

// Dense [x][y][z] (z fastest) into the padded layout (padding zero) and back.
inline void to_padded(const Grid& g, const double* src, std::vector<double>& dst) {
  dst.assign(g.padded, 0.0);
  const int64_t rows = g.N / g.n[g.D - 1], len = g.n[g.D - 1];
  const size_t plen = (size_t)g.dims[g.D - 1];
  for (int64_t r = 0; r < rows; ++r)
    for (int64_t k = 0; k < len; ++k) dst[(size_t)r * plen + k] = src[r * len + k];
}

inline void from_padded(const Grid& g, const std::vector<double>& src, double* dst) {
  const int64_t rows = g.N / g.n[g.D - 1], len = g.n[g.D - 1];
  const size_t plen = (size_t)g.dims[g.D - 1];
  for (int64_t r = 0; r < rows; ++r)
    for (int64_t k = 0; k < len; ++k) dst[r * len + k] = src[(size_t)r * plen + k];
}

}  // namespace lbm
