// probe.hpp -- launcher of the bandwidth probes (probe.hip): the step
// kernel's memory access pattern without its arithmetic.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/metalbm_amd.h"

namespace lbm {

// One probe launch of kind LBM_PROBE_* over the population lattices src ->
// dst of a grid nx x ny x nz (nz = 1 in 2D) with elem bytes per value; the
// contiguous axis must hold a multiple of 16 bytes.
hipError_t launch_bandwidth_probe(int kind, int lattice, const void* src, void* dst,
                                  int nx, int ny, int nz, int elem, hipStream_t st);

}  // namespace lbm
