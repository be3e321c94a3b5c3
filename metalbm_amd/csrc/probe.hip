// probe.hip -- bandwidth probes with the step kernel's memory access pattern
// and none of its arithmetic: the measured ceiling the roofline of the
// vector step kernel (k_bgk_vec) is read against.  No reference counterpart.
//
// Every probe uses k_bgk_vec's work decomposition and cache policies: 16
// bytes per lane, one 64-lane wavefront per row segment of 64 lanes,
// kVecWaves wavefronts per workgroup, the XCD-aware tile order, per-row
// buffer resources, non-temporal loads and stores.  Kinds (LBM_PROBE_*):
//   PULL_COPY  each lane reads its Q rows at the pull sources (x - c_x,
//              y - c_y; the +-1 shift along the contiguous axis is left out:
//              it costs no HBM traffic, its lines are shared with the
//              neighbouring lanes) and writes them to its own row -- the
//              step's Q read and Q write streams;
//   READ       the same Q reads, no stores (one store per lane only if the
//              XOR of everything read equals an improbable pattern, so the
//              loads cannot be removed);
//   WRITE      the Q stores only;
//   COPY       one contiguous copy of the same bytes: each wavefront copies
//              Q consecutive KiB (one read and one write stream).
#include <hip/hip_runtime.h>

#include "lattice.hpp"
#include "physics.hpp"
#include "probe.hpp"
#include "step_kernels.hpp"

namespace lbm {
namespace {

typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
constexpr int kRawBuffer = 0x00020000;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t row_rsrc(const void* base, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes,
                                           kRawBuffer);
}

struct ProbeArgs {
  const char* src;
  char* dst;
  int nx, ny, nk;       // nk: extent of the contiguous axis (cells)
  int elem;             // bytes per stored value
  int segs;             // 64-lane segments per row
  long long pop_bytes;  // bytes of one population array (N * elem)
};

constexpr int kLoadNT = 2;  // the step kernel's load policy (non-temporal)

template <class L, int KIND>
__global__ void __launch_bounds__(64 * kVecWaves) k_probe(const ProbeArgs a) {
  constexpr int Q = L::dimQ, D = L::dimD;
  const int wave = xcd_tile(blockIdx.x, gridDim.x) * kVecWaves +
                   __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  if constexpr (KIND == LBM_PROBE_COPY) {
    // Q consecutive KiB per wavefront
    const long long total = a.pop_bytes * Q;
    const long long base = (long long)wave * Q * 1024;
    if (base >= total) return;
    const __amdgpu_buffer_rsrc_t rs = row_rsrc(a.src + base, (unsigned)min(total - base, (long long)Q * 1024));
    const __amdgpu_buffer_rsrc_t rd = row_rsrc(a.dst + base, (unsigned)min(total - base, (long long)Q * 1024));
    u32x4 v[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i)
      v[i] = __builtin_amdgcn_raw_buffer_load_b128(rs, (unsigned)(i * 1024 + lane * 16), 0,
                                                   kLoadNT);
#pragma unroll
    for (int i = 0; i < Q; ++i)
      __builtin_amdgcn_raw_buffer_store_b128(v[i], rd, (unsigned)(i * 1024 + lane * 16), 0, 2);
    return;
  } else {
    const int row = wave / a.segs;
    const int rows = D == 3 ? a.nx * a.ny : a.nx;
    if (row >= rows) return;
    const int seg = wave - row * a.segs;
    const int lanes_per_row = a.nk * a.elem / 16;
    const int na = min(64, lanes_per_row - seg * 64);
    const bool active = lane < na;
    const unsigned voff = (unsigned)((seg * 64 + (active ? lane : 0)) * 16);
    const unsigned row_bytes = (unsigned)(a.nk * a.elem);
    int x, y;
    if constexpr (D == 3) {
      x = row / a.ny;
      y = row - x * a.ny;
    } else {
      x = row;
      y = 0;
    }
    const int xm = x == 0 ? a.nx - 1 : x - 1, xp = x == a.nx - 1 ? 0 : x + 1;
    const int ym = y == 0 ? a.ny - 1 : y - 1, yp = y == a.ny - 1 ? 0 : y + 1;
    u32x4 v[Q];
    if constexpr (KIND != LBM_PROBE_WRITE) {
      static_for<Q>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        constexpr int cx = cel<L, i, 0>(), cy = cel<L, i, 1>();
        const int xs = cx > 0 ? xm : (cx < 0 ? xp : x);
        int src_row;
        if constexpr (D == 3) {
          const int ys = cy > 0 ? ym : (cy < 0 ? yp : y);
          src_row = xs * a.ny + ys;
        } else {
          src_row = xs;
        }
        const char* p = a.src + (long long)i * a.pop_bytes + (long long)src_row * row_bytes;
        v[i] = __builtin_amdgcn_raw_buffer_load_b128(row_rsrc(p, row_bytes), voff, 0, kLoadNT);
      });
    } else {
#pragma unroll
      for (int i = 0; i < Q; ++i)
        v[i] = u32x4{(unsigned)i, (unsigned)lane, (unsigned)row, 0x3f800000u};
    }
    if constexpr (KIND == LBM_PROBE_READ) {
      unsigned acc = 0u;
#pragma unroll
      for (int i = 0; i < Q; ++i) acc ^= v[i][0] ^ v[i][1] ^ v[i][2] ^ v[i][3];
      if (acc == 0x7fb3c5a1u && active) {
        char* p = a.dst + (long long)row * row_bytes;
        __builtin_amdgcn_raw_buffer_store_b32(acc, row_rsrc(p, row_bytes), voff, 0, 2);
      }
    } else {
      static_for<Q>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        char* p = a.dst + (long long)i * a.pop_bytes + (long long)row * row_bytes;
        if (active)
          __builtin_amdgcn_raw_buffer_store_b128(v[i], row_rsrc(p, row_bytes), voff, 0, 2);
      });
    }
  }
}

template <class L>
hipError_t launch_for(int kind, const ProbeArgs& a, int grid, hipStream_t st) {
  switch (kind) {
    case LBM_PROBE_PULL_COPY:
      hipLaunchKernelGGL((k_probe<L, LBM_PROBE_PULL_COPY>), dim3(grid), dim3(64 * kVecWaves), 0, st,
                         a);
      break;
    case LBM_PROBE_READ:
      hipLaunchKernelGGL((k_probe<L, LBM_PROBE_READ>), dim3(grid), dim3(64 * kVecWaves), 0, st,
                         a);
      break;
    case LBM_PROBE_WRITE:
      hipLaunchKernelGGL((k_probe<L, LBM_PROBE_WRITE>), dim3(grid), dim3(64 * kVecWaves), 0, st,
                         a);
      break;
    default:
      hipLaunchKernelGGL((k_probe<L, LBM_PROBE_COPY>), dim3(grid), dim3(64 * kVecWaves), 0, st,
                         a);
      break;
  }
  return hipGetLastError();
}

}  // namespace

hipError_t launch_bandwidth_probe(int kind, int lattice, const void* src, void* dst,
                                  int nx, int ny, int nz, int elem, hipStream_t st) {
  const int D = lattice == (int)LatticeType::D2Q9 ? 2 : 3;
  const int Q = lattice == (int)LatticeType::D2Q9 ? 9
                : (lattice == (int)LatticeType::D3Q19 ? 19 : 27);
  ProbeArgs a;
  a.src = static_cast<const char*>(src);
  a.dst = static_cast<char*>(dst);
  a.nx = nx;
  a.ny = D == 3 ? ny : 1;
  a.nk = D == 3 ? nz : ny;
  a.elem = elem;
  const int lanes_per_row = a.nk * elem / 16;
  a.segs = (lanes_per_row + 63) / 64;
  a.pop_bytes = (long long)nx * ny * nz * elem;
  const long long rows = D == 3 ? (long long)nx * ny : nx;
  long long waves = rows * a.segs;
  if (kind == LBM_PROBE_COPY) waves = (a.pop_bytes * Q + (long long)Q * 1024 - 1) / ((long long)Q * 1024);
  const int grid = (int)((waves + kVecWaves - 1) / kVecWaves);
  switch (lattice) {
    case (int)LatticeType::D2Q9:
      return launch_for<Lattice<LatticeType::D2Q9>>(kind, a, grid, st);
    case (int)LatticeType::D3Q19:
      return launch_for<Lattice<LatticeType::D3Q19>>(kind, a, grid, st);
    default:
      return launch_for<Lattice<LatticeType::D3Q27>>(kind, a, grid, st);
  }
}

}  // namespace lbm
