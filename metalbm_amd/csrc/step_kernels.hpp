// step_kernels.hpp -- the fused collide-and-stream-pull kernels.
//
// One kernel launch is one reference time step (Algorithm.h:326-358 with the
// per-cell body of Algorithm.h:97-126): every thread owns one destination
// cell, pulls its Q populations from the source lattice (periodic wrap in the
// index math replaces the halo fill of Communication.h:134-180 and
// Boundary.h:45-102), forms the moments, force, equilibrium, forcing source
// and relaxation in registers, and writes Q populations to the destination
// lattice (two-lattice ping-pong, Algorithm.h:336).
//
// Work decomposition (MI355X): a workgroup of BLOCK threads covers BLOCK
// consecutive cells of one row along the contiguous axis (z in 3D, y in 2D),
// so every population is read and written as whole wavefront-wide 256 B
// (FP32) segments.  The +-1 pull shift along the contiguous axis makes those
// loads straddle one extra 128 B line, which the neighbouring wavefront of
// the same row reads anyway, so the line is served from L2 and HBM traffic
// stays at the algorithmic 2 Q sizeof(storage) bytes per cell.  Tiles are
// remapped so each of the 8 XCDs sweeps a contiguous slab of rows (its L2
// then holds the lines neighbouring tiles share).  Destination stores are
// non-temporal: they are never re-read within the step.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "physics.hpp"

namespace lbm {

// The six ELBM counters (small deviation, capped, solved, fallback, G
// evaluations, accepted corrections) are kept in kStatSlots copies, one
// 128-byte line each; workgroup b adds to copy b mod kStatSlots and the host
// sums the copies.  With one copy every workgroup's atomics met on a single
// cache line and serialised there: at 2048^2 D2Q9 (16 k workgroups) that,
// not the arithmetic, set the entropic step's time.
constexpr int kStatSlots = 256;
constexpr int kStatStride = 16;  // counters per slot: 6 used, padded to 128 bytes

struct StepArgs {
  const void* src;
  void* dst;
  int nx, ny, nz;       // nz == 1 in 2D
  int nk;               // extent of the contiguous axis (nz in 3D, ny in 2D)
  int tiles_per_row;    // ceil(nk / blockDim.x)
  int N;                // cells
  double omega_f;       // 1 - 2 beta            (Collision.h:145)
  double omega_eq;      // 2 beta                (Collision.h:147)
  double beta;          // 1 / (2 tau)           (Collision.h:122)
  double guo_coef;      // 1 - 1 / (2 tau0)      (ForcingScheme.h:115)
  const void* ftab[3];  // per-component force tables (compute type) or NULL
  int fax[3];           // axis indexing table d: 0 x, 1 y, 2 z, 3 constant, 4 cell
  int nt_loads;         // BGK cell kernel: non-temporal population loads
  int store;            // write output fields on this step
  double* rho_out;
  double* uh_out;       // [D][N]
  double* alpha_out;
  double* F_out;        // [D][N]
  void* alpha_state;    // ELBM warm-start field (compute type)
  unsigned long long* stats;  // ELBM counters (kStatSlots slots) or NULL
  // Entropic alpha solver: 0 reference Newton; 1 network start, 2 stored
  // alpha start, each with one certified correction (|G| <= residual_tol or
  // Taylor bound), else the reference Newton; 3 plain Newton from the stored
  // alpha until |G| <= residual_tol.
  int alpha_mode;
  double residual_tol;
  const void* mlp_packed;  // network in its LDS layout (alpha mode 1)
  // (features, alpha) sample log of Newton-solved cells (log_cap 0: off)
  float* log_x;
  double* log_alpha;
  unsigned long long* log_count;
  long long log_cap;
  unsigned log_threshold;  // sample when hash(cell, salt) < threshold
  unsigned log_salt;
  // Alpha modes 1 and 2: cells whose certified correction fails are listed
  // here and solved by a second launch running plain Newton, so the
  // correction kernel does not carry the Newton's registers.
  unsigned* defer_count;
  int* defer_list;  // N entries
};

// Storage / compute codes (values of LBM_F64 / LBM_F32 / LBM_F16).
enum : int { TYPE_F64 = 0, TYPE_F32 = 1, TYPE_F16 = 2 };

// Forcing-scheme selector of the kernels: None and Shan-Chen both have a
// zero collision source (ForcingScheme.h:68-78, 141-151).
enum : int { SRC_ZERO = 0, SRC_GUO = 2, SRC_EDM = 4 };

// Wavefronts per workgroup of the vector kernel (k_bgk_vec).  One: the
// pattern of Q read and Q write streams measured fastest with
// single-wavefront groups (pull-copy probe at 256^3 FP32: 5.45 TB/s at one
// wavefront, 5.36 at two, 5.22 at four, 5.21 at eight).
constexpr int kVecWaves = 1;

__device__ __forceinline__ int xcd_tile(int b, int G) {
  // Workgroups are handed to the 8 XCDs round-robin (XCD = b mod 8).
  if ((G & 7) != 0) return b;
  return (b & 7) * (G >> 3) + (b >> 3);
}

// Population loads, non-temporal when NT.  NT pays where one workgroup
// covers a whole row and the lattices far exceed the caches (D3Q19 256^3
// FP64: +5.5%); where several workgroups share a row, the lines the +-1
// shifts along the row share between them are better kept (D3Q27 384^3 with
// two tiles per row: -7% FP16 storage, -5% FP32 with non-temporal loads;
// A/B on one MI355X).
template <bool NT, class S>
__device__ __forceinline__ S ld(const S* p) {
  if constexpr (NT)
    return __builtin_nontemporal_load(p);
  else
    return *p;
}

template <class L, int i, class S, class C>
__device__ __forceinline__ C to_compute(S v) {
  if constexpr (std::is_same_v<S, _Float16>)
    return C(float(v)) * C(kHalfScaleInv<L, i>);
  else
    return C(v);
}

template <class L, int i, class S, class C>
__device__ __forceinline__ void store_pop(S* p, C v) {
  if constexpr (std::is_same_v<S, _Float16>) {
    const _Float16 h = _Float16(float(v * C(kHalfScale<L, i>)));
    __builtin_nontemporal_store(h, p);
  } else {
    __builtin_nontemporal_store(S(v), p);
  }
}

// Coordinates of the thread's cell and its periodic neighbours.
struct CellPos {
  int x, y, z, xm, xp, ym, yp, zm, zp;
  bool active;
};

template <class L, int BLOCK>
__device__ __forceinline__ CellPos cell_of_thread(const StepArgs& a) {
  CellPos c;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int row = tile / a.tiles_per_row;
  const int k = (tile - row * a.tiles_per_row) * (int)blockDim.x + threadIdx.x;
  c.active = k < a.nk;
  if constexpr (L::dimD == 3) {
    c.x = row / a.ny;
    c.y = row - c.x * a.ny;
    c.z = c.active ? k : 0;
  } else {
    c.x = row;
    c.y = c.active ? k : 0;
    c.z = 0;
  }
  c.xm = c.x == 0 ? a.nx - 1 : c.x - 1;
  c.xp = c.x == a.nx - 1 ? 0 : c.x + 1;
  c.ym = c.y == 0 ? a.ny - 1 : c.y - 1;
  c.yp = c.y == a.ny - 1 ? 0 : c.y + 1;
  c.zm = c.z == 0 ? a.nz - 1 : c.z - 1;
  c.zp = c.z == a.nz - 1 ? 0 : c.z + 1;
  return c;
}

// Position of a cell given by its index (the deferred-solve list).
template <class L>
__device__ __forceinline__ CellPos cell_from_index(const StepArgs& a, int cell) {
  CellPos c;
  c.active = true;
  if constexpr (L::dimD == 3) {
    const unsigned u = (unsigned)cell, t = u / (unsigned)a.nz;
    c.z = (int)(u - t * (unsigned)a.nz);
    c.y = (int)(t % (unsigned)a.ny);
    c.x = (int)(t / (unsigned)a.ny);
  } else {
    c.y = cell % a.ny;
    c.x = cell / a.ny;
    c.z = 0;
  }
  c.xm = c.x == 0 ? a.nx - 1 : c.x - 1;
  c.xp = c.x == a.nx - 1 ? 0 : c.x + 1;
  c.ym = c.y == 0 ? a.ny - 1 : c.y - 1;
  c.yp = c.y == a.ny - 1 ? 0 : c.y + 1;
  c.zm = c.z == 0 ? a.nz - 1 : c.z - 1;
  c.zp = c.z == a.nz - 1 ? 0 : c.z + 1;
  return c;
}

// Cell index of the pull source x - c_i (Collision.h:146), periodic.
template <class L, int i>
__device__ __forceinline__ int pull_index(const StepArgs& a, const CellPos& c) {
  constexpr int cx = cel<L, i, 0>(), cy = cel<L, i, 1>(), cz = cel<L, i, 2>();
  const int xs = cx > 0 ? c.xm : (cx < 0 ? c.xp : c.x);
  const int ys = cy > 0 ? c.ym : (cy < 0 ? c.yp : c.y);
  if constexpr (L::dimD == 3) {
    const int zs = cz > 0 ? c.zm : (cz < 0 ? c.zp : c.z);
    return (xs * a.ny + ys) * a.nz + zs;
  } else {
    return xs * a.ny + ys;
  }
}

template <class L>
__device__ __forceinline__ int self_index(const StepArgs& a, const CellPos& c) {
  if constexpr (L::dimD == 3) return (c.x * a.ny + c.y) * a.nz + c.z;
  else return c.x * a.ny + c.y;
}

// Force_::setForce from the host-built tables (Force.h:104-292) or, for the
// spectral forces, from the per-cell force field (fax 4: Force<Generic>::
// setForce, Force.h:39-48).
template <class L, class C>
__device__ __forceinline__ void force_of_cell(const StepArgs& a,
                                              const CellPos& c, C* F) {
  const int pos[5] = {c.x, c.y, c.z, 0, self_index<L>(a, c)};
  F[0] = F[1] = F[2] = C(0);
#pragma unroll
  for (int d = 0; d < L::dimD; ++d) {
    const C* t = static_cast<const C*>(a.ftab[d]);
    if (t) F[d] = t[pos[a.fax[d]]];
  }
}

// Pull all Q populations of the cell into registers.
template <class L, class S, class C, bool NT = false>
__device__ __forceinline__ void pull(const StepArgs& a, const CellPos& c,
                                     C* f) {
  const S* __restrict__ src = static_cast<const S*>(a.src);
  static_for<L::dimQ>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    const S* base = src + (size_t)i * (size_t)a.N;
    f[i] = to_compute<L, i, S, C>(ld<NT>(base + pull_index<L, i>(a, c)));
  });
}

// ---- reference-order operators (FP64, unshifted) --------------------------

// Guo source, reference order (ForcingScheme.h:99-117), raw velocity.
template <class L, int i, class T>
__device__ __forceinline__ T guo_ref(T gcoef, const T* u, const T* F) {
  const T cu = cdot<L, i>(u);
  const T s3 = T(3.0) * cu;
  T r = ((T(cel<L, i, 0>()) - u[0]) + T(cel<L, i, 0>()) * s3) * F[0];
  r += ((T(cel<L, i, 1>()) - u[1]) + T(cel<L, i, 1>()) * s3) * F[1];
  if constexpr (L::dimD == 3)
    r += ((T(cel<L, i, 2>()) - u[2]) + T(cel<L, i, 2>()) * s3) * F[2];
  return gcoef * T(L::W[i]) * T(3.0) * r;
}

template <class L, int EQ, class T>
struct RefEquilibrium {
  // EQ == Exact keeps the per-cell factors; Ma3 is evaluated directly.
  ExactFactors<L, T> ex;
  T rho, u2;
  const T* u;
  __device__ __forceinline__ RefEquilibrium(T rho_, const T* u_, T u2_)
      : ex(u_), rho(rho_), u2(u2_), u(u_) {}
  template <int i>
  __device__ __forceinline__ T feq() const {
    if constexpr (EQ == (int)EquilibriumType::Exact) return ex.template feq<i>(rho);
    else return feq_ma3_ref<L, i>(rho, u, u2);
  }
};

// ---- shifted / fast operators ----------------------------------------------

template <class L, int EQ, class T>
struct FastEquilibrium {
  // p_i = P_i(u) - 1 for either equilibrium.
  Ma3Coefs<T> ma3;
  ExactShifted<L, T> ex;
  const T* u;
  __device__ __forceinline__ FastEquilibrium(const T* u_, T u2)
      : ma3(u2), ex(u_), u(u_) {}
  template <int i>
  __device__ __forceinline__ T p() const {
    if constexpr (EQ == (int)EquilibriumType::Exact) return ex.template p<i>();
    else return ma3.p(cdot<L, i>(u));
  }
};

}  // namespace lbm
