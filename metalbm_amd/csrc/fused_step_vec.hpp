// fused_step_vec.hpp -- vectorised BGK fused step (the bandwidth kernel).
//
// Each lane owns V consecutive cells along the contiguous axis (z in 3D, y in
// 2D), so every population row is read and written with aligned
// V * sizeof(S) = 8..16 byte vector accesses, and a 64-lane wavefront covers
// a segment of 64 V cells of one row.  The +-1 pull shift along the
// contiguous axis is done in registers: the one value a lane needs from its
// neighbour lane is moved with a DPP full-wavefront rotate (wave_ror:1 /
// wave_rol:1), so no load is misaligned.  When the segment is the whole row
// the rotate already wraps periodically; otherwise lane 0 (or 63) takes the
// value across the segment boundary from one extra uniform load.
//
// Same arithmetic as k_bgk per cell (reference order for FP64 unshifted).
#pragma once

#include "fused_step.hpp"

namespace lbm {

// Lane l receives lane (l - 1) mod 64's value (DPP wave_ror:1).
__device__ __forceinline__ int dpp_from_prev(int v) {
  return __builtin_amdgcn_mov_dpp(v, 0x13C, 0xF, 0xF, false);
}
// Lane l receives lane (l + 1) mod 64's value (DPP wave_rol:1).
__device__ __forceinline__ int dpp_from_next(int v) {
  return __builtin_amdgcn_mov_dpp(v, 0x134, 0xF, 0xF, false);
}

template <bool PREV, class S>
__device__ __forceinline__ S lane_shift(S v) {
  if constexpr (sizeof(S) == 4) {
    const int r = PREV ? dpp_from_prev(__builtin_bit_cast(int, v))
                       : dpp_from_next(__builtin_bit_cast(int, v));
    return __builtin_bit_cast(S, r);
  } else if constexpr (sizeof(S) == 8) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = (int)(b & 0xffffffffLL), hi = (int)(b >> 32);
    const int rlo = PREV ? dpp_from_prev(lo) : dpp_from_next(lo);
    const int rhi = PREV ? dpp_from_prev(hi) : dpp_from_next(hi);
    return __builtin_bit_cast(
        S, (long long)(((unsigned long long)(unsigned)rhi << 32) | (unsigned)rlo));
  } else {
    const int b = (int)__builtin_bit_cast(unsigned short, v);
    const int r = PREV ? dpp_from_prev(b) : dpp_from_next(b);
    return __builtin_bit_cast(S, (unsigned short)r);
  }
}

template <class S, int V>
struct VecT {
  typedef S type __attribute__((ext_vector_type(V)));
};

struct WavePos {
  int x, y, xm, xp, ym, yp;  // row coordinates (3D: x, y; 2D: x only)
  int k0;                    // first contiguous-axis cell of this lane
  int seg_lo, seg_hi;        // segment [seg_lo, seg_hi) along the contiguous axis
  int lane;
  bool whole_row;
};

// Pull population i for the lane's V cells (Collision.h:146, periodic).
template <class L, int i, class S, class C, int V>
__device__ __forceinline__ void pull_vec(const StepArgs& a, const WavePos& w,
                                         C (&f)[V]) {
  constexpr int D = L::dimD;
  constexpr int cx = cel<L, i, 0>(), cy = cel<L, i, 1>(), cz = cel<L, i, 2>();
  constexpr int ck = D == 3 ? cz : cy;  // shift along the contiguous axis
  const int xs = cx > 0 ? w.xm : (cx < 0 ? w.xp : w.x);
  int row_off;
  if constexpr (D == 3) {
    const int ys = cy > 0 ? w.ym : (cy < 0 ? w.yp : w.y);
    row_off = (xs * a.ny + ys) * a.nz;
  } else {
    row_off = xs * a.ny;
  }
  const S* __restrict__ rowp =
      static_cast<const S*>(a.src) + (size_t)i * (size_t)a.N + row_off;
  using vt = typename VecT<S, V>::type;
  const vt v = *reinterpret_cast<const vt*>(rowp + w.k0);
  S s[V];
#pragma unroll
  for (int j = 0; j < V; ++j) s[j] = v[j];
  S out[V];
  if constexpr (ck == 0) {
#pragma unroll
    for (int j = 0; j < V; ++j) out[j] = s[j];
  } else if constexpr (ck > 0) {
    // cell k pulls k - 1
    S edge = lane_shift<true>(s[V - 1]);
    if (!w.whole_row) {
      const int kb = w.seg_lo == 0 ? a.nk - 1 : w.seg_lo - 1;
      const S b = rowp[kb];
      if (w.lane == 0) edge = b;
    }
    out[0] = edge;
#pragma unroll
    for (int j = 1; j < V; ++j) out[j] = s[j - 1];
  } else {
    // cell k pulls k + 1
    S edge = lane_shift<false>(s[0]);
    if (!w.whole_row) {
      const int kb = w.seg_hi == a.nk ? 0 : w.seg_hi;
      const S b = rowp[kb];
      if (w.lane == 63) edge = b;
    }
#pragma unroll
    for (int j = 0; j < V - 1; ++j) out[j] = s[j + 1];
    out[V - 1] = edge;
  }
#pragma unroll
  for (int j = 0; j < V; ++j) f[j] = to_compute<L, i, S, C>(out[j]);
}

template <class L, int i, class S, class C, int V>
__device__ __forceinline__ void store_vec(S* p, const C (&v)[V]) {
  using vt = typename VecT<S, V>::type;
  vt o;
#pragma unroll
  for (int j = 0; j < V; ++j) {
    if constexpr (std::is_same_v<S, _Float16>)
      o[j] = _Float16(float(v[j] * C(half_scale<L, i>())));
    else
      o[j] = S(v[j]);
  }
  __builtin_nontemporal_store(o, reinterpret_cast<vt*>(p));
}

template <class L, class S, class C, int FS, int EQ, bool SH, int V>
__global__ void __launch_bounds__(256) k_bgk_vec(const StepArgs a, int FSX) {
  constexpr int Q = L::dimQ, D = L::dimD;
  constexpr bool REF = std::is_same_v<C, double> && !SH;
  const int wave = xcd_tile(blockIdx.x, gridDim.x) * 4 + (int)(threadIdx.x >> 6);
  const int segs = a.tiles_per_row;   // segments of 64 V cells per row
  const int row = wave / segs;
  const int rows = D == 3 ? a.nx * a.ny : a.nx;
  if (row >= rows) return;            // uniform per wavefront
  WavePos w;
  w.lane = threadIdx.x & 63;
  const int seg = wave - row * segs;
  w.seg_lo = seg * 64 * V;
  w.seg_hi = w.seg_lo + 64 * V;
  w.k0 = w.seg_lo + w.lane * V;
  w.whole_row = segs == 1;
  if constexpr (D == 3) {
    w.x = row / a.ny;
    w.y = row - w.x * a.ny;
  } else {
    w.x = row;
    w.y = 0;
  }
  w.xm = w.x == 0 ? a.nx - 1 : w.x - 1;
  w.xp = w.x == a.nx - 1 ? 0 : w.x + 1;
  w.ym = w.y == 0 ? a.ny - 1 : w.y - 1;
  w.yp = w.y == a.ny - 1 ? 0 : w.y + 1;

  C f[Q][V];
  static_for<Q>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    pull_vec<L, i, S, C, V>(a, w, f[i]);
  });

  const int self_row = D == 3 ? (w.x * a.ny + w.y) * a.nz : w.x * a.ny;
  S* __restrict__ dst = static_cast<S*>(a.dst) + self_row + w.k0;
  const C of = C(a.omega_f), oe = C(a.omega_eq), gc = C(a.guo_coef);

  // Per-cell moments and force.
  C rho[V], drho[V], u[V][3], u2[V], F[V][3];
#pragma unroll
  for (int j = 0; j < V; ++j) {
    C fc[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) fc[i] = f[i][j];
    u[j][0] = u[j][1] = u[j][2] = C(0);
    if constexpr (REF) {
      moments_ref<L>(fc, rho[j], u[j], u2[j]);
      drho[j] = rho[j] - C(1);
    } else if constexpr (SH) {
      moments_shifted<L>(fc, drho[j], rho[j], u[j], u2[j]);
    } else {
      moments_ref<L>(fc, rho[j], u[j], u2[j]);
      drho[j] = rho[j] - C(1);
    }
    CellPos c;
    c.x = w.x;
    c.y = D == 3 ? w.y : w.k0 + j;
    c.z = D == 3 ? w.k0 + j : 0;
    force_of_cell<L, C>(a, c, F[j]);
  }

  if constexpr (REF) {
    static_for<Q>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      C out[V];
#pragma unroll
      for (int j = 0; j < V; ++j) {
        const RefEquilibrium<L, EQ, C> E(rho[j], u[j], u2[j]);
        const C fe = E.template feq<i>();
        C s = C(0);
        if constexpr (FS == SRC_GUO) s = guo_ref<L, i>(gc, u[j], F[j]);
        if constexpr (FS == SRC_EDM) {
          C v[3] = {C(0), C(0), C(0)};
          const C k = C(1.0) / rho[j];
          for (int d = 0; d < D; ++d) v[d] = u[j][d] + F[j][d] * k;
          C v2 = v[0] * v[0];
          for (int d = 1; d < D; ++d) v2 += v[d] * v[d];
          const RefEquilibrium<L, EQ, C> E2(rho[j], v, v2);
          s = E2.template feq<i>() - fe;
        }
        out[j] = of * f[i][j] + oe * fe + s;
      }
      store_vec<L, i, S, C, V>(dst + (size_t)i * a.N, out);
    });
  } else {
    C uF[V];
#pragma unroll
    for (int j = 0; j < V; ++j) {
      uF[j] = u[j][0] * F[j][0];
      for (int d = 1; d < D; ++d) uF[j] += u[j][d] * F[j][d];
    }
    static_for<Q>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      const C w_i = C(L::W[i]);
      C out[V];
#pragma unroll
      for (int j = 0; j < V; ++j) {
        const FastEquilibrium<L, EQ, C> E(u[j], u2[j]);
        const C p = E.template p<i>();
        const C eq = SH ? w_i * (drho[j] + rho[j] * p) : w_i * rho[j] * (C(1) + p);
        C s = C(0);
        if constexpr (FS == SRC_GUO) {
          const C cF = cdot<L, i>(F[j]);
          const C cu = cdot<L, i>(u[j]);
          s = (C(3) * gc * w_i) * (cF - uF[j] + C(3) * cu * cF);
        }
        if constexpr (FS == SRC_EDM) {
          C v[3] = {C(0), C(0), C(0)};
          const C k = C(1) / rho[j];
          for (int d = 0; d < D; ++d) v[d] = u[j][d] + F[j][d] * k;
          C v2 = v[0] * v[0];
          for (int d = 1; d < D; ++d) v2 += v[d] * v[d];
          const FastEquilibrium<L, EQ, C> E2(v, v2);
          s = rho[j] * w_i * (E2.template p<i>() - p);
        }
        out[j] = of * f[i][j] + oe * eq + s;
      }
      store_vec<L, i, S, C, V>(dst + (size_t)i * a.N, out);
    });
  }
  if (a.store) {
#pragma unroll
    for (int j = 0; j < V; ++j) {
      const int cell = self_row + w.k0 + j;
      store_fields<L, C>(a, cell, FSX, rho[j], u[j], F[j], 2.0);
    }
  }
}

}  // namespace lbm
