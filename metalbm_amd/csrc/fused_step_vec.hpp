// fused_step_vec.hpp -- vectorised BGK fused step (the bandwidth kernel).
//
// Each lane owns V consecutive cells along the contiguous axis (z in 3D, y in
// 2D), so every population row is read and written with aligned
// V * sizeof(S) = 4..16 byte vector accesses, and a 64-lane wavefront covers
// a segment of up to 64 V cells of one row (the last segment of a row whose
// length is not a multiple of 64 V has fewer active lanes).  The +-1 pull
// shift along the contiguous axis is done in registers: the one value a lane
// needs from its neighbour lane is moved with a cross-lane permute over the
// segment's active lanes, so no load is misaligned.  When the segment is the
// whole row the permute wraps periodically; otherwise the first (or last)
// active lane takes the value across the segment boundary from one extra
// uniform load.
//
// Same arithmetic as k_bgk per cell (reference order for FP64 unshifted).
#pragma once

#include "fused_step.hpp"

namespace lbm {

// Cross-lane move of the one value each lane needs from its neighbour:
// ds_bpermute (LDS crossbar, no LDS allocation) with the source lane index
// precomputed per lane.  DPP wave_ror/wave_rol were tried first and
// rejected: the compiler emits them in place (vdst == vsrc), and a wave64
// rotate across 16-lane rows then reads lanes an earlier pass has already
// overwritten, which corrupted results at full size.
struct LaneXfer {
  // byte addresses of the lanes holding cells k - V and k + V: lanes l - 1 and
  // l + 1 cyclically over the na active lanes of the segment
  int from_prev, from_next;
  __device__ __forceinline__ LaneXfer(int lane, int na)
      : from_prev((lane == 0 ? na - 1 : lane - 1) << 2),
        from_next((lane >= na - 1 ? 0 : lane + 1) << 2) {}
};

template <bool PREV, class S>
__device__ __forceinline__ S lane_shift(const LaneXfer& x, S v) {
  const int addr = PREV ? x.from_prev : x.from_next;
  if constexpr (sizeof(S) == 4) {
    return __builtin_bit_cast(S, __builtin_amdgcn_ds_bpermute(addr, __builtin_bit_cast(int, v)));
  } else if constexpr (sizeof(S) == 8) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_ds_bpermute(addr, (int)(b & 0xffffffffLL));
    const int hi = __builtin_amdgcn_ds_bpermute(addr, (int)(b >> 32));
    return __builtin_bit_cast(
        S, (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
  } else {
    const int r = __builtin_amdgcn_ds_bpermute(
        addr, (int)__builtin_bit_cast(unsigned short, v));
    return __builtin_bit_cast(S, (unsigned short)r);
  }
}

template <class S, int V>
struct VecT {
  typedef S type __attribute__((ext_vector_type(V)));
};

struct WavePos {
  int x, y, xm, xp, ym, yp;  // row coordinates (3D: x, y; 2D: x only)
  LaneXfer xfer{0, 64};
  int k0;                    // first contiguous-axis cell of this lane
  int seg_lo, seg_hi;        // segment [seg_lo, seg_hi) along the contiguous axis
  int lane;
  int last;                  // last active lane (na - 1)
  bool active;               // lane < na: owns cells and stores them
  bool whole_row;
};

// Buffer-resource access (MUBUF): one SGPR resource per population row, its
// 64-bit base formed in scalar registers (row start = lattice + i N + row
// offset) and its extent the row's nk values, plus a single 32-bit lane
// offset shared by all Q loads and stores, so no per-lane 64-bit address
// arithmetic is needed and no 32-bit byte count or offset depends on the
// lattice size (a resource spanning a whole population would overflow its
// 32-bit num_records from 2^32 bytes on, e.g. D3Q19 FP32 at 1024^3).
// Accesses beyond num_records return 0 / are dropped.
constexpr int kBufferDword3 = 0x00020000;  // raw buffer, 32-bit data format

// Cache policy of the population loads: non-temporal (each value is read
// once per step from HBM; the shifted-row lines neighbouring wavefronts
// share still meet in the L2).  Measured on the same probe: +3.8% over the
// default policy, +8.5% together with single-wavefront groups.
constexpr int kLoadPolicy = 2;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* base,
                                                          unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0,
                                           (int)bytes, kBufferDword3);
}

template <int BYTES>
struct BufIO;
template <>
struct BufIO<16> {
  typedef unsigned raw __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ raw load(__amdgpu_buffer_rsrc_t r, unsigned vo, int so) {
    return __builtin_amdgcn_raw_buffer_load_b128(r, vo, so, kLoadPolicy);
  }
  static __device__ __forceinline__ void store(raw v, __amdgpu_buffer_rsrc_t r, unsigned vo, int so) {
    __builtin_amdgcn_raw_buffer_store_b128(v, r, vo, so, 2 /* nt */);
  }
};
template <>
struct BufIO<8> {
  typedef unsigned raw __attribute__((ext_vector_type(2)));
  static __device__ __forceinline__ raw load(__amdgpu_buffer_rsrc_t r, unsigned vo, int so) {
    return __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, kLoadPolicy);
  }
  static __device__ __forceinline__ void store(raw v, __amdgpu_buffer_rsrc_t r, unsigned vo, int so) {
    __builtin_amdgcn_raw_buffer_store_b64(v, r, vo, so, 2);
  }
};
template <>
struct BufIO<4> {
  typedef unsigned raw;
  static __device__ __forceinline__ raw load(__amdgpu_buffer_rsrc_t r, unsigned vo, int so) {
    return __builtin_amdgcn_raw_buffer_load_b32(r, vo, so, kLoadPolicy);
  }
  static __device__ __forceinline__ void store(raw v, __amdgpu_buffer_rsrc_t r, unsigned vo, int so) {
    __builtin_amdgcn_raw_buffer_store_b32(v, r, vo, so, 2);
  }
};

// Pull population i for the lane's V cells (Collision.h:146, periodic).
// SCALED: FP16 values stay in their storage scale (converted, not unscaled).
template <class L, int i, class S, class C, int V, bool SCALED = false>
__device__ __forceinline__ void pull_vec(const StepArgs& a, const WavePos& w,
                                         unsigned voff, C (&f)[V]) {
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
  const S* row = static_cast<const S*>(a.src) + (size_t)i * (size_t)a.N + (size_t)row_off;
  const __amdgpu_buffer_rsrc_t r = make_rsrc(row, (unsigned)a.nk * (unsigned)sizeof(S));
  using vt = typename VecT<S, V>::type;
  using IO = BufIO<V * (int)sizeof(S)>;
  const vt v = __builtin_bit_cast(vt, IO::load(r, voff, 0));
  S s[V];
#pragma unroll
  for (int j = 0; j < V; ++j) s[j] = v[j];
  S out[V];
  if constexpr (ck == 0) {
#pragma unroll
    for (int j = 0; j < V; ++j) out[j] = s[j];
  } else if constexpr (ck > 0) {
    // cell k pulls k - 1
    S edge = lane_shift<true>(w.xfer, s[V - 1]);
    if (!w.whole_row) {
      const int kb = w.seg_lo == 0 ? a.nk - 1 : w.seg_lo - 1;
      const S b = row[kb];
      if (w.lane == 0) edge = b;
    }
    out[0] = edge;
#pragma unroll
    for (int j = 1; j < V; ++j) out[j] = s[j - 1];
  } else {
    // cell k pulls k + 1
    S edge = lane_shift<false>(w.xfer, s[0]);
    if (!w.whole_row) {
      const int kb = w.seg_hi == a.nk ? 0 : w.seg_hi;
      const S b = row[kb];
      if (w.lane == w.last) edge = b;
    }
#pragma unroll
    for (int j = 0; j < V - 1; ++j) out[j] = s[j + 1];
    out[V - 1] = edge;
  }
#pragma unroll
  for (int j = 0; j < V; ++j)
    f[j] = SCALED ? C(float(out[j])) : to_compute<L, i, S, C>(out[j]);
}

// SCALED: v is already in population i's FP16 storage scale.  Only active
// lanes store.
template <class L, int i, class S, class C, int V, bool SCALED = false>
__device__ __forceinline__ void store_vec(const StepArgs& a, const WavePos& w,
                                          int self_row, unsigned voff,
                                          const C (&v)[V]) {
  using vt = typename VecT<S, V>::type;
  using IO = BufIO<V * (int)sizeof(S)>;
  vt o;
#pragma unroll
  for (int j = 0; j < V; ++j) {
    if constexpr (std::is_same_v<S, _Float16>)
      o[j] = SCALED ? _Float16(float(v[j])) : _Float16(float(v[j] * C(kHalfScale<L, i>)));
    else
      o[j] = S(v[j]);
  }
  S* row = static_cast<S*>(a.dst) + (size_t)i * (size_t)a.N + (size_t)self_row;
  // soffset is the constant 0.  With a register soffset the compiler assumes
  // a >8-byte store cannot be corrupted by a VALU write to its data VGPRs
  // right after issue, and emits none of the wait states it otherwise
  // inserts; on gfx950 that overwrote store data at full size.
  if (w.active)
    IO::store(__builtin_bit_cast(typename IO::raw, o),
              make_rsrc(row, (unsigned)a.nk * (unsigned)sizeof(S)), voff, 0);
}

template <class L, class S, class C, int FS, int EQ, bool SH, int V>
__global__ void __launch_bounds__(64 * kVecWaves) k_bgk_vec(const StepArgs a, int FSX) {
  constexpr int Q = L::dimQ, D = L::dimD;
  constexpr bool REF = std::is_same_v<C, double> && !SH;
  // threadIdx.x >> 6 is the wavefront index: uniform, which readfirstlane
  // makes visible to the compiler so all row addressing stays in SGPRs.
  const int wave = xcd_tile(blockIdx.x, gridDim.x) * kVecWaves +
                   __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int segs = a.tiles_per_row;   // segments of 64 V cells per row
  const int row = wave / segs;
  const int rows = D == 3 ? a.nx * a.ny : a.nx;
  if (row >= rows) return;            // uniform per wavefront
  WavePos w;
  w.lane = threadIdx.x & 63;
  const int seg = wave - row * segs;
  w.seg_lo = seg * 64 * V;
  // Active lanes of this segment: 64, or fewer in the last segment of a row
  // whose length is not a multiple of 64 V (each lane owns V whole cells).
  const int na = min(64, (a.nk - w.seg_lo) / V);
  w.last = na - 1;
  w.active = w.lane < na;
  w.xfer = LaneXfer(w.lane, na);
  w.seg_hi = w.seg_lo + na * V;
  // Inactive lanes mirror lane 0's cells so every address they form is valid.
  w.k0 = w.seg_lo + (w.active ? w.lane * V : 0);
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

  const unsigned voff = (unsigned)w.k0 * (unsigned)sizeof(S);
  // FP16 populations stay in their per-population power-of-two storage scale
  // s_i throughout (the scales are folded into the moment and collision
  // constants).
  constexpr bool SC = std::is_same_v<S, _Float16>;
  constexpr bool FOLD = !REF && EQ == (int)EquilibriumType::TruncationMa3;
  C f[Q][V];
  static_for<Q>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    pull_vec<L, i, S, C, V, SC>(a, w, voff, f[i]);
  });

  const int self_row = D == 3 ? (w.x * a.ny + w.y) * a.nz : w.x * a.ny;
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
      moments_shifted_scaled<L, C, SC>(fc, drho[j], rho[j], u[j], u2[j]);
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
      store_vec<L, i, S, C, V>(a, w, self_row, voff, out);
    });
  } else if constexpr (FOLD) {
    // Ma3 with the density folded into the polynomial's coefficients once per
    // cell: t_i = (eq_i - (SH ? w_i : 0)) / w_i in four fused multiply-adds;
    // out = of f + oe eq + S computed in the population's storage scale.
    C uF[V];
    Ma3Folded<C> M[V];
#pragma unroll
    for (int j = 0; j < V; ++j) {
      uF[j] = u[j][0] * F[j][0];
      for (int d = 1; d < D; ++d) uF[j] += u[j][d] * F[j][d];
      M[j] = Ma3Folded<C>(rho[j], SH ? drho[j] : rho[j], u2[j]);
    }
    // EDM: the equilibrium at v = u + F/rho, same folding (ForcingScheme.h:184-197).
    Ma3Folded<C> M2[V];
    C v[V][3];
    if constexpr (FS == SRC_EDM) {
#pragma unroll
      for (int j = 0; j < V; ++j) {
        const C k = C(1) / rho[j];
        v[j][0] = v[j][1] = v[j][2] = C(0);
        for (int d = 0; d < D; ++d) v[j][d] = u[j][d] + F[j][d] * k;
        C v2 = v[j][0] * v[j][0];
        for (int d = 1; d < D; ++d) v2 += v[j][d] * v[j][d];
        M2[j] = Ma3Folded<C>(rho[j], SH ? drho[j] : rho[j], v2);
      }
    }
    static_for<Q>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      const C si = SC ? C(kHalfScale<L, i>) : C(1);
      const C w_i = C(L::W[i]);
      // w_i s_i is exact (s_i is a power of two).  omega_eq is applied per
      // cell: a product constant fl(omega_eq w_i) would carry one fixed
      // rounding error per population into every collision of every cell,
      // which biases the relaxation coherently (measured: 3-5x the
      // kinetic-energy error of a 2000-step Taylor-Green run).
      const C ws = w_i * si;
      C out[V];
#pragma unroll
      for (int j = 0; j < V; ++j) {
        const C cu = cdot<L, i>(u[j]);
        const C t = M[j].t(cu);
        C src = C(0);
        if constexpr (FS == SRC_GUO) {
          const C cF = cdot<L, i>(F[j]);
          src = (C(3) * gc * w_i) * (cF - uF[j] + C(3) * cu * cF);
        }
        if constexpr (FS == SRC_EDM) src = w_i * (M2[j].t(cdot<L, i>(v[j])) - t);
        out[j] = fma(of, f[i][j], fma(oe, ws * t, si * src));
      }
      store_vec<L, i, S, C, V, SC>(a, w, self_row, voff, out);
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
        // f is in storage scale s_i (FP16); eq and s are not.
        const C si = SC ? C(kHalfScale<L, i>) : C(1);
        out[j] = of * f[i][j] + si * (oe * eq + s);
      }
      store_vec<L, i, S, C, V, SC>(a, w, self_row, voff, out);
    });
  }
  if (a.store && w.active) {
#pragma unroll
    for (int j = 0; j < V; ++j) {
      const int cell = self_row + w.k0 + j;
      store_fields<L, C>(a, cell, FSX, rho[j], u[j], F[j], 2.0);
    }
  }
}

}  // namespace lbm
