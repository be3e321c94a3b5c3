// physics.hpp -- per-cell operators of the fused step, named after the
// reference's template plugins (Equilibrium, ForcingScheme, Force, Moment).
//
// Two arithmetic forms exist for every operator:
//   * Reference order (RefOrder): the exact sequence of separately rounded
//     operations of the reference expression.  Used for FP64 compute with
//     unshifted storage, compiled with -ffp-contract=off, so results match
//     the FP64 oracle bit for bit where no libm call is involved.
//   * Shifted form: populations are carried as g_i = f_i - w_i, moments are
//     accumulated on g (rho = 1 + sum g, rho u = sum c g) and the equilibrium
//     is formed as feq_i - w_i = w_i (drho + rho p_i) with p_i = P_i(u) - 1,
//     which avoids the cancellation of f - w in FP32/FP16.  Horner form,
//     FMA-friendly.  Used for every reduced-precision configuration.
#pragma once

#include <type_traits>
#include <utility>

#include "lattice.hpp"

namespace lbm {

// ---- compile-time loop -----------------------------------------------------

template <class F, int... Is>
__host__ __device__ __forceinline__ void static_for_impl(
    F&& f, std::integer_sequence<int, Is...>) {
  (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, class F>
__host__ __device__ __forceinline__ void static_for(F&& f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// c_{i,d} with 0 beyond the lattice dimension.
template <class L, int i, int d>
__host__ __device__ constexpr int cel() {
  if constexpr (d < L::dimD) return L::C[i][d];
  else return 0;
}

// sum_d c_{i,d} v_d over the nonzero c only, in component order.  Equal in
// value to the reference's MathVector::dot (MathVector.h:45-53), whose zero
// terms only add signed zeros.
template <class L, int i, class T>
__host__ __device__ __forceinline__ T cdot(const T* v) {
  constexpr int c0 = cel<L, i, 0>(), c1 = cel<L, i, 1>(), c2 = cel<L, i, 2>();
  if constexpr (c0 != 0) {
    T r = c0 > 0 ? v[0] : -v[0];
    if constexpr (c1 != 0) r = c1 > 0 ? r + v[1] : r - v[1];
    if constexpr (c2 != 0) r = c2 > 0 ? r + v[2] : r - v[2];
    return r;
  } else if constexpr (c1 != 0) {
    T r = c1 > 0 ? v[1] : -v[1];
    if constexpr (c2 != 0) r = c2 > 0 ? r + v[2] : r - v[2];
    return r;
  } else if constexpr (c2 != 0) {
    return c2 > 0 ? v[2] : -v[2];
  } else {
    return T(0);
  }
}

// First population index with c_{i,d} != 0 (for the moment accumulation).
template <class L, int d>
__host__ __device__ constexpr int first_nonzero() {
  for (int i = 0; i < L::dimQ; ++i)
    if (L::C[i][d] != 0) return i;
  return L::dimQ;
}

// Power-of-two scale applied to shifted FP16 populations so that
// (f_i - w_i) * s_i stays in FP16's normal range: s_i = 2^floor(log2(1/w_i)).
template <class L, int i>
__host__ __device__ constexpr double half_scale_of() {
  double inv = 1.0 / L::W[i];
  double s = 1.0;
  while (s * 2.0 <= inv) s *= 2.0;
  return s;
}
// Constant-expression instances: the scale and its inverse are immediates in
// the kernels (a plain constexpr call may be evaluated at run time).
template <class L, int i>
inline constexpr double kHalfScale = half_scale_of<L, i>();
template <class L, int i>
inline constexpr double kHalfScaleInv = 1.0 / half_scale_of<L, i>();
template <class L, int i>
__host__ __device__ constexpr double half_scale() {
  return kHalfScale<L, i>;
}

// ---- Moment (Moment.h:13-47) -------------------------------------------------

// Density and velocity of the pulled populations f[Q] in reference order:
// rho = f_0 + f_1 + ...; u_d = (sum_i c_{i,d} f_i) / rho.  u2 = |u|^2 in
// MathVector::norm2 order (MathVector.h:55-62).
template <class L, class T>
__host__ __device__ __forceinline__ void moments_ref(const T* f, T& rho,
                                                     T* u, T& u2) {
  rho = f[0];
  static_for<L::dimQ - 1>([&](auto ic) { rho += f[decltype(ic)::value + 1]; });
  static_for<L::dimD>([&](auto dc) {
    constexpr int d = decltype(dc)::value;
    constexpr int i0 = first_nonzero<L, d>();
    T acc = L::C[i0][d] > 0 ? f[i0] : -f[i0];
    static_for<L::dimQ>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      if constexpr (i > i0 && L::C[i][d] != 0)
        acc = L::C[i][d] > 0 ? acc + f[i] : acc - f[i];
    });
    u[d] = acc / rho;
  });
  u2 = u[0] * u[0];
  static_for<L::dimD - 1>([&](auto dc) {
    constexpr int d = decltype(dc)::value + 1;
    u2 += u[d] * u[d];
  });
}

// Shifted moments: g_i = f_i - w_i.  drho = sum g, rho = 1 + drho,
// u = (sum c g) / rho (sum_i w_i c_i = 0 and sum_i w_i = 1).
template <class L, class T>
__host__ __device__ __forceinline__ void moments_shifted(const T* g, T& drho,
                                                         T& rho, T* u, T& u2) {
  drho = g[0];
  static_for<L::dimQ - 1>([&](auto ic) { drho += g[decltype(ic)::value + 1]; });
  rho = T(1) + drho;
  const T rrho = T(1) / rho;
  static_for<L::dimD>([&](auto dc) {
    constexpr int d = decltype(dc)::value;
    constexpr int i0 = first_nonzero<L, d>();
    T acc = L::C[i0][d] > 0 ? g[i0] : -g[i0];
    static_for<L::dimQ>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      if constexpr (i > i0 && L::C[i][d] != 0)
        acc = L::C[i][d] > 0 ? acc + g[i] : acc - g[i];
    });
    u[d] = acc * rrho;
  });
  u2 = u[0] * u[0];
  static_for<L::dimD - 1>([&](auto dc) {
    constexpr int d = decltype(dc)::value + 1;
    u2 += u[d] * u[d];
  });
}

// Moments of shifted populations held in a per-population scale:
// g_i = h_i * isc_i with isc_i = 1/s_i (FP16 storage) or 1 (otherwise), each
// term one fused multiply-add, so no separate unscaling pass is needed.
template <class L, class T, bool SCALED>
__device__ __forceinline__ void moments_shifted_scaled(const T* h, T& drho, T& rho,
                                                       T* u, T& u2) {
  auto isc = [](auto ic) {
    constexpr int i = decltype(ic)::value;
    return SCALED ? T(kHalfScaleInv<L, i>) : T(1);
  };
  drho = h[0] * isc(std::integral_constant<int, 0>{});
  static_for<L::dimQ - 1>([&](auto ic) {
    constexpr int i = decltype(ic)::value + 1;
    drho = fma(h[i], isc(std::integral_constant<int, i>{}), drho);
  });
  rho = T(1) + drho;
  // FP32: the hardware reciprocal estimate and one Newton-Raphson correction
  // instead of the division sequence.  The correction matters: the bare
  // estimate is not round-to-nearest, and a biased 1/rho biases every
  // equilibrium velocity, which damps momentum a little at each collision
  // (measured: 4x the kinetic-energy error of a 2000-step Taylor-Green run).
  // FP64 keeps the division.
  T rrho;
  if constexpr (std::is_same_v<T, float>) {
    const T r0 = __builtin_amdgcn_rcpf(rho);
    rrho = fma(r0, fma(-rho, r0, T(1)), r0);
  } else {
    rrho = T(1) / rho;
  }
  static_for<L::dimD>([&](auto dc) {
    constexpr int d = decltype(dc)::value;
    T acc = T(0);
    static_for<L::dimQ>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      if constexpr (L::C[i][d] != 0)
        acc = fma(h[i], T(L::C[i][d]) * isc(std::integral_constant<int, i>{}), acc);
    });
    u[d] = acc * rrho;
  });
  u2 = u[0] * u[0];
  static_for<L::dimD - 1>([&](auto dc) {
    constexpr int d = decltype(dc)::value + 1;
    u2 += u[d] * u[d];
  });
}

// ---- Equilibrium (Equilibrium.h) -------------------------------------------

// TruncationMa3, reference operation order (Equilibrium.h:17-33).
template <class L, int i, class T>
__host__ __device__ __forceinline__ T feq_ma3_ref(T rho, const T* u, T u2) {
  const T cu = cdot<L, i>(u);
  const T fe = T(1.0) + cu * T(3.0) - T(0.5) * u2 * T(3.0) +
               T(0.5) * T(9.0) * cu * cu - T(0.5) * T(9.0) * cu * u2 +
               cu * cu * cu * T(27.0) / T(6.0) + T(0.125) * u2 * u2 * T(9.0) -
               T(0.25) * cu * cu * u2 * T(27.0) +
               cu * cu * cu * cu * T(81.0) / T(24.0);
  return rho * T(L::W[i]) * fe;
}

// Per-cell coefficients of the Ma3 polynomial in c.u:
// P - 1 = a0 + cu (a1 + cu (a2 + cu (4.5 + 3.375 cu))).
template <class T>
struct Ma3Coefs {
  T a0, a1, a2;
  __host__ __device__ __forceinline__ explicit Ma3Coefs(T u2)
      : a0(u2 * (T(-1.5) + T(1.125) * u2)),
        a1(T(3.0) - T(4.5) * u2),
        a2(T(4.5) - T(6.75) * u2) {}
  __host__ __device__ __forceinline__ T p(T cu) const {
    return a0 + cu * (a1 + cu * (a2 + cu * (T(4.5) + T(3.375) * cu)));
  }
};

// TruncationMa3 with the density folded in, for one cell:
//   t(cu) = base + rho (P(cu) - 1) = b0 + cu (b1 + cu (b2 + cu (b3 + b4 cu)))
// with base = drho (shifted: eq - w = w t) or rho (unshifted: eq = w t).
template <class T>
struct Ma3Folded {
  T b0, b1, b2, b3, b4;
  Ma3Folded() = default;
  __device__ __forceinline__ Ma3Folded(T rho, T base, T u2) {
    const Ma3Coefs<T> m(u2);
    b0 = fma(rho, m.a0, base);
    b1 = rho * m.a1;
    b2 = rho * m.a2;
    b3 = rho * T(4.5);
    b4 = rho * T(3.375);
  }
  __device__ __forceinline__ T t(T cu) const {
    return fma(cu, fma(cu, fma(cu, fma(b4, cu, b3), b2), b1), b0);
  }
};

// Exact equilibrium (Equilibrium.h:60-81, 106-126) per-cell factors:
// m[d][c+1] = (2 - s_d) * b_d^c, s_d = sqrt(1 + 3 u_d^2),
// b_d = (2 u_d + s_d) / (1 - u_d), c in {-1, 0, 1} (PowerBase, Helpers.h:61-71).
template <class L, class T>
struct ExactFactors {
  T m[3][3];
  __host__ __device__ __forceinline__ explicit ExactFactors(const T* u) {
    for (int d = 0; d < L::dimD; ++d) {
      const T s = sqrt(T(1.0) + T(3.0) * u[d] * u[d]);
      const T b = (T(2) * u[d] + s) / (T(1.0) - u[d]);
      const T a = T(2.0) - s;
      m[d][0] = a * (T(1.0) / b);
      m[d][1] = a * T(1.0);
      m[d][2] = a * b;
    }
  }
  // Reference order: fEq = ((1 * m0) * m1) * m2, then (rho * w) * fEq.
  template <int i>
  __host__ __device__ __forceinline__ T feq(T rho) const {
    T fe = m[0][cel<L, i, 0>() + 1];
    fe *= m[1][cel<L, i, 1>() + 1];
    if constexpr (L::dimD == 3) fe *= m[2][cel<L, i, 2>() + 1];
    return rho * T(L::W[i]) * fe;
  }
};

// Exact equilibrium in shifted form: e[d][c+1] = phi_d^c - 1 written
// without cancellation (s - 1 = 3u^2 / (1 + s)):
//   c =  0: 1 - s
//   c = +1: (u (5 - 2s) + 2 (s - 1) - 3u^2) / (1 - u)
//   c = -1: (-4u + u s - 2 (s - 1)) / (2u + s)
// and P - 1 = prod_d (1 + e_d) - 1 expanded term by term.
template <class L, class T>
struct ExactShifted {
  T e[3][3];
  __host__ __device__ __forceinline__ explicit ExactShifted(const T* u) {
    for (int d = 0; d < L::dimD; ++d) {
      const T v = u[d];
      const T s = sqrt(T(1) + T(3) * v * v);
      const T sm1 = T(3) * v * v / (T(1) + s);
      e[d][1] = -sm1;
      e[d][2] = (v * (T(5) - T(2) * s) + T(2) * sm1 - T(3) * v * v) / (T(1) - v);
      e[d][0] = (v * s - T(4) * v - T(2) * sm1) / (T(2) * v + s);
    }
  }
  template <int i>
  __host__ __device__ __forceinline__ T p() const {
    const T e0 = e[0][cel<L, i, 0>() + 1], e1 = e[1][cel<L, i, 1>() + 1];
    T t = e0 + e1 + e0 * e1;
    if constexpr (L::dimD == 3) {
      const T e2 = e[2][cel<L, i, 2>() + 1];
      t = t + e2 + t * e2;
    }
    return t;
  }
};

// ---- Force (Force.h:104-292) -------------------------------------------------
// Analytic forces are evaluated on the host once per configuration into
// per-component tables indexed along one axis (Constant: length 1;
// Sinusoidal: F_d indexed by coordinate d; Kolmogorov: F_x indexed by y),
// with the reference's own expression A * sin(unsigned(2 i) * pi / lambda).
// The kernel reads F_d = tab_d[coord[axis_d]], or 0 for a NULL table.

}  // namespace lbm
