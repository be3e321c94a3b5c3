// mlp.hpp -- the alpha predictor of the learned entropic collision.
//
// No reference counterpart: the reference finds alpha by Newton only
// (Collision.h:328-349).  Here a small ReLU MLP maps the cell's normalised
// non-equilibrium moments to a first guess of alpha; the Newton loop below
// (residual_newton) then certifies or corrects it.
//
// Features: m_k = (1/rho) sum_i fneq_i prod_d phi_{a_kd}(c_id), phi_0 = 1,
// phi_1 = c, phi_2 = c^2, over the full tensor basis a_k in {0,1,2}^D (the
// Q = 3^D raw moments of D2Q9 / D3Q27), k = digits a_x a_y [a_z] in base 3.
// The network sees z = (m - mean) * scale and returns y; alpha_hat =
// out_mean + out_scale * y.  Parameter block layout (FP32, shared with the
// training kernels and the FP64 oracle's mlp_ref.c):
//   W1[H1][NIN] b1[H1] W2[H2][H1] b2[H2] W3[H2] b3 | mean[NIN] scale[NIN]
//   out_mean out_scale
#pragma once

#include "physics.hpp"

namespace lbm {

template <class L>
struct MlpShape {
  static constexpr int NIN = 0, H1 = 0, H2 = 0;
};
template <>
struct MlpShape<Lattice<LatticeType::D2Q9>> {
  static constexpr int NIN = 9, H1 = 32, H2 = 32;
};
template <>
struct MlpShape<Lattice<LatticeType::D3Q27>> {
  static constexpr int NIN = 27, H1 = 64, H2 = 64;
};

template <int NIN, int H1, int H2>
struct MlpDims {
  static constexpr int NPARAM = H1 * NIN + H1 + H2 * H1 + H2 + H2 + 1;
  static constexpr int NBLOCK = NPARAM + 2 * NIN + 2;  // + normalisation
  static constexpr int W1 = 0, B1 = H1 * NIN, W2 = B1 + H1, B2 = W2 + H2 * H1,
                       W3 = B2 + H2, B3 = W3 + H2, MEAN = NPARAM,
                       SCALE = NPARAM + NIN, OUT = NPARAM + 2 * NIN;
};

// Value of basis function k at velocity i: prod_d phi_{a_d}(c_d) in {-1,0,1}.
template <class L, int k, int i>
__host__ __device__ constexpr int moment_basis() {
  int r = 1, kk = k;
  for (int d = L::dimD - 1; d >= 0; --d) {
    const int a = kk % 3;
    kk /= 3;
    const int c = L::C[i][d];
    const int phi = a == 0 ? 1 : (a == 1 ? c : c * c);
    r *= phi;
  }
  return r;
}

// Raw non-equilibrium moments divided by rho.
template <class L, class T>
__device__ __forceinline__ void nonequilibrium_moments(const T* fneq, T rrho,
                                                       float* m) {
  constexpr int Q = L::dimQ;
  static_for<Q>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    T s = T(0);
    static_for<Q>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      constexpr int b = moment_basis<L, k, i>();
      if constexpr (b == 1) s += fneq[i];
      if constexpr (b == -1) s -= fneq[i];
    });
    m[k] = float(s * rrho);
  });
}

// One forward pass for one cell, parameters in LDS (uniform reads).
template <int NIN, int H1, int H2>
__device__ __forceinline__ double mlp_alpha(const float* __restrict__ P,
                                            const float* m) {
  using Dm = MlpDims<NIN, H1, H2>;
  float z[NIN];
#pragma unroll
  for (int k = 0; k < NIN; ++k) z[k] = (m[k] - P[Dm::MEAN + k]) * P[Dm::SCALE + k];
  float a1[H1];
#pragma unroll
  for (int j = 0; j < H1; ++j) {
    float s = P[Dm::B1 + j];
#pragma unroll
    for (int k = 0; k < NIN; ++k) s = fmaf(P[Dm::W1 + j * NIN + k], z[k], s);
    a1[j] = fmaxf(s, 0.0f);
  }
  float y = P[Dm::B3];
#pragma unroll
  for (int j = 0; j < H2; ++j) {
    float s = P[Dm::B2 + j];
#pragma unroll
    for (int k = 0; k < H1; ++k) s = fmaf(P[Dm::W2 + j * H1 + k], a1[k], s);
    y = fmaf(P[Dm::W3 + j], fmaxf(s, 0.0f), y);
  }
  return double(P[Dm::OUT]) + double(P[Dm::OUT + 1]) * double(y);
}

// ---- residual Newton with a Taylor certificate ------------------------------
//
// G(a) = H(f) - H(f - a fneq) = sum_i f_i [a x_i A_i - (1 - a x_i) log1p(-a x_i)]
// with x_i = fneq_i / f_i and A_i = ln(f_i / w_i): the reference's H-difference
// (EntropicStep.h:31-45) without its cancellation.  G'(a) = sum fneq (1 + ln(fm/w))
// as in EntropicStep.h:48-62, and G''(a) = -sum f x^2 / (1 - a x) < 0 with |G''|
// convex in a.  After a Newton step s = G/G' from a, Taylor's theorem gives
// |G(a - s)| <= s^2/2 max_[a, a - s] |G''| = s^2/2 max(|G''(a)|, |G''(a - s)|),
// a certificate that needs no logarithm.  A cell is accepted when |G| <= tol at
// an evaluated point, or when the certificate of the next point is <= tol.
// Outcome codes: 2 Newton accepted, 3 fallback (alpha = 2); `first` reports
// acceptance from the starting point's single evaluation (plus certificate).
template <class L>
__device__ __forceinline__ double residual_newton(const double* f,
                                                  const double* n, double amax,
                                                  double start, double tol,
                                                  int& branch, int& evals,
                                                  bool& first) {
  constexpr int Q = L::dimQ;
  double x[Q], A[Q];
  static_for<Q>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    x[i] = n[i] / f[i];
    A[i] = log(f[i] * (1.0 / L::W[i]));
  });
  auto d2 = [&](double a) {
    double s = 0.0;
    static_for<Q>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      s += n[i] * x[i] / (1.0 - a * x[i]);
    });
    return s;  // = -G''(a) >= 0
  };
  double a = (start > 1.0 && start < amax) ? start : 2.0;
  evals = 0;
  first = false;
  for (int it = 0; it < 50; ++it) {
    double g = 0.0, dg = 0.0, h = 0.0;
    static_for<Q>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      const double t = a * x[i];
      const double l1 = log1p(-t);
      g += f[i] * (t * A[i] - (1.0 - t) * l1);
      dg += n[i] * (1.0 + A[i] + l1);
      h += n[i] * x[i] / (1.0 - t);
    });
    ++evals;
    if (fabs(g) <= tol) {
      branch = 2;
      first = it == 0;
      return a;
    }
    const double s = g / dg;
    const double an = a - s;
    if (!(an > 1.0 && an < amax)) break;
    const double bound = 0.5 * s * s * fmax(h, d2(an)) * (1.0 + 1e-6);
    if (bound <= tol) {
      branch = 2;
      first = it == 0;
      return an;
    }
    a = an;
  }
  branch = 3;
  return 2.0;
}

}  // namespace lbm
