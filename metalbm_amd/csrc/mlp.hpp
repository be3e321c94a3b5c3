// mlp.hpp -- the alpha predictor of the learned entropic collision.
//
// No reference counterpart: the reference finds alpha by Newton only
// (Collision.h:328-349).  Here a small ReLU MLP maps the cell's normalised
// non-equilibrium moments to a first guess of alpha; the Newton loop below
// (residual_correction) then certifies or corrects it, and the reference
// Newton takes over when it cannot.
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

// Network shape per lattice and the matrix-core operand type of its
// inference (D2Q9: f16, 11-bit significand; D3Q27: bf16 weights as specified
// for the 27-64-64-1 network).  Master weights and training stay FP32.
template <class L>
struct MlpShape {
  static constexpr int NIN = 0, H1 = 0, H2 = 0;
  using Op = _Float16;
};
template <>
struct MlpShape<Lattice<LatticeType::D2Q9>> {
  static constexpr int NIN = 9, H1 = 32, H2 = 32;
  using Op = _Float16;
};
template <>
struct MlpShape<Lattice<LatticeType::D3Q27>> {
  static constexpr int NIN = 27, H1 = 64, H2 = 64;
  using Op = __bf16;
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

// ---- wavefront inference on the matrix cores --------------------------------
//
// The fused kernel runs the network for all 64 cells of a wavefront with
// v_mfma_f32_16x16x16 (f16 or bf16 operands, FP32 accumulation).  Cells are the N
// dimension, so a layer's accumulator (lane l: cell l%16 of the tile, rows
// 4(l/16)+j) is already the next layer's B operand (lane l: column l%16,
// k = 4(l/16)+j): H^T = relu(W . X^T + b) chains without a transpose.
// A workgroup holds one copy of the packed network in LDS, [0, NET), staged
// by all its threads before one barrier; every wavefront then owns a tile
// region of TILE bytes at NET + wave TILE for its features and outputs.
//   network: W1h[H1][KP+4] W2h[H2][H1+4] W3h[16][H2+4] (operand type, zero
//            padded; W3h row 0 = W3), b1[H1] b2[H2] b3 mean[NIN] scale[NIN]
//            out[2] (f32)
//   X[64][KP+4] (standardised features of the wavefront's cells), Y[64] (f32)
template <int NIN, int H1, int H2>
struct MfmaLayout {
  static constexpr int KP = NIN <= 16 ? 16 : 32;
  // Row strides in operand elements, padded by four (two dwords) so that the
  // 16 rows one lane group reads (8 bytes each) fall in distinct LDS banks:
  // unpadded, a stride of 16 or 32 dwords put every row on the same bank
  // pair (measured: 1.6e9 bank-conflict cycles per D3Q27 256^3 step).
  static constexpr int KPS = KP + 4, H1S = H1 + 4, H2S = H2 + 4;
  static constexpr int W1H = 0;
  static constexpr int W2H = W1H + H1 * KPS * 2;
  static constexpr int W3H = W2H + H2 * H1S * 2;
  static constexpr int F32 = W3H + 16 * H2S * 2;
  static constexpr int B1 = F32, B2 = B1 + 4 * H1, B3 = B2 + 4 * H2, MEAN = B3 + 4,
                       SCALE = MEAN + 4 * NIN, OUT = SCALE + 4 * NIN;
  static constexpr int NET = ((OUT + 8 + 15) / 16) * 16;
  // offsets inside a wavefront's tile region
  static constexpr int X = 0;
  static constexpr int Y = X + 64 * KPS * 2;
  static constexpr int TILE = ((Y + 64 * 4 + 15) / 16) * 16;
  static constexpr int BYTES = NET + 4 * TILE;  // a workgroup of four wavefronts
};

typedef float mfma_f4 __attribute__((ext_vector_type(4)));

// One 16x16x16 matrix-core product, D = A B + C, for operand type T.
template <class T>
struct Mfma;
template <>
struct Mfma<_Float16> {
  typedef _Float16 v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ mfma_f4 mac(v4 a, v4 b, mfma_f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
  }
};
template <>
struct Mfma<__bf16> {
  typedef __bf16 v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ mfma_f4 mac(v4 a, v4 b, mfma_f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
  }
};

// Writes the FP32 parameter block in the LDS layout (bytes [0, NET)), the
// weights rounded to the operand type T (round to nearest even).  Run once per
// parameter change into a global "packed" buffer; kernels then stage it with
// copy_packed_network.  Cooperative over the threads of one workgroup.
template <int NIN, int H1, int H2, class T>
__device__ __forceinline__ void stage_mfma_network(const float* __restrict__ P,
                                                   char* lds) {
  using Dm = MlpDims<NIN, H1, H2>;
  using M = MfmaLayout<NIN, H1, H2>;
  T* w1 = reinterpret_cast<T*>(lds + M::W1H);
  T* w2 = reinterpret_cast<T*>(lds + M::W2H);
  T* w3 = reinterpret_cast<T*>(lds + M::W3H);
  float* fl = reinterpret_cast<float*>(lds + M::F32);
  for (int t = threadIdx.x; t < H1 * M::KPS; t += blockDim.x) {
    const int r = t / M::KPS, k = t - r * M::KPS;
    w1[t] = k < NIN ? T(P[Dm::W1 + r * NIN + k]) : T(0.0f);
  }
  for (int t = threadIdx.x; t < H2 * M::H1S; t += blockDim.x) {
    const int r = t / M::H1S, k = t - r * M::H1S;
    w2[t] = k < H1 ? T(P[Dm::W2 + r * H1 + k]) : T(0.0f);
  }
  for (int t = threadIdx.x; t < 16 * M::H2S; t += blockDim.x)
    w3[t] = t < H2 ? T(P[Dm::W3 + t]) : T(0.0f);
  for (int t = threadIdx.x; t < H1; t += blockDim.x) fl[t] = P[Dm::B1 + t];
  for (int t = threadIdx.x; t < H2; t += blockDim.x) fl[H1 + t] = P[Dm::B2 + t];
  for (int t = threadIdx.x; t < 2 * NIN + 3; t += blockDim.x)
    fl[H1 + H2 + t] = P[Dm::B3 + t];  // b3, mean[NIN], scale[NIN], out[2]
}

// Copy of the packed network (M::NET bytes, 16-aligned) into the
// workgroup's LDS by all its threads; the caller makes it visible with a
// workgroup barrier before the first mlp_alpha_wave.
template <int NIN, int H1, int H2>
__device__ __forceinline__ void copy_packed_network(const void* __restrict__ src,
                                                    char* lds) {
  constexpr int n = MfmaLayout<NIN, H1, H2>::NET / 16;
  const uint4* s = static_cast<const uint4*>(src);
  uint4* d = reinterpret_cast<uint4*>(lds);
  for (int t = threadIdx.x; t < n; t += blockDim.x) d[t] = s[t];
}

// Orders one wavefront's LDS writes before its later LDS reads by other lanes
// (the LDS executes a wavefront's operations in order; the fences stop the
// compiler from moving accesses across).
__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// alpha_hat for every lane's cell (m: raw features; lanes without a cell pass
// zeros), from the workgroup's staged network `net` (copy_packed_network and
// a barrier) and the wavefront's own tile region `tile`, which it
// synchronises with itself only; all 64 lanes of the wavefront must be active.
template <int NIN, int H1, int H2, class T>
__device__ __forceinline__ double mlp_alpha_wave(const char* net, char* tile, const float* m) {
  using M = MfmaLayout<NIN, H1, H2>;
  using V4 = typename Mfma<T>::v4;
  constexpr int KP = M::KP, KPS = M::KPS, H1S = M::H1S, H2S = M::H2S;
  const int lane = threadIdx.x & 63;
  const float* fl = reinterpret_cast<const float*>(net + M::F32);
  const float* b1 = fl;
  const float* b2 = fl + H1;
  const float b3 = fl[H1 + H2];
  const float* mean = fl + H1 + H2 + 1;
  const float* scale = mean + NIN;
  T* X = reinterpret_cast<T*>(tile + M::X);
  float* Y = reinterpret_cast<float*>(tile + M::Y);
#pragma unroll
  for (int k4 = 0; k4 < KP; k4 += 4) {
    V4 z;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = k4 + j;
      z[j] = k < NIN ? T((m[k] - mean[k]) * scale[k]) : T(0.0f);
    }
    *reinterpret_cast<V4*>(X + lane * KPS + k4) = z;
  }
  wave_sync_lds();
  const T* w1 = reinterpret_cast<const T*>(net + M::W1H);
  const T* w2 = reinterpret_cast<const T*>(net + M::W2H);
  const T* w3 = reinterpret_cast<const T*>(net + M::W3H);
  const int col = lane & 15, kq = 4 * (lane >> 4);
  auto frag = [](const T* p) { return *reinterpret_cast<const V4*>(p); };
#pragma unroll
  for (int t = 0; t < 4; ++t) {  // 16-cell tiles of the wavefront
    V4 h1[H1 / 16];
#pragma unroll
    for (int rt = 0; rt < H1 / 16; ++rt) {
      mfma_f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int kb = 0; kb < KP / 16; ++kb)
        acc = Mfma<T>::mac(frag(w1 + (rt * 16 + col) * KPS + kb * 16 + kq),
                           frag(X + (t * 16 + col) * KPS + kb * 16 + kq), acc);
#pragma unroll
      for (int j = 0; j < 4; ++j) h1[rt][j] = T(fmaxf(acc[j] + b1[rt * 16 + kq + j], 0.0f));
    }
    V4 h2[H2 / 16];
#pragma unroll
    for (int rt = 0; rt < H2 / 16; ++rt) {
      mfma_f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int kb = 0; kb < H1 / 16; ++kb)
        acc = Mfma<T>::mac(frag(w2 + (rt * 16 + col) * H1S + kb * 16 + kq), h1[kb], acc);
#pragma unroll
      for (int j = 0; j < 4; ++j) h2[rt][j] = T(fmaxf(acc[j] + b2[rt * 16 + kq + j], 0.0f));
    }
    mfma_f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb < H2 / 16; ++kb)
      acc = Mfma<T>::mac(frag(w3 + col * H2S + kb * 16 + kq), h2[kb], acc);
    if (lane < 16) Y[t * 16 + lane] = acc[0];  // row 0 = the output unit
  }
  wave_sync_lds();
  const float y = Y[lane] + b3;
  return double(fl[H1 + H2 + 1 + 2 * NIN]) + double(fl[H1 + H2 + 2 + 2 * NIN]) * double(y);
}

// ---- certified Newton correction ------------------------------------------
//
// G(a) = H(f) - H(f - a fneq) = sum_i f_i [a x_i A_i - (1 - a x_i) log1p(-a x_i)]
// with x_i = fneq_i / f_i and A_i = ln(f_i / w_i): the reference's H-difference
// (EntropicStep.h:31-45) without its cancellation.  G'(a) = sum fneq (1 + ln(fm/w))
// as in EntropicStep.h:48-62, and G''(a) = -sum f x^2 / (1 - a x) < 0 with |G''|
// convex in a.  After a Newton step s = G/G' from a, Taylor's theorem gives
// |G(a - s)| <= s^2/2 max_[a, a - s] |G''| = s^2/2 max(|G''(a)|, |G''(a - s)|),
// a certificate that needs no logarithm.  A start is accepted when |G| <= tol
// there, or its Newton point when the certificate is <= tol.

// log1p(y) by its Taylor series to degree D (Horner: D - 1 FMAs and a
// multiply instead of a library call); truncation < |y|^(D+1) / ((D+1)(1 - |y|)).
// The tiers below keep it under ~1e-19 |y|, a thousandth of an FP64 ulp:
// degree 6 for |y| <= kTinyMax, 8 for <= kSmallMax, 12 for <= kSeriesMax.
constexpr double kTinyMax = 9e-4;    // 9e-4^6 / 7 < 7.6e-20
constexpr double kSmallMax = 5e-3;   // 5e-3^8 / 9 < 4.4e-20
constexpr double kSeriesMax = 0.03;  // 0.03^12 / 13 < 4.1e-20
template <int D>
__device__ __forceinline__ double log1p_series(double y) {
  double p = (D % 2 ? 1.0 : -1.0) / D;
  static_for<D - 1>([&](auto jc) {
    constexpr int k = D - 1 - decltype(jc)::value;
    p = fma(p, y, (k % 2 ? 1.0 : -1.0) / k);
  });
  return p * y;
}
// 1/d for d > 0 in the normal range: hardware reciprocal estimate plus two
// Newton-Raphson corrections (each squares the relative error), a few ulp.
__device__ __forceinline__ double recip_refined(double d) {
  double r = __builtin_amdgcn_rcp(d);
  r = fma(r, fma(-d, r, 1.0), r);
  return fma(r, fma(-d, r, 1.0), r);
}

// n / f for f > 0 in the normal range: refined reciprocal and a final
// residual correction of the quotient (error of a few ulp, well inside the
// solver's rounding allowance).
__device__ __forceinline__ double div_refined(double n, double f) {
  const double r = recip_refined(f);
  const double q = n * r;
  return fma(r, fma(-f, q, n), q);
}

// log1p(y) for |y| <= kMidMax as 2 atanh(z), z = y / (2 + y), |z| <= 1/7:
// 2 z sum_{k=0..9} z^{2k} / (2k + 1), truncation < 2 |z| (z^2)^10 / 21 / (1 - z^2)
// < 2e-18 |z|.
constexpr double kMidMax = 0.25;
__device__ __forceinline__ double log1p_mid(double y) {
  const double z = div_refined(y, 2.0 + y);
  const double p = z * z;
  double s = 1.0 / 19.0;
  s = fma(s, p, 1.0 / 17.0);
  s = fma(s, p, 1.0 / 15.0);
  s = fma(s, p, 1.0 / 13.0);
  s = fma(s, p, 1.0 / 11.0);
  s = fma(s, p, 1.0 / 9.0);
  s = fma(s, p, 1.0 / 7.0);
  s = fma(s, p, 1.0 / 5.0);
  s = fma(s, p, 1.0 / 3.0);
  s = fma(s, p, 1.0);
  return 2.0 * z * s;
}

// Runs body(log1p_of) with the cheapest log1p variant that is exact (to the
// bounds above) for every active lane of the wavefront: the lanes' largest
// |y| decides, so the wavefront runs one variant rather than one per group
// of lanes.  A NaN bound takes the library.
template <class F>
__device__ __forceinline__ void with_log1p(double ymax, F&& body) {
  if (!__any(!(ymax <= kTinyMax)))
    body([](double y) { return log1p_series<6>(y); });
  else if (!__any(!(ymax <= kSmallMax)))
    body([](double y) { return log1p_series<8>(y); });
  else if (!__any(!(ymax <= kSeriesMax)))
    body([](double y) { return log1p_series<12>(y); });
  else if (!__any(!(ymax <= kMidMax)))
    body([](double y) { return log1p_mid(y); });
  else
    body([](double y) { return log1p(y); });
}

// x_i = n_i / f_i for the certified correction: held per term for small Q,
// recomputed at each use for D3Q27, where holding 27 FP64 values would cost
// occupancy (the solve is latency-bound).
template <class L, class V>
struct XTerms {
  static constexpr bool KEEP = L::dimQ <= 9;
  const V& p;
  double x[KEEP ? L::dimQ : 1];
  double xmax = 0.0;  // max_i |x_i|
  double xpos = 0.0;  // max_i x_i, 0 if none positive: f - a n > 0 iff a xpos < 1 (a >= 0)
  __device__ __forceinline__ explicit XTerms(const V& p_) : p(p_) {
    static_for<L::dimQ>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      const double xi = div_refined(p.template n<i>(), p.template f<i>());
      xmax = fmax(xmax, fabs(xi));
      xpos = fmax(xpos, xi);
      if constexpr (KEEP) x[i] = xi;
    });
  }
  template <int i>
  __device__ __forceinline__ double get() const {
    if constexpr (KEEP)
      return x[i];
    else
      return div_refined(p.template n<i>(), p.template f<i>());
  }
};

// G, G' and -G'' at alpha = a > 0 from t_i = a x_i, with S = sum n_i A_i and
// Sn = sum n_i precomputed (A_i = ln(f_i / w_i)):
//   G = a S - sum (f_i - a n_i) log1p(-t_i),  G' = Sn + S + sum n_i log1p(-t_i),
//   -G'' = sum n_i x_i / (1 - t_i)
// (f - a n = f (1 - t), ln((f - a n)/w) = A + log1p(-t)).  log1p(-t) by the
// Taylor series when max |t_i| = a max |x_i| <= kSeriesMax (degree by tier),
// by the atanh series when <= kMidMax, else by the library; the choice is
// made once per evaluation and wavefront, not per term (with_log1p).  1/(1 - t) by the refined reciprocal.
// FRESH (the iterating plain Newton) converts the populations afresh in
// every evaluation, so the compiler holds no 2Q converted doubles across the
// loop: 157 -> 109 VGPRs, four wavefronts per SIMD without spills.
template <class L, bool FRESH = false, class V>
__device__ __forceinline__ void residual_terms(const V& p, const XTerms<L, V>& X,
                                               double S, double Sn, double a,
                                               double& g, double& dg, double& h) {
  constexpr int Q = L::dimQ;
  const double tmax = a * X.xmax;
  double gs = 0.0, ds = 0.0;
  h = 0.0;
  auto terms = [&](auto log1p_of) {
    static_for<Q>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      double ni, fi;
      if constexpr (FRESH) {
        ni = p.template n_fresh<i>();
        fi = p.template f_fresh<i>();
      } else {
        ni = p.template n<i>();
        fi = p.template f<i>();
      }
      const double xi = X.template get<i>();
      const double t = a * xi;
      const double l1 = log1p_of(-t);
      gs += (fi - a * ni) * l1;
      ds += ni * l1;
      h += ni * xi * recip_refined(1.0 - t);
    });
  };
  with_log1p(tmax, terms);
  g = a * S - gs;
  dg = Sn + S + ds;
}

// What a certified evaluation leaves behind for reference_reaches: the point
// e it evaluated, G'(e), h = -G''(e) = sum n^2 / (f - e n) and an
// r >= max |n| / (f - e n).
struct EvalCurv {
  double e, dg, h, r;
};

// True only if the reference Newton (EntropicStep.h:111-140), started at the
// stored alpha s, provably converges to the root accepted at a_acc -- the
// condition under which a solve from another start (the network's) keeps the
// reference's outcome.  G is concave (G'' = -sum n^2 / (f - a n) < 0) with
// G(0) = 0, so (0, amax) holds at most one other root alpha*, on the branch
// where G' < 0, and G' decreases.  From the data of one evaluation at e, for
// every a with |a - e| r < 1: f - a n >= (f - e n)(1 - |a - e| r) > 0 (every
// logarithm defined) and |G''(a)| <= h / (1 - |a - e| r).  With H that bound
// over the points within D of e (D covers s, a_acc and a margin for delta):
//  * G'(s) <= G'(e) + max(e - s, 0) H =: gps, and likewise gpa >= G'(a_acc).
//    gps < 0 puts s right of G's maximum, so the reference's iterates do not
//    head for the trivial root (from the left of the maximum the first step
//    lands at or below 0 and the solve falls back to alpha = 2,
//    Collision.h:344-346).
//  * |G(a_acc)| <= tol and |G'| >= min(-gps, -gpa) between a_acc and alpha*
//    wherever it matters below, so |a_acc - alpha*| <= delta = tol / that.
//  * s >= a_acc + delta gives G(s) <= tol + (s - a_acc) gpa < 0: s is right
//    of the root, and the iterates decrease monotonically to it inside
//    [alpha*, s], where every log is defined.
//  * otherwise, with d = alpha* - s > 0, Newton's error identity
//    s1 - alpha* = d^2 (G''(x2) - G''(x1)/2) / G'(s) bounds the first step's
//    overshoot by d^2 H / |G'(s)|; the iterates then decrease monotonically,
//    and the logs stay defined if the overshoot point is within 1/r of e.
// Quadratic convergence then meets |step| <= 1e-8 well inside the 50
// iterations.  Relative allowances of 1e-6 cover the rounding of the bounds.
__device__ __forceinline__ bool reference_reaches(const EvalCurv& c, double a_acc, double s,
                                                  double tol) {
  if (!(c.dg < 0.0)) return false;
  // The bounds hold within D = D0 + m of e, m (at most 1) the margin left for
  // a_acc +- delta, checked once delta is known.  Quotients by the refined
  // reciprocal (a few ulp), well inside the 1e-6 allowances.
  const double D0 = fmax(fabs(s - c.e), fabs(a_acc - c.e));
  const double lim = c.r > 0.0 ? (1.0 - 1e-3) * recip_refined(c.r) : 2.0;
  if (!(D0 < lim)) return false;
  const double m = fmin(0.5 * (lim - D0), 1.0);
  const double H = c.h * recip_refined(1.0 - (D0 + m) * c.r) * (1.0 + 1e-6);
  const double gps = c.dg + fmax(c.e - s, 0.0) * H;
  const double gpa = c.dg + fmax(c.e - a_acc, 0.0) * H;
  if (!(gps < 0.0 && gpa < 0.0)) return false;
  const double delta = tol * recip_refined(fmin(-gps, -gpa)) * (1.0 + 1e-6);
  if (!(delta <= m)) return false;
  if (s >= a_acc + delta) return true;
  const double d = fmax(a_acc - s, 0.0) + delta;
  const double far = fabs(a_acc + delta + d * d * H * recip_refined(-gps) - c.e);
  return far * c.r < 1.0 - 1e-3;
}

// One residual-checked Newton correction of a starting alpha (G in the log1p
// form of EntropicStep.h:31-45).  Accepts the start if |G(start)| <= tol,
// else the Newton point a1 = start - G/G' if it lies in (1, amax) and its
// Taylor certificate 1/2 s^2 max |G''| over the step is <= tol, with a
// relative rounding allowance; max |G''| is bounded from the curvature at the
// start (see below).  Returns false -- the caller then defers the cell to
// plain Newton from the stored alpha -- when the start is outside (1, amax)
// or neither test passes.  evals: G evaluations made (0 or 1).  p views f and fneq in FP64.
template <class L, class V>
__device__ __forceinline__ bool residual_correction(const V& p, double amax,
                                                    double start, double tol,
                                                    double& alpha, int& evals,
                                                    EvalCurv& cv) {
  constexpr int Q = L::dimQ;
  evals = 0;
  if (!(start > 1.0 && start < amax)) return false;
  const XTerms<L, V> X(p);
  double ymax = 0.0;
  static_for<Q>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    ymax = fmax(ymax, fabs(fma(p.template f<i>(), 1.0 / L::W[i], -1.0)));
  });
  double S = 0.0, Sn = 0.0;  // sum n ln(f/w), sum n
  auto sums = [&](auto log1p_of) {
    static_for<Q>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      const double ni = p.template n<i>();
      S += ni * log1p_of(fma(p.template f<i>(), 1.0 / L::W[i], -1.0));
      Sn += ni;
    });
  };
  with_log1p(ymax, sums);
  double g, dg, h;
  residual_terms<L>(p, X, S, Sn, start, g, dg, h);
  evals = 1;
  {
    // |n| / (f - a n) = |x| / (1 - a x) <= xmax / (1 - a xmax)
    const double den = 1.0 - start * X.xmax;
    cv = EvalCurv{start, dg, h, den > 0.0 ? X.xmax / den : __builtin_inf()};
  }
  if (fabs(g) <= tol) {
    alpha = start;
    return true;
  }
  const double s = g / dg;
  const double a1 = start - s;
  if (!(a1 > 1.0 && a1 < amax)) return false;  // also a non-finite step
  // -G'' over the step, without a second pass over the populations: every
  // term n_i x_i / (1 - a x_i) is >= 0 (n_i x_i = n_i^2 / f_i), and for a
  // between start and a1, 1/(1 - a x_i) <= (1 + |s| xmax / (1 - am xmax)) /
  // (1 - start x_i) with am = max(start, a1) -- so -G'' <= h(start) rho.
  const double am = fmax(start, a1);
  const double room = 1.0 - am * X.xmax;
  if (!(room > 0.0)) return false;
  const double rho = 1.0 + fabs(s) * X.xmax / room;
  const double bound = 0.5 * s * s * h * rho * (1.0 + 1e-6);
  if (!(bound <= tol)) return false;
  alpha = a1;
  return true;
}

// Residual Newton with G and G' in the reference's logarithmic form
// (EntropicStep.h:31-62): G(a) = sum f ln(f/w) - sum (f - a n) ln((f - a n)/w),
// G'(a) = sum n (1 + ln((f - a n)/w)), one library logarithm per term, the
// constant sum taken once.  Up to max_evals evaluations from `start`:
// accepts the first iterate with |G| <= tol inside (1, amax) -- if STRICT,
// also a Newton point whose Taylor remainder bound is <= tol, without
// evaluating G there (see below); returns false
// when an iterate leaves the evaluation interval -- (1, amax) if STRICT (the
// certified solvers), else a > 0 with every f - a n > 0 (outside, a
// logarithm is NaN and so is G: the reference's loop fails there too) -- or
// max_evals evaluations do not reach tol.
// FP64 rounding of G is ~1e-15 absolute here (terms of order 1e-2), far below
// the tolerances used.  Fewer live values than the log1p form of
// residual_terms (no x_i, no series variants, no curvature), which keeps the
// D3Q27 kernels within their register budget; D2Q9 keeps the cheaper series
// form.  evals: G evaluations made.  FRESH converts the stored populations
// to FP64 anew in every evaluation (V::f_fresh / n_fresh) instead of letting
// the compiler keep 2Q converted doubles live across the iterations: with
// the certificate's terms that held the two-evaluation D3Q27 solver at 61
// spilled registers (7 with FRESH); the one-launch plain Newton, with fewer
// live terms, runs faster without it.
template <class L, bool STRICT, bool FRESH, class V>
__device__ __forceinline__ bool residual_newton_log(const V& p, double amax, double start,
                                                    double tol, int max_evals,
                                                    double& alpha, int& evals,
                                                    EvalCurv& cv) {
  constexpr int Q = L::dimQ;
  evals = 0;
  double SA = 0.0;
  static_for<Q>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    constexpr double inv_w = 1.0 / L::W[i];
    SA += p.template f<i>() * log(p.template f<i>() * inv_w);
  });
  double a = start;
  for (int it = 0; it < max_evals; ++it) {
    // STRICT keeps its iterates inside (1, amax); the plain Newton, like the
    // reference's loop, bounds a only at convergence: an iterate outside the
    // logarithms' domain gives a NaN G and ends in the fallback.
    if (STRICT ? !(a > 1.0 && a < amax) : !(a > 0.0)) return false;
    // STRICT: each evaluation also gathers the curvature terms of the
    // certificate below, h = sum n^2 / fm and r = max |n| / fm, in the same
    // pass over the populations.
    constexpr bool cert = STRICT;
    double g = SA, dg = 0.0, h = 0.0, r = 0.0;
    static_for<Q>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      constexpr double inv_w = 1.0 / L::W[i];
      double ni, fi;
      if constexpr (FRESH) {
        ni = p.template n_fresh<i>();
        fi = p.template f_fresh<i>();
      } else {
        ni = p.template n<i>();
        fi = p.template f<i>();
      }
      const double fm = fi - a * ni;
      const double lg = log(fm * inv_w);
      g -= fm * lg;
      dg += ni * (1.0 + lg);
      if constexpr (cert) {
        const double rf = recip_refined(fm);
        h += ni * ni * rf;
        r = fmax(r, fabs(ni) * rf);
      }
    });
    evals = it + 1;
    cv = EvalCurv{a, dg, h, r};  // h, r: STRICT only
    if (fabs(g) <= tol) {
      if (!(a > 1.0 && a < amax)) return false;
      alpha = a;
      return true;
    }
    const double step = g / dg;
    if constexpr (cert) {
      // Certificate of the Newton point a1 = a - step without evaluating G
      // there: G(a1) = 1/2 step^2 G''(xi) for some xi between a and a1, and
      // f - xi n = fm + (a - xi) n >= fm (1 - |step| r) > 0, so
      // |G''(xi)| = sum n^2 / (f - xi n) <= h / (1 - |step| r); a relative
      // allowance of 1e-6 covers the rounding of the bound.
      const double a1 = a - step;
      const double room = 1.0 - fabs(step) * r;
      if (a1 > 1.0 && a1 < amax && room > 0.0 &&
          0.5 * step * step * h / room * (1.0 + 1e-6) <= tol) {
        alpha = a1;
        return true;
      }
    }
    a -= step;
  }
  return false;
}

// Plain Newton to the residual tolerance: from `start` (the stored alpha),
// alpha <- alpha - G/G' until |G(alpha)| <= tol, at most 50 evaluations, G
// in the same log1p form as the certified correction.  Accepted only inside
// the reference's bounds (1, amax); otherwise -- no convergence, a step
// leaving the logarithm's domain (f - alpha fneq <= 0 for some i), or a root
// outside the bounds -- the reference's fallback alpha = 2
// (Collision.h:328-349) applies.  This is the reference Newton
// (EntropicStep.h:111-140) with the H-residual as its stopping rule instead of
// |step| <= 1e-8.  branch: 2 solved, 3 fallback; evals: G evaluations.
template <class L, class V>
__device__ __forceinline__ double newton_residual(const V& p, double amax, double start,
                                                  double tol, int& branch, int& evals) {
  constexpr int Q = L::dimQ;
  evals = 0;
  branch = 3;
  const XTerms<L, V> X(p);
  double ymax = 0.0;
  static_for<Q>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    ymax = fmax(ymax, fabs(fma(p.template f<i>(), 1.0 / L::W[i], -1.0)));
  });
  double S = 0.0, Sn = 0.0;  // sum n ln(f/w), sum n
  auto sums = [&](auto log1p_of) {
    static_for<Q>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      const double ni = p.template n<i>();
      S += ni * log1p_of(fma(p.template f<i>(), 1.0 / L::W[i], -1.0));
      Sn += ni;
    });
  };
  with_log1p(ymax, sums);
  double a = start;
  for (int it = 1; it <= 50; ++it) {
    // Every logarithm needs f_i - a fneq_i > 0: a xpos < 1 for a >= 0.  An
    // iterate outside -- where the reference's G is NaN, or a < 0, from where
    // its iterates converge to the trivial root -- ends in the fallback, as
    // the reference's does; no bound on a before convergence, like the
    // reference's loop.
    if (!(a >= 0.0 && a * X.xpos < 1.0)) break;
    double g, dg, h;
    residual_terms<L, true>(p, X, S, Sn, a, g, dg, h);
    evals = it;
    if (fabs(g) <= tol) {
      if (a > 1.0 && a < amax) {
        branch = 2;
        return a;
      }
      break;
    }
    a -= g / dg;
  }
  return 2.0;
}

}  // namespace lbm
