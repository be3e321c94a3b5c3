// mlp_train.hip -- training and batch inference of the alpha predictor.
//
// Deterministic by construction: one thread computes one sample's gradient
// and writes it parameter-major (G[p][b]); a second kernel sums each
// parameter's column in a fixed order; a third applies Adam (Kingma & Ba
// 2015).  Same network, loss (1/2B sum (y - t)^2) and parameter layout as the
// FP64 reference in oracle/mlp_ref.c, in FP32.
#include "mlp_train.hpp"

#include "mlp.hpp"

namespace lbm {

namespace {

template <int NIN, int H1, int H2>
__global__ void __launch_bounds__(64)
k_mlp_sample_grad(const float* __restrict__ P, const float* __restrict__ X,
                  const float* __restrict__ T, const int* __restrict__ idx,
                  int B, float* __restrict__ G, float* __restrict__ loss) {
  using Dm = MlpDims<NIN, H1, H2>;
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const int s = idx ? idx[b] : b;
  float z[NIN];
  for (int k = 0; k < NIN; ++k)
    z[k] = (X[(size_t)s * NIN + k] - P[Dm::MEAN + k]) * P[Dm::SCALE + k];
  float a1[H1], a2[H2];
  for (int j = 0; j < H1; ++j) {
    float acc = P[Dm::B1 + j];
    for (int k = 0; k < NIN; ++k) acc = fmaf(P[Dm::W1 + j * NIN + k], z[k], acc);
    a1[j] = fmaxf(acc, 0.0f);
  }
  float y = P[Dm::B3];
  for (int j = 0; j < H2; ++j) {
    float acc = P[Dm::B2 + j];
    for (int k = 0; k < H1; ++k) acc = fmaf(P[Dm::W2 + j * H1 + k], a1[k], acc);
    a2[j] = fmaxf(acc, 0.0f);
    y = fmaf(P[Dm::W3 + j], a2[j], y);
  }
  const float r = y - T[s];
  loss[b] = r * r;
  const float dy = r / float(B);
  auto put = [&](int p, float v) { G[(size_t)p * B + b] = v; };
  put(Dm::B3, dy);
  float d1[H1];
  for (int j = 0; j < H1; ++j) d1[j] = 0.0f;
  for (int k = 0; k < H2; ++k) {
    put(Dm::W3 + k, dy * a2[k]);
    const float d2 = a2[k] > 0.0f ? dy * P[Dm::W3 + k] : 0.0f;
    put(Dm::B2 + k, d2);
    for (int j = 0; j < H1; ++j) {
      put(Dm::W2 + k * H1 + j, d2 * a1[j]);
      d1[j] = fmaf(P[Dm::W2 + k * H1 + j], d2, d1[j]);
    }
  }
  for (int j = 0; j < H1; ++j) {
    const float d = a1[j] > 0.0f ? d1[j] : 0.0f;
    put(Dm::B1 + j, d);
    for (int k = 0; k < NIN; ++k) put(Dm::W1 + j * NIN + k, d * z[k]);
  }
}

__global__ void __launch_bounds__(256)
k_mlp_sum(const float* __restrict__ G, int B, int np, float* __restrict__ g) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const float* row = G + (size_t)p * B;
  float s = 0.0f;
  for (int b = 0; b < B; ++b) s += row[b];
  g[p] = s;
}

__global__ void __launch_bounds__(256)
k_mlp_adam(float* __restrict__ P, const float* __restrict__ g,
           float* __restrict__ m, float* __restrict__ v, int np, float lr,
           float b1, float b2, float eps, float c1, float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const float gp = g[p];
  const float mp = b1 * m[p] + (1.0f - b1) * gp;
  const float vp = b2 * v[p] + (1.0f - b2) * gp * gp;
  m[p] = mp;
  v[p] = vp;
  P[p] -= lr * (mp / c1) / (sqrtf(vp / c2) + eps);
}

// Batch inference with the step kernel's own matrix-core code (f16 operands,
// FP32 accumulation), so predictions equal what the fused step starts from.
template <int NIN, int H1, int H2, class T>
__global__ void __launch_bounds__(256) k_mlp_pack(const float* __restrict__ P,
                                                  char* __restrict__ packed) {
  stage_mfma_network<NIN, H1, H2, T>(P, packed);
}

template <int NIN, int H1, int H2, class T>
__global__ void __launch_bounds__(256)
k_mlp_forward(const void* __restrict__ packed, const float* __restrict__ X, int n,
              double* __restrict__ alpha) {
  extern __shared__ __align__(16) char s_lds[];
  char* wave_lds = s_lds + (threadIdx.x >> 6) * MfmaLayout<NIN, H1, H2>::WAVE;
  copy_packed_network<NIN, H1, H2>(packed, wave_lds);
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  float m[NIN];
  for (int k = 0; k < NIN; ++k) m[k] = s < n ? X[(size_t)s * NIN + k] : 0.0f;
  const double y = mlp_alpha_wave<NIN, H1, H2, T>(wave_lds, m);
  if (s < n) alpha[s] = y;
}

template <class F>
hipError_t by_lattice(int lattice, F&& f) {
  if (lattice == (int)LatticeType::D2Q9) {
    using S = MlpShape<Lattice<LatticeType::D2Q9>>;
    return f(std::integral_constant<int, S::NIN>{}, std::integral_constant<int, S::H1>{},
             std::integral_constant<int, S::H2>{});
  }
  if (lattice == (int)LatticeType::D3Q27) {
    using S = MlpShape<Lattice<LatticeType::D3Q27>>;
    return f(std::integral_constant<int, S::NIN>{}, std::integral_constant<int, S::H1>{},
             std::integral_constant<int, S::H2>{});
  }
  return hipErrorInvalidValue;
}

}  // namespace

int mlp_shape(int lattice, int* nin, int* h1, int* h2, int* nparam, int* nblock) {
  auto set = [&](auto S) {
    using Sh = decltype(S);
    using Dm = MlpDims<Sh::NIN, Sh::H1, Sh::H2>;
    *nin = Sh::NIN;
    *h1 = Sh::H1;
    *h2 = Sh::H2;
    *nparam = Dm::NPARAM;
    *nblock = Dm::NBLOCK;
    return 0;
  };
  if (lattice == (int)LatticeType::D2Q9) return set(MlpShape<Lattice<LatticeType::D2Q9>>{});
  if (lattice == (int)LatticeType::D3Q27) return set(MlpShape<Lattice<LatticeType::D3Q27>>{});
  return -1;
}

hipError_t launch_mlp_sample_grad(int lattice, const float* P, const float* X,
                                  const float* T, const int* idx, int B,
                                  float* G, float* loss, hipStream_t st) {
  return by_lattice(lattice, [&](auto nin, auto h1, auto h2) {
    hipLaunchKernelGGL((k_mlp_sample_grad<decltype(nin)::value, decltype(h1)::value,
                                          decltype(h2)::value>),
                       dim3((B + 63) / 64), dim3(64), 0, st, P, X, T, idx, B, G, loss);
    return hipGetLastError();
  });
}

hipError_t launch_mlp_sum(const float* G, int B, int np, float* g, hipStream_t st) {
  hipLaunchKernelGGL(k_mlp_sum, dim3((np + 255) / 256), dim3(256), 0, st, G, B, np, g);
  return hipGetLastError();
}

hipError_t launch_mlp_adam(float* P, const float* g, float* m, float* v, int np,
                           float lr, float b1, float b2, float eps, float c1,
                           float c2, hipStream_t st) {
  hipLaunchKernelGGL(k_mlp_adam, dim3((np + 255) / 256), dim3(256), 0, st, P, g, m,
                     v, np, lr, b1, b2, eps, c1, c2);
  return hipGetLastError();
}

template <int NIN>
using OpOf = std::conditional_t<NIN == 9, MlpShape<Lattice<LatticeType::D2Q9>>::Op,
                                MlpShape<Lattice<LatticeType::D3Q27>>::Op>;

size_t mlp_packed_bytes(int lattice) {
  size_t b = 0;
  by_lattice(lattice, [&](auto nin, auto h1, auto h2) {
    b = MfmaLayout<decltype(nin)::value, decltype(h1)::value, decltype(h2)::value>::NET;
    return hipSuccess;
  });
  return b;
}

hipError_t launch_mlp_pack(int lattice, const float* P, void* packed, hipStream_t st) {
  return by_lattice(lattice, [&](auto nin, auto h1, auto h2) {
    constexpr int I = decltype(nin)::value;
    hipLaunchKernelGGL((k_mlp_pack<I, decltype(h1)::value, decltype(h2)::value, OpOf<I>>),
                       dim3(1), dim3(256), 0, st, P, static_cast<char*>(packed));
    return hipGetLastError();
  });
}

hipError_t launch_mlp_forward(int lattice, const void* packed, const float* X, int n,
                              double* alpha, hipStream_t st) {
  return by_lattice(lattice, [&](auto nin, auto h1, auto h2) {
    using M = MfmaLayout<decltype(nin)::value, decltype(h1)::value, decltype(h2)::value>;
    using T = OpOf<decltype(nin)::value>;
    hipLaunchKernelGGL((k_mlp_forward<decltype(nin)::value, decltype(h1)::value,
                                      decltype(h2)::value, T>),
                       dim3((n + 255) / 256), dim3(256), M::BYTES, st, packed, X, n, alpha);
    return hipGetLastError();
  });
}

}  // namespace lbm
