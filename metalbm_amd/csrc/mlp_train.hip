// mlp_train.hip -- training and batch inference of the alpha predictor.
//
// Deterministic by construction: each workgroup forms the gradient of its
// chunk of samples with every sum in sample order, a second kernel sums the
// chunk partials in chunk order and applies Adam (Kingma & Ba 2015).  Same
// network, loss (1/2B sum (y - t)^2) and parameter layout as the FP64
// reference in oracle/mlp_ref.c, in FP32.
#include "mlp_train.hpp"

#include "mlp.hpp"

namespace lbm {

namespace {

// Batch gradient, one workgroup per chunk of kSamplesPerBlock samples, as
// GEMM-shaped products of LDS tiles: forward A1 = relu(Z W1^T + b1),
// A2 = relu(A1 W2^T + b2), y = A2 W3 + b3; backward dy = (y - t) / B,
// dW3 = A2^T dy, D2 = (dy W3) o [A2 > 0], dW2 = D2^T A1, D1 = (D2 W2) o
// [A1 > 0], dW1 = D1^T Z (biases: column sums).  Each output of a product is
// one thread's sum over the chunk's samples in sample order, and the chunk
// partials are summed in chunk order by k_mlp_reduce_adam, so the gradient
// is the same bits whatever the timing.
constexpr int kSamplesPerBlock = 64;
constexpr int kTrainThreads = 256;

template <int NIN, int H1, int H2>
__global__ void __launch_bounds__(kTrainThreads)
k_mlp_block_grad(const float* __restrict__ P, const float* __restrict__ X,
                 const float* __restrict__ T, const int* __restrict__ idx, int B,
                 float* __restrict__ partial, float* __restrict__ loss_partial) {
  using Dm = MlpDims<NIN, H1, H2>;
  constexpr int S = kSamplesPerBlock, NT = kTrainThreads;
  __shared__ float sZ[S][NIN];
  __shared__ float sA1[S][H1];   // A1, then D1
  __shared__ float sA2[S][H2];   // A2, then D2
  __shared__ float sdy[S];
  __shared__ float sr2[S];
  const int b0 = blockIdx.x * S;
  const int nb = min(S, B - b0);  // samples of this chunk (the last may be short)
  float* g = partial + (size_t)blockIdx.x * Dm::NPARAM;
  const int t = threadIdx.x;
  // normalised features
  for (int o = t; o < S * NIN; o += NT) {
    const int b = o / NIN, k = o - b * NIN;
    float z = 0.0f;
    if (b < nb) {
      const int s = idx ? idx[b0 + b] : b0 + b;
      z = (X[(size_t)s * NIN + k] - P[Dm::MEAN + k]) * P[Dm::SCALE + k];
    }
    sZ[b][k] = z;
  }
  __syncthreads();
  for (int o = t; o < S * H1; o += NT) {
    const int b = o / H1, j = o - b * H1;
    float acc = P[Dm::B1 + j];
#pragma unroll
    for (int k = 0; k < NIN; ++k) acc = fmaf(P[Dm::W1 + j * NIN + k], sZ[b][k], acc);
    sA1[b][j] = fmaxf(acc, 0.0f);
  }
  __syncthreads();
  for (int o = t; o < S * H2; o += NT) {
    const int b = o / H2, k = o - b * H2;
    float acc = P[Dm::B2 + k];
#pragma unroll 8
    for (int j = 0; j < H1; ++j) acc = fmaf(P[Dm::W2 + k * H1 + j], sA1[b][j], acc);
    sA2[b][k] = fmaxf(acc, 0.0f);
  }
  __syncthreads();
  if (t < S) {
    float y = P[Dm::B3];
#pragma unroll 8
    for (int k = 0; k < H2; ++k) y = fmaf(P[Dm::W3 + k], sA2[t][k], y);
    float r = 0.0f;
    if (t < nb) {
      const int s = idx ? idx[b0 + t] : b0 + t;
      r = y - T[s];
    }
    sr2[t] = r * r;
    sdy[t] = r / float(B);
  }
  __syncthreads();
  // dW3, db3 and the chunk's squared error
  for (int k = t; k < H2 + 2; k += NT) {
    float acc = 0.0f;
    if (k < H2) {
      for (int b = 0; b < nb; ++b) acc = fmaf(sdy[b], sA2[b][k], acc);
      g[Dm::W3 + k] = acc;
    } else if (k == H2) {
      for (int b = 0; b < nb; ++b) acc += sdy[b];
      g[Dm::B3] = acc;
    } else {
      for (int b = 0; b < nb; ++b) acc += sr2[b];
      loss_partial[blockIdx.x] = acc;
    }
  }
  __syncthreads();
  // D2 in place of A2
  for (int o = t; o < S * H2; o += NT) {
    const int b = o / H2, k = o - b * H2;
    sA2[b][k] = sA2[b][k] > 0.0f ? sdy[b] * P[Dm::W3 + k] : 0.0f;
  }
  __syncthreads();
  // dW2 = D2^T A1, db2
  for (int o = t; o < H2 * H1 + H2; o += NT) {
    float acc = 0.0f;
    if (o < H2 * H1) {
      const int k = o / H1, j = o - k * H1;
      for (int b = 0; b < nb; ++b) acc = fmaf(sA2[b][k], sA1[b][j], acc);
      g[Dm::W2 + o] = acc;
    } else {
      const int k = o - H2 * H1;
      for (int b = 0; b < nb; ++b) acc += sA2[b][k];
      g[Dm::B2 + k] = acc;
    }
  }
  __syncthreads();
  // D1 = (D2 W2) o [A1 > 0], in place of A1 (each element read and written by
  // one thread; every reader of A1 above has passed the barrier)
  for (int o = t; o < S * H1; o += NT) {
    const int b = o / H1, j = o - b * H1;
    float d = 0.0f;
    if (sA1[b][j] > 0.0f) {
#pragma unroll 8
      for (int k = 0; k < H2; ++k) d = fmaf(P[Dm::W2 + k * H1 + j], sA2[b][k], d);
    }
    sA1[b][j] = d;
  }
  __syncthreads();
  // dW1 = D1^T Z, db1
  for (int o = t; o < H1 * NIN + H1; o += NT) {
    float acc = 0.0f;
    if (o < H1 * NIN) {
      const int j = o / NIN, k = o - j * NIN;
      for (int b = 0; b < nb; ++b) acc = fmaf(sA1[b][j], sZ[b][k], acc);
      g[Dm::W1 + o] = acc;
    } else {
      const int j = o - H1 * NIN;
      for (int b = 0; b < nb; ++b) acc += sA1[b][j];
      g[Dm::B1 + j] = acc;
    }
  }
}

// g[p] = sum of the chunk partials in chunk order; then, if adam, one Adam
// step (Kingma & Ba 2015) with c1 = 1 - b1^t, c2 = 1 - b2^t.  Thread 0 also
// sums the chunk losses in order.
__global__ void __launch_bounds__(256)
k_mlp_reduce_adam(const float* __restrict__ partial, const float* __restrict__ loss_partial,
                  int nchunks, int np, float* __restrict__ g, float* __restrict__ loss,
                  int adam, float* __restrict__ P, float* __restrict__ m,
                  float* __restrict__ v, float lr, float b1, float b2, float eps, float c1,
                  float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p == 0) {
    float s = 0.0f;
    for (int c = 0; c < nchunks; ++c) s += loss_partial[c];
    *loss = s;
  }
  if (p >= np) return;
  float gp = 0.0f;
  for (int c = 0; c < nchunks; ++c) gp += partial[(size_t)c * np + p];
  g[p] = gp;
  if (!adam) return;
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
  using M = MfmaLayout<NIN, H1, H2>;
  char* wave_tile = s_lds + M::NET + (threadIdx.x >> 6) * M::TILE;
  copy_packed_network<NIN, H1, H2>(packed, s_lds);
  __syncthreads();
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  float m[NIN];
  for (int k = 0; k < NIN; ++k) m[k] = s < n ? X[(size_t)s * NIN + k] : 0.0f;
  const double y = mlp_alpha_wave<NIN, H1, H2, T>(s_lds, wave_tile, m);
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

int mlp_grad_chunks(int B) { return (B + kSamplesPerBlock - 1) / kSamplesPerBlock; }

hipError_t launch_mlp_grad(int lattice, const float* P, const float* X, const float* T,
                           const int* idx, int B, float* partial, float* loss_partial,
                           hipStream_t st) {
  return by_lattice(lattice, [&](auto nin, auto h1, auto h2) {
    hipLaunchKernelGGL((k_mlp_block_grad<decltype(nin)::value, decltype(h1)::value,
                                         decltype(h2)::value>),
                       dim3(mlp_grad_chunks(B)), dim3(kTrainThreads), 0, st, P, X, T, idx, B,
                       partial, loss_partial);
    return hipGetLastError();
  });
}

hipError_t launch_mlp_reduce_adam(const float* partial, const float* loss_partial,
                                  int nchunks, int np, float* g, float* loss, int adam,
                                  float* P, float* m, float* v, float lr, float b1, float b2,
                                  float eps, float c1, float c2, hipStream_t st) {
  hipLaunchKernelGGL(k_mlp_reduce_adam, dim3((np + 255) / 256), dim3(256), 0, st, partial,
                     loss_partial, nchunks, np, g, loss, adam, P, m, v, lr, b1, b2, eps, c1,
                     c2);
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
