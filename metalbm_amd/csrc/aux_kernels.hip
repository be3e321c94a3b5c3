// aux_kernels.hip -- initial conditions, population conversion and the FP64
// diagnostics reduction.  Built with -ffp-contract=off: the equilibrium
// initialisation reproduces the reference's FP64 operation order
// (Initialize.h:106-117 via Equilibrium.h).
#include "aux_kernels.hpp"

#include "step_kernels.hpp"

namespace lbm {

namespace {

template <class S>
__device__ __forceinline__ void put(S* p, double v) {
  if constexpr (std::is_same_v<S, _Float16>) *p = _Float16(float(v));
  else *p = S(v);
}

template <class L, int EQ, class S, bool SH>
__global__ void k_init_eq(S* f, const double* rho, double rho0,
                          const double* u, int N) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  const double r = rho ? rho[c] : rho0;
  double v[3] = {0.0, 0.0, 0.0};
  for (int d = 0; d < L::dimD; ++d) v[d] = u[(size_t)d * N + c];
  double v2 = v[0] * v[0];
  for (int d = 1; d < L::dimD; ++d) v2 += v[d] * v[d];
  const RefEquilibrium<L, EQ, double> E(r, v, v2);
  static_for<L::dimQ>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    const double fe = E.template feq<i>();
    double val = SH ? fe - L::W[i] : fe;
    if constexpr (std::is_same_v<S, _Float16>) val *= kHalfScale<L, i>;
    put(f + (size_t)i * N + c, val);
  });
}

template <class L, int EQ>
hipError_t init_eq_l(int storage, int shifted, void* f, const double* rho,
                     double rho0, const double* u, int N, hipStream_t st) {
  const dim3 b(256), g((N + 255) / 256);
  if (storage == TYPE_F64 && !shifted)
    hipLaunchKernelGGL((k_init_eq<L, EQ, double, false>), g, b, 0, st,
                       (double*)f, rho, rho0, u, N);
  else if (storage == TYPE_F64)
    hipLaunchKernelGGL((k_init_eq<L, EQ, double, true>), g, b, 0, st,
                       (double*)f, rho, rho0, u, N);
  else if (storage == TYPE_F32 && !shifted)
    hipLaunchKernelGGL((k_init_eq<L, EQ, float, false>), g, b, 0, st,
                       (float*)f, rho, rho0, u, N);
  else if (storage == TYPE_F32)
    hipLaunchKernelGGL((k_init_eq<L, EQ, float, true>), g, b, 0, st,
                       (float*)f, rho, rho0, u, N);
  else
    hipLaunchKernelGGL((k_init_eq<L, EQ, _Float16, true>), g, b, 0, st,
                       (_Float16*)f, rho, rho0, u, N);
  return hipGetLastError();
}

template <class L>
hipError_t init_eq_lat(int eq, int storage, int shifted, void* f,
                       const double* rho, double rho0, const double* u, int N,
                       hipStream_t st) {
  if constexpr (L::Type != LatticeType::D3Q19) {
    if (eq == (int)EquilibriumType::Exact)
      return init_eq_l<L, (int)EquilibriumType::Exact>(storage, shifted, f, rho,
                                                        rho0, u, N, st);
  }
  return init_eq_l<L, (int)EquilibriumType::TruncationMa3>(storage, shifted, f,
                                                          rho, rho0, u, N, st);
}

__device__ __forceinline__ uint64_t splitmix64(uint64_t seed, uint64_t k) {
  uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

__global__ void k_noise(double* u, int N, int D, uint64_t seed, double amp) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  for (int d = 0; d < D; ++d) {
    const uint64_t r = splitmix64(seed, (uint64_t)c * D + d);
    const double unit = (double)(r >> 11) * 0x1.0p-53;
    u[(size_t)d * N + c] = amp * (2.0 * unit - 1.0);
  }
}

__global__ void k_taylor_green(double* u, int nx, int ny, int nz, int D,
                               double u0, const double* sx, const double* cx,
                               const double* sy, const double* cy,
                               const double* cz) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int N = nx * ny * nz;
  if (c >= N) return;
  const int z = c % nz, y = (c / nz) % ny, x = c / (nz * ny);
  const double czv = D == 3 ? cz[z] : 1.0;
  u[c] = u0 * sx[x] * cy[y] * czv;
  u[(size_t)N + c] = -u0 * cx[x] * sy[y] * czv;
  if (D == 3) u[(size_t)2 * N + c] = 0.0;
}

// Spectral initialisation: three separable inverse-DFT passes over a
// Hermitian mode cube uh[M][M][Mz][D] (M = 2K+1, Mz = 2Kz+1), each pass a
// sum in ascending wavenumber with twiddles e^{2 pi i t/n} from cos/sin
// tables, in the operation order of the FP64 oracle.
__device__ __forceinline__ int twiddle(int k, int p, int n) {
  return ((k * p) % n + n) % n;
}

__global__ void k_spec_z(const double* ur, const double* ui, int M, int Mz, int Kz,
                         int D, int nz, const double* tc, const double* ts,
                         double* t1r, double* t1i) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M * M * nz) return;
  const int ixy = i / nz, z = i - ixy * nz;
  for (int d = 0; d < D; ++d) {
    double re = 0.0, im = 0.0;
    for (int iz = 0; iz < Mz; ++iz) {
      const int t = twiddle(iz - Kz, z, nz);
      const double c = tc[t], sv = ts[t];
      const size_t p = ((size_t)ixy * Mz + iz) * D + d;
      const double a = ur[p], b = ui[p];
      const double tr = a * c - b * sv;
      const double ti = a * sv + b * c;
      re += tr;
      im += ti;
    }
    const size_t o = ((size_t)ixy * nz + z) * D + d;
    t1r[o] = re;
    t1i[o] = im;
  }
}

__global__ void k_spec_y(const double* t1r, const double* t1i, int M, int K, int D,
                         int ny, int nz, const double* tc, const double* ts,
                         double* t2r, double* t2i) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M * ny * nz) return;
  const int z = i % nz, y = (i / nz) % ny, ix = i / (nz * ny);
  for (int d = 0; d < D; ++d) {
    double re = 0.0, im = 0.0;
    for (int iy = 0; iy < M; ++iy) {
      const int t = twiddle(iy - K, y, ny);
      const double c = tc[t], sv = ts[t];
      const size_t p = (((size_t)ix * M + iy) * nz + z) * D + d;
      const double a = t1r[p], b = t1i[p];
      const double tr = a * c - b * sv;
      const double ti = a * sv + b * c;
      re += tr;
      im += ti;
    }
    const size_t o = (((size_t)ix * ny + y) * nz + z) * D + d;
    t2r[o] = re;
    t2i[o] = im;
  }
}

__global__ void k_spec_x(const double* t2r, const double* t2i, int M, int K, int D,
                         int nx, int ny, int nz, const double* tc, const double* ts,
                         double gamma, double* u) {
  const int N = nx * ny * nz;
  const int c0 = blockIdx.x * blockDim.x + threadIdx.x;
  if (c0 >= N) return;
  const int z = c0 % nz, y = (c0 / nz) % ny, x = c0 / (nz * ny);
  for (int d = 0; d < D; ++d) {
    double re = 0.0;
    for (int ix = 0; ix < M; ++ix) {
      const int t = twiddle(ix - K, x, nx);
      const size_t p = (((size_t)ix * ny + y) * nz + z) * D + d;
      const double tr = t2r[p] * tc[t] - t2i[p] * ts[t];
      re += tr;
    }
    u[(size_t)d * N + c0] = gamma * re;
  }
}

template <class S>
__global__ void k_pop_from_f64(S* dst, const double* src, int N, double w,
                               double scale) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  put(dst + c, (src[c] - w) * scale);
}

template <class S>
__global__ void k_pop_to_f64(double* dst, const S* src, int N, double w,
                             double inv_scale) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  double v;
  if constexpr (std::is_same_v<S, _Float16>) v = double(float(src[c]));
  else v = double(src[c]);
  dst[c] = v * inv_scale + w;
}

template <class S>
__global__ void k_fill(S* dst, double value, int N) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < N) put(dst + c, value);
}

// ---- deterministic reduction ----------------------------------------------

constexpr int RED_THREADS = 256;
constexpr int RED_CELLS = 8192;  // cells per block, fixed: result depends on N only

__device__ __forceinline__ int wrapi(int v, int n) {
  return v < 0 ? v + n : (v >= n ? v - n : v);
}

__global__ void __launch_bounds__(RED_THREADS)
k_reduce_partial(const double* rho, const double* u, int nx, int ny, int nz,
                 int D, double* partial) {
  __shared__ double sh[3][RED_THREADS];
  const int N = nx * ny * nz;
  const int begin = blockIdx.x * RED_CELLS;
  const int end = min(begin + RED_CELLS, N);
  double m = 0.0, e = 0.0, w = 0.0;
  for (int c = begin + threadIdx.x; c < end; c += RED_THREADS) {
    const int z = c % nz, y = (c / nz) % ny, x = c / (nz * ny);
    m += rho[c];
    double s = 0.0;
    for (int d = 0; d < D; ++d) s += 0.5 * rho[c] * u[(size_t)d * N + c] * u[(size_t)d * N + c];
    e += s;
    auto U = [&](int d, int X, int Y, int Z) {
      return u[(size_t)d * N + ((size_t)X * ny + Y) * nz + Z];
    };
    const int xp = wrapi(x + 1, nx), xm = wrapi(x - 1, nx);
    const int yp = wrapi(y + 1, ny), ym = wrapi(y - 1, ny);
    if (D == 2) {
      const double om = 0.5 * (U(1, xp, y, z) - U(1, xm, y, z)) -
                        0.5 * (U(0, x, yp, z) - U(0, x, ym, z));
      w += 0.5 * om * om;
    } else {
      const int zp = wrapi(z + 1, nz), zm = wrapi(z - 1, nz);
      const double ox = 0.5 * (U(2, x, yp, z) - U(2, x, ym, z)) -
                        0.5 * (U(1, x, y, zp) - U(1, x, y, zm));
      const double oy = 0.5 * (U(0, x, y, zp) - U(0, x, y, zm)) -
                        0.5 * (U(2, xp, y, z) - U(2, xm, y, z));
      const double oz = 0.5 * (U(1, xp, y, z) - U(1, xm, y, z)) -
                        0.5 * (U(0, x, yp, z) - U(0, x, ym, z));
      w += 0.5 * (ox * ox + oy * oy + oz * oz);
    }
  }
  sh[0][threadIdx.x] = m;
  sh[1][threadIdx.x] = e;
  sh[2][threadIdx.x] = w;
  __syncthreads();
  for (int s = RED_THREADS / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int k = 0; k < 3; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < 3) partial[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = sh[threadIdx.x][0];
}

__global__ void __launch_bounds__(RED_THREADS)
k_reduce_final(const double* partial, int nb, double* out) {
  __shared__ double sh[3][RED_THREADS];
  for (int k = 0; k < 3; ++k) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += RED_THREADS) s += partial[(size_t)k * nb + b];
    sh[k][threadIdx.x] = s;
  }
  __syncthreads();
  for (int s = RED_THREADS / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int k = 0; k < 3; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < 3) out[threadIdx.x] = sh[threadIdx.x][0];
}

template <class L>
void tables_of(double* w, double* scale) {
  static_for<L::dimQ>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    w[i] = L::W[i];
    scale[i] = kHalfScale<L, i>;
  });
}

}  // namespace

hipError_t launch_init_equilibrium(int lattice, int eq, int storage,
                                   int shifted, void* f, const double* rho,
                                   double rho0, const double* u, int N,
                                   hipStream_t st) {
  switch (lattice) {
    case (int)LatticeType::D2Q9:
      return init_eq_lat<Lattice<LatticeType::D2Q9>>(eq, storage, shifted, f, rho, rho0, u, N, st);
    case (int)LatticeType::D3Q19:
      return init_eq_lat<Lattice<LatticeType::D3Q19>>(eq, storage, shifted, f, rho, rho0, u, N, st);
    default:
      return init_eq_lat<Lattice<LatticeType::D3Q27>>(eq, storage, shifted, f, rho, rho0, u, N, st);
  }
}

hipError_t launch_noise(double* u, int N, int D, uint64_t seed, double amp,
                        hipStream_t st) {
  hipLaunchKernelGGL(k_noise, dim3((N + 255) / 256), dim3(256), 0, st, u, N, D,
                     seed, amp);
  return hipGetLastError();
}

hipError_t launch_taylor_green(double* u, int nx, int ny, int nz, int D,
                               double u0, const double* sx, const double* cx,
                               const double* sy, const double* cy,
                               const double* cz, hipStream_t st) {
  const int N = nx * ny * nz;
  hipLaunchKernelGGL(k_taylor_green, dim3((N + 255) / 256), dim3(256), 0, st, u,
                     nx, ny, nz, D, u0, sx, cx, sy, cy, cz);
  return hipGetLastError();
}

hipError_t launch_spectral(const double* ur, const double* ui, int K, int D,
                           int nx, int ny, int nz, const double* tab, double gamma,
                           double* t1r, double* t1i, double* t2r, double* t2i,
                           double* u, hipStream_t st) {
  const int M = 2 * K + 1, Kz = D == 3 ? K : 0, Mz = 2 * Kz + 1;
  const double *cx = tab, *sx = cx + nx, *cy = sx + nx, *sy = cy + ny,
               *cz = sy + ny, *sz = cz + nz;
  const int n1 = M * M * nz, n2 = M * ny * nz, n3 = nx * ny * nz;
  hipLaunchKernelGGL(k_spec_z, dim3((n1 + 255) / 256), dim3(256), 0, st, ur, ui, M,
                     Mz, Kz, D, nz, cz, sz, t1r, t1i);
  hipLaunchKernelGGL(k_spec_y, dim3((n2 + 255) / 256), dim3(256), 0, st, t1r, t1i,
                     M, K, D, ny, nz, cy, sy, t2r, t2i);
  hipLaunchKernelGGL(k_spec_x, dim3((n3 + 255) / 256), dim3(256), 0, st, t2r, t2i,
                     M, K, D, nx, ny, nz, cx, sx, gamma, u);
  return hipGetLastError();
}

hipError_t launch_pop_from_f64(int storage, void* dst, const double* src,
                               int N, double w, double scale, hipStream_t st) {
  const dim3 b(256), g((N + 255) / 256);
  if (storage == TYPE_F64)
    hipLaunchKernelGGL(k_pop_from_f64<double>, g, b, 0, st, (double*)dst, src, N, w, scale);
  else if (storage == TYPE_F32)
    hipLaunchKernelGGL(k_pop_from_f64<float>, g, b, 0, st, (float*)dst, src, N, w, scale);
  else
    hipLaunchKernelGGL(k_pop_from_f64<_Float16>, g, b, 0, st, (_Float16*)dst, src, N, w, scale);
  return hipGetLastError();
}

hipError_t launch_pop_to_f64(int storage, double* dst, const void* src, int N,
                             double w, double inv_scale, hipStream_t st) {
  const dim3 b(256), g((N + 255) / 256);
  if (storage == TYPE_F64)
    hipLaunchKernelGGL(k_pop_to_f64<double>, g, b, 0, st, dst, (const double*)src, N, w, inv_scale);
  else if (storage == TYPE_F32)
    hipLaunchKernelGGL(k_pop_to_f64<float>, g, b, 0, st, dst, (const float*)src, N, w, inv_scale);
  else
    hipLaunchKernelGGL(k_pop_to_f64<_Float16>, g, b, 0, st, dst, (const _Float16*)src, N, w, inv_scale);
  return hipGetLastError();
}

hipError_t launch_scalar_from_f64(int type, void* dst, const double* src,
                                  int N, hipStream_t st) {
  return launch_pop_from_f64(type, dst, src, N, 0.0, 1.0, st);
}

hipError_t launch_scalar_to_f64(int type, double* dst, const void* src, int N,
                                hipStream_t st) {
  return launch_pop_to_f64(type, dst, src, N, 0.0, 1.0, st);
}

hipError_t launch_fill(int type, void* dst, double value, int N,
                       hipStream_t st) {
  const dim3 b(256), g((N + 255) / 256);
  if (type == TYPE_F64)
    hipLaunchKernelGGL(k_fill<double>, g, b, 0, st, (double*)dst, value, N);
  else if (type == TYPE_F32)
    hipLaunchKernelGGL(k_fill<float>, g, b, 0, st, (float*)dst, value, N);
  else
    hipLaunchKernelGGL(k_fill<_Float16>, g, b, 0, st, (_Float16*)dst, value, N);
  return hipGetLastError();
}

// Restart fields (Initialize.h:124-145): density = sum_i f_i and velocity =
// sum_i c_i f_i -- the reference does not divide by the density there --
// accumulated one population at a time in index order.
__global__ void k_restart_fields(double* __restrict__ rho, double* __restrict__ u,
                                 const double* __restrict__ f, int N, int D, int cx, int cy,
                                 int cz, int first) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= N) return;
  const double v = f[x];
  const int c[3] = {cx, cy, cz};
  rho[x] = first ? v : rho[x] + v;
  for (int d = 0; d < D; ++d) {
    const double m = (double)c[d] * v;
    u[(size_t)d * N + x] = first ? m : u[(size_t)d * N + x] + m;
  }
}

hipError_t launch_restart_fields(int lattice, double* rho, double* u, const double* f, int i,
                                 int N, hipStream_t st) {
  int c[3] = {0, 0, 0}, D = 3;
  switch (lattice) {
    case (int)LatticeType::D2Q9:
      D = 2;
      for (int d = 0; d < 3; ++d) c[d] = Lattice<LatticeType::D2Q9>::C[i][d];
      break;
    case (int)LatticeType::D3Q19:
      for (int d = 0; d < 3; ++d) c[d] = Lattice<LatticeType::D3Q19>::C[i][d];
      break;
    default:
      for (int d = 0; d < 3; ++d) c[d] = Lattice<LatticeType::D3Q27>::C[i][d];
      break;
  }
  hipLaunchKernelGGL(k_restart_fields, dim3((N + 255) / 256), dim3(256), 0, st, rho, u, f, N, D,
                     c[0], c[1], c[2], i == 0 ? 1 : 0);
  return hipGetLastError();
}

int reduce_blocks(int N) { return (N + RED_CELLS - 1) / RED_CELLS; }

hipError_t launch_reduce(const double* rho, const double* u, int nx, int ny,
                         int nz, int D, double* partial, double* out,
                         hipStream_t st) {
  const int nb = reduce_blocks(nx * ny * nz);
  hipLaunchKernelGGL(k_reduce_partial, dim3(nb), dim3(RED_THREADS), 0, st, rho,
                     u, nx, ny, nz, D, partial);
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(RED_THREADS), 0, st,
                     partial, nb, out);
  return hipGetLastError();
}

int lattice_tables(int lattice, double* w, double* scale, int* q, int* d) {
  switch (lattice) {
    case (int)LatticeType::D2Q9:
      tables_of<Lattice<LatticeType::D2Q9>>(w, scale);
      *q = 9; *d = 2;
      return 0;
    case (int)LatticeType::D3Q19:
      tables_of<Lattice<LatticeType::D3Q19>>(w, scale);
      *q = 19; *d = 3;
      return 0;
    case (int)LatticeType::D3Q27:
      tables_of<Lattice<LatticeType::D3Q27>>(w, scale);
      *q = 27; *d = 3;
      return 0;
    default:
      return -1;
  }
}

}  // namespace lbm
