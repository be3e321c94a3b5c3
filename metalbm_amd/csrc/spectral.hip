// spectral.hip -- spectral diagnostics of the stored fields: the curl of the
// hydrodynamic velocity and the shell-binned energy and forcing spectra.
//
// Reference: Curl<double, CPU, OneD, 2|3> (Transformer.h:114-295) called as
// executeSpace + normalize after every storing step (Routine.h:98-101,
// 129-132), TotalEnstrophy (Analysis.h:68-98), and SpectralAnalysisList::
// writeAnalyses with PowerSpectra (AnalysisList.h:98-171, Analysis.h:100-176).
// The reference runs FFTW-MPI on the host; here the real-to-complex and
// complex-to-real transforms are hipFFT plans over the device fields, and the
// Fourier-space work (curl, shell binning) is this file's kernels.
//
// Two scalings (lbm_vorticity / lbm_spectra):
//  REFERENCE -- the reference's numbers: integer wavenumbers k_d (index i <=
//    n/2 -> i, else i - n), vorticity divided by N twice (the backward
//    transform's 1/N, Transformer.h:100-111, then Curl::normalize,
//    Transformer.h:179-185, 284-294), the 2D curl's imaginary part set to
//    zero (Transformer.h:161-165); shells floor(|k|) (int MathVector::norm,
//    MathVector.h:65); weight 1/2 on the k_last = 0 plane (Analysis.h:
//    150-153); the energy spectrum divided by N, the forcing spectrum not
//    normalised (AnalysisList.h:195 normalises energy only).
//  PHYSICAL -- wavenumbers 2 pi k_d / n_d, one 1/N, the full complex 2D
//    curl; shells round(|k|); weight 1/2 on the k_last = 0 and (even n)
//    k_last = n/2 planes; both spectra divided by N^2, so that
//    sum_s E(s) = (1/N) sum_x 1/2 |u|^2 when every mode lies below the last
//    shell.
// Nyquist indices (n_d/2 of an even axis): PHYSICAL takes the derivative
// there as zero.  REFERENCE keeps the reference's k = +n_d/2 except on the
// self-paired planes k_last = 0 and k_last = n_last/2, where each mode's
// conjugate partner is stored too: there the reference's +n/2 makes the
// spectrum non-Hermitian, and the complex-to-real inverse returns its
// Hermitian part, in which the Nyquist derivative is zero -- so it is zero
// here as well.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "capi_internal.hpp"

namespace {

constexpr int kBlock = 256;
constexpr int kReduceBlocks = 1024;

struct State {
  int D = 0, C = 0, n[3] = {1, 1, 1};
  int64_t N = 0, M = 0;   // real cells, half-spectrum modes
  int nshells = 0;
  hipfftHandle fwd = 0, inv = 0;
  bool have_fwd = false, have_inv = false;
  double* real = nullptr;     // D N: transform input copy
  double2* hat = nullptr;     // D M
  double2* what = nullptr;    // C M
  double* omega = nullptr;    // C N
  double* spec = nullptr;     // 2 nshells
  double* partial = nullptr;  // kReduceBlocks + 1
  double* bins = nullptr;     // shell partials: shell_blocks x min(nshells, window)
  int shell_blocks = 0;
  const double* hat_src = nullptr;  // field whose spectrum hat holds, and its version
  uint64_t hat_version = 0;
  hipfftHandle inv_d = 0;     // inverse transform of D components (force field)
  bool have_inv_d = false;
  // direct transform of a small removal shell (spectral forces)
  bool direct_checked = false, direct = false;
  int nmodes = 0, dft_blocks = 0;
  int npadded = 0;                // nmodes rounded up to whole passes
  int dft_rows = 0;               // 2D small shells: row groups R of the column walk
  bool small_k = false;           // every |k_d| of the modes <= 2
  int4* modes = nullptr;          // (kx, ky, kz, Hermitian weight), zero-padded
  double2* tw = nullptr;          // e^{-2 pi i m / n} per axis: [nx] [ny] [nz]
  double2* dft_partial = nullptr; // [dft_blocks][nmodes][D]
  double2* coef = nullptr;        // [npadded][D], zero beyond nmodes
  double2* dense = nullptr;       // 2D small shells: [5][3][2] coefficient grid
  double* finj = nullptr;         // injection field [D][N] (Turbulent2D)
};

void free_state(State* s) {
  if (!s) return;
  if (s->have_fwd) hipfftDestroy(s->fwd);
  if (s->have_inv) hipfftDestroy(s->inv);
  if (s->have_inv_d) hipfftDestroy(s->inv_d);
  for (void* p : {(void*)s->real, (void*)s->hat, (void*)s->what, (void*)s->omega,
                  (void*)s->spec, (void*)s->partial, (void*)s->bins, (void*)s->modes,
                  (void*)s->tw, (void*)s->dft_partial, (void*)s->coef, (void*)s->dense,
                  (void*)s->finj})
    if (p) (void)hipFree(p);
  delete s;
}

const char* fft_text(hipfftResult r) {
  switch (r) {
    case HIPFFT_SUCCESS: return "success";
    case HIPFFT_ALLOC_FAILED: return "allocation failed";
    case HIPFFT_INVALID_VALUE: return "invalid value";
    case HIPFFT_INVALID_SIZE: return "invalid size";
    case HIPFFT_EXEC_FAILED: return "execution failed";
    case HIPFFT_SETUP_FAILED: return "setup failed";
    default: return "hipFFT error";
  }
}

int fft_fail(lbm_handle* h, hipfftResult r, const char* what) {
  return lbm::record_error(h, r == HIPFFT_ALLOC_FAILED ? LBM_E_NOMEM : LBM_E_HIP,
                           std::string(what) + ": " + fft_text(r));
}

// Wavenumber of index i along an axis of n points (FourierDomain.h /
// AnalysisList.h:148-156: i <= n/2 -> i, else i - n).
__host__ __device__ inline int wavenumber(int i, int n) { return i <= n / 2 ? i : i - n; }

// Index decomposition of half-spectrum mode m, last axis halved.
struct ModeIndex {
  int i[3];
};
// 32-bit unsigned arithmetic: M <= N < 2^31 (64-bit division is emulated
// on the GPU and dominated the per-mode kernels).
__host__ __device__ inline ModeIndex mode_index(int64_t m64, int D, const int* n) {
  ModeIndex r{{0, 0, 0}};
  const unsigned m = (unsigned)m64;
  const unsigned nh = (unsigned)(n[D - 1] / 2 + 1);
  r.i[D - 1] = (int)(m % nh);
  unsigned rest = m / nh;
  if (D == 3) {
    r.i[1] = (int)(rest % (unsigned)n[1]);
    rest /= (unsigned)n[1];
  }
  r.i[0] = (int)rest;
  return r;
}

struct Dims {
  int D, n[3];
  int64_t M;
};

// Derivative factor along an axis: integer k (REFERENCE) or 2 pi k / n
// (PHYSICAL); at the Nyquist index of an even axis zero in PHYSICAL and on
// the self-paired planes, +n/2 otherwise.
__device__ inline double dfactor(int i, int n, bool physical, bool self_paired) {
  if ((n & 1) == 0 && i == n / 2 && (physical || self_paired)) return 0.0;
  const int k = wavenumber(i, n);
  return physical ? (2.0 * M_PI / n) * k : double(k);
}

// i a for a = (re, im): (-im, re)
__device__ inline double2 times_i(double2 a) { return make_double2(-a.y, a.x); }

__global__ void k_curl_hat(const double2* __restrict__ uh, double2* __restrict__ wh, Dims g,
                           int physical, double scale) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= g.M) return;
  const ModeIndex ix = mode_index(m, g.D, g.n);
  const int nl = g.n[g.D - 1], l = ix.i[g.D - 1];
  const bool self_paired = l == 0 || ((nl & 1) == 0 && l == nl / 2);
  double k[3] = {0.0, 0.0, 0.0};
  for (int d = 0; d < g.D; ++d) k[d] = dfactor(ix.i[d], g.n[d], physical != 0, self_paired);
  if (g.D == 2) {
    const double2 ux = uh[m], uy = uh[g.M + m];
    // omega_z = i (kx uy - ky ux)
    double2 w = times_i(make_double2(k[0] * uy.x - k[1] * ux.x, k[0] * uy.y - k[1] * ux.y));
    if (!physical) w.y = 0.0;  // Transformer.h:161-165
    wh[m] = make_double2(w.x * scale, w.y * scale);
  } else {
    const double2 ux = uh[m], uy = uh[g.M + m], uz = uh[2 * g.M + m];
    const double2 wx = times_i(make_double2(k[1] * uz.x - k[2] * uy.x, k[1] * uz.y - k[2] * uy.y));
    const double2 wy = times_i(make_double2(k[2] * ux.x - k[0] * uz.x, k[2] * ux.y - k[0] * uz.y));
    const double2 wz = times_i(make_double2(k[0] * uy.x - k[1] * ux.x, k[0] * uy.y - k[1] * ux.y));
    wh[m] = make_double2(wx.x * scale, wx.y * scale);
    wh[g.M + m] = make_double2(wy.x * scale, wy.y * scale);
    wh[2 * g.M + m] = make_double2(wz.x * scale, wz.y * scale);
  }
}

// Fixed-order block sum (deterministic for a fixed launch shape).
__device__ inline double block_sum(double v, double* lds) {
  lds[threadIdx.x] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) lds[threadIdx.x] += lds[threadIdx.x + s];
    __syncthreads();
  }
  return lds[0];
}

// partial[b] = sum over the cells of block b's threads (grid-stride) of
// 1/2 sum_c w_c^2, one component after another.  Eight grid-stride loads are
// issued before they are summed (in order), so each lane keeps eight reads in
// flight instead of one.
__global__ void k_half_square_partial(const double* __restrict__ w, int C, int64_t N,
                                      double* __restrict__ partial) {
  __shared__ double lds[kBlock];
  constexpr int kU = 8;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  double acc = 0.0;
  for (int c = 0; c < C; ++c) {
    const double* __restrict__ wc = w + c * N;
    int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; x + (kU - 1) * stride < N; x += kU * stride) {
      double v[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u) v[u] = wc[x + u * stride];
#pragma unroll
      for (int u = 0; u < kU; ++u) acc += 0.5 * v[u] * v[u];
    }
    for (; x < N; x += stride) {
      const double v = wc[x];
      acc += 0.5 * v * v;
    }
  }
  const double t = block_sum(acc, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

__global__ void k_final_sum(const double* __restrict__ partial, int n, double scale,
                            double* __restrict__ out) {
  __shared__ double lds[kBlock];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) acc += partial[i];
  const double t = block_sum(acc, lds);
  if (threadIdx.x == 0) *out = t * scale;
}

// Shell of an integer |k|^2, exactly: REFERENCE truncates
// (unsigned(int(sqrt(int))), MathVector.h:65, AnalysisList.h:157),
// PHYSICAL rounds to the nearest integer (|k|^2 is an integer, so |k| is
// never a half-integer: s + 1 when k2 - s^2 > s).
__device__ inline int shell_of(long long k2, bool physical) {
  long long s = (long long)sqrt((double)k2);
  while (s * s > k2) --s;
  while ((s + 1) * (s + 1) <= k2) ++s;
  if (physical && k2 - s * s > s) ++s;
  return (int)s;
}

constexpr int kShellWindow = 2048;  // shells binned per pass (LDS: 4 waves x 16 KB)
constexpr int kShellBlocks = 4096;  // 1024 left the binning latency-bound (171 us at 256^3)
constexpr int kShellBatch = 4;     // grid-stride iterations whose loads are issued together

// Shell sums of c(m) sum_d |hat_d(m)|^2 over the shells [s_lo, s_lo + cnt),
// c = 1/2 on the k_last = 0 plane and, when physical, on the k_last = n/2
// plane of an even last axis.  Modes are read in storage order (coalesced);
// each wavefront sums its runs of equal shells by a segmented scan and adds
// the run sums into its own LDS bins one lane at a time in lane order, so the
// sums do not depend on timing; the block then
// writes the sum of its wavefronts' bins (in wavefront order) to
// partial[block][cnt].
__global__ void __launch_bounds__(kBlock) k_shell_bins(const double2* __restrict__ hat, Dims g,
                                                       int physical, int s_lo, int cnt,
                                                       double* __restrict__ partial) {
  extern __shared__ double bins[];  // [kBlock / 64][cnt]
  constexpr int kWaves = kBlock / 64;
  for (int i = threadIdx.x; i < kWaves * cnt; i += blockDim.x) bins[i] = 0.0;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  double* mine = bins + (threadIdx.x >> 6) * cnt;
  const int nlast = g.D == 3 ? g.n[2] : g.n[1];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  // kShellBatch grid-stride iterations at a time: every lane issues the loads
  // of all of them before the first scan, so the scans and LDS adds of one
  // iteration overlap the reads of the next.  Iterations are still binned in
  // grid-stride order, so the sums are those of a one-at-a-time loop.
  for (int64_t base0 = (int64_t)blockIdx.x * blockDim.x; base0 < g.M;
       base0 += kShellBatch * stride) {
    int keys[kShellBatch];
    double wts[kShellBatch];
    int rows[kShellBatch];  // row of the last axis (mode index / (n_last/2 + 1))
    double2 hv[kShellBatch][3];
    // loads first (from an in-range mode for lanes past the end or outside
    // the window, whose values are then discarded), no use of them here
#pragma unroll
    for (int u = 0; u < kShellBatch; ++u) {
      const int64_t m = base0 + u * stride + threadIdx.x;
      keys[u] = -1;
      wts[u] = 0.0;
      rows[u] = -1;
      const int64_t ml = m < g.M ? m : 0;
      if (m < g.M) {
        // mode_index and |k|^2 spelled out per dimension: indexing g.n and
        // the index triple by a runtime axis put them in scratch memory
        const unsigned mu = (unsigned)m, nh = (unsigned)(nlast / 2 + 1);
        const int l = (int)(mu % nh);
        const unsigned rest = mu / nh;
        rows[u] = (int)rest;
        long long k2;
        if (g.D == 3) {
          const long long k0 = wavenumber((int)(rest / (unsigned)g.n[1]), g.n[0]);
          const long long k1 = wavenumber((int)(rest % (unsigned)g.n[1]), g.n[1]);
          const long long k2l = wavenumber(l, nlast);
          k2 = k0 * k0 + k1 * k1 + k2l * k2l;
        } else {
          const long long k0 = wavenumber((int)rest, g.n[0]);
          const long long k1 = wavenumber(l, nlast);
          k2 = k0 * k0 + k1 * k1;
        }
        const int b = shell_of(k2, physical != 0) - s_lo;
        if (b >= 0 && b < cnt) {
          keys[u] = b;
          wts[u] = (l == 0 || (physical && (nlast & 1) == 0 && l == nlast / 2)) ? 0.5 : 1.0;
        }
      }
      hv[u][0] = hat[ml];
      hv[u][1] = hat[g.M + ml];
      hv[u][2] = g.D == 3 ? hat[2 * g.M + ml] : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int u = 0; u < kShellBatch; ++u) {
      if (base0 + u * stride >= g.M) break;  // block-uniform
      const int key = keys[u];
      double e = 0.0;
      e += hv[u][0].x * hv[u][0].x;
      e += hv[u][0].y * hv[u][0].y;
      e += hv[u][1].x * hv[u][1].x;
      e += hv[u][1].y * hv[u][1].y;
      if (g.D == 3) {
        e += hv[u][2].x * hv[u][2].x;
        e += hv[u][2].y * hv[u][2].y;
      }
      double val = key >= 0 ? wts[u] * e : 0.0;
      // Consecutive modes along the last axis fall in nondecreasing shells,
      // so the wavefront's keys form runs: a segmented scan (fixed shuffle
      // order) sums each run into its last lane, and only those lanes add to
      // the bins.
      const int prev_key = __shfl_up(key, 1);
      const bool head = lane == 0 || prev_key != key;
      // the run's first lane: the highest head at or below this lane (lane 0
      // is always a head; 2 << 63 wraps to 0, so lane 63 keeps every bit)
      const unsigned long long heads = __ballot(head) & ((2ull << lane) - 1);
      const int start = 63 - __clzll((long long)heads);
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const double o = __shfl_up(val, off);
        if (lane - off >= start) val += o;
      }
      const int next_key = __shfl_down(key, 1);
      const bool tail = (lane == 63 || next_key != key) && key >= 0;
      // Within one row of the last axis the keys are nondecreasing, so its
      // run tails hold distinct shells: each row's tails add in one LDS
      // instruction, rows in lane order -- every bin receives its adds in the
      // same order as one lane at a time.
      unsigned long long todo = __ballot(tail);
      while (todo) {
        const int t = __ffsll((long long)todo) - 1;
        const int row_t = __builtin_amdgcn_readlane(rows[u], t);
        const bool mine_now = tail && rows[u] == row_t;
        todo &= ~__ballot(mine_now);
        if (mine_now) atomicAdd(&mine[key], val);
      }
    }
  }
  __syncthreads();
  for (int b = threadIdx.x; b < cnt; b += blockDim.x) {
    double acc = 0.0;
    for (int w = 0; w < kWaves; ++w) acc += bins[w * cnt + b];
    partial[(int64_t)blockIdx.x * cnt + b] = acc;
  }
}

// One block per shell b: out[b] = scale * sum over blocks of
// partial[block][b], in a fixed order (strided per thread, then a tree).
__global__ void k_shell_final(const double* __restrict__ partial, int blocks, int cnt,
                              double scale, double* __restrict__ out) {
  __shared__ double lds[kBlock];
  const int b = blockIdx.x;
  double acc = 0.0;
  for (int k = threadIdx.x; k < blocks; k += blockDim.x) acc += partial[(int64_t)k * cnt + b];
  const double t = block_sum(acc, lds);
  if (threadIdx.x == 0) out[b] = t * scale;
}

// out[d][x] = rho[x] u[d][x]: the momentum EnergyRemoval transforms
// (Force.h:416-431).
__global__ void k_momentum(const double* __restrict__ rho, const double* __restrict__ u, int D,
                           int64_t N, double* __restrict__ out) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= N) return;
  const double r = rho[x];
  for (int d = 0; d < D; ++d) out[d * N + x] = r * u[d * N + x];
}

struct ForceSpec {
  int injection, kmin2, kmax2;  // ConstantShell shell (integer |k|^2 bounds)
  double A[3];                  // its psi_hat amplitude per component
  int removal, rmin2, rmax2;    // EnergyRemoval shell
  double R[3];                  // its amplitude per component
  double scale;                 // the backward transform's 1/N
};

// Fourier-space force, in place on hat (read as the momentum spectrum when
// removal is on): EnergyRemoval -R_d p_hat_d on its shell (Force.h:436-497),
// plus ConstantShell's MakeIncompressible of psi_hat = A on its shell
// (Force.h:327-393; Transformer.h:320-334 in 2D: F = (i ky psi, -i kx psi);
// in 3D the Curl of Transformer.h:216-275: F = i k x psi), times 1/N.
__global__ void k_force_hat(double2* __restrict__ hat, Dims g, ForceSpec f) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= g.M) return;
  const ModeIndex ix = mode_index(m, g.D, g.n);
  long long k2 = 0;
  for (int d = 0; d < g.D; ++d) {
    const long long k = wavenumber(ix.i[d], g.n[d]);
    k2 += k * k;
  }
  double2 F[3] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
  if (f.removal && k2 >= f.rmin2 && k2 <= f.rmax2)
    for (int d = 0; d < g.D; ++d) {
      const double2 p = hat[(int64_t)d * g.M + m];
      F[d] = make_double2(-f.R[d] * p.x, -f.R[d] * p.y);
    }
  if (f.injection && k2 >= f.kmin2 && k2 <= f.kmax2) {
    const int nl = g.n[g.D - 1], l = ix.i[g.D - 1];
    const bool self_paired = l == 0 || ((nl & 1) == 0 && l == nl / 2);
    double k[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < g.D; ++d) k[d] = dfactor(ix.i[d], g.n[d], false, self_paired);
    double c[3];  // psi_hat is real, so i (k x psi) is imaginary: F += i c
    if (g.D == 2) {
      c[0] = k[1] * f.A[0];
      c[1] = -k[0] * f.A[0];
      c[2] = 0.0;
    } else {
      c[0] = k[1] * f.A[2] - k[2] * f.A[1];
      c[1] = k[2] * f.A[0] - k[0] * f.A[2];
      c[2] = k[0] * f.A[1] - k[1] * f.A[0];
    }
    for (int d = 0; d < g.D; ++d) F[d].y += c[d];
  }
  for (int d = 0; d < g.D; ++d)
    hat[(int64_t)d * g.M + m] = make_double2(F[d].x * f.scale, F[d].y * f.scale);
}

__global__ void k_to_f32(const double* __restrict__ src, int64_t n, float* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (float)src[i];
}

// -- direct transform of a few shell modes ---------------------------------
// When the removal shell holds few modes (8 for shells 1..2 in 2D), its
// Fourier coefficients are summed directly over the cells and the removal
// field synthesised from them: two passes over the fields instead of a
// forward and an inverse transform of the whole grid.
// modes per pass of the reduction (accumulators live in registers; 16 per
// pass, 162 VGPRs, measured 41% slower than 8, 132 VGPRs, for D2Q9 2048^2)
template <int D>
constexpr int modes_per_pass() { return 8; }
constexpr int kMaxDirectModes = 32;
constexpr int kDftBlocks = 1024;  // 4096 measured 33% slower at D2Q9 2048^2

__device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

constexpr int kMaxPow = 6;  // largest |k_d| of a directly transformed mode

// e^{-i k theta} = e1^|k| (conjugated for k < 0) by repeated products; k is
// uniform across the wavefront (one mode at a time), so is the loop; the
// relative error grows like |k| eps for the few small |k| of the shell.
__device__ inline double2 power_of(double2 e1, int k) {
  const int a = k < 0 ? -k : k;
  double2 r = make_double2(1.0, 0.0);
  for (int t = 0; t < a; ++t) r = cmul(r, e1);
  return k < 0 ? make_double2(r.x, -r.y) : r;
}

// First and second harmonics e^{-2 pi i x_d / n_d}, e^{-4 pi i x_d / n_d} of
// a cell from tw = [nx | ny | nz], shared by all the modes.
struct CellPhases {
  double2 e[3], e2[3];
};
__device__ inline void cell_phases(const double2* __restrict__ tw, const Dims& g, const int* i,
                                   CellPhases& c) {
  c.e[0] = tw[i[0]];
  c.e[1] = tw[g.n[0] + i[1]];
  c.e[2] = g.D == 3 ? tw[g.n[0] + g.n[1] + i[2]] : make_double2(1.0, 0.0);
#pragma unroll
  for (int d = 0; d < 3; ++d) c.e2[d] = cmul(c.e[d], c.e[d]);
}

// e^{-i k theta} along one axis: the shared harmonics for |k| <= 2 (uniform
// branches: one mode at a time), repeated products beyond.
__device__ inline double2 axis_phase(double2 e1, double2 e2, int k) {
  const int a = k < 0 ? -k : k;
  const double2 r = a == 0 ? make_double2(1.0, 0.0)
                           : (a == 1 ? e1 : (a == 2 ? e2 : power_of(e1, a)));
  return k < 0 ? make_double2(r.x, -r.y) : r;
}

// The same for |k| <= 2 by selects on the uniform k, no branches.
__device__ inline double2 axis_phase_small(double2 e1, double2 e2, int k) {
  const int a = k < 0 ? -k : k;
  double2 r = a == 1 ? e1 : e2;
  r = a == 0 ? make_double2(1.0, 0.0) : r;
  return make_double2(r.x, k < 0 ? -r.y : r.y);
}

// e^{-i k.x} of mode m (signed wavenumbers) at the cell; SMALL when every
// |k_d| of the modes is at most 2.
template <bool SMALL>
__device__ inline double2 phase_of(const CellPhases& c, const Dims& g, int4 m) {
  auto ax = [](double2 e1, double2 e2, int k) {
    if constexpr (SMALL)
      return axis_phase_small(e1, e2, k);
    else
      return axis_phase(e1, e2, k);
  };
  double2 p = cmul(ax(c.e[0], c.e2[0], m.x), ax(c.e[1], c.e2[1], m.y));
  if (g.D == 3) p = cmul(p, ax(c.e[2], c.e2[2], m.z));
  return p;
}

__device__ inline void cell_coords(int64_t x64, const Dims& g, int* i) {
  const unsigned x = (unsigned)x64;  // N < 2^31
  if (g.D == 3) {
    i[2] = (int)(x % (unsigned)g.n[2]);
    const unsigned t = x / (unsigned)g.n[2];
    i[1] = (int)(t % (unsigned)g.n[1]);
    i[0] = (int)(t / (unsigned)g.n[1]);
  } else {
    i[2] = 0;
    i[1] = (int)(x % (unsigned)g.n[1]);
    i[0] = (int)(x / (unsigned)g.n[1]);
  }
}

// Wavefront sums of N values per lane (N a power of two, 2 <= N <= 64) by
// recursive halving: at the step with lane offset o a lane keeps the half of
// its values selected by its lane bit o, adding its partner's copy of that
// half, and sends the other half; after log2 N steps each lane holds one
// value summed over the lanes that differ from it in those bits, and xor
// steps over the remaining low bits finish the sums.  Lane l returns the
// wavefront sum of value l >> (6 - log2 N): N - 1 + 6 - log2 N shuffles
// instead of 6 N.  Fixed order, so the sums are deterministic.
template <int N>
__device__ inline double wave_sums_halving(double (&v)[N], int lane) {
  static_assert(N >= 2 && N <= 64 && (N & (N - 1)) == 0, "N: power of two <= 64");
  int o = 32;
#pragma unroll
  for (int n = N; n > 1; n >>= 1, o >>= 1) {
    const bool up = (lane & o) != 0;
#pragma unroll
    for (int s = 0; s < n / 2; ++s) {
      const double lo = v[s], hi = v[s + n / 2];
      const double send = up ? lo : hi;
      const double keep = up ? hi : lo;
      v[s] = keep + __shfl_xor(send, o);
    }
  }
  double r = v[0];
#pragma unroll
  for (int oo = o; oo > 0; oo >>= 1) r += __shfl_xor(r, oo);
  return r;
}

// The block's sums of acc[j][d] (complex) for j < nm, written to
// partial[block][j0 + j][d]: wavefront sums by recursive halving over the
// MP * D * 2 reals (zero-padded to a power of two), then the wavefronts in
// order.
template <int MP, int D>
__device__ inline void block_partials(const double2 (&acc)[MP][D], int nm, int nmodes, int j0,
                                      double2* __restrict__ partial) {
  constexpr int kVals = MP * D * 2;
  constexpr int kN = kVals <= 32 ? 32 : 64;
  static_assert(kVals <= 64, "at most 64 reals per lane");
  constexpr int kShift = kN == 32 ? 1 : 0;  // lane l holds value l >> kShift
  double v[kN];
#pragma unroll
  for (int j = 0; j < MP; ++j)
#pragma unroll
    for (int d = 0; d < D; ++d) {
      v[(j * D + d) * 2] = acc[j][d].x;
      v[(j * D + d) * 2 + 1] = acc[j][d].y;
    }
#pragma unroll
  for (int i = kVals; i < kN; ++i) v[i] = 0.0;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const double r = wave_sums_halving<kN>(v, lane);
  __shared__ double red[kBlock / 64][kN];
  if ((lane & ((1 << kShift) - 1)) == 0) red[wave][lane >> kShift] = r;
  __syncthreads();
  const int t = threadIdx.x;
  if (t < nm * D) {
    const int j = t / D, d = t - j * D;
    double2 sum = make_double2(red[0][(j * D + d) * 2], red[0][(j * D + d) * 2 + 1]);
#pragma unroll
    for (int w = 1; w < kBlock / 64; ++w) {
      sum.x += red[w][(j * D + d) * 2];
      sum.y += red[w][(j * D + d) * 2 + 1];
    }
    partial[((int64_t)blockIdx.x * nmodes + j0 + j) * D + d] = sum;
  }
}

// partial[block][j][d] = sum over the block's cells of rho u_d e^{-i k_j.x},
// modes j of this pass (blockIdx.y); fixed order per launch shape.  All
// accumulator indices are compile-time constants (registers, not scratch).
// The mode list is padded to whole passes (k = 0, weight 0), so every pass
// computes MP modes without a per-mode guard and writes the real ones.
template <int D, bool SMALL, int MP = modes_per_pass<D>()>
__global__ void __launch_bounds__(kBlock) k_mode_dft_partial(
    const double* __restrict__ rho, const double* __restrict__ u, Dims g, int64_t N,
    const int4* __restrict__ modes, int nmodes, const double2* __restrict__ tw,
    double2* __restrict__ partial) {
  const int j0 = blockIdx.y * MP;
  const int nm = min(MP, nmodes - j0);
  double2 acc[MP][D];
#pragma unroll
  for (int j = 0; j < MP; ++j)
#pragma unroll
    for (int d = 0; d < D; ++d) acc[j][d] = make_double2(0.0, 0.0);
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < N;
       x += (int64_t)gridDim.x * blockDim.x) {
    int i[3];
    cell_coords(x, g, i);
    CellPhases cp;
    cell_phases(tw, g, i, cp);
    const double r = rho[x];
    double p[D];
#pragma unroll
    for (int d = 0; d < D; ++d) p[d] = r * u[d * N + x];
#pragma unroll
    for (int j = 0; j < MP; ++j) {
      if (j < nm) {
        const double2 e = phase_of<SMALL>(cp, g, modes[j0 + j]);
#pragma unroll
        for (int d = 0; d < D; ++d) {
          acc[j][d].x += p[d] * e.x;
          acc[j][d].y += p[d] * e.y;
        }
      }
    }
  }
  // wavefront sums by recursive halving (no barrier), then the block's
  // wavefronts in order: one barrier instead of one per accumulator
  block_partials<MP, D>(acc, nm, nmodes, j0, partial);
}

// 2D, every |k_d| <= 2: the same partial sums, accumulated per column.
// Thread t < n1 * R walks column c = t % n1 over the rows t / n1, + R, ...;
// on a column e^{-i ky theta1} is fixed, so the thread sums
// B_kx,d = sum p_d e^{-i kx theta0} for kx = 0, 1, 2 only (p_d = rho u_d is
// real, so B_-kx = conj(B_kx)) and multiplies by e^{-i ky theta1} per mode
// at the end: 14 FP64 operations per cell instead of a phase product and
// four FMAs per mode.  Consecutive threads read consecutive cells.
template <int MP = modes_per_pass<2>()>
__global__ void __launch_bounds__(kBlock) k_mode_dft_partial_2d_small(
    const double* __restrict__ rho, const double* __restrict__ u, Dims g, int64_t N, int R,
    const int4* __restrict__ modes, int nmodes, const double2* __restrict__ tw,
    double2* __restrict__ partial) {
  const int j0 = blockIdx.y * MP;
  const int nm = min(MP, nmodes - j0);
  const int n1 = g.n[1];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  double b0[2] = {0.0, 0.0};
  double2 b1[2] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
  double2 b2[2] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
  int c = 0;
  if (t < n1 * R) {
    c = t % n1;
    // four rows' loads issued together (rows still summed in order; rows
    // past the end reload row r0 and are not summed)
    constexpr int kRows = 4;
    for (int r0 = t / n1; r0 < g.n[0]; r0 += kRows * R) {
      double2 X1[kRows];
      double rh[kRows], u0[kRows], u1[kRows];
#pragma unroll
      for (int q = 0; q < kRows; ++q) {
        const int r = r0 + q * R < g.n[0] ? r0 + q * R : r0;
        const int64_t x = (int64_t)r * n1 + c;
        X1[q] = tw[r];
        rh[q] = rho[x];
        u0[q] = u[x];
        u1[q] = u[N + x];
      }
#pragma unroll
      for (int q = 0; q < kRows; ++q) {
        if (r0 + q * R >= g.n[0]) break;
        const double2 X2 = cmul(X1[q], X1[q]);
        const double p[2] = {rh[q] * u0[q], rh[q] * u1[q]};
#pragma unroll
        for (int d = 0; d < 2; ++d) {
          b0[d] += p[d];
          b1[d].x += p[d] * X1[q].x;
          b1[d].y += p[d] * X1[q].y;
          b2[d].x += p[d] * X2.x;
          b2[d].y += p[d] * X2.y;
        }
      }
    }
  }
  const double2 Y1 = tw[g.n[0] + c];
  const double2 Y2 = cmul(Y1, Y1);
  double2 acc[MP][2];
#pragma unroll
  for (int j = 0; j < MP; ++j) {
    const int4 m = modes[j0 + j];  // padded: in range for every j
    const int a = m.x < 0 ? -m.x : m.x;
    const double2 Y = m.y == 0 ? make_double2(1.0, 0.0) : (m.y == 1 ? Y1 : Y2);
#pragma unroll
    for (int d = 0; d < 2; ++d) {
      double2 B = a == 0 ? make_double2(b0[d], 0.0) : (a == 1 ? b1[d] : b2[d]);
      if (m.x < 0) B.y = -B.y;
      acc[j][d] = t < n1 * R ? cmul(B, Y) : make_double2(0.0, 0.0);
    }
  }
  block_partials<MP, 2>(acc, nm, nmodes, j0, partial);
}

// coef[j][d] = -R_d w_j p_hat_j,d / N: one block per (j, d), fixed order.
// With dense non-null (2D, every |k_d| <= 2) the coefficient is also stored
// at dense[((kx + 2) * 3 + ky) * 2 + d], ky >= 0 being the halved axis.
__global__ void k_mode_dft_final(const double2* __restrict__ partial, int blocks, int nmodes,
                                 int D, const int4* __restrict__ modes, double R0, double R1,
                                 double R2, double invN, double2* __restrict__ coef,
                                 double2* __restrict__ dense) {
  __shared__ double lds[kBlock];
  const int t = blockIdx.x, j = t / D, d = t % D;
  double re = 0.0, im = 0.0;
  for (int b = threadIdx.x; b < blocks; b += blockDim.x) {
    const double2 v = partial[((int64_t)b * nmodes + j) * D + d];
    re += v.x;
    im += v.y;
  }
  re = block_sum(re, lds);
  __syncthreads();
  im = block_sum(im, lds);
  if (threadIdx.x == 0) {
    const double R = d == 0 ? R0 : (d == 1 ? R1 : R2);
    const double w = -R * (double)modes[j].w * invN;
    coef[t] = make_double2(w * re, w * im);
    if (dense) dense[((modes[j].x + 2) * 3 + modes[j].y) * 2 + d] = coef[t];
  }
}

// F_d(x) = finj_d(x) + sum_j Re(coef_j,d e^{+i k_j.x}); the weight in coef
// (1 on the self-paired planes, 2 elsewhere) gives the complex-to-real
// inverse of the half spectrum.
// The mode list and coef are padded to whole groups of modes_per_pass (zero
// coefficients), so each group of modes is synthesised without a guard.
template <class OUT, bool SMALL>
__global__ void k_force_synth(const double* __restrict__ finj, const double2* __restrict__ coef,
                              const int4* __restrict__ modes, int nmodes,
                              const double2* __restrict__ tw, Dims g, int64_t N,
                              OUT* __restrict__ F) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= N) return;
  int i[3];
  cell_coords(x, g, i);
  CellPhases cp;
  cell_phases(tw, g, i, cp);
  double f[3] = {0.0, 0.0, 0.0};
  if (finj)
    for (int d = 0; d < g.D; ++d) f[d] = finj[d * N + x];
  constexpr int G = modes_per_pass<3>();
  for (int j0 = 0; j0 < nmodes; j0 += G) {
#pragma unroll
    for (int j = j0; j < j0 + G; ++j) {
      const double2 e = phase_of<SMALL>(cp, g, modes[j]);  // e^{-i k.x}
#pragma unroll
      for (int d = 0; d < 3; ++d)
        if (d < g.D) {
          const double2 c = coef[j * g.D + d];
          f[d] += c.x * e.x + c.y * e.y;  // Re(c e^{+i k.x})
        }
    }
  }
  for (int d = 0; d < g.D; ++d) F[d * N + x] = (OUT)f[d];
}

// 2D, every |k_d| <= 2: F_d = finj_d + Re sum_kx e^{+i kx theta0}
// sum_ky C[kx][ky][d] e^{+i ky theta1} over the dense grid (zero where the
// shell has no mode): compile-time harmonics, no per-mode selection.  The
// kx and -kx terms are paired: with E = e^{+i k theta0}, G_k = sum_ky C Y,
// Re(E G_k + conj(E) G_-k) = E.x Re(G_k + G_-k) - E.y Im(G_k - G_-k), and
// G_k +- G_-k = sum_ky (C_k +- C_-k) Y, so each pair costs two real sums.
template <class OUT>
__global__ void k_force_synth_2d_small(const double* __restrict__ finj,
                                       const double2* __restrict__ dense,
                                       const double2* __restrict__ tw, Dims g, int64_t N,
                                       OUT* __restrict__ F) {
  // two cells per lane, a block width apart: the pairing of the uniform
  // coefficients below is shared by both
  constexpr int kC = 2;
  const int64_t x0 = (int64_t)blockIdx.x * (kC * blockDim.x) + threadIdx.x;
  if (x0 >= N) return;
  bool live[kC];
  double2 X1[kC], X2[kC], Y1[kC], Y2[kC];
  double f[kC][2];
#pragma unroll
  for (int q = 0; q < kC; ++q) {
    const int64_t xq = x0 + q * blockDim.x;
    live[q] = xq < N;
    const int64_t x = live[q] ? xq : x0;
    int i[3];
    cell_coords(x, g, i);
    // tw holds e^{-i theta}; the synthesis needs e^{+i k theta}
    const double2 a1 = tw[i[0]], b1 = tw[g.n[0] + i[1]];
    X1[q] = make_double2(a1.x, -a1.y);
    Y1[q] = make_double2(b1.x, -b1.y);
    X2[q] = cmul(X1[q], X1[q]);
    Y2[q] = cmul(Y1[q], Y1[q]);
    f[q][0] = finj ? finj[x] : 0.0;
    f[q][1] = finj ? finj[N + x] : 0.0;
  }
  auto C = [&](int kx, int ky, int d) { return dense[((kx + 2) * 3 + ky) * 2 + d]; };
#pragma unroll
  for (int d = 0; d < 2; ++d) {
    // kx = 0: Re G_0
    const double2 c00 = C(0, 0, d), c01 = C(0, 1, d), c02 = C(0, 2, d);
    double acc[kC];
#pragma unroll
    for (int q = 0; q < kC; ++q) {
      acc[q] = c00.x;
      acc[q] += c01.x * Y1[q].x - c01.y * Y1[q].y;
      acc[q] += c02.x * Y2[q].x - c02.y * Y2[q].y;
    }
#pragma unroll
    for (int k = 1; k <= 2; ++k) {
      double re[kC], im[kC];  // Re(G_k + G_-k), Im(G_k - G_-k)
#pragma unroll
      for (int q = 0; q < kC; ++q) re[q] = im[q] = 0.0;
#pragma unroll
      for (int ky = 0; ky < 3; ++ky) {
        const double2 cp = C(k, ky, d), cm = C(-k, ky, d);
        const double px = cp.x + cm.x, py = cp.y + cm.y;  // C_k + C_-k
        const double mx = cp.x - cm.x, my = cp.y - cm.y;  // C_k - C_-k
#pragma unroll
        for (int q = 0; q < kC; ++q) {
          if (ky == 0) {
            re[q] += px;
            im[q] += my;
          } else {
            const double2 Y = ky == 1 ? Y1[q] : Y2[q];
            re[q] += px * Y.x - py * Y.y;
            im[q] += mx * Y.y + my * Y.x;
          }
        }
      }
#pragma unroll
      for (int q = 0; q < kC; ++q) {
        const double2 E = k == 1 ? X1[q] : X2[q];
        acc[q] += E.x * re[q] - E.y * im[q];
      }
    }
#pragma unroll
    for (int q = 0; q < kC; ++q) f[q][d] += acc[q];
  }
#pragma unroll
  for (int q = 0; q < kC; ++q)
    if (live[q]) {
      const int64_t x = x0 + q * blockDim.x;
      F[x] = (OUT)f[q][0];
      F[N + x] = (OUT)f[q][1];
    }
}

unsigned grid_of(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

Dims dims_of(const State& s) {
  Dims g;
  g.D = s.D;
  for (int d = 0; d < 3; ++d) g.n[d] = s.n[d];
  g.M = s.M;
  return g;
}

// Creates (once) the state of a handle: buffers and the forward plan.
int state_of(lbm_handle* h, const lbm::DeviceFields& f, State*& out) {
  void*& slot = lbm::spectral_slot(h);
  if (slot) {
    out = static_cast<State*>(slot);
    return LBM_OK;
  }
  State* s = new State;
  s->D = f.D;
  s->C = f.D == 3 ? 3 : 1;  // 2 D - 3 components
  for (int d = 0; d < 3; ++d) s->n[d] = f.n[d];
  s->N = f.N;
  s->M = f.N / s->n[f.D - 1] * (s->n[f.D - 1] / 2 + 1);
  int nmax = 1;
  for (int d = 0; d < f.D; ++d) nmax = std::max(nmax, s->n[d]);
  s->nshells = nmax / 2;  // gFD::maxWaveNumber (FourierDomain.h:76-78)
  auto alloc = [&](void** p, size_t bytes) {
    const hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return lbm::record_error(h, LBM_E_NOMEM, "spectral buffers: out of memory");
    return (int)LBM_OK;
  };
  int rc = alloc((void**)&s->real, sizeof(double) * s->D * s->N);
  if (!rc) rc = alloc((void**)&s->hat, sizeof(double2) * s->D * s->M);
  if (!rc) rc = alloc((void**)&s->what, sizeof(double2) * s->C * s->M);
  if (!rc) rc = alloc((void**)&s->omega, sizeof(double) * s->C * s->N);
  if (!rc) rc = alloc((void**)&s->spec, sizeof(double) * 2 * std::max(1, s->nshells));
  if (!rc) rc = alloc((void**)&s->partial, sizeof(double) * (kReduceBlocks + 1));
  s->shell_blocks = (int)std::min<int64_t>(kShellBlocks, (s->M + kBlock - 1) / kBlock);
  if (!rc)
    rc = alloc((void**)&s->bins, sizeof(double) * s->shell_blocks *
                                     std::max(1, std::min(s->nshells, kShellWindow)));
  if (!rc) {
    int dims[3] = {s->n[0], s->n[1], s->n[2]};
    hipfftResult r = hipfftPlanMany(&s->fwd, s->D, dims, nullptr, 1, (int)s->N, nullptr, 1,
                                    (int)s->M, HIPFFT_D2Z, s->D);
    if (r != HIPFFT_SUCCESS) rc = fft_fail(h, r, "forward transform plan");
    else s->have_fwd = true;
    if (!rc) {
      r = hipfftPlanMany(&s->inv, s->D, dims, nullptr, 1, (int)s->M, nullptr, 1, (int)s->N,
                         HIPFFT_Z2D, s->C);
      if (r != HIPFFT_SUCCESS) rc = fft_fail(h, r, "inverse transform plan");
      else s->have_inv = true;
    }
  }
  if (rc) {
    free_state(s);
    return rc;
  }
  slot = s;
  out = s;
  return LBM_OK;
}

// hat = forward transform of the D components at src (copied first: the
// transform may use its input as scratch).
int forward(lbm_handle* h, State& s, const double* src, uint64_t version, hipStream_t st) {
  // vorticity and energy spectrum of the same stored fields share one transform
  if (s.hat_src == src && s.hat_version == version) return LBM_OK;
  s.hat_src = nullptr;
  int rc = lbm::record_hip(h, hipMemcpyAsync(s.real, src, sizeof(double) * s.D * s.N,
                                             hipMemcpyDeviceToDevice, st), "field copy");
  if (rc) return rc;
  hipfftResult r = hipfftSetStream(s.fwd, st);
  if (r == HIPFFT_SUCCESS)
    r = hipfftExecD2Z(s.fwd, (hipfftDoubleReal*)s.real, (hipfftDoubleComplex*)s.hat);
  if (r != HIPFFT_SUCCESS) return fft_fail(h, r, "forward transform");
  s.hat_src = src;
  s.hat_version = version;
  return LBM_OK;
}

int check_scaling(lbm_handle* h, int scaling) {
  if (scaling != LBM_SPECTRAL_REFERENCE && scaling != LBM_SPECTRAL_PHYSICAL)
    return lbm::record_error(h, LBM_E_INVALID, "scaling must be LBM_SPECTRAL_REFERENCE or "
                                               "LBM_SPECTRAL_PHYSICAL");
  return LBM_OK;
}

}  // namespace

void lbm::release_spectral(void* state) { free_state(static_cast<State*>(state)); }

namespace {

// The force field by full transforms into out (FP64 [D][N]).
int fft_force(lbm_handle* h, State& s, const lbm::ForceField& ff, const ForceSpec& spec,
              double* out) {
  const hipStream_t st = ff.stream;
  s.hat_src = nullptr;  // hat becomes the force spectrum
  if (!s.have_inv_d) {
    int dims[3] = {s.n[0], s.n[1], s.n[2]};
    const hipfftResult r = hipfftPlanMany(&s.inv_d, s.D, dims, nullptr, 1, (int)s.M, nullptr, 1,
                                          (int)s.N, HIPFFT_Z2D, s.D);
    if (r != HIPFFT_SUCCESS) return fft_fail(h, r, "force transform plan");
    s.have_inv_d = true;
  }
  if (spec.removal) {
    hipLaunchKernelGGL(k_momentum, dim3(grid_of(s.N)), dim3(kBlock), 0, st, ff.rho, ff.u, s.D,
                       s.N, s.real);
    if (int rc = lbm::record_hip(h, hipGetLastError(), "momentum kernel")) return rc;
    hipfftResult r = hipfftSetStream(s.fwd, st);
    if (r == HIPFFT_SUCCESS)
      r = hipfftExecD2Z(s.fwd, (hipfftDoubleReal*)s.real, (hipfftDoubleComplex*)s.hat);
    if (r != HIPFFT_SUCCESS) return fft_fail(h, r, "momentum transform");
  }
  hipLaunchKernelGGL(k_force_hat, dim3(grid_of(s.M)), dim3(kBlock), 0, st, s.hat, dims_of(s),
                     spec);
  if (int rc = lbm::record_hip(h, hipGetLastError(), "force spectrum kernel")) return rc;
  hipfftResult r = hipfftSetStream(s.inv_d, st);
  if (r == HIPFFT_SUCCESS)
    r = hipfftExecZ2D(s.inv_d, (hipfftDoubleComplex*)s.hat, (hipfftDoubleReal*)out);
  return r == HIPFFT_SUCCESS ? LBM_OK : fft_fail(h, r, "force transform");
}

// Chooses the direct path when the removal shell has at most
// kMaxDirectModes half-spectrum modes; builds its mode list, twiddle tables
// and, for Turbulent2D, the constant injection field.
int direct_setup(lbm_handle* h, State& s, const lbm::ForceField& ff, const ForceSpec& spec) {
  s.direct_checked = true;
  if (!spec.removal) return LBM_OK;
  std::vector<int4> modes;
  const int nl = s.n[s.D - 1];
  for (int64_t m = 0; m < s.M && (int)modes.size() <= kMaxDirectModes; ++m) {
    const ModeIndex ix = mode_index(m, s.D, s.n);
    int k[3] = {0, 0, 0};
    long long k2 = 0;
    for (int d = 0; d < s.D; ++d) {
      k[d] = wavenumber(ix.i[d], s.n[d]);
      k2 += (long long)k[d] * k[d];
    }
    if (k2 < spec.rmin2 || k2 > spec.rmax2) continue;
    const int l = ix.i[s.D - 1];
    const int w = (l == 0 || ((nl & 1) == 0 && l == nl / 2)) ? 1 : 2;
    modes.push_back(make_int4(k[0], k[1], k[2], w));
    for (int d = 0; d < s.D; ++d)
      if (std::abs(k[d]) > kMaxPow) return LBM_OK;  // beyond the phase powers
  }
  if (modes.empty() || (int)modes.size() > kMaxDirectModes) return LBM_OK;
  s.nmodes = (int)modes.size();
  s.small_k = true;
  for (const int4& m : modes)
    if (std::abs(m.x) > 2 || std::abs(m.y) > 2 || std::abs(m.z) > 2) s.small_k = false;
  const int mp = modes_per_pass<3>();
  s.npadded = (s.nmodes + mp - 1) / mp * mp;
  modes.resize(s.npadded, make_int4(0, 0, 0, 0));
  if (s.D == 2 && s.small_k) {
    // n1 * R column walkers, about kDftBlocks blocks' worth, R <= n0
    s.dft_rows = (int)std::max<int64_t>(
        1, std::min<int64_t>(s.n[0], (int64_t)kDftBlocks * kBlock / s.n[1]));
    s.dft_blocks = (int)(((int64_t)s.n[1] * s.dft_rows + kBlock - 1) / kBlock);
  }
  if (!(s.D == 2 && s.small_k))
    s.dft_blocks = (int)std::min<int64_t>(kDftBlocks, (s.N + kBlock - 1) / kBlock);
  std::vector<double2> tw;
  for (int d = 0; d < s.D; ++d)
    for (int m = 0; m < s.n[d]; ++m) {
      const double a = 2.0 * M_PI * m / s.n[d];
      tw.push_back(make_double2(std::cos(a), -std::sin(a)));
    }
  auto alloc = [&](void** p, size_t bytes) {
    return hipMalloc(p, bytes) == hipSuccess
               ? (int)LBM_OK
               : lbm::record_error(h, LBM_E_NOMEM, "spectral force buffers: out of memory");
  };
  int rc = alloc((void**)&s.modes, sizeof(int4) * modes.size());
  if (!rc) rc = alloc((void**)&s.tw, sizeof(double2) * tw.size());
  if (!rc) rc = alloc((void**)&s.dft_partial, sizeof(double2) * s.dft_blocks * s.nmodes * s.D);
  if (!rc) rc = alloc((void**)&s.coef, sizeof(double2) * s.npadded * s.D);
  if (!rc)
    rc = lbm::record_hip(h, hipMemset(s.coef, 0, sizeof(double2) * s.npadded * s.D),
                         "coefficient clear");
  if (!rc && s.D == 2 && s.small_k) {
    rc = alloc((void**)&s.dense, sizeof(double2) * 30);
    if (!rc)
      rc = lbm::record_hip(h, hipMemset(s.dense, 0, sizeof(double2) * 30), "coefficient clear");
  }
  if (!rc && spec.injection) rc = alloc((void**)&s.finj, sizeof(double) * s.D * s.N);
  if (!rc)
    rc = lbm::record_hip(h, hipMemcpy(s.modes, modes.data(), sizeof(int4) * modes.size(),
                                      hipMemcpyHostToDevice), "mode upload");
  if (!rc)
    rc = lbm::record_hip(h, hipMemcpy(s.tw, tw.data(), sizeof(double2) * tw.size(),
                                      hipMemcpyHostToDevice), "twiddle upload");
  if (!rc && spec.injection) {
    ForceSpec inj = spec;
    inj.removal = 0;
    rc = fft_force(h, s, ff, inj, s.finj);
  }
  if (!rc) s.direct = true;
  return rc;
}

}  // namespace

int lbm::update_spectral_force(lbm_handle* h) {
  lbm::ForceField ff;
  if (int rc = lbm::force_field(h, ff)) return rc;
  lbm::DeviceFields f{ff.rho, ff.u, nullptr, ff.D, {ff.n[0], ff.n[1], ff.n[2]}, ff.N, ff.stream, 0};
  State* s = nullptr;
  if (int rc = state_of(h, f, s)) return rc;
  const hipStream_t st = ff.stream;
  const lbm_config& c = ff.cfg;
  ForceSpec spec{};
  spec.injection = c.force == LBM_FORCE_CONSTANT_SHELL || c.force == LBM_FORCE_TURBULENT_2D;
  spec.removal = c.force == LBM_FORCE_ENERGY_REMOVAL || c.force == LBM_FORCE_TURBULENT_2D;
  spec.kmin2 = c.force_kmin * c.force_kmin;
  spec.kmax2 = c.force_kmax * c.force_kmax;
  spec.rmin2 = c.removal_kmin * c.removal_kmin;
  spec.rmax2 = c.removal_kmax * c.removal_kmax;
  for (int d = 0; d < 3; ++d) {
    spec.A[d] = c.amplitude[d];
    spec.R[d] = c.removal_amplitude[d];
  }
  spec.scale = 1.0 / (double)s->N;
  if (!s->direct_checked)
    if (int rc = direct_setup(h, *s, ff, spec)) return rc;
  if (s->direct) {
    const Dims g = dims_of(*s);
    const int mp = s->D == 3 ? modes_per_pass<3>() : modes_per_pass<2>();
    const unsigned passes = (unsigned)((s->nmodes + mp - 1) / mp);
    if (s->D == 2 && s->small_k) {
      hipLaunchKernelGGL(k_mode_dft_partial_2d_small<>, dim3(s->dft_blocks, passes), dim3(kBlock),
                         0, st, ff.rho, ff.u, g, s->N, s->dft_rows, s->modes, s->nmodes, s->tw,
                         s->dft_partial);
    } else {
      auto partial = s->D == 3 ? (s->small_k ? k_mode_dft_partial<3, true>
                                             : k_mode_dft_partial<3, false>)
                               : k_mode_dft_partial<2, false>;
      hipLaunchKernelGGL(partial, dim3(s->dft_blocks, passes), dim3(kBlock), 0, st, ff.rho, ff.u,
                         g, s->N, s->modes, s->nmodes, s->tw, s->dft_partial);
    }
    hipLaunchKernelGGL(k_mode_dft_final, dim3(s->nmodes * s->D), dim3(kBlock), 0, st,
                       s->dft_partial, s->dft_blocks, s->nmodes, s->D, s->modes, spec.R[0],
                       spec.R[1], spec.R[2], spec.scale, s->coef, s->dense);
    if (s->dense) {
      if (ff.f32)
        hipLaunchKernelGGL(k_force_synth_2d_small<float>, dim3(grid_of((s->N + 1) / 2)),
                           dim3(kBlock), 0, st, s->finj, s->dense, s->tw, g, s->N, ff.f32);
      else
        hipLaunchKernelGGL(k_force_synth_2d_small<double>, dim3(grid_of((s->N + 1) / 2)),
                           dim3(kBlock), 0, st, s->finj, s->dense, s->tw, g, s->N, ff.f64);
    } else if (ff.f32) {
      auto synth = s->small_k ? k_force_synth<float, true> : k_force_synth<float, false>;
      hipLaunchKernelGGL(synth, dim3(grid_of(s->N)), dim3(kBlock), 0, st, s->finj, s->coef,
                         s->modes, s->npadded, s->tw, g, s->N, ff.f32);
    } else {
      auto synth = s->small_k ? k_force_synth<double, true> : k_force_synth<double, false>;
      hipLaunchKernelGGL(synth, dim3(grid_of(s->N)), dim3(kBlock), 0, st, s->finj, s->coef,
                         s->modes, s->npadded, s->tw, g, s->N, ff.f64);
    }
    return lbm::record_hip(h, hipGetLastError(), "direct force kernels");
  }
  if (int rc = fft_force(h, *s, ff, spec, ff.f64)) return rc;
  if (ff.f32) {
    hipLaunchKernelGGL(k_to_f32, dim3(grid_of(s->D * s->N)), dim3(kBlock), 0, st, ff.f64,
                       (int64_t)s->D * s->N, ff.f32);
    if (int rc = lbm::record_hip(h, hipGetLastError(), "force conversion")) return rc;
  }
  return LBM_OK;
}

extern "C" {

int lbm_spectral_shells(const lbm_handle* h, int* nshells) {
  if (!h || !nshells) return LBM_E_INVALID;
  lbm_config c;
  if (int rc = lbm_get_config(h, &c)) return rc;
  int nmax = std::max(c.nx, c.ny);
  if (c.lattice != LBM_D2Q9) nmax = std::max(nmax, c.nz);
  *nshells = nmax / 2;
  return LBM_OK;
}

int lbm_vorticity(lbm_handle* h, int scaling, double* omega, double* enstrophy) {
  if (!h) return LBM_E_INVALID;
  if (int rc = check_scaling(h, scaling)) return rc;
  lbm::DeviceFields f;
  if (int rc = lbm::device_fields(h, f)) return rc;
  State* s = nullptr;
  if (int rc = state_of(h, f, s)) return rc;
  const hipStream_t st = f.stream;
  const bool physical = scaling == LBM_SPECTRAL_PHYSICAL;
  if (int rc = forward(h, *s, f.u, f.version, st)) return rc;
  const double invN = 1.0 / (double)s->N;
  hipLaunchKernelGGL(k_curl_hat, dim3(grid_of(s->M)), dim3(kBlock), 0, st, s->hat, s->what,
                     dims_of(*s), physical ? 1 : 0, physical ? invN : invN * invN);
  if (int rc = lbm::record_hip(h, hipGetLastError(), "curl kernel")) return rc;
  hipfftResult r = hipfftSetStream(s->inv, st);
  if (r == HIPFFT_SUCCESS)
    r = hipfftExecZ2D(s->inv, (hipfftDoubleComplex*)s->what, (hipfftDoubleReal*)s->omega);
  if (r != HIPFFT_SUCCESS) return fft_fail(h, r, "inverse transform");
  if (enstrophy) {
    // TotalEnstrophy: (1/N) sum 1/2 |omega|^2 (Analysis.h:86-93, 27)
    hipLaunchKernelGGL(k_half_square_partial, dim3(kReduceBlocks), dim3(kBlock), 0, st,
                       s->omega, s->C, s->N, s->partial);
    hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(kBlock), 0, st, s->partial, kReduceBlocks,
                       invN, s->partial + kReduceBlocks);
    if (int rc = lbm::record_hip(h, hipGetLastError(), "enstrophy kernels")) return rc;
    if (int rc = lbm::record_hip(h, hipMemcpyAsync(enstrophy, s->partial + kReduceBlocks,
                                                   sizeof(double), hipMemcpyDeviceToHost, st),
                                 "enstrophy copy"))
      return rc;
  }
  if (omega) {
    if (int rc = lbm::record_hip(h, hipMemcpyAsync(omega, s->omega, sizeof(double) * s->C * s->N,
                                                   hipMemcpyDeviceToHost, st),
                                 "vorticity copy"))
      return rc;
  }
  return lbm::record_hip(h, hipStreamSynchronize(st), "vorticity");
}

int lbm_spectra(lbm_handle* h, int scaling, double* energy, double* forcing) {
  if (!h) return LBM_E_INVALID;
  if (int rc = check_scaling(h, scaling)) return rc;
  lbm::DeviceFields f;
  if (int rc = lbm::device_fields(h, f)) return rc;
  State* s = nullptr;
  if (int rc = state_of(h, f, s)) return rc;
  const hipStream_t st = f.stream;
  const bool physical = scaling == LBM_SPECTRAL_PHYSICAL;
  const int ns = s->nshells;
  const double invN = 1.0 / (double)s->N;
  // REFERENCE: energy / N (Analysis.h:115-119, AnalysisList.h:195), forcing
  // unnormalised; PHYSICAL: both / N^2
  const double scale[2] = {physical ? invN * invN : invN, physical ? invN * invN : 1.0};
  const double* src[2] = {f.u, f.force};
  double* dst[2] = {energy, forcing};
  for (int which = 0; which < 2; ++which) {
    if (!dst[which] || ns == 0) continue;
    if (int rc = forward(h, *s, src[which], f.version, st)) return rc;
    for (int lo = 0; lo < ns; lo += kShellWindow) {
      const int cnt = std::min(kShellWindow, ns - lo);
      const size_t lds = sizeof(double) * (kBlock / 64) * cnt;
      hipLaunchKernelGGL(k_shell_bins, dim3(s->shell_blocks), dim3(kBlock), lds, st, s->hat,
                         dims_of(*s), physical ? 1 : 0, lo, cnt, s->bins);
      hipLaunchKernelGGL(k_shell_final, dim3(cnt), dim3(kBlock), 0, st,
                         s->bins, s->shell_blocks, cnt, scale[which], s->spec + which * ns + lo);
      if (int rc = lbm::record_hip(h, hipGetLastError(), "shell kernels")) return rc;
    }
    if (int rc = lbm::record_hip(h, hipMemcpyAsync(dst[which], s->spec + which * ns,
                                                   sizeof(double) * ns, hipMemcpyDeviceToHost,
                                                   st),
                                 "spectrum copy"))
      return rc;
  }
  return lbm::record_hip(h, hipStreamSynchronize(st), "spectra");
}

}  // extern "C"
