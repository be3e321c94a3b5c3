// aux_kernels.hpp -- launchers of the non-step kernels: initial conditions,
// population format conversion and the FP64 diagnostics reduction.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace lbm {

// f_i = feq_i(rho, u) evaluated in FP64 reference order (Initialize.h:106-117),
// then stored as (f_i - w_i) * scale_i when shifted.  rho[N], u[D][N] device.
hipError_t launch_init_equilibrium(int lattice, int eq, int storage,
                                   int shifted, void* f, const double* rho,
                                   double rho0, const double* u, int N,
                                   hipStream_t st);

// Seeded velocity noise into u[D][N] (SplitMix64 counter stream).
hipError_t launch_noise(double* u, int N, int D, uint64_t seed, double amp,
                        hipStream_t st);

// Taylor-Green field from per-axis sin/cos tables (host-evaluated).
hipError_t launch_taylor_green(double* u, int nx, int ny, int nz, int D,
                               double u0, const double* sx, const double* cx,
                               const double* sy, const double* cy,
                               const double* cz, hipStream_t st);

// Spectral field u[D][N] = gamma Re sum_k uh(k) e^{2 pi i k.x/n} from the
// Hermitian mode cube (ur, ui)[2K+1][2K+1][2Kz+1][D] (Kz = K in 3D, 0 in 2D)
// by three separable passes.  tab = cos/sin tables per axis: cx[nx] sx[nx]
// cy[ny] sy[ny] cz[nz] sz[nz].  t1: (2K+1)^2 nz D, t2: (2K+1) ny nz D doubles.
hipError_t launch_spectral(const double* ur, const double* ui, int K, int D,
                           int nx, int ny, int nz, const double* tab, double gamma,
                           double* t1r, double* t1i, double* t2r, double* t2i,
                           double* u, hipStream_t st);

// One population between FP64 (unshifted) and storage format.
hipError_t launch_pop_from_f64(int storage, void* dst, const double* src,
                               int N, double w, double scale, hipStream_t st);
hipError_t launch_pop_to_f64(int storage, double* dst, const void* src, int N,
                             double w, double inv_scale, hipStream_t st);

// Scalar field conversions (alpha state) and fills.
hipError_t launch_scalar_from_f64(int type, void* dst, const double* src,
                                  int N, hipStream_t st);
hipError_t launch_scalar_to_f64(int type, double* dst, const void* src, int N,
                                hipStream_t st);
hipError_t launch_fill(int type, void* dst, double value, int N,
                       hipStream_t st);

// Restart fields from population i (FP64, unshifted) of a restart:
// rho (=|+=) f_i, u_d (=|+=) c_id f_i, starting over at i = 0
// (Initialize.h:124-145).
hipError_t launch_restart_fields(int lattice, double* rho, double* u, const double* f, int i,
                                 int N, hipStream_t st);

// Deterministic FP64 diagnostics: out[0] = sum rho, out[1] = sum 1/2 rho
// |u|^2, out[2] = sum 1/2 |curl u|^2.  partial must hold 3 * reduce_blocks(N)
// doubles.
int reduce_blocks(int N);
hipError_t launch_reduce(const double* rho, const double* u, int nx, int ny,
                         int nz, int D, double* partial, double* out,
                         hipStream_t st);

// Host tables of w_i and the FP16 storage scale s_i for a lattice.
int lattice_tables(int lattice, double* w, double* scale, int* q, int* d);

}  // namespace lbm
