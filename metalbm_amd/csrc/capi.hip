// capi.hip -- the C ABI (include/metalbm_amd.h) over the HIP kernels.
//
// A handle owns one periodic domain on the current device: two population
// lattices (ping-pong, Algorithm.h:336), the ELBM alpha field, host-built
// force tables, and lazily allocated output fields and staging buffers.
// Errors are returned as codes with a message; nothing exits.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/metalbm_amd.h"
#include "aux_kernels.hpp"
#include "registry.hpp"

using namespace lbm;

namespace lbm {
const Registry& registry() {
  static const Registry r = [] {
    Registry x;
    register_step_f64(x);
    register_step_f32(x);
    return x;
  }();
  return r;
}
}  // namespace lbm

struct lbm_handle {
  lbm_config cfg{};
  int Q = 0, D = 0, N = 0;
  int nk = 0, rows = 0, block = 256, tiles_per_row = 1;
  int vec = 0;              // cells per lane of the vector kernel (0: cell kernel)
  size_t elem = 8;          // bytes per stored population value
  int ctype = TYPE_F64;     // compute type
  void* f[2] = {nullptr, nullptr};
  int cur = 0;              // lattice holding the current populations
  void* alpha = nullptr;    // ELBM warm-start field (compute type)
  double* fields = nullptr; // rho[N], uh[D N], alpha[N], F[D N]
  bool fields_valid = false;
  double* staging = nullptr;   // N doubles
  double* reduce_buf = nullptr;
  void* ftab[3] = {nullptr, nullptr, nullptr};
  int fax[3] = {3, 3, 3};
  unsigned long long* stats = nullptr;  // 6 counters
  hipStream_t stream = nullptr;
  bool own_stream = false;
  StepFn step = nullptr;
  int64_t steps = 0;
  double w[27], scale[27];
  std::string err;
};

namespace {

thread_local std::string g_create_error;

int fail(lbm_handle* h, int code, const std::string& msg) {
  if (h) h->err = msg;
  else g_create_error = msg;
  return code;
}

int hip_fail(lbm_handle* h, hipError_t e, const char* what) {
  return fail(h, LBM_E_HIP,
              std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(h, expr)                                   \
  do {                                                     \
    hipError_t e_ = (expr);                                \
    if (e_ != hipSuccess) return hip_fail((h), e_, #expr); \
  } while (0)

int validate(const lbm_config* c, std::string& why) {
  const int lat = c->lattice;
  if (lat != LBM_D2Q9 && lat != LBM_D3Q19 && lat != LBM_D3Q27) {
    why = "lattice must be D2Q9 (3), D3Q19 (9) or D3Q27 (10)";
    return LBM_E_INVALID;
  }
  if (c->collision != LBM_BGK && c->collision != LBM_ELBM) {
    why = "collision must be BGK (1) or ELBM (2)";
    return LBM_E_INVALID;
  }
  if (c->forcing_scheme < LBM_FS_NONE || c->forcing_scheme > LBM_FS_EDM) {
    why = "forcing_scheme must be None (1), Guo (2), ShanChen (3) or EDM (4)";
    return LBM_E_INVALID;
  }
  if (c->force != LBM_FORCE_NONE && c->force != LBM_FORCE_CONSTANT &&
      c->force != LBM_FORCE_SINUSOIDAL && c->force != LBM_FORCE_KOLMOGOROV) {
    why = "force must be None (0), Constant (4), Sinusoidal (5) or Kolmogorov (6)";
    return LBM_E_INVALID;
  }
  if (c->equilibrium != LBM_EQ_MA3 && c->equilibrium != LBM_EQ_EXACT) {
    why = "equilibrium must be TruncationMa3 (3) or Exact (1)";
    return LBM_E_INVALID;
  }
  if (c->equilibrium == LBM_EQ_EXACT && lat == LBM_D3Q19) {
    why = "there is no Exact equilibrium for D3Q19";
    return LBM_E_INVALID;
  }
  if (c->nx <= 0 || c->ny <= 0 || c->nz <= 0) {
    why = "nx, ny, nz must be positive";
    return LBM_E_INVALID;
  }
  if (lat == LBM_D2Q9 && c->nz != 1) {
    why = "nz must be 1 for D2Q9";
    return LBM_E_INVALID;
  }
  const double cells = (double)c->nx * c->ny * c->nz;
  if (cells >= 2147483648.0) {
    why = "nx*ny*nz must be below 2^31";
    return LBM_E_INVALID;
  }
  if (!(c->tau > 0.0) || !std::isfinite(c->tau)) {
    why = "tau must be positive and finite";
    return LBM_E_INVALID;
  }
  const int D = lat == LBM_D2Q9 ? 2 : 3;
  if (c->force == LBM_FORCE_SINUSOIDAL)
    for (int d = 0; d < D; ++d)
      if (c->wavelength[d] == 0.0) {
        why = "Sinusoidal force needs nonzero wavelength in every dimension";
        return LBM_E_INVALID;
      }
  if (c->force == LBM_FORCE_KOLMOGOROV && c->wavelength[0] == 0.0) {
    why = "Kolmogorov force needs a nonzero wavelength[0]";
    return LBM_E_INVALID;
  }
  if (c->storage < LBM_F64 || c->storage > LBM_F16 || c->compute < LBM_F64 ||
      c->compute > LBM_F32) {
    why = "storage must be F64/F32/F16 and compute F64/F32";
    return LBM_E_INVALID;
  }
  if (c->storage == LBM_F16 && !c->shifted) {
    why = "F16 storage requires shifted populations (f - w)";
    return LBM_E_INVALID;
  }
  if (c->compute == LBM_F64 && c->storage != LBM_F64) {
    why = "F64 compute requires F64 storage";
    return LBM_E_INVALID;
  }
  if (c->compute == LBM_F32 && c->storage == LBM_F64 && !c->shifted) {
    why = "F32 compute with F64 storage requires shifted populations";
    return LBM_E_INVALID;
  }
  if (c->kernel < LBM_KERNEL_AUTO || c->kernel > LBM_KERNEL_VECTOR) {
    why = "kernel must be AUTO (0), CELL (1) or VECTOR (2)";
    return LBM_E_INVALID;
  }
  if (c->kernel == LBM_KERNEL_VECTOR && c->collision != LBM_BGK) {
    why = "the vector kernel implements BGK only";
    return LBM_E_INVALID;
  }
  if (c->collision == LBM_ELBM && c->alpha_mode != LBM_ALPHA_NEWTON &&
      c->alpha_mode != LBM_ALPHA_LEARNED) {
    why = "alpha_mode must be NEWTON (0) or LEARNED (1)";
    return LBM_E_INVALID;
  }
  return LBM_OK;
}

int src_of(int fs) {
  if (fs == LBM_FS_GUO) return SRC_GUO;
  if (fs == LBM_FS_EDM) return SRC_EDM;
  return SRC_ZERO;
}

int pick_block(int nk) {
  if (nk <= 256) return ((nk + 63) / 64) * 64;
  for (int b : {256, 192, 128})
    if (nk % b == 0) return b;
  return 256;
}

// Force tables (Force.h:154-159, 208-215, 262-267), FP64 on the host with the
// reference's own expression, stored in the compute type.
int build_force_tables(lbm_handle* h) {
  const lbm_config& c = h->cfg;
  const int n[3] = {c.nx, c.ny, c.nz};
  std::vector<double> tab[3];
  int axis[3] = {3, 3, 3};
  switch (c.force) {
    case LBM_FORCE_CONSTANT:
      for (int d = 0; d < h->D; ++d) tab[d].assign(1, c.amplitude[d]);
      break;
    case LBM_FORCE_SINUSOIDAL:
      for (int d = 0; d < h->D; ++d) {
        axis[d] = d;
        tab[d].resize(n[d]);
        for (int i = 0; i < n[d]; ++i)
          tab[d][i] = c.amplitude[d] *
                      std::sin((double)((unsigned)i * 2u) * M_PI / c.wavelength[d]);
      }
      break;
    case LBM_FORCE_KOLMOGOROV:
      axis[0] = 1;
      tab[0].resize(c.ny);
      for (int y = 0; y < c.ny; ++y)
        tab[0][y] = c.amplitude[0] *
                    std::sin((double)((unsigned)y * 2u) * M_PI / c.wavelength[0]);
      break;
    default:
      break;
  }
  for (int d = 0; d < 3; ++d) {
    h->fax[d] = axis[d];
    if (tab[d].empty()) continue;
    const size_t n_ = tab[d].size();
    if (h->ctype == TYPE_F64) {
      HIP_TRY(h, hipMalloc(&h->ftab[d], n_ * sizeof(double)));
      HIP_TRY(h, hipMemcpy(h->ftab[d], tab[d].data(), n_ * sizeof(double),
                           hipMemcpyHostToDevice));
    } else {
      std::vector<float> t(tab[d].begin(), tab[d].end());
      HIP_TRY(h, hipMalloc(&h->ftab[d], n_ * sizeof(float)));
      HIP_TRY(h, hipMemcpy(h->ftab[d], t.data(), n_ * sizeof(float),
                           hipMemcpyHostToDevice));
    }
  }
  return LBM_OK;
}

int ensure_staging(lbm_handle* h) {
  if (!h->staging)
    HIP_TRY(h, hipMalloc(&h->staging, (size_t)h->N * (1 + h->D) * sizeof(double)));
  return LBM_OK;
}

int ensure_fields(lbm_handle* h) {
  if (!h->fields)
    HIP_TRY(h, hipMalloc(&h->fields, (size_t)h->N * (2 + 2 * h->D) * sizeof(double)));
  return LBM_OK;
}

void release(lbm_handle* h) {
  for (void* p : {h->f[0], h->f[1], h->alpha, (void*)h->fields,
                  (void*)h->staging, (void*)h->reduce_buf, h->ftab[0],
                  h->ftab[1], h->ftab[2], (void*)h->stats})
    if (p) (void)hipFree(p);
  if (h->own_stream && h->stream) (void)hipStreamDestroy(h->stream);
}

StepArgs make_args(const lbm_handle* h) {
  StepArgs a{};
  const lbm_config& c = h->cfg;
  a.nx = c.nx;
  a.ny = c.ny;
  a.nz = c.nz;
  a.nk = h->nk;
  a.tiles_per_row = h->tiles_per_row;
  a.N = h->N;
  const double beta = 1.0 / (2.0 * c.tau);    // Collision.h:122
  a.beta = beta;
  a.omega_f = 1. - 2. * beta;                 // Collision.h:145
  a.omega_eq = 2. * beta;                     // Collision.h:147
  a.guo_coef = 1.0 - 1.0 / (2.0 * c.tau);     // ForcingScheme.h:115
  for (int d = 0; d < 3; ++d) {
    a.ftab[d] = h->ftab[d];
    a.fax[d] = h->fax[d];
  }
  a.alpha_state = h->alpha;
  a.stats = h->stats;
  return a;
}

}  // namespace

extern "C" {

int lbm_abi_version(void) { return LBM_ABI_VERSION; }

int lbm_config_default(lbm_config* c) {
  if (!c) return LBM_E_INVALID;
  std::memset(c, 0, sizeof(*c));
  c->lattice = LBM_D2Q9;
  c->collision = LBM_BGK;
  c->forcing_scheme = LBM_FS_GUO;
  c->force = LBM_FORCE_NONE;
  c->equilibrium = LBM_EQ_MA3;
  c->nx = 16;
  c->ny = 16;
  c->nz = 1;
  c->tau = 0.6;
  for (int d = 0; d < 3; ++d) c->wavelength[d] = 1.0;
  c->storage = LBM_F64;
  c->compute = LBM_F64;
  c->shifted = 0;
  c->alpha_mode = LBM_ALPHA_NEWTON;
  c->residual_tol = 1e-10;
  c->kernel = LBM_KERNEL_AUTO;
  return LBM_OK;
}

const char* lbm_last_error(const lbm_handle* h) {
  return h ? h->err.c_str() : g_create_error.c_str();
}

int lbm_create(const lbm_config* cfg, lbm_handle** out) {
  if (!cfg || !out) return fail(nullptr, LBM_E_INVALID, "null argument");
  *out = nullptr;
  std::string why;
  if (validate(cfg, why) != LBM_OK) return fail(nullptr, LBM_E_INVALID, why);
  if (cfg->collision == LBM_ELBM && cfg->alpha_mode == LBM_ALPHA_LEARNED)
    return fail(nullptr, LBM_E_INVALID,
                "LEARNED alpha needs a trained network (not yet loaded)");
  lbm_handle* h = new lbm_handle();
  h->cfg = *cfg;
  if (lattice_tables(cfg->lattice, h->w, h->scale, &h->Q, &h->D) != 0) {
    delete h;
    return fail(nullptr, LBM_E_INVALID, "unknown lattice");
  }
  h->N = cfg->nx * cfg->ny * cfg->nz;
  h->ctype = cfg->compute == LBM_F64 ? TYPE_F64 : TYPE_F32;
  h->elem = cfg->storage == LBM_F64 ? 8 : (cfg->storage == LBM_F32 ? 4 : 2);
  h->nk = h->D == 3 ? cfg->nz : cfg->ny;
  h->rows = h->D == 3 ? cfg->nx * cfg->ny : cfg->nx;
  h->block = pick_block(h->nk);
  h->tiles_per_row = (h->nk + h->block - 1) / h->block;
  KernelKey key{cfg->lattice, cfg->collision, src_of(cfg->forcing_scheme),
                cfg->equilibrium, cfg->storage, h->ctype, cfg->shifted ? 1 : 0};
  h->step = registry().find(key);
  if (cfg->collision == LBM_BGK && cfg->kernel != LBM_KERNEL_CELL) {
    // Widest vector whose 64-lane segment tiles the contiguous axis.
    static const int cands[3][3] = {{2, 0, 0}, {4, 2, 0}, {8, 4, 2}};
    for (int V : cands[cfg->storage]) {
      if (V == 0 || h->nk % (64 * V) != 0) continue;
      KernelKey kv = key;
      kv.vec = V;
      if (StepFn fn = registry().find(kv)) {
        h->step = fn;
        h->vec = V;
        h->block = 256;
        h->tiles_per_row = h->nk / (64 * V);
        break;
      }
    }
  }
  auto bail = [&](int code) {
    g_create_error = h->err;
    release(h);
    delete h;
    return code;
  };
  if (!h->step) {
    h->err = "no kernel instantiated for this storage/compute combination";
    return bail(LBM_E_INVALID);
  }
  if (cfg->kernel == LBM_KERNEL_VECTOR && !h->vec) {
    h->err = "the vector kernel needs nz (3D) or ny (2D) divisible by 64 V "
             "(V = 2 for F64, 2 or 4 for F32, 2, 4 or 8 for F16)";
    return bail(LBM_E_INVALID);
  }
  const size_t bytes = (size_t)h->N * h->Q * h->elem;
  for (int k = 0; k < 2; ++k) {
    hipError_t e = hipMalloc(&h->f[k], bytes);
    if (e != hipSuccess) {
      h->err = std::string("population lattice allocation: ") + hipGetErrorString(e);
      return bail(LBM_E_NOMEM);
    }
  }
  hipError_t e = hipStreamCreate(&h->stream);
  if (e != hipSuccess) {
    h->err = std::string("hipStreamCreate: ") + hipGetErrorString(e);
    return bail(LBM_E_HIP);
  }
  h->own_stream = true;
  if (int rc = build_force_tables(h)) return bail(rc);
  if (cfg->collision == LBM_ELBM) {
    const size_t ab = (size_t)h->N * (h->ctype == TYPE_F64 ? 8 : 4);
    e = hipMalloc(&h->alpha, ab);
    if (e == hipSuccess) e = launch_fill(h->ctype, h->alpha, 2.0, h->N, h->stream);  // Initialize.h:83-89
    if (e == hipSuccess) e = hipMalloc(&h->stats, 6 * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemsetAsync(h->stats, 0, 6 * sizeof(unsigned long long), h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) {
      h->err = std::string("alpha field: ") + hipGetErrorString(e);
      return bail(LBM_E_HIP);
    }
  }
  *out = h;
  return LBM_OK;
}

int lbm_destroy(lbm_handle* h) {
  if (!h) return LBM_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  release(h);
  delete h;
  return LBM_OK;
}

int lbm_sizes(const lbm_handle* h, int64_t* cells, int* q, int* d) {
  if (!h) return LBM_E_INVALID;
  if (cells) *cells = h->N;
  if (q) *q = h->Q;
  if (d) *d = h->D;
  return LBM_OK;
}

int lbm_init_equilibrium(lbm_handle* h, const double* rho, const double* u) {
  if (!h || !rho || !u) return h ? fail(h, LBM_E_INVALID, "null array") : LBM_E_INVALID;
  if (int rc = ensure_staging(h)) return rc;
  double* drho = h->staging;
  double* du = h->staging + h->N;
  HIP_TRY(h, hipMemcpyAsync(drho, rho, (size_t)h->N * sizeof(double),
                            hipMemcpyHostToDevice, h->stream));
  HIP_TRY(h, hipMemcpyAsync(du, u, (size_t)h->N * h->D * sizeof(double),
                            hipMemcpyHostToDevice, h->stream));
  HIP_TRY(h, launch_init_equilibrium(h->cfg.lattice, h->cfg.equilibrium,
                                     h->cfg.storage, h->cfg.shifted,
                                     h->f[h->cur], drho, 0.0, du, h->N, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return LBM_OK;
}

int lbm_init_noise(lbm_handle* h, double rho0, double amp, uint64_t seed) {
  if (!h) return LBM_E_INVALID;
  if (int rc = ensure_staging(h)) return rc;
  double* du = h->staging + h->N;
  HIP_TRY(h, launch_noise(du, h->N, h->D, seed, amp, h->stream));
  HIP_TRY(h, launch_init_equilibrium(h->cfg.lattice, h->cfg.equilibrium,
                                     h->cfg.storage, h->cfg.shifted,
                                     h->f[h->cur], nullptr, rho0, du, h->N,
                                     h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return LBM_OK;
}

int lbm_init_taylor_green(lbm_handle* h, double rho0, double u0) {
  if (!h) return LBM_E_INVALID;
  if (int rc = ensure_staging(h)) return rc;
  const lbm_config& c = h->cfg;
  const double two_pi = 2.0 * M_PI;
  const int n[3] = {c.nx, c.ny, c.nz};
  std::vector<double> host;
  size_t off[6];
  for (int d = 0; d < 3; ++d) {
    off[2 * d] = host.size();
    for (int i = 0; i < n[d]; ++i) host.push_back(std::sin(two_pi * i / n[d]));
    off[2 * d + 1] = host.size();
    for (int i = 0; i < n[d]; ++i) host.push_back(std::cos(two_pi * i / n[d]));
  }
  double* tab = nullptr;
  HIP_TRY(h, hipMalloc(&tab, host.size() * sizeof(double)));
  hipError_t e = hipMemcpy(tab, host.data(), host.size() * sizeof(double),
                           hipMemcpyHostToDevice);
  double* du = h->staging + h->N;
  if (e == hipSuccess)
    e = launch_taylor_green(du, c.nx, c.ny, c.nz, h->D, u0, tab + off[0],
                            tab + off[1], tab + off[2], tab + off[3],
                            tab + off[5], h->stream);
  if (e == hipSuccess)
    e = launch_init_equilibrium(c.lattice, c.equilibrium, c.storage, c.shifted,
                                h->f[h->cur], nullptr, rho0, du, h->N, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  (void)hipFree(tab);
  if (e != hipSuccess) return hip_fail(h, e, "Taylor-Green initialisation");
  return LBM_OK;
}

int lbm_set_populations(lbm_handle* h, const double* f) {
  if (!h || !f) return h ? fail(h, LBM_E_INVALID, "null array") : LBM_E_INVALID;
  if (int rc = ensure_staging(h)) return rc;
  char* dst = static_cast<char*>(h->f[h->cur]);
  for (int i = 0; i < h->Q; ++i) {
    HIP_TRY(h, hipMemcpyAsync(h->staging, f + (size_t)i * h->N,
                              (size_t)h->N * sizeof(double),
                              hipMemcpyHostToDevice, h->stream));
    const double w = h->cfg.shifted ? h->w[i] : 0.0;
    const double s = h->cfg.storage == LBM_F16 ? h->scale[i] : 1.0;
    HIP_TRY(h, launch_pop_from_f64(h->cfg.storage, dst + (size_t)i * h->N * h->elem,
                                   h->staging, h->N, w, s, h->stream));
  }
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return LBM_OK;
}

int lbm_get_populations(lbm_handle* h, double* f) {
  if (!h || !f) return h ? fail(h, LBM_E_INVALID, "null array") : LBM_E_INVALID;
  if (int rc = ensure_staging(h)) return rc;
  const char* src = static_cast<const char*>(h->f[h->cur]);
  for (int i = 0; i < h->Q; ++i) {
    const double w = h->cfg.shifted ? h->w[i] : 0.0;
    const double s = h->cfg.storage == LBM_F16 ? 1.0 / h->scale[i] : 1.0;
    HIP_TRY(h, launch_pop_to_f64(h->cfg.storage, h->staging,
                                 src + (size_t)i * h->N * h->elem, h->N, w, s,
                                 h->stream));
    HIP_TRY(h, hipMemcpyAsync(f + (size_t)i * h->N, h->staging,
                              (size_t)h->N * sizeof(double),
                              hipMemcpyDeviceToHost, h->stream));
  }
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return LBM_OK;
}

int lbm_set_alpha(lbm_handle* h, const double* alpha) {
  if (!h || !alpha) return h ? fail(h, LBM_E_INVALID, "null array") : LBM_E_INVALID;
  if (!h->alpha) return fail(h, LBM_E_STATE, "alpha field exists only for ELBM");
  if (int rc = ensure_staging(h)) return rc;
  HIP_TRY(h, hipMemcpyAsync(h->staging, alpha, (size_t)h->N * sizeof(double),
                            hipMemcpyHostToDevice, h->stream));
  HIP_TRY(h, launch_scalar_from_f64(h->ctype, h->alpha, h->staging, h->N, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return LBM_OK;
}

int lbm_get_alpha(lbm_handle* h, double* alpha) {
  if (!h || !alpha) return h ? fail(h, LBM_E_INVALID, "null array") : LBM_E_INVALID;
  if (!h->alpha) return fail(h, LBM_E_STATE, "alpha field exists only for ELBM");
  if (int rc = ensure_staging(h)) return rc;
  HIP_TRY(h, launch_scalar_to_f64(h->ctype, h->staging, h->alpha, h->N, h->stream));
  HIP_TRY(h, hipMemcpyAsync(alpha, h->staging, (size_t)h->N * sizeof(double),
                            hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return LBM_OK;
}

int lbm_step(lbm_handle* h, int nsteps, int store_last) {
  if (!h) return LBM_E_INVALID;
  if (nsteps < 0) return fail(h, LBM_E_INVALID, "nsteps must be >= 0");
  if (nsteps == 0) return LBM_OK;
  if (store_last)
    if (int rc = ensure_fields(h)) return rc;
  StepArgs a = make_args(h);
  // Vector kernel: one wavefront per 64 V-cell segment, 4 per workgroup.
  const int grid = h->vec ? (h->rows * h->tiles_per_row + 3) / 4
                          : h->rows * h->tiles_per_row;
  for (int s = 0; s < nsteps; ++s) {
    a.src = h->f[h->cur];
    a.dst = h->f[1 - h->cur];
    a.store = (store_last && s == nsteps - 1) ? 1 : 0;
    if (a.store) {
      a.rho_out = h->fields;
      a.uh_out = h->fields + h->N;
      a.alpha_out = h->fields + (size_t)h->N * (1 + h->D);
      a.F_out = h->fields + (size_t)h->N * (2 + h->D);
    }
    hipError_t e = h->step(a, h->cfg.forcing_scheme, h->block, grid, h->stream);
    if (e != hipSuccess) return hip_fail(h, e, "fused step launch");
    h->cur = 1 - h->cur;
    h->steps += 1;
  }
  if (store_last) h->fields_valid = true;
  return LBM_OK;
}

int lbm_get_fields(lbm_handle* h, double* rho, double* u_hydro, double* alpha,
                   double* force) {
  if (!h) return LBM_E_INVALID;
  if (!h->fields_valid)
    return fail(h, LBM_E_STATE, "no storing step has run (lbm_step store_last)");
  const size_t n = (size_t)h->N;
  struct {
    double* dst;
    size_t off, count;
  } parts[4] = {{rho, 0, n},
                {u_hydro, n, n * h->D},
                {alpha, n * (1 + h->D), n},
                {force, n * (2 + h->D), n * h->D}};
  for (auto& p : parts)
    if (p.dst)
      HIP_TRY(h, hipMemcpyAsync(p.dst, h->fields + p.off, p.count * sizeof(double),
                                hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return LBM_OK;
}

int lbm_reduce(lbm_handle* h, double* mass, double* energy, double* enstrophy) {
  if (!h) return LBM_E_INVALID;
  if (!h->fields_valid)
    return fail(h, LBM_E_STATE, "no storing step has run (lbm_step store_last)");
  if (!h->reduce_buf)
    HIP_TRY(h, hipMalloc(&h->reduce_buf, (3 * (size_t)reduce_blocks(h->N) + 3) * sizeof(double)));
  double* out_dev = h->reduce_buf + 3 * (size_t)reduce_blocks(h->N);
  HIP_TRY(h, launch_reduce(h->fields, h->fields + h->N, h->cfg.nx, h->cfg.ny,
                           h->cfg.nz, h->D, h->reduce_buf, out_dev, h->stream));
  double out[3];
  HIP_TRY(h, hipMemcpyAsync(out, out_dev, sizeof(out), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  if (mass) *mass = out[0];
  if (energy) *energy = out[1] / (double)h->N;
  if (enstrophy) *enstrophy = out[2] / (double)h->N;
  return LBM_OK;
}

int lbm_get_elbm_stats(lbm_handle* h, lbm_elbm_stats* out) {
  if (!h || !out) return LBM_E_INVALID;
  std::memset(out, 0, sizeof(*out));
  if (!h->stats) return LBM_OK;
  unsigned long long s[6];
  HIP_TRY(h, hipMemcpyAsync(s, h->stats, sizeof(s), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  out->small = (int64_t)s[0];
  out->capped = (int64_t)s[1];
  out->newton = (int64_t)s[2];
  out->fallback = (int64_t)s[3];
  out->newton_iterations = (int64_t)s[4];
  out->learned_accepted = (int64_t)s[5];
  return LBM_OK;
}

int lbm_reset_elbm_stats(lbm_handle* h) {
  if (!h) return LBM_E_INVALID;
  if (h->stats)
    HIP_TRY(h, hipMemsetAsync(h->stats, 0, 6 * sizeof(unsigned long long), h->stream));
  return LBM_OK;
}

int lbm_set_stream(lbm_handle* h, void* stream) {
  if (!h) return LBM_E_INVALID;
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  if (h->own_stream) (void)hipStreamDestroy(h->stream);
  h->own_stream = false;
  h->stream = static_cast<hipStream_t>(stream);
  return LBM_OK;
}

int lbm_synchronize(lbm_handle* h) {
  if (!h) return LBM_E_INVALID;
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return LBM_OK;
}

int lbm_device_populations(lbm_handle* h, void** ptr, int* elem_bytes) {
  if (!h) return LBM_E_INVALID;
  if (ptr) *ptr = h->f[h->cur];
  if (elem_bytes) *elem_bytes = (int)h->elem;
  return LBM_OK;
}

int64_t lbm_steps_done(const lbm_handle* h) { return h ? h->steps : -1; }

}  // extern "C"
