// capi.hip -- the C ABI (include/metalbm_amd.h) over the HIP kernels.
//
// A handle owns one periodic domain on the current device: two population
// lattices (ping-pong, Algorithm.h:336), the ELBM alpha field, host-built
// force tables, and lazily allocated output fields and staging buffers.
// Errors are returned as codes with a message; nothing exits.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/metalbm_amd.h"
#include "aux_kernels.hpp"
#include "capi_internal.hpp"
#include "config_rules.hpp"
#include "mlp_train.hpp"
#include "probe.hpp"
#include "registry.hpp"

using namespace lbm;

namespace lbm {
const Registry& registry() {
  static const Registry r = [] {
    Registry x;
    register_step_f64_d2q9(x);
    register_step_f64_d3q19(x);
    register_step_f64_d3q27(x);
    register_step_f32_d2q9(x);
    register_step_f32_d3q19(x);
    register_step_f32_d3q27(x);
    return x;
  }();
  return r;
}
}  // namespace lbm

struct lbm_handle {
  lbm_config cfg{};
  int device = 0;           // the HIP device current at lbm_create
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
  uint64_t fields_version = 0;  // bumped whenever the fields are rewritten
  double* staging = nullptr;   // N doubles
  double* reduce_buf = nullptr;
  void* ftab[3] = {nullptr, nullptr, nullptr};
  int fax[3] = {3, 3, 3};
  unsigned long long* stats = nullptr;  // kStatSlots x kStatStride counters
  hipStream_t stream = nullptr;
  bool own_stream = false;
  StepFn step = nullptr;
  int64_t steps = 0;
  double w[27], scale[27];
  std::string err;
  // entropic solver
  int alpha_mode = LBM_ALPHA_NEWTON;
  double residual_tol = 1e-10;
  // learned alpha: parameter block (params | mean | scale | out_mean, out_scale)
  int mlp_nin = 0, mlp_h1 = 0, mlp_h2 = 0, mlp_np = 0, mlp_nb = 0;
  float* mlp = nullptr;
  float* mlp_m = nullptr;   // Adam first moments
  float* mlp_v = nullptr;   // Adam second moments
  void* mlp_packed = nullptr;  // inference image of mlp (see mlp_packed_bytes)
  bool mlp_dirty = true;    // mlp changed since the last pack
  int adam_t = 0;
  bool mlp_ready = false;   // normalisation set (trained or loaded)
  bool mlp_has_params = false;  // initialised or loaded (else all zero)
  // (features, alpha) sample log
  float* log_x = nullptr;
  double* log_alpha = nullptr;
  unsigned long long* log_count = nullptr;
  long long log_cap = 0;
  unsigned log_threshold = 0, log_salt = 0;
  bool log_on = false;
  void* spectral = nullptr;  // spectral diagnostics state (spectral.hip)
  // spectral forces: the force field [D][N] in FP64 and, for FP32 compute,
  // its FP32 copy; the step kernels read it through ftab (fax = 4)
  double* ffield = nullptr;
  float* ffield32 = nullptr;
  bool ffield_built = false;
  uint64_t ffield_version = 0;  // fields_version the force field was built from
  // alpha modes 1 and 2: count and list of the cells the correction defers
  unsigned* defer_count = nullptr;
  int* defer_list = nullptr;
  // datasets of the field file besides density and velocity
  // (writeAlpha, writeForce, writeVorticity; all on in Input_prod.in:68-72)
  int out_alpha = 1, out_force = 1, out_vorticity = 1;
};

namespace {

thread_local std::string g_create_error;

int fail(lbm_handle* h, int code, const std::string& msg) {
  if (h) h->err = msg;
  else g_create_error = msg;
  return code;
}

int hip_fail(lbm_handle* h, hipError_t e, const char* what) {
  // The error is reported through the return code; take it out of HIP's
  // per-thread last-error state too, so that the launch checks of later calls
  // (which read hipGetLastError) report only their own launches.
  (void)hipGetLastError();
  return fail(h, LBM_E_HIP,
              std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(h, expr)                                   \
  do {                                                     \
    hipError_t e_ = (expr);                                \
    if (e_ != hipSuccess) return hip_fail((h), e_, #expr); \
  } while (0)

int src_of(int fs) {
  if (fs == LBM_FS_GUO) return SRC_GUO;
  if (fs == LBM_FS_EDM) return SRC_EDM;
  return SRC_ZERO;
}


// The step kernel for a validated configuration, and in *vec the vector
// kernel's cells per lane V (0: one cell per lane).  Host-only.
StepFn choose_kernel(const lbm_config* cfg, int* vec) {
  const int ctype = cfg->compute == LBM_F64 ? TYPE_F64 : TYPE_F32;
  const KernelKey key{cfg->lattice, cfg->collision, src_of(cfg->forcing_scheme),
                      cfg->equilibrium, cfg->storage, ctype, cfg->shifted ? 1 : 0};
  *vec = pick_vector_width(cfg, [&](int V) {
    KernelKey kv = key;
    kv.vec = V;
    return registry().find(kv) != nullptr;
  });
  KernelKey kv = key;
  kv.vec = *vec;
  return registry().find(kv);
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
    case LBM_FORCE_CONSTANT_SHELL:
    case LBM_FORCE_ENERGY_REMOVAL:
    case LBM_FORCE_TURBULENT_2D: {
      // a per-cell field (Force<Generic>::setForce, Force.h:39-48), filled
      // by the spectral-force update before the first step
      const size_t n = (size_t)h->D * h->N;
      HIP_TRY(h, hipMalloc(&h->ffield, n * sizeof(double)));
      HIP_TRY(h, hipMemset(h->ffield, 0, n * sizeof(double)));
      if (h->ctype != TYPE_F64) {
        HIP_TRY(h, hipMalloc(&h->ffield32, n * sizeof(float)));
        HIP_TRY(h, hipMemset(h->ffield32, 0, n * sizeof(float)));
      }
      for (int d = 0; d < h->D; ++d) {
        h->fax[d] = 4;
        h->ftab[d] = h->ctype == TYPE_F64 ? (void*)(h->ffield + (size_t)d * h->N)
                                           : (void*)(h->ffield32 + (size_t)d * h->N);
      }
      // EnergyRemoval reads rho and u of the stored fields (the initial
      // fields until a storing step): zero until an initialiser writes them
      HIP_TRY(h, hipMalloc(&h->fields, (size_t)h->N * (2 + 2 * h->D) * sizeof(double)));
      HIP_TRY(h, hipMemset(h->fields, 0, (size_t)h->N * (2 + 2 * h->D) * sizeof(double)));
      return LBM_OK;
    }
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

// SplitMix64 (Steele, Lea, Flood 2014): output k of the stream seeded with
// `seed`; the seeded draws of the network initialisation, the minibatch
// shuffle and the spectral initialiser.
uint64_t mix64(uint64_t seed, uint64_t k) {
  uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

int pack_mlp(lbm_handle* h);  // learned-alpha section below

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

// State every initialiser starts from.  The entropic alpha field restarts at
// 2 (initAlpha, Initialize.h:81-88), so a re-initialised handle matches a
// fresh one.  The reference's field list holds the initial density and
// velocity until the first storing step (Initialize.h:20-70,
// Algorithm.h:150-160); the EnergyRemoval and Turbulent2D forces read them
// (Force.h:416-431).  drho may be NULL for a uniform rho0; du is the device
// velocity [D][N].
hipError_t seed_state(lbm_handle* h, const double* drho, double rho0, const double* du) {
  if (h->alpha) {
    const hipError_t e = launch_fill(h->ctype, h->alpha, 2.0, h->N, h->stream);
    if (e != hipSuccess) return e;
  }
  if (h->cfg.force != LBM_FORCE_ENERGY_REMOVAL && h->cfg.force != LBM_FORCE_TURBULENT_2D)
    return hipSuccess;
  ++h->fields_version;
  hipError_t e = drho ? hipMemcpyAsync(h->fields, drho, (size_t)h->N * sizeof(double),
                                       hipMemcpyDeviceToDevice, h->stream)
                      : launch_fill(TYPE_F64, h->fields, rho0, h->N, h->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(h->fields + h->N, du, (size_t)h->N * h->D * sizeof(double),
                       hipMemcpyDeviceToDevice, h->stream);
  return e;
}

void release(lbm_handle* h) {
  lbm::release_spectral(h->spectral);
  h->spectral = nullptr;
  if (h->ffield) {  // ftab points into the force field
    for (void*& t : h->ftab) t = nullptr;
    (void)hipFree(h->ffield);
    if (h->ffield32) (void)hipFree(h->ffield32);
  }
  for (void* p : {h->f[0], h->f[1], h->alpha, (void*)h->fields,
                  (void*)h->staging, (void*)h->reduce_buf, h->ftab[0],
                  h->ftab[1], h->ftab[2], (void*)h->stats, (void*)h->mlp, h->mlp_packed,
                  (void*)h->mlp_m, (void*)h->mlp_v, (void*)h->defer_count,
                  (void*)h->defer_list, (void*)h->log_x,
                  (void*)h->log_alpha, (void*)h->log_count})
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
  // Non-temporal loads in the BGK cell kernel where one workgroup covers a
  // whole row and the two lattices far exceed the caches (>= 1 GiB).
  a.nt_loads = (h->tiles_per_row == 1 &&
                2.0 * h->Q * (double)h->N * (double)h->elem >= 1073741824.0) ? 1 : 0;
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
  a.defer_count = h->defer_count;
  a.defer_list = h->defer_list;
  a.alpha_mode = h->alpha_mode;
  a.residual_tol = h->residual_tol;
  a.mlp_packed = h->mlp_packed;
  if (h->log_on) {
    a.log_x = h->log_x;
    a.log_alpha = h->log_alpha;
    a.log_count = h->log_count;
    a.log_cap = h->log_cap;
    a.log_threshold = h->log_threshold;
  }
  return a;
}

}  // namespace

int lbm::record_error(lbm_handle* h, int code, const std::string& msg) {
  return fail(h, code, msg);
}

int lbm::record_hip(lbm_handle* h, hipError_t e, const char* what) {
  return e == hipSuccess ? LBM_OK : hip_fail(h, e, what);
}

int lbm::reset_alpha(lbm_handle* h) {
  if (!h->alpha) return LBM_OK;
  HIP_TRY(h, launch_fill(h->ctype, h->alpha, 2.0, h->N, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return LBM_OK;
}

void lbm::field_output(const lbm_handle* h, int* alpha, int* force, int* vorticity) {
  *alpha = h->out_alpha;
  *force = h->out_force;
  *vorticity = h->out_vorticity;
}

int lbm::device_fields(lbm_handle* h, DeviceFields& out) {
  if (!h->fields_valid)
    return fail(h, LBM_E_STATE, "no storing step has run (lbm_step store_last)");
  const size_t n = (size_t)h->N;
  out.rho = h->fields;
  out.u = h->fields + n;
  out.force = h->fields + n * (2 + h->D);
  out.D = h->D;
  out.n[0] = h->cfg.nx;
  out.n[1] = h->cfg.ny;
  out.n[2] = h->cfg.nz;
  out.N = h->N;
  out.stream = h->stream;
  out.version = h->fields_version;
  return LBM_OK;
}

void*& lbm::spectral_slot(lbm_handle* h) { return h->spectral; }

int lbm::force_field(lbm_handle* h, ForceField& out) {
  if (!h->ffield) return fail(h, LBM_E_STATE, "the configured force has no force field");
  out.f64 = h->ffield;
  out.f32 = h->ffield32;
  out.rho = h->fields;
  out.u = h->fields + h->N;
  out.D = h->D;
  out.n[0] = h->cfg.nx;
  out.n[1] = h->cfg.ny;
  out.n[2] = h->cfg.nz;
  out.N = h->N;
  out.stream = h->stream;
  out.cfg = h->cfg;
  return LBM_OK;
}

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
  c->force_kmin = 1;  // Input_prod.in:52-58
  c->force_kmax = 2;
  for (int d = 0; d < 3; ++d) c->removal_amplitude[d] = 1e-6;
  c->removal_kmin = 1;
  c->removal_kmax = 2;
  return LBM_OK;
}

const char* lbm_last_error(const lbm_handle* h) {
  return h ? h->err.c_str() : g_create_error.c_str();
}

int lbm_get_config(const lbm_handle* h, lbm_config* out) {
  if (!h || !out) return LBM_E_INVALID;
  *out = h->cfg;
  out->alpha_mode = h->alpha_mode;
  out->residual_tol = h->residual_tol;
  out->kernel = h->vec ? LBM_KERNEL_VECTOR : LBM_KERNEL_CELL;
  return LBM_OK;
}

int lbm_set_field_output(lbm_handle* h, int write_alpha, int write_force, int write_vorticity) {
  if (!h) return LBM_E_INVALID;
  h->out_alpha = write_alpha ? 1 : 0;
  h->out_force = write_force ? 1 : 0;
  h->out_vorticity = write_vorticity ? 1 : 0;
  return LBM_OK;
}

int lbm_kernel_shape(const lbm_config* cfg, int* shape, int* vec) {
  if (!cfg) return fail(nullptr, LBM_E_INVALID, "null argument");
  std::string why;
  if (validate_config(cfg, why) != LBM_OK) return fail(nullptr, LBM_E_INVALID, why);
  int V = 0;
  if (!choose_kernel(cfg, &V))
    return fail(nullptr, LBM_E_INVALID,
                "no kernel instantiated for this storage/compute combination");
  if (cfg->kernel == LBM_KERNEL_VECTOR && !V)
    return fail(nullptr, LBM_E_INVALID,
                "the vector kernel needs nz (3D) or ny (2D) divisible by V "
                "(V = 2 for F64, 2 or 4 for F32, 2, 4 or 8 for F16)");
  if (shape) *shape = V ? LBM_KERNEL_VECTOR : LBM_KERNEL_CELL;
  if (vec) *vec = V;
  return LBM_OK;
}

int lbm_create(const lbm_config* cfg, lbm_handle** out) {
  if (!cfg || !out) return fail(nullptr, LBM_E_INVALID, "null argument");
  *out = nullptr;
  std::string why;
  if (validate_config(cfg, why) != LBM_OK) return fail(nullptr, LBM_E_INVALID, why);
  lbm_handle* h = new lbm_handle();
  h->cfg = *cfg;
  if (hipGetDevice(&h->device) != hipSuccess) {
    (void)hipGetLastError();
    delete h;
    return fail(nullptr, LBM_E_HIP, "hipGetDevice failed: no usable HIP device");
  }
  h->alpha_mode = cfg->collision == LBM_ELBM ? cfg->alpha_mode : LBM_ALPHA_NEWTON;
  h->residual_tol = cfg->residual_tol > 0.0 ? cfg->residual_tol : 1e-10;
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
  int V = 0;
  h->step = choose_kernel(cfg, &V);
  if (V) {
    h->vec = V;
    h->block = 256;
    h->tiles_per_row = (h->nk + 64 * V - 1) / (64 * V);
  }
  auto bail = [&](int code) {
    (void)hipGetLastError();  // reported here; see hip_fail
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
    h->err = "the vector kernel needs nz (3D) or ny (2D) divisible by V "
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
    const size_t sb = (size_t)kStatSlots * kStatStride * sizeof(unsigned long long);
    if (e == hipSuccess) e = hipMalloc(&h->stats, sb);
    if (e == hipSuccess) e = hipMemsetAsync(h->stats, 0, sb, h->stream);
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
  HIP_TRY(h, seed_state(h, drho, 0.0, du));
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
  HIP_TRY(h, seed_state(h, nullptr, rho0, du));
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
  if (e == hipSuccess) e = seed_state(h, nullptr, rho0, du);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  (void)hipFree(tab);
  if (e != hipSuccess) return hip_fail(h, e, "Taylor-Green initialisation");
  return LBM_OK;
}

namespace {

int shell_of(int k2) {
  int s = (int)std::sqrt((double)k2);
  while ((2 * s + 1) * (2 * s + 1) <= 4 * k2) ++s;
  while (s > 0 && (2 * s - 1) * (2 * s - 1) > 4 * k2) --s;
  return s;
}

void cross3(const double* a, const double* b, double* r) {
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}

// Hermitian mode cube of the spectral initialiser (specification in
// lbm_init_spectral's declaration).  Returns gamma.
double spectral_modes(int D, int K, double urms, double kpeak, uint64_t seed,
                      std::vector<double>& ur, std::vector<double>& ui) {
  const int M = 2 * K + 1, Kz = D == 3 ? K : 0, Mz = 2 * Kz + 1;
  const size_t nm = (size_t)M * M * Mz;
  auto active_half = [&](int kx, int ky, int kz, int& s) {
    s = shell_of(kx * kx + ky * ky + kz * kz);
    const bool half = kx > 0 || (kx == 0 && (ky > 0 || (ky == 0 && kz > 0)));
    return std::make_pair(s >= 1 && s <= K, half);
  };
  std::vector<int> cnt(K + 1, 0);
  size_t nh = 0;
  for (int ix = 0; ix < M; ++ix)
    for (int iy = 0; iy < M; ++iy)
      for (int iz = 0; iz < Mz; ++iz) {
        int s;
        auto [act, half] = active_half(ix - K, iy - K, iz - Kz, s);
        if (act) cnt[s]++;
        if (act && half) nh++;
      }
  std::vector<double> amp(K + 1, 0.0);
  double esum = 0.0;
  for (int s = 1; s <= K; ++s) {
    const double r = s / kpeak;
    const double e = (double)s * s * s * s * std::exp(-2.0 * r * r);
    esum += e;
    amp[s] = cnt[s] > 0 ? std::sqrt(e / cnt[s]) : 0.0;
  }
  // Angles, then their cosines and sines in separate loops, so that no
  // compiler fuses a sin/cos pair into one sincos call (rounded differently).
  std::vector<double> ang(3 * nh), cs(3 * nh), sn(3 * nh);
  for (size_t j = 0; j < 3 * nh; ++j) {
    ang[j] = 2.0 * M_PI * ((double)(mix64(seed, j) >> 11) * 0x1.0p-53);
  }
  for (size_t j = 0; j < 3 * nh; ++j) cs[j] = std::cos(ang[j]);
  for (size_t j = 0; j < 3 * nh; ++j) sn[j] = std::sin(ang[j]);
  ur.assign(nm * D, 0.0);
  ui.assign(nm * D, 0.0);
  size_t m = 0;
  for (int ix = 0; ix < M; ++ix)
    for (int iy = 0; iy < M; ++iy)
      for (int iz = 0; iz < Mz; ++iz) {
        const int kx = ix - K, ky = iy - K, kz = iz - Kz;
        int s;
        auto [act, half] = active_half(kx, ky, kz, s);
        if (!(act && half)) continue;
        const double c1 = cs[3 * m], s1 = sn[3 * m];
        const double c2 = cs[3 * m + 1], s2 = sn[3 * m + 1];
        const double cp = cs[3 * m + 2], sp = sn[3 * m + 2];
        ++m;
        double re[3] = {0, 0, 0}, im[3] = {0, 0, 0};
        if (D == 3) {
          const double k[3] = {(double)kx, (double)ky, (double)kz};
          int a = 0;
          if (std::fabs(k[1]) < std::fabs(k[a])) a = 1;
          if (std::fabs(k[2]) < std::fabs(k[a])) a = 2;
          double ea[3] = {0, 0, 0}, e1[3], e2[3], w[3];
          ea[a] = 1.0;
          cross3(k, ea, w);
          double nw = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
          for (int d = 0; d < 3; ++d) e1[d] = w[d] / nw;
          cross3(k, e1, w);
          nw = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
          for (int d = 0; d < 3; ++d) e2[d] = w[d] / nw;
          for (int d = 0; d < 3; ++d) {
            re[d] = amp[s] * (cp * c1 * e1[d] + sp * c2 * e2[d]);
            im[d] = amp[s] * (cp * s1 * e1[d] + sp * s2 * e2[d]);
          }
        } else {
          const double nk = std::sqrt((double)(kx * kx + ky * ky));
          const double e1[2] = {-ky / nk, kx / nk};
          for (int d = 0; d < 2; ++d) {
            re[d] = amp[s] * (c1 * e1[d]);
            im[d] = amp[s] * (s1 * e1[d]);
          }
        }
        const size_t p = ((size_t)ix * M + iy) * Mz + iz;
        const size_t q = ((size_t)(M - 1 - ix) * M + (M - 1 - iy)) * Mz + (Mz - 1 - iz);
        for (int d = 0; d < D; ++d) {
          ur[p * D + d] = re[d];
          ui[p * D + d] = im[d];
          ur[q * D + d] = re[d];
          ui[q * D + d] = -im[d];
        }
      }
  return urms * std::sqrt((double)D / esum);
}

}  // namespace

int lbm_init_spectral(lbm_handle* h, double rho0, double urms, int kmax,
                      double kpeak, uint64_t seed) {
  if (!h) return LBM_E_INVALID;
  const lbm_config& c = h->cfg;
  const int n[3] = {c.nx, c.ny, c.nz};
  if (kmax < 1) return fail(h, LBM_E_INVALID, "spectral init: kmax must be >= 1");
  for (int d = 0; d < h->D; ++d)
    if (2 * kmax >= n[d])
      return fail(h, LBM_E_INVALID, "spectral init: 2 kmax must be below every extent");
  if (!(kpeak > 0.0)) return fail(h, LBM_E_INVALID, "spectral init: kpeak must be > 0");
  if (int rc = ensure_staging(h)) return rc;
  std::vector<double> ur, ui;
  const double gamma = spectral_modes(h->D, kmax, urms, kpeak, seed, ur, ui);
  std::vector<double> tab;
  const double two_pi = 2.0 * M_PI;
  for (int d = 0; d < 3; ++d) {  // cos table, then sin table, per axis
    for (int i = 0; i < n[d]; ++i) tab.push_back(std::cos(two_pi * i / n[d]));
    for (int i = 0; i < n[d]; ++i) tab.push_back(std::sin(two_pi * i / n[d]));
  }
  const int M = 2 * kmax + 1;
  const size_t nt1 = (size_t)M * M * c.nz * h->D, nt2 = (size_t)M * c.ny * c.nz * h->D;
  const size_t total = 2 * ur.size() + tab.size() + 2 * nt1 + 2 * nt2;
  double* buf = nullptr;
  HIP_TRY(h, hipMalloc(&buf, total * sizeof(double)));
  double* dur = buf;
  double* dui = dur + ur.size();
  double* dtab = dui + ui.size();
  double* t1r = dtab + tab.size();
  double* t1i = t1r + nt1;
  double* t2r = t1i + nt1;
  double* t2i = t2r + nt2;
  hipError_t e = hipMemcpy(dur, ur.data(), ur.size() * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(dui, ui.data(), ui.size() * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(dtab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice);
  double* du = h->staging + h->N;
  if (e == hipSuccess)
    e = launch_spectral(dur, dui, kmax, h->D, c.nx, c.ny, c.nz, dtab, gamma, t1r, t1i,
                        t2r, t2i, du, h->stream);
  if (e == hipSuccess)
    e = launch_init_equilibrium(c.lattice, c.equilibrium, c.storage, c.shifted,
                                h->f[h->cur], nullptr, rho0, du, h->N, h->stream);
  if (e == hipSuccess) e = seed_state(h, nullptr, rho0, du);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  (void)hipFree(buf);
  if (e != hipSuccess) return hip_fail(h, e, "spectral initialisation");
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
    // the fields a restart derives from the populations (Initialize.h:124-145),
    // read by EnergyRemoval until the first storing step
    if (h->cfg.force == LBM_FORCE_ENERGY_REMOVAL || h->cfg.force == LBM_FORCE_TURBULENT_2D) {
      if (i == 0) ++h->fields_version;
      HIP_TRY(h, launch_restart_fields(h->cfg.lattice, h->fields, h->fields + h->N, h->staging,
                                       i, h->N, h->stream));
    }
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
  if (h->cfg.collision == LBM_ELBM && h->alpha_mode == LBM_ALPHA_LEARNED && !h->mlp_ready)
    return fail(h, LBM_E_STATE,
                "learned alpha needs a network: lbm_mlp_train or lbm_mlp_set_params + "
                "lbm_mlp_set_normalization");
  if (h->cfg.collision == LBM_ELBM && h->alpha_mode == LBM_ALPHA_LEARNED)
    if (int rc = pack_mlp(h)) return rc;
  if (h->cfg.collision == LBM_ELBM &&
      (h->alpha_mode == LBM_ALPHA_LEARNED || h->alpha_mode == LBM_ALPHA_WARM_RESIDUAL)) {
    // each buffer on its own, so a failed second allocation leaks nothing
    // when the next call retries
    if (!h->defer_count) HIP_TRY(h, hipMalloc(&h->defer_count, sizeof(unsigned)));
    if (!h->defer_list) HIP_TRY(h, hipMalloc(&h->defer_list, (size_t)h->N * sizeof(int)));
  }
  StepArgs a = make_args(h);
  // Vector kernel: one wavefront per 64 V-cell segment, kVecWaves per workgroup.
  const int grid = h->vec ? (h->rows * h->tiles_per_row + kVecWaves - 1) / kVecWaves
                          : h->rows * h->tiles_per_row;
  const bool field_force = h->ffield != nullptr;
  const bool time_dependent = h->cfg.force == LBM_FORCE_ENERGY_REMOVAL ||
                              h->cfg.force == LBM_FORCE_TURBULENT_2D;
  for (int s = 0; s < nsteps; ++s) {
    // Collision::update before the step (Algorithm.h:338, Collision.h:96-99);
    // ConstantShell is set once (Collision.h:51-54).  EnergyRemoval and
    // Turbulent2D read only the fields of the last storing step, which change
    // once per storing step; the update is deterministic, so recomputing it
    // from unchanged fields would rewrite the same bits and is skipped.
    if (field_force && (!h->ffield_built ||
                        (time_dependent && h->ffield_version != h->fields_version))) {
      if (int rc = lbm::update_spectral_force(h)) return rc;
      h->ffield_built = true;
      h->ffield_version = h->fields_version;
    }
    a.src = h->f[h->cur];
    a.dst = h->f[1 - h->cur];
    a.store = (store_last && s == nsteps - 1) ? 1 : 0;
    a.log_salt = h->log_salt + (unsigned)h->steps * 0x9E3779B9u;
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
  if (store_last) {
    h->fields_valid = true;
    ++h->fields_version;
  }
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

}  // extern "C"

// ---- reduction without allocation --------------------------------------------
namespace {

// Partial sums of lbm_reduce_noalloc: 3 per reduction block for up to
// kNoAllocBlocks blocks of 8192 cells (grids of up to 2^29 cells), then the
// 3 results -- 1.5 MiB.  Every GPU the library runs on holds its own copy
// until the process ends.  In host code `scratch` only names the symbol: its
// device address comes from hipGetSymbolAddress, resolved once per device.
constexpr int kNoAllocBlocks = 65536;
static __device__ __align__(256) double scratch[3 * kNoAllocBlocks + 3];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

// The device address of the handle's GPU's scratch array (resolved on the
// first call for that device, with the device current).
int scratch_of(lbm_handle* h, double** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  const int dev = h->device;
  if (dev < 0 || dev >= kMaxDevices)
    return fail(h, LBM_E_INVALID, "lbm_reduce_noalloc: device index out of range");
  if (!g_scratch[dev]) {
    int cur = -1;
    HIP_TRY(h, hipGetDevice(&cur));
    if (cur != dev) HIP_TRY(h, hipSetDevice(dev));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (e != hipSuccess) return hip_fail(h, e, "hipGetSymbolAddress(scratch)");
    if (!p) return fail(h, LBM_E_HIP, "hipGetSymbolAddress(scratch) returned NULL");
    g_scratch[dev] = static_cast<double*>(p);
  }
  *out = g_scratch[dev];
  return LBM_OK;
}

}  // namespace

extern "C" {

int lbm_reduce_noalloc(lbm_handle* h, double* mass, double* energy, double* enstrophy) {
  if (!h) return LBM_E_INVALID;
  const int nb = reduce_blocks(h->N);
  if (nb > kNoAllocBlocks)
    return fail(h, LBM_E_INVALID,
                "lbm_reduce_noalloc: the grid has more than 2^29 cells, beyond its static "
                "scratch of 65536 reduction blocks (use lbm_reduce)");
  if (!h->fields_valid)
    return fail(h, LBM_E_STATE, "no storing step has run (lbm_step store_last)");
  double* base = nullptr;
  if (int rc = scratch_of(h, &base)) return rc;
  // The partial sums are all rewritten by the first kernel before the second
  // reads them, so nothing left from an earlier call needs clearing.
  double* out_dev = base + 3 * (size_t)nb;
  HIP_TRY(h, launch_reduce(h->fields, h->fields + h->N, h->cfg.nx, h->cfg.ny,
                           h->cfg.nz, h->D, base, out_dev, h->stream));
  double out[3];
  HIP_TRY(h, hipMemcpyAsync(out, out_dev, sizeof(out), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  if (mass) *mass = out[0];
  if (energy) *energy = out[1] / (double)h->N;
  if (enstrophy) *enstrophy = out[2] / (double)h->N;
  return LBM_OK;
}

int lbm_bandwidth_probe(lbm_handle* h, int kind, int nsteps) {
  if (!h) return LBM_E_INVALID;
  if (kind < LBM_PROBE_PULL_COPY || kind > LBM_PROBE_COPY)
    return fail(h, LBM_E_INVALID, "probe kind must be PULL_COPY (0), READ (1), WRITE (2) or COPY (3)");
  if (nsteps < 0) return fail(h, LBM_E_INVALID, "nsteps must be >= 0");
  if ((size_t)h->nk * h->elem % 16 != 0)
    return fail(h, LBM_E_INVALID, "the probes need a contiguous extent of a multiple of 16 bytes");
  for (int s = 0; s < nsteps; ++s) {
    HIP_TRY(h, launch_bandwidth_probe(kind, h->cfg.lattice, h->f[h->cur], h->f[1 - h->cur],
                                      h->cfg.nx, h->cfg.ny, h->cfg.nz, (int)h->elem, h->stream));
    h->cur = 1 - h->cur;
  }
  return LBM_OK;
}

int lbm_get_elbm_stats(lbm_handle* h, lbm_elbm_stats* out) {
  if (!h || !out) return LBM_E_INVALID;
  std::memset(out, 0, sizeof(*out));
  if (!h->stats) return LBM_OK;
  std::vector<unsigned long long> slots((size_t)kStatSlots * kStatStride);
  HIP_TRY(h, hipMemcpyAsync(slots.data(), h->stats, slots.size() * sizeof(unsigned long long),
                            hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  unsigned long long s[6] = {0, 0, 0, 0, 0, 0};
  for (int k = 0; k < kStatSlots; ++k)
    for (int c = 0; c < 6; ++c) s[c] += slots[(size_t)k * kStatStride + c];
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
  // a fill kernel rather than a memset, so a caller may capture this call in
  // a graph; +0.0 is the all-zero bit pattern of the 64-bit counters
  if (h->stats)
    HIP_TRY(h, launch_fill(TYPE_F64, h->stats, 0.0, kStatSlots * kStatStride, h->stream));
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

// ---- learned alpha ----------------------------------------------------------

namespace {

int need_mlp_lattice(lbm_handle* h) {
  if (!h) return LBM_E_INVALID;
  if (h->cfg.collision != LBM_ELBM)
    return fail(h, LBM_E_STATE, "learned alpha belongs to the ELBM collision");
  if (h->mlp_nin == 0 &&
      mlp_shape(h->cfg.lattice, &h->mlp_nin, &h->mlp_h1, &h->mlp_h2, &h->mlp_np,
                &h->mlp_nb) != 0)
    return fail(h, LBM_E_INVALID, "learned alpha is defined for D2Q9 and D3Q27");
  return LBM_OK;
}

int ensure_mlp(lbm_handle* h) {
  if (int rc = need_mlp_lattice(h)) return rc;
  if (!h->mlp) {
    // all four buffers or none: the handle never holds a partial set
    float *p = nullptr, *m = nullptr, *v = nullptr;
    void* packed = nullptr;
    hipError_t e = hipMalloc(&p, sizeof(float) * h->mlp_nb);
    if (e == hipSuccess) e = hipMalloc(&m, sizeof(float) * h->mlp_np);
    if (e == hipSuccess) e = hipMalloc(&v, sizeof(float) * h->mlp_np);
    if (e == hipSuccess) e = hipMalloc(&packed, mlp_packed_bytes(h->cfg.lattice));
    if (e == hipSuccess) e = hipMemset(p, 0, sizeof(float) * h->mlp_nb);
    if (e == hipSuccess) e = hipMemset(m, 0, sizeof(float) * h->mlp_np);
    if (e == hipSuccess) e = hipMemset(v, 0, sizeof(float) * h->mlp_np);
    if (e != hipSuccess) {
      for (void* q : {(void*)p, (void*)m, (void*)v, packed}) (void)hipFree(q);
      return hip_fail(h, e, "network buffers");
    }
    h->mlp = p;
    h->mlp_m = m;
    h->mlp_v = v;
    h->mlp_packed = packed;
    h->mlp_dirty = true;
  }
  return LBM_OK;
}

// Refreshes the packed inference image after any change of the parameters.
int pack_mlp(lbm_handle* h) {
  if (int rc = ensure_mlp(h)) return rc;
  if (h->mlp_dirty) {
    HIP_TRY(h, launch_mlp_pack(h->cfg.lattice, h->mlp, h->mlp_packed, h->stream));
    h->mlp_dirty = false;
  }
  return LBM_OK;
}

double unit01(uint64_t seed, uint64_t k) { return (double)(mix64(seed, k) >> 11) * 0x1.0p-53; }

}  // namespace

int lbm_mlp_shape(const lbm_handle* h, int* nin, int* h1, int* h2, int* nparams) {
  if (!h) return LBM_E_INVALID;
  int a, b, c, d, e;
  if (mlp_shape(h->cfg.lattice, &a, &b, &c, &d, &e) != 0) return LBM_E_INVALID;
  if (nin) *nin = a;
  if (h1) *h1 = b;
  if (h2) *h2 = c;
  if (nparams) *nparams = d;
  return LBM_OK;
}

int lbm_mlp_init(lbm_handle* h, uint64_t seed) {
  if (int rc = ensure_mlp(h)) return rc;
  std::vector<float> p(h->mlp_nb, 0.0f);
  const int nin = h->mlp_nin, h1 = h->mlp_h1, h2 = h->mlp_h2;
  uint64_t k = 0;
  auto fill = [&](int off, int count, int fan_in) {
    const double lim = std::sqrt(6.0 / fan_in);   // He-uniform for ReLU
    for (int i = 0; i < count; ++i) p[off + i] = (float)(lim * (2.0 * unit01(seed, k++) - 1.0));
  };
  fill(0, h1 * nin, nin);
  fill(h1 * nin + h1, h2 * h1, h1);
  fill(h1 * nin + h1 + h2 * h1 + h2, h2, h2);
  for (int i = 0; i < nin; ++i) p[h->mlp_np + nin + i] = 1.0f;  // identity normalisation
  p[h->mlp_np + 2 * nin] = 2.0f;
  p[h->mlp_np + 2 * nin + 1] = 1.0f;
  HIP_TRY(h, hipMemcpy(h->mlp, p.data(), sizeof(float) * h->mlp_nb, hipMemcpyHostToDevice));
  h->mlp_dirty = true;
  HIP_TRY(h, hipMemset(h->mlp_m, 0, sizeof(float) * h->mlp_np));
  HIP_TRY(h, hipMemset(h->mlp_v, 0, sizeof(float) * h->mlp_np));
  h->adam_t = 0;
  h->mlp_has_params = true;
  return LBM_OK;
}

int lbm_mlp_set_params(lbm_handle* h, const float* params) {
  if (!params) return fail(h, LBM_E_INVALID, "null params");
  if (int rc = ensure_mlp(h)) return rc;
  HIP_TRY(h, hipMemcpy(h->mlp, params, sizeof(float) * h->mlp_np, hipMemcpyHostToDevice));
  h->mlp_dirty = true;
  h->mlp_has_params = true;
  return LBM_OK;
}

int lbm_mlp_get_params(lbm_handle* h, float* params) {
  if (!params) return fail(h, LBM_E_INVALID, "null params");
  if (int rc = ensure_mlp(h)) return rc;
  HIP_TRY(h, hipMemcpy(params, h->mlp, sizeof(float) * h->mlp_np, hipMemcpyDeviceToHost));
  return LBM_OK;
}

int lbm_mlp_set_normalization(lbm_handle* h, const float* in_mean, const float* in_scale,
                              double out_mean, double out_scale) {
  if (!in_mean || !in_scale) return fail(h, LBM_E_INVALID, "null normalisation");
  if (int rc = ensure_mlp(h)) return rc;
  std::vector<float> t(2 * h->mlp_nin + 2);
  for (int i = 0; i < h->mlp_nin; ++i) {
    t[i] = in_mean[i];
    t[h->mlp_nin + i] = in_scale[i];
  }
  t[2 * h->mlp_nin] = (float)out_mean;
  t[2 * h->mlp_nin + 1] = (float)out_scale;
  HIP_TRY(h, hipMemcpy(h->mlp + h->mlp_np, t.data(), sizeof(float) * t.size(),
                       hipMemcpyHostToDevice));
  h->mlp_dirty = true;
  h->mlp_ready = true;
  return LBM_OK;
}

int lbm_mlp_get_normalization(lbm_handle* h, float* in_mean, float* in_scale,
                              double* out_mean, double* out_scale) {
  if (int rc = ensure_mlp(h)) return rc;
  std::vector<float> t(2 * h->mlp_nin + 2);
  HIP_TRY(h, hipMemcpy(t.data(), h->mlp + h->mlp_np, sizeof(float) * t.size(),
                       hipMemcpyDeviceToHost));
  for (int i = 0; i < h->mlp_nin; ++i) {
    if (in_mean) in_mean[i] = t[i];
    if (in_scale) in_scale[i] = t[h->mlp_nin + i];
  }
  if (out_mean) *out_mean = t[2 * h->mlp_nin];
  if (out_scale) *out_scale = t[2 * h->mlp_nin + 1];
  return LBM_OK;
}

namespace {

// Device buffers for one gradient / training pass over n samples.
struct TrainBufs {
  float *X = nullptr, *T = nullptr, *partial = nullptr, *g = nullptr, *loss = nullptr;
  float* loss_partial = nullptr;
  int* perm = nullptr;
  ~TrainBufs() {
    for (void* p : {(void*)X, (void*)T, (void*)partial, (void*)g, (void*)loss,
                    (void*)loss_partial, (void*)perm})
      if (p) (void)hipFree(p);
  }
};

int targets_of(lbm_handle* h, const double* alpha, size_t n, std::vector<float>& t) {
  std::vector<float> norm(2 * h->mlp_nin + 2);
  HIP_TRY(h, hipMemcpy(norm.data(), h->mlp + h->mlp_np, sizeof(float) * norm.size(),
                       hipMemcpyDeviceToHost));
  const double om = norm[2 * h->mlp_nin], os = norm[2 * h->mlp_nin + 1];
  t.resize(n);
  for (size_t i = 0; i < n; ++i) t[i] = (float)((alpha[i] - om) / os);
  return LBM_OK;
}

}  // namespace

int lbm_mlp_gradient(lbm_handle* h, const float* x, const double* alpha, int n,
                     float* grad, double* loss) {
  if (!x || !alpha || !grad || n <= 0) return fail(h, LBM_E_INVALID, "bad samples");
  if (int rc = ensure_mlp(h)) return rc;
  std::vector<float> t;
  if (int rc = targets_of(h, alpha, (size_t)n, t)) return rc;
  TrainBufs b;
  const int np = h->mlp_np;
  const int nc = mlp_grad_chunks(n);
  HIP_TRY(h, hipMalloc(&b.X, sizeof(float) * (size_t)n * h->mlp_nin));
  HIP_TRY(h, hipMalloc(&b.T, sizeof(float) * (size_t)n));
  HIP_TRY(h, hipMalloc(&b.partial, sizeof(float) * (size_t)nc * np));
  HIP_TRY(h, hipMalloc(&b.loss_partial, sizeof(float) * (size_t)nc));
  HIP_TRY(h, hipMalloc(&b.g, sizeof(float) * (size_t)np));
  HIP_TRY(h, hipMalloc(&b.loss, sizeof(float)));
  HIP_TRY(h, hipMemcpy(b.X, x, sizeof(float) * (size_t)n * h->mlp_nin, hipMemcpyHostToDevice));
  HIP_TRY(h, hipMemcpy(b.T, t.data(), sizeof(float) * (size_t)n, hipMemcpyHostToDevice));
  HIP_TRY(h, launch_mlp_grad(h->cfg.lattice, h->mlp, b.X, b.T, nullptr, n, b.partial,
                             b.loss_partial, h->stream));
  HIP_TRY(h, launch_mlp_reduce_adam(b.partial, b.loss_partial, nc, np, b.g, b.loss, 0,
                                    nullptr, nullptr, nullptr, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f,
                                    h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  HIP_TRY(h, hipMemcpy(grad, b.g, sizeof(float) * np, hipMemcpyDeviceToHost));
  if (loss) {
    float l = 0.0f;
    HIP_TRY(h, hipMemcpy(&l, b.loss, sizeof(float), hipMemcpyDeviceToHost));
    *loss = (double)l / (2.0 * n);
  }
  return LBM_OK;
}

int lbm_mlp_predict(lbm_handle* h, const float* x, int n, double* alpha) {
  if (!x || !alpha || n <= 0) return fail(h, LBM_E_INVALID, "bad samples");
  if (int rc = ensure_mlp(h)) return rc;
  float* dx = nullptr;
  double* da = nullptr;
  hipError_t e = hipMalloc(&dx, sizeof(float) * (size_t)n * h->mlp_nin);
  if (e == hipSuccess) e = hipMalloc(&da, sizeof(double) * (size_t)n);
  if (e == hipSuccess)
    e = hipMemcpy(dx, x, sizeof(float) * (size_t)n * h->mlp_nin, hipMemcpyHostToDevice);
  if (e == hipSuccess && pack_mlp(h) != LBM_OK) e = hipErrorUnknown;
  if (e == hipSuccess)
    e = launch_mlp_forward(h->cfg.lattice, h->mlp_packed, dx, n, da, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess) e = hipMemcpy(alpha, da, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost);
  if (dx) (void)hipFree(dx);
  if (da) (void)hipFree(da);
  if (e != hipSuccess) return hip_fail(h, e, "mlp predict");
  return LBM_OK;
}

int lbm_sample_log_enable(lbm_handle* h, int64_t capacity, double fraction, uint32_t salt) {
  if (int rc = need_mlp_lattice(h)) return rc;
  if (capacity <= 0 || !(fraction > 0.0) || fraction > 1.0)
    return fail(h, LBM_E_INVALID, "capacity > 0 and 0 < fraction <= 1 required");
  // Logging stays off until every buffer of the new capacity exists, so a
  // failed allocation below leaves the steps logging nothing (make_args hands
  // the kernels log buffers only while log_on).
  h->log_on = false;
  const bool resize = h->log_cap != capacity;
  h->log_cap = 0;
  if (resize || !h->log_x || !h->log_alpha) {
    // also frees what a failed earlier attempt may have left of the pair
    (void)hipFree(h->log_x);
    (void)hipFree(h->log_alpha);
    h->log_x = nullptr;
    h->log_alpha = nullptr;
    HIP_TRY(h, hipMalloc(&h->log_x, sizeof(float) * (size_t)capacity * h->mlp_nin));
    HIP_TRY(h, hipMalloc(&h->log_alpha, sizeof(double) * (size_t)capacity));
  }
  if (!h->log_count) HIP_TRY(h, hipMalloc(&h->log_count, sizeof(unsigned long long)));
  HIP_TRY(h, launch_fill(TYPE_F64, h->log_count, 0.0, 1, h->stream));
  h->log_cap = capacity;
  h->log_threshold = fraction >= 1.0 ? 0xFFFFFFFFu : (unsigned)(fraction * 4294967296.0);
  h->log_salt = salt;
  h->log_on = true;
  return LBM_OK;
}

int lbm_sample_log_disable(lbm_handle* h) {
  if (!h) return LBM_E_INVALID;
  h->log_on = false;
  return LBM_OK;
}

int lbm_sample_log_count(lbm_handle* h, int64_t* n) {
  if (!h || !n) return LBM_E_INVALID;
  *n = 0;
  if (!h->log_count) return LBM_OK;
  unsigned long long c = 0;
  HIP_TRY(h, hipMemcpyAsync(&c, h->log_count, sizeof(c), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  *n = (int64_t)std::min<unsigned long long>(c, (unsigned long long)h->log_cap);
  return LBM_OK;
}

int lbm_sample_log_copy(lbm_handle* h, float* x, double* alpha, int64_t n) {
  int64_t have = 0;
  if (int rc = lbm_sample_log_count(h, &have)) return rc;
  if (n > have) return fail(h, LBM_E_INVALID, "more samples requested than logged");
  if (x) HIP_TRY(h, hipMemcpy(x, h->log_x, sizeof(float) * (size_t)n * h->mlp_nin,
                              hipMemcpyDeviceToHost));
  if (alpha) HIP_TRY(h, hipMemcpy(alpha, h->log_alpha, sizeof(double) * (size_t)n,
                                  hipMemcpyDeviceToHost));
  return LBM_OK;
}

int lbm_mlp_train(lbm_handle* h, int iterations, int batch, double lr, uint64_t seed,
                  double* loss_first, double* loss_last) {
  if (int rc = ensure_mlp(h)) return rc;
  int64_t n = 0;
  if (int rc = lbm_sample_log_count(h, &n)) return rc;
  if (n < 2 || batch <= 0 || iterations <= 0)
    return fail(h, LBM_E_STATE, "training needs logged samples (lbm_sample_log_enable + steps)");
  // A network that was never initialised or loaded is all zeros, where ReLU
  // units are dead: initialise it from the training seed.
  if (!h->mlp_has_params)
    if (int rc = lbm_mlp_init(h, seed)) return rc;
  const int nin = h->mlp_nin, np = h->mlp_np;
  // Normalisation from the samples: features standardised (scale floored at
  // 1e-3 of the largest feature spread so conserved, ~zero moments stay
  // quiet); alpha = 2 + std(alpha - 2) y.
  std::vector<float> X((size_t)n * nin);
  std::vector<double> al((size_t)n);
  HIP_TRY(h, hipMemcpy(X.data(), h->log_x, sizeof(float) * X.size(), hipMemcpyDeviceToHost));
  HIP_TRY(h, hipMemcpy(al.data(), h->log_alpha, sizeof(double) * al.size(), hipMemcpyDeviceToHost));
  {
    // Canonical sample order (by alpha, then features): the step kernel fills
    // the log in whatever order its wavefronts reach the slot counter, and
    // training (normalisation sums, seeded minibatches) must not depend on it.
    std::vector<int64_t> ord((size_t)n);
    for (int64_t i = 0; i < n; ++i) ord[i] = i;
    std::sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) {
      if (al[a] != al[b]) return al[a] < al[b];
      for (int k = 0; k < nin; ++k) {
        const float xa = X[(size_t)a * nin + k], xb = X[(size_t)b * nin + k];
        if (xa != xb) return xa < xb;
      }
      return false;
    });
    std::vector<float> Xs(X.size());
    std::vector<double> as(al.size());
    for (int64_t i = 0; i < n; ++i) {
      as[i] = al[ord[i]];
      std::copy(X.begin() + ord[i] * nin, X.begin() + (ord[i] + 1) * nin, Xs.begin() + i * nin);
    }
    X.swap(Xs);
    al.swap(as);
    HIP_TRY(h, hipMemcpy(h->log_x, X.data(), sizeof(float) * X.size(), hipMemcpyHostToDevice));
    HIP_TRY(h, hipMemcpy(h->log_alpha, al.data(), sizeof(double) * al.size(),
                         hipMemcpyHostToDevice));
  }
  std::vector<double> mean(nin, 0.0), sd(nin, 0.0);
  for (int64_t s = 0; s < n; ++s)
    for (int k = 0; k < nin; ++k) mean[k] += X[(size_t)s * nin + k];
  for (int k = 0; k < nin; ++k) mean[k] /= (double)n;
  for (int64_t s = 0; s < n; ++s)
    for (int k = 0; k < nin; ++k) {
      const double d = X[(size_t)s * nin + k] - mean[k];
      sd[k] += d * d;
    }
  double sdmax = 0.0;
  for (int k = 0; k < nin; ++k) {
    sd[k] = std::sqrt(sd[k] / (double)n);
    sdmax = std::max(sdmax, sd[k]);
  }
  std::vector<float> fm(nin), fs(nin);
  for (int k = 0; k < nin; ++k) {
    fm[k] = (float)mean[k];
    fs[k] = (float)(1.0 / std::max(sd[k], 1e-3 * sdmax + 1e-300));
  }
  double sa = 0.0;
  for (double v : al) sa += (v - 2.0) * (v - 2.0);
  const double out_scale = std::max(std::sqrt(sa / (double)n), 1e-12);
  if (int rc = lbm_mlp_set_normalization(h, fm.data(), fs.data(), 2.0, out_scale)) return rc;
  std::vector<float> t;
  if (int rc = targets_of(h, al.data(), (size_t)n, t)) return rc;
  const int B = (int)std::min<int64_t>(batch, n);
  const int nc = mlp_grad_chunks(B);
  TrainBufs b;
  HIP_TRY(h, hipMalloc(&b.T, sizeof(float) * (size_t)n));
  HIP_TRY(h, hipMalloc(&b.partial, sizeof(float) * (size_t)nc * np));
  HIP_TRY(h, hipMalloc(&b.loss_partial, sizeof(float) * (size_t)nc));
  HIP_TRY(h, hipMalloc(&b.g, sizeof(float) * (size_t)np));
  HIP_TRY(h, hipMalloc(&b.loss, sizeof(float)));
  HIP_TRY(h, hipMalloc(&b.perm, sizeof(int) * (size_t)n));
  HIP_TRY(h, hipMemcpy(b.T, t.data(), sizeof(float) * (size_t)n, hipMemcpyHostToDevice));
  // Seeded Fisher-Yates permutation, consumed B at a time; each new
  // permutation is copied to the device once and the minibatches index it.
  std::vector<int> perm((size_t)n);
  for (int64_t i = 0; i < n; ++i) perm[i] = (int)i;
  uint64_t draw = 0;
  auto shuffle = [&]() -> int {
    for (int64_t i = n - 1; i > 0; --i) {
      const int64_t j = (int64_t)(mix64(seed, draw++) % (uint64_t)(i + 1));
      std::swap(perm[i], perm[j]);
    }
    // the previous permutation may still be read by queued kernels
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    HIP_TRY(h, hipMemcpy(b.perm, perm.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice));
    return LBM_OK;
  };
  if (int rc = shuffle()) return rc;
  int64_t pos = 0;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int it = 0; it < iterations; ++it) {
    if (pos + B > n) {
      if (int rc = shuffle()) return rc;
      pos = 0;
    }
    HIP_TRY(h, launch_mlp_grad(h->cfg.lattice, h->mlp, h->log_x, b.T, b.perm + pos, B,
                               b.partial, b.loss_partial, h->stream));
    pos += B;
    h->adam_t += 1;
    const float c1 = (float)(1.0 - std::pow(0.9, h->adam_t));
    const float c2 = (float)(1.0 - std::pow(0.999, h->adam_t));
    h->mlp_dirty = true;
    HIP_TRY(h, launch_mlp_reduce_adam(b.partial, b.loss_partial, nc, np, b.g, b.loss, 1,
                                      h->mlp, h->mlp_m, h->mlp_v, (float)lr, b1, b2, eps, c1,
                                      c2, h->stream));
    if ((it == 0 && loss_first) || (it == iterations - 1 && loss_last)) {
      float l = 0.0f;
      HIP_TRY(h, hipMemcpyAsync(&l, b.loss, sizeof(float), hipMemcpyDeviceToHost, h->stream));
      HIP_TRY(h, hipStreamSynchronize(h->stream));
      if (it == 0 && loss_first) *loss_first = (double)l / (2.0 * B);
      if (it == iterations - 1 && loss_last) *loss_last = (double)l / (2.0 * B);
    }
  }
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return LBM_OK;
}

int lbm_set_alpha_mode(lbm_handle* h, int mode, double residual_tol) {
  if (!h) return LBM_E_INVALID;
  if (h->cfg.collision != LBM_ELBM) return fail(h, LBM_E_STATE, "alpha mode belongs to ELBM");
  if (mode != LBM_ALPHA_NEWTON && mode != LBM_ALPHA_LEARNED && mode != LBM_ALPHA_WARM_RESIDUAL &&
      mode != LBM_ALPHA_NEWTON_RESIDUAL)
    return fail(h, LBM_E_INVALID, "alpha_mode must be NEWTON (0), LEARNED (1), WARM_RESIDUAL (2) "
                                  "or NEWTON_RESIDUAL (3)");
  if (mode == LBM_ALPHA_LEARNED) {
    if (int rc = need_mlp_lattice(h)) return rc;
  }
  if (mode != LBM_ALPHA_NEWTON && !(residual_tol > 0.0))
    return fail(h, LBM_E_INVALID, "residual_tol must be positive");
  h->alpha_mode = mode;
  if (residual_tol > 0.0) h->residual_tol = residual_tol;
  return LBM_OK;
}

}  // extern "C"
