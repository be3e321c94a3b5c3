// capi_internal.hpp -- what the C ABI's translation units share besides the
// public header: error recording and read access to a handle's device state.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/metalbm_amd.h"

namespace lbm {

// Records msg as the handle's last error and returns code.
int record_error(lbm_handle* h, int code, const std::string& msg);
// LBM_E_HIP with the HIP error text, unless e is hipSuccess (then LBM_OK).
int record_hip(lbm_handle* h, hipError_t e, const char* what);

// Datasets the field file writes besides density and velocity.
void field_output(const lbm_handle* h, int* alpha, int* force, int* vorticity);

// Sets the entropic alpha field to 2 (initAlpha, Initialize.h:81-88); no-op
// for BGK.
int reset_alpha(lbm_handle* h);

// Device view of the fields of the last storing step.
struct DeviceFields {
  const double* rho;    // [N]
  const double* u;      // [D][N] hydrodynamic velocity
  const double* force;  // [D][N]
  int D, n[3];          // extents; n[2] = 1 in 2D
  int64_t N;
  hipStream_t stream;   // the handle's stream
  uint64_t version;     // changes whenever the stored fields are rewritten
};
// LBM_E_STATE (with message) when no storing step has run.
int device_fields(lbm_handle* h, DeviceFields& out);

// Per-handle state of the spectral diagnostics (owned by spectral.hip,
// released with the handle).
void*& spectral_slot(lbm_handle* h);
void release_spectral(void* state);

// The force field of the spectral forces and what its update reads.
struct ForceField {
  double* f64;          // [D][N] force field, FP64
  float* f32;           // its FP32 copy for FP32 compute, else NULL
  const double* rho;    // [N] density of the last storing step / initialiser
  const double* u;      // [D][N] velocity of the same
  int D, n[3];
  int64_t N;
  hipStream_t stream;
  lbm_config cfg;
};
int force_field(lbm_handle* h, ForceField& out);
// Recomputes the force field (ConstantShell, EnergyRemoval, Turbulent2D);
// defined in spectral.hip.
int update_spectral_force(lbm_handle* h);

}  // namespace lbm
