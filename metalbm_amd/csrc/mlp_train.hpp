// mlp_train.hpp -- launchers of the alpha-predictor training kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace lbm {

// Network shape for a lattice (D2Q9: 9-32-32-1, D3Q27: 27-64-64-1);
// nblock = parameters + normalisation floats.  Returns -1 if unsupported.
int mlp_shape(int lattice, int* nin, int* h1, int* h2, int* nparam, int* nblock);

// Batch gradient of the samples idx[0..B) (idx may be NULL: samples 0..B):
// partial[c][p] for chunk c = 0..mlp_grad_chunks(B)-1 and the chunk's summed
// squared errors loss_partial[c]; T holds the normalised targets
// (alpha - out_mean) / out_scale.
int mlp_grad_chunks(int B);
hipError_t launch_mlp_grad(int lattice, const float* P, const float* X, const float* T,
                           const int* idx, int B, float* partial, float* loss_partial,
                           hipStream_t st);
// g[p] = sum_c partial[c][p] in chunk order and *loss = sum_c loss_partial[c];
// then, if adam, one Adam step on P (c1 = 1 - b1^t, c2 = 1 - b2^t).
hipError_t launch_mlp_reduce_adam(const float* partial, const float* loss_partial,
                                  int nchunks, int np, float* g, float* loss, int adam,
                                  float* P, float* m, float* v, float lr, float b1, float b2,
                                  float eps, float c1, float c2, hipStream_t st);
// Bytes of the packed network (the inference kernels' LDS image of the
// parameter block, weights in the matrix-core operand type).
size_t mlp_packed_bytes(int lattice);
// Packs the FP32 parameter block P into `packed` (mlp_packed_bytes bytes).
hipError_t launch_mlp_pack(int lattice, const float* P, void* packed, hipStream_t st);
// alpha_hat for raw feature rows X[n][nin] with the step kernel's inference
// code, from a packed network.
hipError_t launch_mlp_forward(int lattice, const void* packed, const float* X, int n,
                              double* alpha, hipStream_t st);

}  // namespace lbm
