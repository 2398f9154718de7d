// Host launchers of the rank-predictor training kernels (mlp.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hcamd {

int64_t mlp_workspace_floats(int64_t n);
float *mlp_loss_partials(float *ws, int64_t n);
int64_t mlp_loss_blocks(int64_t n);
hipError_t launch_mlp_forward(hipStream_t st, const float *P, const float *X, int64_t n, float *ws);
hipError_t launch_mlp_backward(hipStream_t st, const float *P, const float *X, const float *y, int64_t n, float *ws,
                               float *G);
hipError_t launch_mlp_adam(hipStream_t st, float *P, const float *G, float *M, float *V, float lr, float b1, float b2,
                           float eps, int64_t step);
hipError_t launch_mlp_pack_bf16(hipStream_t st, const float *P, uint16_t *out);
hipError_t launch_mlp_predict_bf16(hipStream_t st, const uint16_t *P, const float *X, int64_t n, float *out);

}  // namespace hcamd
