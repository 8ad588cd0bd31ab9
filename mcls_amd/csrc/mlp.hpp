// Internal interface between the importance network (mlp.hip) and the
// solver context (context.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

struct mcls_mlp;

namespace mcls {

// Records the message returned by mcls_gpu_last_error() and returns code.
int set_error(int code, const std::string& msg);

// Device of a network and its master parameters (MCLS_MLP_PARAMS floats).
int mlp_device(const mcls_mlp* m);
const float* mlp_params_device(const mcls_mlp* m);

// out[r] = network output o for rows x 8 FP32 inputs (device pointers).
hipError_t launch_mlp_forward(const float* params, const float* X, int64_t rows, float* out, hipStream_t s);

}  // namespace mcls
