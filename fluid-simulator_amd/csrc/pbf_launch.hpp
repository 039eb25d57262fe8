// Launch and error helpers shared by the kernel translation units.
#pragma once

#include <hip/hip_runtime.h>

#include "pbf_amd.h"

// Records a thread-local message and returns the status unchanged.
pbf_status pbf_fail(pbf_status status, const char* message);
// PBF_OK for hipSuccess; otherwise PBF_ERR_DEVICE with the HIP error string.
pbf_status pbf_hip_check(hipError_t error);

// Launches and reports launch errors the way `<<<>>>` plus cudaGetLastError()
// does in the reference.
#define PBF_LAUNCH(kernel, grid, block, stream, ...)                                        \
    do {                                                                                    \
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0,                              \
                           static_cast<hipStream_t>(stream), __VA_ARGS__);                  \
        return pbf_hip_check(hipGetLastError());                                            \
    } while (0)
