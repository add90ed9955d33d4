// fs_host_common.h — host helpers shared by the single-tank and the batched
// C-ABI: error reporting, parameter checks, step constants.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "fluidstep.h"
#include "fs_device.h"

namespace fs {

extern thread_local std::string g_last_error;
// Records msg as the fs_last_error() text and returns code.
int fail(int code, const std::string& msg);

// Params uniform fill (src/gpu/fluid.rs:289-318) + live clamps of the
// setters (src/gpu/fluid.rs:1295-1371).
void build_step_params(const fs_params& q, uint32_t n, StepParams& P);
bool all_finite(const float* a, size_t n);
// Creation-time checks of one tank's parameters and particles.
int check_params(const fs_params& q, uint32_t n, const float* pos_xyz, const float* vel_xyz);
// The persistent grids' block count from the occupancy figure.
int half_resident_blocks(int max_blocks);
int persistent_grid_blocks(int max_blocks, int cus);

}  // namespace fs

#define FS_HIP(call)                                                                             \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return ::fs::fail(e_ == hipErrorOutOfMemory ? FS_ERR_OUT_OF_MEMORY : FS_ERR_DEVICE,  \
                              std::string(#call) + ": " + hipGetErrorString(e_));                \
    } while (0)
