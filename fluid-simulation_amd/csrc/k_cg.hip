// k_cg.hip — the matrix-free 7-point CG pressure solve.
//
// Reference-structured form: one kernel per WGSL pass (sh/cg_init.wgsl,
// sh/cg_reduce.wgsl, sh/cg_reduce_final.wgsl, sh/cg_set_rsold.wgsl,
// sh/cg_spmv.wgsl, sh/cg_alpha.wgsl, sh/cg_update.wgsl, sh/cg_beta.wgsl,
// sh/cg_dir.wgsl), f32 scalars in device memory and the same 256-element
// reduction trees, launched 9 times per iteration by the host.  The fused
// mixed-precision solve is in k_cg_fused.hip.
#include "fs_kernels.h"
#include "fs_face.h"

namespace fs {

namespace {
inline unsigned grid_for(size_t n, int per_block) { return (unsigned)((n + per_block - 1) / per_block); }
}  // namespace

// 7-point operator on one liquid cell, neighbour order -x,+x,-y,+y,-z,+z with
// a float counter (sh/cg_spmv.wgsl:46-68, sh/cg_init.wgsl:24-48).
__device__ __forceinline__ float apply_cell(const uint8_t* __restrict__ ct, const float* __restrict__ x,
                                            uint32_t c, uint32_t nx, uint32_t nxny, float xc) {
    float cnt = 0.0f, sum = 0.0f;
    const uint32_t nb[6] = {c - 1u, c + 1u, c - nx, c + nx, c - nxny, c + nxny};
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        const uint8_t t = ct[nb[s]];
        if (t != kSolid) {
            cnt += 1.0f;
            if (t == kLiquid) sum += x[nb[s]];
        }
    }
    return cnt * xc - sum;
}

// ---------------------------------------------------------------------------
// Reference-structured passes.  scal: 0 rs_old, 1 dot, 2 alpha, 3 beta,
// 4 rs_initial, 5 active, 6 tol_sq (sh/cg_init.wgsl:20-22).
__global__ void k_cg_init_ref(StepParams P, const float* __restrict__ div, const uint8_t* __restrict__ ct,
                              float* __restrict__ p, float* __restrict__ r, float* __restrict__ d,
                              float* __restrict__ scal) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c == 0u) {
        scal[0] = 0.0f; scal[1] = 0.0f; scal[2] = 0.0f; scal[3] = 0.0f; scal[4] = 0.0f; scal[5] = 1.0f;
    }
    if (c >= P.C) return;
    if (ct[c] == kLiquid) {
        const float scale = P.rho * P.h * P.h / P.dt;
        const float rhs = -scale * div[c];
        float ap = 0.0f;
        if (P.warm) ap = apply_cell(ct, p, c, P.nx, P.nx * P.ny, p[c]);
        else p[c] = 0.0f;
        const float rc = rhs - ap;
        r[c] = rc;
        d[c] = rc;
    } else {
        p[c] = 0.0f;
        r[c] = 0.0f;
        d[c] = 0.0f;
    }
}

__global__ __launch_bounds__(256) void k_cg_reduce_ref(StepParams P, const float* __restrict__ a,
                                                        const float* __restrict__ b,
                                                        float* __restrict__ partials,
                                                        const float* __restrict__ scal) {
    __shared__ float s[256];
    const uint32_t li = threadIdx.x;
    const uint32_t idx = blockIdx.x * 256u + li;
    const bool active = scal[5] != 0.0f;
    s[li] = (active && idx < P.C) ? a[idx] * b[idx] : 0.0f;
    __syncthreads();
    for (uint32_t d = 128; d >= 1; d >>= 1) {
        if (li < d) s[li] += s[li + d];
        __syncthreads();
    }
    if (li == 0) partials[blockIdx.x] = s[0];
}

__global__ __launch_bounds__(256) void k_cg_reduce_final_ref(StepParams P, const float* __restrict__ partials,
                                                              float* __restrict__ scal) {
    __shared__ float s[256];
    const uint32_t li = threadIdx.x;
    const bool active = scal[5] != 0.0f;
    const uint32_t num_wg = (P.C + 255u) / 256u;
    float acc = 0.0f;
    if (active)
        for (uint32_t idx = li; idx < num_wg; idx += 256u) acc += partials[idx];
    s[li] = acc;
    __syncthreads();
    for (uint32_t d = 128; d >= 1; d >>= 1) {
        if (li < d) s[li] += s[li + d];
        __syncthreads();
    }
    if (li == 0) scal[1] = s[0];
}

__global__ void k_cg_set_rsold_ref(float* scal) {
    scal[0] = scal[1];
    scal[4] = scal[1];
    scal[5] = 1.0f;
}

__global__ void k_cg_spmv_ref(StepParams P, const uint8_t* __restrict__ ct, const float* __restrict__ d,
                              float* __restrict__ q, const float* __restrict__ scal) {
    if (scal[5] == 0.0f) return;
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= P.C) return;
    if (ct[c] != kLiquid) {
        q[c] = 0.0f;
        return;
    }
    q[c] = apply_cell(ct, d, c, P.nx, P.nx * P.ny, d[c]);
}

__global__ void k_cg_alpha_ref(float* scal) {
    if (scal[5] == 0.0f) {
        scal[2] = 0.0f;
        return;
    }
    const float dq = scal[1];
    scal[2] = fabsf(dq) > 1e-30f ? scal[0] / dq : 0.0f;
}

__global__ void k_cg_update_ref(StepParams P, const float* __restrict__ scal, const float* __restrict__ d,
                                const float* __restrict__ q, float* __restrict__ p, float* __restrict__ r) {
    if (scal[5] == 0.0f) return;
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= P.C) return;
    const float a = scal[2];
    p[c] = p[c] + a * d[c];
    r[c] = r[c] - a * q[c];
}

__global__ void k_cg_beta_ref(float* scal) {
    if (scal[5] == 0.0f) {
        scal[3] = 0.0f;
        return;
    }
    const float rs_old = scal[0], rs_new = scal[1];
    scal[3] = rs_old > 1e-30f ? rs_new / rs_old : 0.0f;
    scal[0] = rs_new;
    const float tol_sq = scal[6], rs_init = scal[4];
    if (tol_sq > 0.0f && rs_new <= tol_sq * rs_init) {
        scal[5] = 0.0f;
        scal[3] = 0.0f;
    }
}

__global__ void k_cg_dir_ref(StepParams P, const float* __restrict__ scal, const float* __restrict__ r,
                             float* __restrict__ d) {
    if (scal[5] == 0.0f) return;
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= P.C) return;
    const float b = scal[3];
    d[c] = r[c] + b * d[c];
}

// Iteration counter (diagnostic only): counts spmv launches that found the
// solve active; launched right before each spmv.
__global__ void k_cg_iter_count_ref(const float* scal, uint32_t* count) {
    if (scal[5] != 0.0f) *count += 1u;
}

hipError_t launch_cg_init_ref(const StepParams& P, const float* div, const uint8_t* ct, float* p, float* r,
                              float* d, float* scal, hipStream_t s) {
    FS_LAUNCH(k_cg_init_ref, dim3(grid_for(P.C, 64)), dim3(64), 0, s, P, div, ct, p, r, d, scal);
    return hipGetLastError();
}
hipError_t launch_cg_reduce_ref(const StepParams& P, const float* a, const float* b, float* partials,
                                const float* scal, hipStream_t s) {
    FS_LAUNCH(k_cg_reduce_ref, dim3(grid_for(P.C, 256)), dim3(256), 0, s, P, a, b, partials, scal);
    return hipGetLastError();
}
hipError_t launch_cg_reduce_final_ref(const StepParams& P, const float* partials, float* scal, hipStream_t s) {
    FS_LAUNCH(k_cg_reduce_final_ref, dim3(1), dim3(256), 0, s, P, partials, scal);
    return hipGetLastError();
}
hipError_t launch_cg_set_rsold_ref(float* scal, hipStream_t s) {
    FS_LAUNCH(k_cg_set_rsold_ref, dim3(1), dim3(1), 0, s, scal);
    return hipGetLastError();
}
hipError_t launch_cg_spmv_ref(const StepParams& P, const uint8_t* ct, const float* d, float* q, const float* scal,
                              hipStream_t s) {
    FS_LAUNCH(k_cg_spmv_ref, dim3(grid_for(P.C, 64)), dim3(64), 0, s, P, ct, d, q, scal);
    return hipGetLastError();
}
hipError_t launch_cg_alpha_ref(float* scal, hipStream_t s) {
    FS_LAUNCH(k_cg_alpha_ref, dim3(1), dim3(1), 0, s, scal);
    return hipGetLastError();
}
hipError_t launch_cg_update_ref(const StepParams& P, const float* scal, const float* d, const float* q, float* p,
                                float* r, hipStream_t s) {
    FS_LAUNCH(k_cg_update_ref, dim3(grid_for(P.C, 64)), dim3(64), 0, s, P, scal, d, q, p, r);
    return hipGetLastError();
}
hipError_t launch_cg_beta_ref(float* scal, hipStream_t s) {
    FS_LAUNCH(k_cg_beta_ref, dim3(1), dim3(1), 0, s, scal);
    return hipGetLastError();
}
hipError_t launch_cg_dir_ref(const StepParams& P, const float* scal, const float* r, float* d, hipStream_t s) {
    FS_LAUNCH(k_cg_dir_ref, dim3(grid_for(P.C, 64)), dim3(64), 0, s, P, scal, r, d);
    return hipGetLastError();
}
hipError_t launch_cg_iter_count_ref(const float* scal, uint32_t* count, hipStream_t s) {
    FS_LAUNCH(k_cg_iter_count_ref, dim3(1), dim3(1), 0, s, scal, count);
    return hipGetLastError();
}


}  // namespace fs
