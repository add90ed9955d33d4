// k_cg.hip — the matrix-free 7-point CG pressure solve.
//
// Reference-structured form: one kernel per WGSL pass (sh/cg_init.wgsl,
// sh/cg_reduce.wgsl, sh/cg_reduce_final.wgsl, sh/cg_set_rsold.wgsl,
// sh/cg_spmv.wgsl, sh/cg_alpha.wgsl, sh/cg_update.wgsl, sh/cg_beta.wgsl,
// sh/cg_dir.wgsl), f32 scalars in device memory and the same 256-element
// reduction trees, launched 9 times per iteration by the host.
//
// Fused mixed-precision form: divergence + CG init in one cell pass, the
// liquid tiles compacted into a list, then ONE persistent kernel runs the
// whole solve: per iteration a fused SpMV (with the direction update
// d = r + beta d folded in, evaluated identically at every stencil point) and
// a fused update; each ends in a grid-wide deterministic f64 reduction
// (per-block partials summed in block order by the last block to arrive).
// Vectors stay f32; dots, alpha/beta numerators and the stop test are f64,
// alpha and beta are rounded to f32 as src/sim/pressure.rs:282,296 do.
#include "fs_kernels.h"
#include "fs_face.h"

namespace fs {

namespace {
inline unsigned grid_for(size_t n, int per_block) { return (unsigned)((n + per_block - 1) / per_block); }
constexpr int kTile = 64;          // cells per CG tile (one wave)
constexpr int kCgThreads = 256;    // persistent CG workgroup
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
    hipLaunchKernelGGL(k_cg_init_ref, dim3(grid_for(P.C, 64)), dim3(64), 0, s, P, div, ct, p, r, d, scal);
    return hipGetLastError();
}
hipError_t launch_cg_reduce_ref(const StepParams& P, const float* a, const float* b, float* partials,
                                const float* scal, hipStream_t s) {
    hipLaunchKernelGGL(k_cg_reduce_ref, dim3(grid_for(P.C, 256)), dim3(256), 0, s, P, a, b, partials, scal);
    return hipGetLastError();
}
hipError_t launch_cg_reduce_final_ref(const StepParams& P, const float* partials, float* scal, hipStream_t s) {
    hipLaunchKernelGGL(k_cg_reduce_final_ref, dim3(1), dim3(256), 0, s, P, partials, scal);
    return hipGetLastError();
}
hipError_t launch_cg_set_rsold_ref(float* scal, hipStream_t s) {
    hipLaunchKernelGGL(k_cg_set_rsold_ref, dim3(1), dim3(1), 0, s, scal);
    return hipGetLastError();
}
hipError_t launch_cg_spmv_ref(const StepParams& P, const uint8_t* ct, const float* d, float* q, const float* scal,
                              hipStream_t s) {
    hipLaunchKernelGGL(k_cg_spmv_ref, dim3(grid_for(P.C, 64)), dim3(64), 0, s, P, ct, d, q, scal);
    return hipGetLastError();
}
hipError_t launch_cg_alpha_ref(float* scal, hipStream_t s) {
    hipLaunchKernelGGL(k_cg_alpha_ref, dim3(1), dim3(1), 0, s, scal);
    return hipGetLastError();
}
hipError_t launch_cg_update_ref(const StepParams& P, const float* scal, const float* d, const float* q, float* p,
                                float* r, hipStream_t s) {
    hipLaunchKernelGGL(k_cg_update_ref, dim3(grid_for(P.C, 64)), dim3(64), 0, s, P, scal, d, q, p, r);
    return hipGetLastError();
}
hipError_t launch_cg_beta_ref(float* scal, hipStream_t s) {
    hipLaunchKernelGGL(k_cg_beta_ref, dim3(1), dim3(1), 0, s, scal);
    return hipGetLastError();
}
hipError_t launch_cg_dir_ref(const StepParams& P, const float* scal, const float* r, float* d, hipStream_t s) {
    hipLaunchKernelGGL(k_cg_dir_ref, dim3(grid_for(P.C, 64)), dim3(64), 0, s, P, scal, r, d);
    return hipGetLastError();
}
hipError_t launch_cg_iter_count_ref(const float* scal, uint32_t* count, hipStream_t s) {
    hipLaunchKernelGGL(k_cg_iter_count_ref, dim3(1), dim3(1), 0, s, scal, count);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Fused divergence + CG init (sh/divergence.wgsl then sh/cg_init.wgsl per
// cell); b = rhs is kept for the f64 true-residual check.  tile_flags[t] = 1
// when the 64-cell tile t holds a liquid cell.
__device__ __forceinline__ float cell_div(const StepParams& P, const float* __restrict__ u,
                                          const float* __restrict__ v, const float* __restrict__ w,
                                          const uint32_t* __restrict__ occ, uint32_t c) {
    const uint32_t nx = P.nx, ny = P.ny;
    const uint32_t i = c % nx, j = (c / nx) % ny, k = c / (nx * ny);
    const float u_hi = u[(i + 1u) + (nx + 1u) * (j + ny * k)];
    const float u_lo = u[i + (nx + 1u) * (j + ny * k)];
    const float v_hi = v[i + nx * ((j + 1u) + (ny + 1u) * k)];
    const float v_lo = v[i + nx * (j + (ny + 1u) * k)];
    const float w_hi = w[i + nx * (j + ny * (k + 1u))];
    const float w_lo = w[i + nx * (j + ny * k)];
    float div = ((u_hi - u_lo) + (v_hi - v_lo) + (w_hi - w_lo)) * P.inv_h;
    if (P.stiff > 0.0f && P.rest > 0.0f) {
        const float over = fmaxf(0.0f, (float)occ[c] - P.rest) / P.rest;
        const float drift = fminf(P.stiff * over, P.drift);
        div = div - drift;
    }
    return div;
}

__global__ void k_div_cginit(StepParams P, const float* __restrict__ u, const float* __restrict__ v,
                             const float* __restrict__ w, const uint8_t* __restrict__ ct,
                             const uint32_t* __restrict__ occ, float* __restrict__ div, float* __restrict__ b,
                             float* __restrict__ p, float* __restrict__ r, float* __restrict__ d,
                             uint32_t* __restrict__ tile_flags) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    bool liquid = false;
    if (c < P.C) {
        liquid = ct[c] == kLiquid;
        if (liquid) {
            const float dv = cell_div(P, u, v, w, occ, c);
            const float scale = P.rho * P.h * P.h / P.dt;
            const float rhs = -scale * dv;
            float ap = 0.0f;
            if (P.warm) ap = apply_cell(ct, p, c, P.nx, P.nx * P.ny, p[c]);
            else p[c] = 0.0f;
            const float rc = rhs - ap;
            div[c] = dv;
            b[c] = rhs;
            r[c] = rc;
            d[c] = rc;
        } else {
            div[c] = 0.0f;
            b[c] = 0.0f;
            p[c] = 0.0f;
            r[c] = 0.0f;
            d[c] = 0.0f;
        }
    }
    const uint64_t m = __ballot(liquid);
    const uint32_t tile = c / kTile;
    if (lane_id() == 0 && (size_t)tile * kTile < P.C) tile_flags[tile] = m ? 1u : 0u;
}

__global__ void k_compact_tiles(uint32_t ntiles, const uint32_t* __restrict__ flags, const uint32_t* __restrict__ scan,
                                uint32_t* __restrict__ tiles, uint32_t* __restrict__ n_active) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    if (flags[t]) tiles[scan[t]] = t;
    if (t == ntiles - 1u) *n_active = scan[t] + flags[t];
}

hipError_t launch_div_cginit(const StepParams& P, float* const vel[3], const uint8_t* ct, const uint32_t* occ,
                             float* div, float* b, float* p, float* r, float* d, const CgWork& w,
                             uint32_t* scan_ws, hipStream_t s) {
    // blockDim is a multiple of 64 so each wave is exactly one tile.
    hipLaunchKernelGGL(k_div_cginit, dim3(grid_for(P.C, 256)), dim3(256), 0, s, P, vel[0], vel[1], vel[2], ct, occ,
                       div, b, p, r, d, w.tile_flags);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const uint32_t ntiles = (P.C + kTile - 1) / kTile;
    e = launch_exclusive_scan(w.tile_flags, w.tile_scan, ntiles, scan_ws, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_compact_tiles, dim3(grid_for(ntiles, 256)), dim3(256), 0, s, ntiles, w.tile_flags,
                       w.tile_scan, w.tiles, w.n_tiles);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Grid-wide deterministic f64 sum.  Each block adds its waves in wave order,
// writes one partial, and arrives on a counter; the last block to arrive sums
// the partials in block order (fixed tree), publishes the total and bumps a
// generation word the others spin on.  A spin that outlives ~4 s raises the
// abort flag so every block leaves the kernel instead of hanging.
namespace {
constexpr long long kSpinTimeoutTicks = 400000000LL;  // wall_clock64 runs at 100 MHz
}

__device__ __forceinline__ uint32_t ld_acquire(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ bool grid_sum(double v, const CgWork& w, double& out) {
    __shared__ double s_w[kCgThreads / 64];
    __shared__ double s_tree[kCgThreads];
    __shared__ int s_last;
    __shared__ int s_abort;
    __shared__ uint32_t s_gen;
    const uint32_t lane = lane_id(), wv = threadIdx.x >> 6;
    v = wave_sum(v);
    if (lane == 0) s_w[wv] = v;
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        double bs = 0.0;
        for (int i = 0; i < kCgThreads / 64; ++i) bs += s_w[i];
        w.partials[blockIdx.x] = bs;
        const uint32_t g0 = ld_acquire(w.bar_gen);
        __threadfence();
        const uint32_t arrived =
            __hip_atomic_fetch_add(w.bar_count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        s_last = arrived == gridDim.x - 1u;
        s_gen = g0;
        s_abort = 0;
    }
    __syncthreads();
    if (s_last) {
        __threadfence();
        double acc = 0.0;
        for (uint32_t i = threadIdx.x; i < gridDim.x; i += kCgThreads)
            acc += __hip_atomic_load(&w.partials[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_tree[threadIdx.x] = acc;
        __syncthreads();
        for (int d = kCgThreads / 2; d >= 1; d >>= 1) {
            if ((int)threadIdx.x < d) s_tree[threadIdx.x] += s_tree[threadIdx.x + d];
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            __hip_atomic_store(w.result, s_tree[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(w.bar_count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __threadfence();
            __hip_atomic_store(w.bar_gen, s_gen + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
    } else if (threadIdx.x == 0) {
        const long long t0 = wall_clock64();
        while (ld_acquire(w.bar_gen) == s_gen) {
            if (ld_acquire(w.abort_flag)) {
                s_abort = 1;
                break;
            }
            if (wall_clock64() - t0 > kSpinTimeoutTicks) {
                __hip_atomic_store(w.abort_flag, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                s_abort = 1;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    __syncthreads();
    __threadfence();
    if (s_abort) return false;
    out = __hip_atomic_load(w.result, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return true;
}

// The persistent solve.  Block b (XCD-remapped) owns a contiguous slice of the
// active-tile list; wave wv of the block takes every 4th tile of that slice.
__global__ __launch_bounds__(kCgThreads) void k_cg_persistent(StepParams P, const uint8_t* __restrict__ ct,
                                                               const float* __restrict__ b, float* __restrict__ p,
                                                               float* __restrict__ r, float* __restrict__ d0,
                                                               float* __restrict__ d1, float* __restrict__ q,
                                                               CgWork w) {
    const uint32_t G = gridDim.x;
    const uint32_t lb = xcd_remap(blockIdx.x, G);
    const uint32_t nt = *w.n_tiles;
    const uint32_t t_begin = (uint32_t)(((uint64_t)nt * lb) / G);
    const uint32_t t_end = (uint32_t)(((uint64_t)nt * (lb + 1u)) / G);
    const uint32_t lane = lane_id(), wv = threadIdx.x >> 6;
    const uint32_t nx = P.nx, nxny = P.nx * P.ny, C = P.C;

    // rs_initial = r . r
    double acc = 0.0;
    for (uint32_t ti = t_begin + wv; ti < t_end; ti += kCgThreads / 64) {
        const uint32_t c = w.tiles[ti] * kTile + lane;
        if (c < C) {
            const double rv = (double)r[c];
            acc += rv * rv;
        }
    }
    double rs0;
    if (!grid_sum(acc, w, rs0)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) w.state->error = 1u;
        return;
    }
    double rs_old = rs0;
    const double rs_init = rs0;
    const double tol = (double)P.tol;
    const double tol_sq = tol * tol;
    float beta = 0.0f;
    float* dprev = d0;   // direction of the previous iteration
    float* dnext = d1;   // direction written this iteration
    uint32_t iters = 0;
    bool converged = false;
    for (uint32_t it = 0; it < P.iters; ++it) {
        ++iters;
        // Phase A: dn = (it == 0) ? d : r + beta * d_prev; q = A dn; acc = dn . q
        float* dn_buf = it == 0 ? dprev : dnext;
        acc = 0.0;
        for (uint32_t ti = t_begin + wv; ti < t_end; ti += kCgThreads / 64) {
            const uint32_t c = w.tiles[ti] * kTile + lane;
            if (c >= C) continue;
            const uint8_t tc = ct[c];
            float dn;
            if (it == 0) dn = dprev[c];
            else dn = r[c] + beta * dprev[c];
            float qc = 0.0f;
            if (tc == kLiquid) {
                float cnt = 0.0f, sum = 0.0f;
                const uint32_t nb[6] = {c - 1u, c + 1u, c - nx, c + nx, c - nxny, c + nxny};
#pragma unroll
                for (int s = 0; s < 6; ++s) {
                    const uint8_t t = ct[nb[s]];
                    if (t != kSolid) {
                        cnt += 1.0f;
                        if (t == kLiquid) sum += it == 0 ? dprev[nb[s]] : r[nb[s]] + beta * dprev[nb[s]];
                    }
                }
                qc = cnt * dn - sum;
            }
            if (it != 0) dnext[c] = dn;
            q[c] = qc;
            acc += (double)dn * (double)qc;
        }
        double dq;
        if (!grid_sum(acc, w, dq)) {
            if (blockIdx.x == 0 && threadIdx.x == 0) w.state->error = 1u;
            return;
        }
        const float alpha = fabs(dq) > 1e-30 ? (float)(rs_old / dq) : 0.0f;
        // Phase B: p += alpha dn; r -= alpha q; acc = r . r
        acc = 0.0;
        for (uint32_t ti = t_begin + wv; ti < t_end; ti += kCgThreads / 64) {
            const uint32_t c = w.tiles[ti] * kTile + lane;
            if (c >= C) continue;
            const float dn = dn_buf[c];
            p[c] = p[c] + alpha * dn;
            const float rn = r[c] - alpha * q[c];
            r[c] = rn;
            acc += (double)rn * (double)rn;
        }
        double rr;
        if (!grid_sum(acc, w, rr)) {
            if (blockIdx.x == 0 && threadIdx.x == 0) w.state->error = 1u;
            return;
        }
        beta = rs_old > 1e-30 ? (float)(rr / rs_old) : 0.0f;
        rs_old = rr;
        if (tol > 0.0 && rr <= tol_sq * rs_init) {
            converged = true;
            break;
        }
        if (it != 0) {
            float* t = dprev;
            dprev = dnext;
            dnext = t;
        }
    }
    (void)converged;
    // f64 true residual |b - A p|^2 over liquid cells.
    acc = 0.0;
    for (uint32_t ti = t_begin + wv; ti < t_end; ti += kCgThreads / 64) {
        const uint32_t c = w.tiles[ti] * kTile + lane;
        if (c >= C || ct[c] != kLiquid) continue;
        double cnt = 0.0, sum = 0.0;
        const uint32_t nb[6] = {c - 1u, c + 1u, c - nx, c + nx, c - nxny, c + nxny};
#pragma unroll
        for (int s = 0; s < 6; ++s) {
            const uint8_t t = ct[nb[s]];
            if (t != kSolid) {
                cnt += 1.0;
                if (t == kLiquid) sum += (double)p[nb[s]];
            }
        }
        const double e = (double)b[c] - (cnt * (double)p[c] - sum);
        acc += e * e;
    }
    double tr;
    if (!grid_sum(acc, w, tr)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) w.state->error = 1u;
        return;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        w.state->rs_init = rs_init;
        w.state->rs_old = rs_old;
        w.state->true_res_sq = tr;
        w.state->iterations = iters;
        w.state->active = 0u;
        w.state->error = 0u;
    }
}

int cg_persistent_max_blocks() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_cg_persistent, kCgThreads, 0) != hipSuccess)
        return 0;
    return cus * per_cu;
}

hipError_t launch_cg_persistent(const StepParams& P, const uint8_t* ct, const float* b, float* p, float* r,
                                float* d0, float* d1, float* q, const CgWork& w, int grid_blocks, hipStream_t s) {
    void* args[] = {(void*)&P, (void*)&ct, (void*)&b, (void*)&p, (void*)&r,
                    (void*)&d0, (void*)&d1, (void*)&q, (void*)&w};
    return hipLaunchCooperativeKernel((const void*)k_cg_persistent, dim3(grid_blocks), dim3(kCgThreads), args, 0, s);
}

}  // namespace fs
