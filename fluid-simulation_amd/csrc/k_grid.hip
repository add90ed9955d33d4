// k_grid.hip — cell typing, P2G, the face passes, divergence and gradient.
//
// Two forms of each pass live here:
//   * reference-structured kernels, one per WGSL pass and axis
//     (sh/classify.wgsl, sh/scatter.wgsl, sh/normalize.wgsl, sh/save_vel.wgsl,
//     sh/forces.wgsl, sh/boundaries.wgsl, sh/divergence.wgsl, sh/gradient.wgsl);
//   * fused kernels: the sorted-segment P2G gather per brick of cells with the
//     per-face epilogue (normalize -> save -> forces -> enforce) in the same
//     kernel, and gradient + enforce as one pass over all faces.
// Both forms evaluate the per-face functions of fs_face.h, so they agree bit
// for bit.
#include "fs_kernels.h"
#include "fs_face.h"

namespace fs {

namespace {
inline unsigned grid_for(size_t n, int per_block) { return (unsigned)((n + per_block - 1) / per_block); }
}

// sh/classify.wgsl:25-61 (u8 cell types; stats[0] counts liquid cells).
__global__ void k_classify(StepParams P, const uint32_t* __restrict__ occ, uint8_t* __restrict__ ct,
                           uint32_t* __restrict__ stats) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    bool liquid = false;
    if (c < P.C) {
        const uint32_t nx = P.nx, ny = P.ny, nz = P.nz;
        const uint32_t i = c % nx, j = (c / nx) % ny, k = c / (nx * ny);
        const bool boundary = i == 0u || j == 0u || k == 0u || i == nx - 1u || j == ny - 1u || k == nz - 1u;
        if (boundary) {
            ct[c] = kSolid;
        } else {
            const uint32_t t = P.thresh;
            liquid = occ[c] >= t;
            if (!liquid && P.dilation) {
                liquid = occ[c - 1u] >= t || occ[c + 1u] >= t || occ[c - nx] >= t || occ[c + nx] >= t ||
                         occ[c - nx * ny] >= t || occ[c + nx * ny] >= t;
            }
            ct[c] = liquid ? kLiquid : kAir;
        }
    }
    // One atomic per block (the count is an integer sum, so grouping is free).
    __shared__ uint32_t s_cnt;
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    const uint64_t m = __ballot(liquid);
    if (lane_id() == 0 && m) atomicAdd(&s_cnt, (uint32_t)__popcll(m));
    __syncthreads();
    if (threadIdx.x == 0 && s_cnt) atomicAdd(stats, s_cnt);
}

hipError_t launch_classify(const StepParams& P, const uint32_t* occ, uint8_t* ct, uint32_t* stats,
                           hipStream_t s) {
    hipLaunchKernelGGL(k_classify, dim3(grid_for(P.C, 1024)), dim3(1024), 0, s, P, occ, ct, stats);
    return hipGetLastError();
}

// sh/scatter.wgsl:78-118: per particle, three axes, global i32 atomics.
template <int BF16>
__global__ void k_p2g_atomic(StepParams P, ParticleSoA sp, int32_t* __restrict__ un, int32_t* __restrict__ ud,
                             int32_t* __restrict__ vn, int32_t* __restrict__ vd, int32_t* __restrict__ wn,
                             int32_t* __restrict__ wd, const uint32_t* __restrict__ guard) {
    if (guard && *guard == 0u) return;
    int32_t* nums[3] = {un, vn, wn};
    int32_t* dens[3] = {ud, vd, wd};
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < P.N; p += stride) {
        const float pos[3] = {(sp.pos[0][p] - P.origin[0]) * P.inv_h, (sp.pos[1][p] - P.origin[1]) * P.inv_h,
                              (sp.pos[2][p] - P.origin[2]) * P.inv_h};
        const float pv[3] = {load_vel<BF16>(sp.vel[0], p), load_vel<BF16>(sp.vel[1], p),
                             load_vel<BF16>(sp.vel[2], p)};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            uint32_t dim[3];
            face_dims(P, a, dim);
            const float g[3] = {pos[0] + stagger(a, 0), pos[1] + stagger(a, 1), pos[2] + stagger(a, 2)};
            int32_t* nm = nums[a];
            int32_t* dn = dens[a];
            p2g_taps(g, dim, pv[a], P.scale, [&](int32_t ii, int32_t jj, int32_t kk, int32_t tn, int32_t td) {
                const uint32_t idx = (uint32_t)ii + dim[0] * ((uint32_t)jj + dim[1] * (uint32_t)kk);
                atomicAdd(&nm[idx], tn);
                atomicAdd(&dn[idx], td);
            });
        }
    }
}

// A guarded launch is the fused engine's fallback and is usually a no-op, so
// it uses a small grid-stride grid; unguarded launches take one particle per
// thread like sh/scatter.wgsl.
hipError_t launch_p2g_atomic(const StepParams& P, const ParticleSoA& src, int bf16, int32_t* const num[3],
                             int32_t* const den[3], const uint32_t* guard, hipStream_t s) {
    if (P.N == 0) return hipSuccess;
    unsigned g = grid_for(P.N, 64);
    if (guard && g > 4096) g = 4096;
    if (bf16)
        hipLaunchKernelGGL(k_p2g_atomic<1>, dim3(g), dim3(64), 0, s, P, src, num[0], den[0], num[1], den[1], num[2],
                           den[2], guard);
    else
        hipLaunchKernelGGL(k_p2g_atomic<0>, dim3(g), dim3(64), 0, s, P, src, num[0], den[0], num[1], den[1], num[2],
                           den[2], guard);
    return hipGetLastError();
}

// ---- reference-structured per-axis passes ----
__global__ void k_normalize(uint32_t n, const int32_t* __restrict__ num, const int32_t* __restrict__ den,
                            float* __restrict__ vel) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n) return;
    const int32_t d = den[f];
    vel[f] = d > 0 ? (float)num[f] / (float)d : 0.0f;
}

__global__ void k_copy(const float* __restrict__ src, float* __restrict__ dst, size_t n) {
    const size_t f = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f < n) dst[f] = src[f];
}

__global__ void k_forces(StepParams P, int a, const uint8_t* __restrict__ ct, float* __restrict__ vel) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= P.F[a]) return;
    uint32_t d[3], i, j, k;
    face_dims(P, a, d);
    face_ijk(f, d, i, j, k);
    if (face_wet(P, a, ct, i, j, k)) vel[f] = vel[f] + P.grav[a] * P.dt;
}

__global__ void k_enforce(StepParams P, int a, const uint8_t* __restrict__ ct, float* __restrict__ vel) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= P.F[a]) return;
    uint32_t d[3], i, j, k;
    face_dims(P, a, d);
    face_ijk(f, d, i, j, k);
    if (face_solid(P, a, ct, i, j, k)) vel[f] = 0.0f;
}

__global__ void k_gradient(StepParams P, int a, const float* __restrict__ pr, const uint8_t* __restrict__ ct,
                           float* __restrict__ vel) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= P.F[a]) return;
    uint32_t d[3], i, j, k;
    face_dims(P, a, d);
    face_ijk(f, d, i, j, k);
    vel[f] = face_gradient(P, a, pr, ct, i, j, k, vel[f]);
}

hipError_t launch_normalize(const StepParams& P, int a, const int32_t* num, const int32_t* den, float* vel,
                            hipStream_t s) {
    hipLaunchKernelGGL(k_normalize, dim3(grid_for(P.F[a], 64)), dim3(64), 0, s, P.F[a], num, den, vel);
    return hipGetLastError();
}
hipError_t launch_copy_f32(const float* src, float* dst, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_copy, dim3(grid_for(n, 64)), dim3(64), 0, s, src, dst, n);
    return hipGetLastError();
}
hipError_t launch_forces(const StepParams& P, int a, const uint8_t* ct, float* vel, hipStream_t s) {
    hipLaunchKernelGGL(k_forces, dim3(grid_for(P.F[a], 64)), dim3(64), 0, s, P, a, ct, vel);
    return hipGetLastError();
}
hipError_t launch_enforce(const StepParams& P, int a, const uint8_t* ct, float* vel, hipStream_t s) {
    hipLaunchKernelGGL(k_enforce, dim3(grid_for(P.F[a], 64)), dim3(64), 0, s, P, a, ct, vel);
    return hipGetLastError();
}
hipError_t launch_gradient(const StepParams& P, int a, const float* p, const uint8_t* ct, float* vel,
                           hipStream_t s) {
    hipLaunchKernelGGL(k_gradient, dim3(grid_for(P.F[a], 64)), dim3(64), 0, s, P, a, p, ct, vel);
    return hipGetLastError();
}

// sh/divergence.wgsl:22-59.
__device__ __forceinline__ float cell_divergence(const StepParams& P, const float* __restrict__ u,
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

__global__ void k_divergence(StepParams P, const float* __restrict__ u, const float* __restrict__ v,
                             const float* __restrict__ w, const uint8_t* __restrict__ ct,
                             const uint32_t* __restrict__ occ, float* __restrict__ div) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= P.C) return;
    div[c] = ct[c] != kLiquid ? 0.0f : cell_divergence(P, u, v, w, occ, c);
}

hipError_t launch_divergence(const StepParams& P, float* const vel[3], const uint8_t* ct, const uint32_t* occ,
                             float* div, hipStream_t s) {
    hipLaunchKernelGGL(k_divergence, dim3(grid_for(P.C, 64)), dim3(64), 0, s, P, vel[0], vel[1], vel[2], ct,
                       occ, div);
    return hipGetLastError();
}

// ---- fused face passes over the concatenated u|v|w face index space ----
__device__ __forceinline__ bool split_face(const StepParams& P, uint32_t g, int& a, uint32_t& f) {
    if (g < P.F[0]) { a = 0; f = g; return true; }
    g -= P.F[0];
    if (g < P.F[1]) { a = 1; f = g; return true; }
    g -= P.F[1];
    if (g < P.F[2]) { a = 2; f = g; return true; }
    return false;
}

__global__ void k_face_epilogue(StepParams P, const int32_t* __restrict__ n0, const int32_t* __restrict__ n1,
                                const int32_t* __restrict__ n2, const int32_t* __restrict__ d0,
                                const int32_t* __restrict__ d1, const int32_t* __restrict__ d2,
                                const uint8_t* __restrict__ ct, float* __restrict__ v0, float* __restrict__ v1,
                                float* __restrict__ v2, float* __restrict__ s0, float* __restrict__ s1,
                                float* __restrict__ s2, const uint32_t* __restrict__ guard) {
    if (guard && *guard == 0u) return;
    const uint32_t tot = P.F[0] + P.F[1] + P.F[2];
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t gi = blockIdx.x * blockDim.x + threadIdx.x; gi < tot; gi += stride) {
        int a;
        uint32_t f;
        split_face(P, gi, a, f);
        const int32_t* num = a == 0 ? n0 : (a == 1 ? n1 : n2);
        const int32_t* den = a == 0 ? d0 : (a == 1 ? d1 : d2);
        float* vel = a == 0 ? v0 : (a == 1 ? v1 : v2);
        float* sav = a == 0 ? s0 : (a == 1 ? s1 : s2);
        uint32_t d[3], i, j, k;
        face_dims(P, a, d);
        face_ijk(f, d, i, j, k);
        float vv, ss;
        face_epilogue(P, a, ct, i, j, k, num[f], den[f], vv, ss);
        vel[f] = vv;
        sav[f] = ss;
    }
}

hipError_t launch_face_epilogue(const StepParams& P, const int32_t* const num[3], const int32_t* const den[3],
                                const uint8_t* ct, float* const vel[3], float* const saved[3], const uint32_t* guard,
                                hipStream_t s) {
    const size_t tot = (size_t)P.F[0] + P.F[1] + P.F[2];
    unsigned g = grid_for(tot, 256);
    if (guard && g > 4096) g = 4096;  // usually a no-op fallback: small grid-stride grid
    hipLaunchKernelGGL(k_face_epilogue, dim3(g), dim3(256), 0, s, P, num[0], num[1], num[2], den[0], den[1], den[2],
                       ct, vel[0], vel[1], vel[2], saved[0], saved[1], saved[2], guard);
    return hipGetLastError();
}

__global__ void k_gradient_enforce(StepParams P, const float* __restrict__ pr, const uint8_t* __restrict__ ct,
                                   float* __restrict__ v0, float* __restrict__ v1, float* __restrict__ v2) {
    int a;
    uint32_t f;
    if (!split_face(P, blockIdx.x * blockDim.x + threadIdx.x, a, f)) return;
    float* vel = a == 0 ? v0 : (a == 1 ? v1 : v2);
    uint32_t d[3], i, j, k;
    face_dims(P, a, d);
    face_ijk(f, d, i, j, k);
    float v = face_gradient(P, a, pr, ct, i, j, k, vel[f]);
    if (face_solid(P, a, ct, i, j, k)) v = 0.0f;
    vel[f] = v;
}

hipError_t launch_gradient_enforce(const StepParams& P, const float* p, const uint8_t* ct, float* const vel[3],
                                   hipStream_t s) {
    const size_t tot = (size_t)P.F[0] + P.F[1] + P.F[2];
    hipLaunchKernelGGL(k_gradient_enforce, dim3(grid_for(tot, 256)), dim3(256), 0, s, P, p, ct, vel[0], vel[1],
                       vel[2]);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Sorted-segment P2G gather.  One workgroup owns a brick of BX x BY x BZ cells
// and every MAC face whose cell coordinate (own-axis coordinate clamped to
// n-1) lies in the brick.  A face's taps come from particles whose cell is in
// the brick grown by one cell on every side; after the sort those particles
// are, per (y,z) row of that halo region, one contiguous segment of the
// sorted arrays.  The workgroup walks the segments, evaluates each particle's
// 24 taps exactly as sh/scatter.wgsl does, and adds the ones it owns into LDS
// with integer atomics (order-free, so the sums equal the reference's i32
// atomics bit for bit).  The epilogue normalizes, saves, applies gravity and
// enforces walls per owned face, and stores vel/saved once.
//
// Requires every particle's P2G cell floor((pos-origin)*inv_h) to equal its
// sort cell (clamped floor((pos-origin)/h)); the host checks the count of
// particles breaking that (k_misfit_count) and uses the atomic form otherwise.
namespace {
constexpr int kGBX = 16, kGBY = 8, kGBZ = 8;
constexpr int kGThreads = 256;
constexpr int kGRowsMax = (kGBY + 2) * (kGBZ + 2);
constexpr int kGU = (kGBX + 1) * kGBY * kGBZ;
constexpr int kGV = kGBX * (kGBY + 1) * kGBZ;
constexpr int kGW = kGBX * kGBY * (kGBZ + 1);
}  // namespace

template <int BF16>
__global__ __launch_bounds__(kGThreads) void k_p2g_gather(StepParams P, ParticleSoA sp,
                                                           const uint32_t* __restrict__ occ,
                                                           const uint32_t* __restrict__ off,
                                                           const uint8_t* __restrict__ ct, float* __restrict__ v0,
                                                           float* __restrict__ v1, float* __restrict__ v2,
                                                           float* __restrict__ s0, float* __restrict__ s1,
                                                           float* __restrict__ s2, uint32_t nbx, uint32_t nby,
                                                           uint32_t nbz, const uint32_t* __restrict__ misfit) {
    if (*misfit != 0u) return;  // the atomic form handles this step
    __shared__ int32_t s_num[kGU + kGV + kGW];
    __shared__ int32_t s_den[kGU + kGV + kGW];
    __shared__ uint32_t s_row_start[kGRowsMax];
    __shared__ uint32_t s_row_pref[kGRowsMax + 1];

    const uint32_t G = nbx * nby * nbz;
    const uint32_t b = xcd_remap(blockIdx.x, G);
    const uint32_t bxi = b % nbx, byi = (b / nbx) % nby, bzi = b / (nbx * nby);
    const int32_t x0 = (int32_t)bxi * kGBX, y0 = (int32_t)byi * kGBY, z0 = (int32_t)bzi * kGBZ;
    const int32_t nx = (int32_t)P.nx, ny = (int32_t)P.ny, nz = (int32_t)P.nz;

    for (int t = threadIdx.x; t < kGU + kGV + kGW; t += kGThreads) {
        s_num[t] = 0;
        s_den[t] = 0;
    }
    // Halo cell range (inclusive), clamped to the grid.
    const int32_t xs = max(x0 - 1, 0), xe = min(x0 + kGBX, nx - 1);
    const int32_t ys = max(y0 - 1, 0), ye = min(y0 + kGBY, ny - 1);
    const int32_t zs = max(z0 - 1, 0), ze = min(z0 + kGBZ, nz - 1);
    const int32_t rows_y = ye - ys + 1;
    const int32_t rows = rows_y * (ze - zs + 1);
    if (threadIdx.x < 64) {
        // Row segments and their prefix sums (one wave, rows <= 100).
        uint32_t carry = 0;
        for (int r0 = 0; r0 < rows; r0 += 64) {
            const int r = r0 + (int)threadIdx.x;
            uint32_t len = 0, start = 0;
            if (r < rows) {
                const int32_t cy = ys + r % rows_y, cz = zs + r / rows_y;
                const uint32_t cf = (uint32_t)xs + P.nx * ((uint32_t)cy + P.ny * (uint32_t)cz);
                const uint32_t cl = (uint32_t)xe + P.nx * ((uint32_t)cy + P.ny * (uint32_t)cz);
                start = off[cf];
                len = off[cl] + occ[cl] - start;
                s_row_start[r] = start;
            }
            uint32_t inc = len;
            for (int o = 1; o < 64; o <<= 1) {
                uint32_t t = __shfl_up(inc, o, 64);
                if ((int)threadIdx.x >= o) inc += t;
            }
            if (r < rows) s_row_pref[r + 1] = carry + inc;
            carry += __shfl(inc, 63, 64);
        }
        if (threadIdx.x == 0) s_row_pref[0] = 0;
    }
    __syncthreads();
    const uint32_t total = s_row_pref[rows];
    const int32_t ownx_hi = (x0 + kGBX >= nx) ? kGBX : kGBX - 1;  // own-axis local max index per axis
    const int32_t owny_hi = (y0 + kGBY >= ny) ? kGBY : kGBY - 1;
    const int32_t ownz_hi = (z0 + kGBZ >= nz) ? kGBZ : kGBZ - 1;
    int row = 0;
    for (uint32_t base = 0; base < total; base += kGThreads) {
        // Each wave takes every 4th item of the chunk, so its lanes fall in
        // different cells (sorted particles of one cell share faces, and
        // same-address LDS atomics within a wave serialize).
        const uint32_t t = base + (threadIdx.x & 63u) * (kGThreads / 64) + (threadIdx.x >> 6);
        if (t >= total) continue;
        // Find the row holding flattened item t (rows are few; linear advance).
        while (s_row_pref[row + 1] <= t) ++row;
        const uint32_t p = s_row_start[row] + (t - s_row_pref[row]);
        const float pos[3] = {(sp.pos[0][p] - P.origin[0]) * P.inv_h, (sp.pos[1][p] - P.origin[1]) * P.inv_h,
                              (sp.pos[2][p] - P.origin[2]) * P.inv_h};
        const float pv[3] = {load_vel<BF16>(sp.vel[0], p), load_vel<BF16>(sp.vel[1], p),
                             load_vel<BF16>(sp.vel[2], p)};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            uint32_t dim[3];
            face_dims(P, a, dim);
            const float g[3] = {pos[0] + stagger(a, 0), pos[1] + stagger(a, 1), pos[2] + stagger(a, 2)};
            const int32_t hx = a == 0 ? ownx_hi : kGBX - 1;
            const int32_t hy = a == 1 ? owny_hi : kGBY - 1;
            const int32_t hz = a == 2 ? ownz_hi : kGBZ - 1;
            const int32_t lx_dim = a == 0 ? kGBX + 1 : kGBX;
            const int32_t ly_dim = a == 1 ? kGBY + 1 : kGBY;
            const int32_t base = a == 0 ? 0 : (a == 1 ? kGU : kGU + kGV);
            p2g_taps(g, dim, pv[a], P.scale, [&](int32_t ii, int32_t jj, int32_t kk, int32_t tn, int32_t td) {
                const int32_t lx = ii - x0, ly = jj - y0, lz = kk - z0;
                if (lx < 0 || lx > hx || ly < 0 || ly > hy || lz < 0 || lz > hz) return;
                const int32_t li = base + lx + lx_dim * (ly + ly_dim * lz);
                atomicAdd(&s_num[li], tn);
                atomicAdd(&s_den[li], td);
            });
        }
    }
    __syncthreads();
    // Epilogue over owned faces of the three axes.
    for (int a = 0; a < 3; ++a) {
        const int32_t lxd = a == 0 ? kGBX + 1 : kGBX;
        const int32_t lyd = a == 1 ? kGBY + 1 : kGBY;
        const int32_t lzd = a == 2 ? kGBZ + 1 : kGBZ;
        const int32_t hx = a == 0 ? ownx_hi : kGBX - 1;
        const int32_t hy = a == 1 ? owny_hi : kGBY - 1;
        const int32_t hz = a == 2 ? ownz_hi : kGBZ - 1;
        const int32_t base = a == 0 ? 0 : (a == 1 ? kGU : kGU + kGV);
        float* vel = a == 0 ? v0 : (a == 1 ? v1 : v2);
        float* sav = a == 0 ? s0 : (a == 1 ? s1 : s2);
        uint32_t dim[3];
        face_dims(P, a, dim);
        for (int t = threadIdx.x; t < lxd * lyd * lzd; t += kGThreads) {
            const int32_t lx = t % lxd, ly = (t / lxd) % lyd, lz = t / (lxd * lyd);
            if (lx > hx || ly > hy || lz > hz) continue;
            const int32_t i = x0 + lx, j = y0 + ly, k = z0 + lz;
            if (i >= (int32_t)dim[0] || j >= (int32_t)dim[1] || k >= (int32_t)dim[2]) continue;
            float vv, ss;
            face_epilogue(P, a, ct, (uint32_t)i, (uint32_t)j, (uint32_t)k, s_num[base + t], s_den[base + t], vv, ss);
            const uint32_t f = (uint32_t)i + dim[0] * ((uint32_t)j + dim[1] * (uint32_t)k);
            vel[f] = vv;
            sav[f] = ss;
        }
    }
}

hipError_t launch_p2g_gather(const StepParams& P, const ParticleSoA& sorted, int bf16, const uint32_t* occ,
                             const uint32_t* cell_offset, const uint8_t* ct, float* const vel[3],
                             float* const saved[3], const uint32_t* misfit, hipStream_t s) {
    const uint32_t nbx = (P.nx + kGBX - 1) / kGBX, nby = (P.ny + kGBY - 1) / kGBY, nbz = (P.nz + kGBZ - 1) / kGBZ;
    const uint32_t G = nbx * nby * nbz;
    if (bf16)
        hipLaunchKernelGGL(k_p2g_gather<1>, dim3(G), dim3(kGThreads), 0, s, P, sorted, occ, cell_offset, ct, vel[0],
                           vel[1], vel[2], saved[0], saved[1], saved[2], nbx, nby, nbz, misfit);
    else
        hipLaunchKernelGGL(k_p2g_gather<0>, dim3(G), dim3(kGThreads), 0, s, P, sorted, occ, cell_offset, ct, vel[0],
                           vel[1], vel[2], saved[0], saved[1], saved[2], nbx, nby, nbz, misfit);
    return hipGetLastError();
}

}  // namespace fs
