// k_g2p.hip — grid-to-particle transfer with the FLIP/PIC blend, CFL clamp,
// wall friction, advection and wall clamp (sh/g2p.wgsl:131-198), and the
// uniform velocity impulse (sh/impulse.wgsl:26-36).
//
// Particles are read and written in place on the live ping-pong side (SoA).
// After the sort, consecutive lanes hold particles of the same or adjacent
// cells, so the 48 face loads per particle hit the same cache lines.
#include "fs_kernels.h"
#include "fs_face.h"

namespace fs {

namespace {
inline unsigned grid_for(size_t n, int per_block) { return (unsigned)((n + per_block - 1) / per_block); }
}

// sample_u/v/w of sh/g2p.wgsl:56-129: wall-aware trilinear gather of
// (final, saved) with renormalized weights.
__device__ __forceinline__ void sample_axis(const StepParams& P, int a, const float Pp[3],
                                            const float* __restrict__ vel, const float* __restrict__ sav,
                                            float& ofin, float& osav) {
    uint32_t dim[3];
    face_dims(P, a, dim);
    float g[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) g[c] = (Pp[c] - P.origin[c]) * P.inv_h + stagger(a, c);
    const int32_t bx = sat_i32(floorf(g[0])), by = sat_i32(floorf(g[1])), bz = sat_i32(floorf(g[2]));
    const float tx = g[0] - (float)bx, ty = g[1] - (float)by, tz = g[2] - (float)bz;
    const int32_t nx = (int32_t)P.nx, ny = (int32_t)P.ny, nz = (int32_t)P.nz;
    const int32_t lx = a == 0 ? 1 : 0, ly = a == 1 ? 1 : 0, lz = a == 2 ? 1 : 0;
    float fin = 0.0f, sv = 0.0f, wsum = 0.0f;
#pragma unroll
    for (int dk = 0; dk < 2; ++dk) {
        const int32_t kk = bz + dk;
        if (kk < 0 || kk >= (int32_t)dim[2]) continue;
        const float wz = dk == 1 ? tz : 1.0f - tz;
#pragma unroll
        for (int dj = 0; dj < 2; ++dj) {
            const int32_t jj = by + dj;
            if (jj < 0 || jj >= (int32_t)dim[1]) continue;
            const float wy = dj == 1 ? ty : 1.0f - ty;
#pragma unroll
            for (int di = 0; di < 2; ++di) {
                const int32_t ii = bx + di;
                if (ii < 0 || ii >= (int32_t)dim[0]) continue;
                // *_face_touches_static_solid (sh/g2p.wgsl:35-54)
                if (ii <= lx || ii >= nx - 1 || jj <= ly || jj >= ny - 1 || kk <= lz || kk >= nz - 1) continue;
                const float wx = di == 1 ? tx : 1.0f - tx;
                const float w = wx * wy * wz;
                const uint32_t idx = (uint32_t)ii + dim[0] * ((uint32_t)jj + dim[1] * (uint32_t)kk);
                fin += w * vel[idx];
                sv += w * sav[idx];
                wsum += w;
            }
        }
    }
    if (wsum > 0.0f) {
        ofin = fin / wsum;
        osav = sv / wsum;
    } else {
        ofin = 0.0f;
        osav = 0.0f;
    }
}

template <int BF16>
__global__ __launch_bounds__(256) void k_g2p(StepParams P, ParticleSoA sp, const float* __restrict__ u,
                                              const float* __restrict__ v, const float* __restrict__ w,
                                              const float* __restrict__ us, const float* __restrict__ vs,
                                              const float* __restrict__ ws) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P.N) return;
    const float h = P.h, dt = P.dt, alpha = P.flip;
    const float Pp[3] = {sp.pos[0][p], sp.pos[1][p], sp.pos[2][p]};
    const float pv[3] = {load_vel<BF16>(sp.vel[0], p), load_vel<BF16>(sp.vel[1], p), load_vel<BF16>(sp.vel[2], p)};
    float pic[3], sav[3];
    sample_axis(P, 0, Pp, u, us, pic[0], sav[0]);
    sample_axis(P, 1, Pp, v, vs, pic[1], sav[1]);
    sample_axis(P, 2, Pp, w, ws, pic[2], sav[2]);
    float nv[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float flip = pv[a] + (pic[a] - sav[a]);
        nv[a] = pic[a] * (1.0f - alpha) + flip * alpha;  // WGSL mix(pic, flip, alpha)
    }
    const float cfl = fmaxf(P.cfl, 1.0f);
    const float maxs = cfl * h / dt;
    const float sp2 = sqrtf(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
    if (sp2 > maxs && sp2 > 0.0f) {
        const float s = maxs / sp2;
#pragma unroll
        for (int a = 0; a < 3; ++a) nv[a] = nv[a] * s;
    }
    const float hband = h * 1.05f;
    const float ext[3] = {(float)P.nx * h, (float)P.ny * h, (float)P.nz * h};
    float lo[3], hi[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        lo[a] = P.origin[a] + hband;
        hi[a] = P.origin[a] + ext[a] - hband;
    }
    const float friction = P.friction;
    if (friction > 0.0f) {
        const float band = h * 2.5f;
        float nr[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const float dlo = (Pp[a] - lo[a]) / band;
            const float dhi = (hi[a] - Pp[a]) / band;
            nr[a] = clampf(1.0f - fminf(dlo, dhi), 0.0f, 1.0f);
        }
        nv[0] = nv[0] * (1.0f - friction * fmaxf(nr[1], nr[2]));
        nv[1] = nv[1] * (1.0f - friction * fmaxf(nr[0], nr[2]));
        nv[2] = nv[2] * (1.0f - friction * fmaxf(nr[0], nr[1]));
    }
    float np[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        np[a] = Pp[a] + nv[a] * dt;
        if (np[a] < lo[a]) { np[a] = lo[a]; nv[a] = 0.0f; }
        if (np[a] > hi[a]) { np[a] = hi[a]; nv[a] = 0.0f; }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        sp.pos[a][p] = np[a];
        store_vel<BF16>(sp.vel[a], p, nv[a]);
    }
}

hipError_t launch_g2p(const StepParams& P, const ParticleSoA& live, int bf16, const float* const vel[3],
                      const float* const saved[3], hipStream_t s) {
    if (P.N == 0) return hipSuccess;
    if (bf16)
        hipLaunchKernelGGL(k_g2p<1>, dim3(grid_for(P.N, 256)), dim3(256), 0, s, P, live, vel[0], vel[1], vel[2],
                           saved[0], saved[1], saved[2]);
    else
        hipLaunchKernelGGL(k_g2p<0>, dim3(grid_for(P.N, 256)), dim3(256), 0, s, P, live, vel[0], vel[1], vel[2],
                           saved[0], saved[1], saved[2]);
    return hipGetLastError();
}

template <int BF16>
__global__ void k_impulse(uint32_t n, ParticleSoA sp, float dx, float dy, float dz) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const float dv[3] = {dx, dy, dz};
#pragma unroll
    for (int a = 0; a < 3; ++a) store_vel<BF16>(sp.vel[a], p, load_vel<BF16>(sp.vel[a], p) + dv[a]);
}

hipError_t launch_impulse(const StepParams& P, const ParticleSoA& live, int bf16, float dx, float dy, float dz,
                          hipStream_t s) {
    if (P.N == 0) return hipSuccess;
    if (bf16) hipLaunchKernelGGL(k_impulse<1>, dim3(grid_for(P.N, 256)), dim3(256), 0, s, P.N, live, dx, dy, dz);
    else hipLaunchKernelGGL(k_impulse<0>, dim3(grid_for(P.N, 256)), dim3(256), 0, s, P.N, live, dx, dy, dz);
    return hipGetLastError();
}

}  // namespace fs
