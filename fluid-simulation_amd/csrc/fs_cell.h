// fs_cell.h — per-cell device functions (classification, divergence, CG
// set-up) shared by the single-tank and the batched kernels.
#pragma once
#include "fs_cg.h"

namespace fs {

// sh/classify.wgsl:25-61: solid boundary shell, liquid where the occupancy
// reaches the threshold (or, with dilation, a face neighbour's does), air
// otherwise.
__device__ __forceinline__ uint8_t classify_cell(const StepParams& P, const uint32_t* __restrict__ occ, uint32_t c) {
    const uint32_t nx = P.nx, ny = P.ny, nz = P.nz;
    const uint32_t i = c % nx, j = (c / nx) % ny, k = c / (nx * ny);
    if (i == 0u || j == 0u || k == 0u || i == nx - 1u || j == ny - 1u || k == nz - 1u) return kSolid;
    const uint32_t t = P.thresh;
    bool liquid = occ[c] >= t;
    if (!liquid && P.dilation)
        liquid = occ[c - 1u] >= t || occ[c + 1u] >= t || occ[c - nx] >= t || occ[c + nx] >= t ||
                 occ[c - nx * ny] >= t || occ[c + nx * ny] >= t;
    return liquid ? kLiquid : kAir;
}

// sh/divergence.wgsl:22-59 for a liquid cell, with the volume-drift term.
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

// Divergence (sh/divergence.wgsl) and CG initialisation (sh/cg_init.wgsl:24-48)
// of cell c, plus its stencil mask.  Returns whether the cell is liquid.
__device__ __forceinline__ bool div_cginit_cell(const StepParams& P, const float* __restrict__ u,
                                                const float* __restrict__ v, const float* __restrict__ w,
                                                const uint8_t* __restrict__ ct, const uint32_t* __restrict__ occ,
                                                uint32_t c, float* __restrict__ div, float* __restrict__ b,
                                                float* __restrict__ p, float* __restrict__ r, float* __restrict__ d,
                                                uint16_t* __restrict__ nmask) {
    if (ct[c] != kLiquid) {
        div[c] = 0.0f;
        b[c] = 0.0f;
        p[c] = 0.0f;
        r[c] = 0.0f;
        d[c] = 0.0f;
        nmask[c] = 0;
        return false;
    }
    const uint32_t nx = P.nx, nxny = P.nx * P.ny;
    const uint32_t nb[6] = {c - 1u, c + 1u, c - nx, c + nx, c - nxny, c + nxny};
    uint16_t m = kMaskLiquid;
    uint32_t cnt = 0;
    float fcnt = 0.0f, sum = 0.0f;
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        const uint8_t t = ct[nb[s]];
        if (t != kSolid) {
            ++cnt;
            fcnt += 1.0f;
            if (t == kLiquid) {
                m |= (uint16_t)(1u << (3 + s));
                if (P.warm) sum += p[nb[s]];
            }
        }
    }
    m |= (uint16_t)cnt;
    const float dv = cell_div(P, u, v, w, occ, c);
    const float scale = P.rho * P.h * P.h / P.dt;
    const float rhs = -scale * dv;
    float ap = 0.0f;
    if (P.warm) ap = fcnt * p[c] - sum;
    else p[c] = 0.0f;
    const float rc = rhs - ap;
    div[c] = dv;
    b[c] = rhs;
    r[c] = rc;
    d[c] = rc;
    nmask[c] = m;
    return true;
}

}  // namespace fs
