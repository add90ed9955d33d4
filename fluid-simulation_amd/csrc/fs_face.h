// fs_face.h — per-face and per-particle device functions shared by the
// reference-structured kernels and the fused kernels, so both evaluate the
// same expressions in the same order.
#pragma once
#include "fs_device.h"

namespace fs {

__device__ __forceinline__ void face_ijk(uint32_t f, const uint32_t d[3], uint32_t& i, uint32_t& j,
                                         uint32_t& k) {
    i = f % d[0];
    j = (f / d[0]) % d[1];
    k = f / (d[0] * d[1]);
}

__device__ __forceinline__ uint32_t axis_n(const StepParams& P, int a) {
    return a == 0 ? P.nx : (a == 1 ? P.ny : P.nz);
}

// Face is wet if its lo or hi in-range cell is Liquid (sh/forces.wgsl:56-73).
// cell(ci, cj, ck) returns the type of an in-grid cell.
template <typename CellFn>
__device__ __forceinline__ bool face_wet_by(const StepParams& P, int a, uint32_t i, uint32_t j, uint32_t k,
                                            CellFn&& cell) {
    const uint32_t comp = a == 0 ? i : (a == 1 ? j : k);
    bool liquid = false;
    if (comp > 0u) {
        uint32_t ci = i, cj = j, ck = k;
        if (a == 0) ci -= 1u; else if (a == 1) cj -= 1u; else ck -= 1u;
        if (cell(ci, cj, ck) == kLiquid) liquid = true;
    }
    if (comp < axis_n(P, a)) {
        if (cell(i, j, k) == kLiquid) liquid = true;
    }
    return liquid;
}

// Face touches a Solid cell or the domain edge (sh/boundaries.wgsl:39-58).
template <typename CellFn>
__device__ __forceinline__ bool face_solid_by(const StepParams& P, int a, uint32_t i, uint32_t j, uint32_t k,
                                              CellFn&& cell) {
    const uint32_t comp = a == 0 ? i : (a == 1 ? j : k);
    bool lo_solid = comp == 0u;
    bool hi_solid = comp == axis_n(P, a);
    if (!lo_solid) {
        uint32_t ci = i, cj = j, ck = k;
        if (a == 0) ci -= 1u; else if (a == 1) cj -= 1u; else ck -= 1u;
        if (cell(ci, cj, ck) == kSolid) lo_solid = true;
    }
    if (!hi_solid) {
        uint32_t ci = i < P.nx - 1u ? i : P.nx - 1u;
        uint32_t cj = j < P.ny - 1u ? j : P.ny - 1u;
        uint32_t ck = k < P.nz - 1u ? k : P.nz - 1u;
        if (cell(ci, cj, ck) == kSolid) hi_solid = true;
    }
    return lo_solid || hi_solid;
}

// normalize -> save -> forces -> enforce for one face (sh/normalize.wgsl:21-26,
// sh/save_vel.wgsl:18, sh/forces.wgsl:75-83, sh/boundaries.wgsl:58-60).
template <typename CellFn>
__device__ __forceinline__ void face_epilogue_by(const StepParams& P, int a, uint32_t i, uint32_t j, uint32_t k,
                                                 int32_t num, int32_t den, float& vel, float& saved, CellFn&& cell) {
    float v = den > 0 ? (float)num / (float)den : 0.0f;
    saved = v;
    if (face_wet_by(P, a, i, j, k, cell)) v = v + grav_axis(P, a) * P.dt;
    if (face_solid_by(P, a, i, j, k, cell)) v = 0.0f;
    vel = v;
}

// The same with the cell types read from the grid's array.
__device__ __forceinline__ bool face_wet(const StepParams& P, int a, const uint8_t* __restrict__ ct,
                                         uint32_t i, uint32_t j, uint32_t k) {
    return face_wet_by(P, a, i, j, k,
                       [&](uint32_t ci, uint32_t cj, uint32_t ck) { return ct[ci + P.nx * (cj + P.ny * ck)]; });
}
__device__ __forceinline__ bool face_solid(const StepParams& P, int a, const uint8_t* __restrict__ ct,
                                           uint32_t i, uint32_t j, uint32_t k) {
    return face_solid_by(P, a, i, j, k,
                         [&](uint32_t ci, uint32_t cj, uint32_t ck) { return ct[ci + P.nx * (cj + P.ny * ck)]; });
}
__device__ __forceinline__ void face_epilogue(const StepParams& P, int a, const uint8_t* __restrict__ ct,
                                              uint32_t i, uint32_t j, uint32_t k, int32_t num, int32_t den,
                                              float& vel, float& saved) {
    face_epilogue_by(P, a, i, j, k, num, den, vel, saved,
                     [&](uint32_t ci, uint32_t cj, uint32_t ck) { return ct[ci + P.nx * (cj + P.ny * ck)]; });
}

// Gradient subtraction for one face (sh/gradient.wgsl:41-66); returns the
// updated velocity (unchanged when the face is skipped).
__device__ __forceinline__ float face_gradient(const StepParams& P, int a, const float* __restrict__ pr,
                                               const uint8_t* __restrict__ ct, uint32_t i, uint32_t j,
                                               uint32_t k, float vel) {
    const uint32_t comp = a == 0 ? i : (a == 1 ? j : k);
    if (comp == 0u || comp == axis_n(P, a)) return vel;
    const uint32_t hi = i + P.nx * (j + P.ny * k);
    const uint32_t lo = a == 0 ? hi - 1u : (a == 1 ? hi - P.nx : hi - P.nx * P.ny);
    const uint8_t lt = ct[lo], ht = ct[hi];
    if (lt == kSolid || ht == kSolid) return vel;
    if (lt != kLiquid && ht != kLiquid) return vel;
    const float coeff = (P.dt / P.rho) / P.h;
    return vel - coeff * (pr[hi] - pr[lo]);
}

// The 8 P2G taps of one particle on one axis (sh/scatter.wgsl:52-75), in the
// reference's loop order; add(ii, jj, kk, num_term, den_term) per kept tap.
template <typename AddFn>
__device__ __forceinline__ void p2g_taps(const float g[3], const uint32_t dim[3], float vcomp, float scale,
                                         AddFn&& add) {
    const int32_t bx = sat_i32(floorf(g[0])), by = sat_i32(floorf(g[1])), bz = sat_i32(floorf(g[2]));
    const float tx = g[0] - (float)bx, ty = g[1] - (float)by, tz = g[2] - (float)bz;
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
                const float wx = di == 1 ? tx : 1.0f - tx;
                const float w = wx * wy * wz;
                if (w <= 0.0f) continue;
                add(ii, jj, kk, sat_i32(rintf(w * vcomp * scale)), sat_i32(rintf(w * scale)));
            }
        }
    }
}

template <int BF16>
__device__ __forceinline__ float load_vel(const void* p, uint32_t i) {
    if (BF16) return bf16_to_f32(ld_at((const uint16_t*)p, i));
    return ld_at((const float*)p, i);
}
template <int BF16>
__device__ __forceinline__ void store_vel(void* p, uint32_t i, float v) {
    if (BF16) st_at((uint16_t*)p, i, f32_to_bf16(v));
    else st_at((float*)p, i, v);
}

// Map block b of a G-block launch so that the blocks one XCD runs (b % 8)
// cover a contiguous range of logical ids (blocks go to XCDs round-robin).
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t G) {
    const uint32_t X = 8;
    if (G % X != 0u) return b;
    return (b % X) * (G / X) + b / X;
}

// Splits an index of the concatenated u|v|w face space into (axis, face).
__device__ __forceinline__ bool split_face(const StepParams& P, uint32_t g, int& a, uint32_t& f) {
    if (g < P.F[0]) { a = 0; f = g; return true; }
    g -= P.F[0];
    if (g < P.F[1]) { a = 1; f = g; return true; }
    g -= P.F[1];
    if (g < P.F[2]) { a = 2; f = g; return true; }
    return false;
}

}  // namespace fs
