// fs_particle.h — per-particle device functions (G2P and atomic P2G) shared
// by the single-tank and the batched kernels.
#pragma once
#include "fs_face.h"
#include "fs_kernels.h"

namespace fs {

// Two f32 at 4-byte alignment (gfx950 loads them with one 8-byte access).
typedef float f32x2_a4 __attribute__((ext_vector_type(2), aligned(4)));

// sample_u/v/w of sh/g2p.wgsl:56-129: wall-aware trilinear gather of
// (final, saved) with renormalized weights, in three parts: the taps
// (axis_taps), the eight (final, saved) face values (loaded by the caller,
// from global memory or from a staged copy), and the blend (axis_blend).
//
// Tap (di, dj, dk) is used when its face is in the grid and does not touch a
// static solid (*_face_touches_static_solid, sh/g2p.wgsl:35-54).  Both tests
// are per dimension: 0 <= i < dim and !(i <= l || i >= n-1) reduce to
// l < i < n-1.  All loads are issued first (a skipped row is loaded from an
// in-bounds place), then the taps are accumulated in the loop order of the
// reference; a skipped tap leaves the sums unchanged (a select, not an added
// zero), so the result equals the branchy form bit for bit.
struct AxisTaps {
    int32_t bx, by, bz;  // base face of the 2x2x2 stencil
    float tx, ty, tz;
    bool ok[8];          // tap t = (t & 1, (t >> 1) & 1, t >> 2) used
};

__device__ __forceinline__ void axis_taps(const StepParams& P, int a, const float Pp[3], AxisTaps& T) {
    float g[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) g[c] = (Pp[c] - P.origin[c]) * P.inv_h + stagger(a, c);
    T.bx = sat_i32(floorf(g[0]));
    T.by = sat_i32(floorf(g[1]));
    T.bz = sat_i32(floorf(g[2]));
    T.tx = g[0] - (float)T.bx;
    T.ty = g[1] - (float)T.by;
    T.tz = g[2] - (float)T.bz;
    const int32_t nx = (int32_t)P.nx, ny = (int32_t)P.ny, nz = (int32_t)P.nz;
    const int32_t lx = a == 0 ? 1 : 0, ly = a == 1 ? 1 : 0, lz = a == 2 ? 1 : 0;
    bool okx[2], oky[2], okz[2];
#pragma unroll
    for (int d = 0; d < 2; ++d) {
        okx[d] = T.bx + d > lx && T.bx + d < nx - 1;
        oky[d] = T.by + d > ly && T.by + d < ny - 1;
        okz[d] = T.bz + d > lz && T.bz + d < nz - 1;
    }
#pragma unroll
    for (int rw = 0; rw < 4; ++rw) {
        T.ok[2 * rw] = okx[0] && oky[rw & 1] && okz[rw >> 1];
        T.ok[2 * rw + 1] = okx[1] && oky[rw & 1] && okz[rw >> 1];
    }
}

__device__ __forceinline__ void axis_blend(const AxisTaps& T, const float fv[8], const float sw[8], float& ofin,
                                           float& osav) {
    const float wx[2] = {1.0f - T.tx, T.tx}, wy[2] = {1.0f - T.ty, T.ty}, wz[2] = {1.0f - T.tz, T.tz};
    float wxy[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) wxy[t] = wx[t & 1] * wy[t >> 1];
    float fin = 0.0f, sv = 0.0f, wsum = 0.0f;
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        const float w = wxy[t & 3] * wz[t >> 2];  // wx * wy * wz
        fin = T.ok[t] ? fin + w * fv[t] : fin;
        sv = T.ok[t] ? sv + w * sw[t] : sv;
        wsum = T.ok[t] ? wsum + w : wsum;
    }
    if (wsum > 0.0f) {
        ofin = fin / wsum;
        osav = sv / wsum;
    } else {
        ofin = 0.0f;
        osav = 0.0f;
    }
}

// The eight face values of the stencil from global memory.  The two x-taps of
// a (y, z) row are adjacent faces: one 8-byte load each from vel and sav.  A
// row is loaded when either of its taps is used; then both of its faces are
// in the grid (a used tap at bx + 1 means bx >= l >= 0, one at bx means
// bx + 1 <= n - 1).  Offsets are 32-bit bytes from the array base.
__device__ __forceinline__ void axis_load_global(const StepParams& P, int a, const AxisTaps& T,
                                                 const float* __restrict__ vel, const float* __restrict__ sav,
                                                 float fv[8], float sw[8]) {
    uint32_t dim[3];
    face_dims(P, a, dim);
    const uint32_t base = (uint32_t)T.bx + dim[0] * ((uint32_t)T.by + dim[1] * (uint32_t)T.bz);
    const uint32_t sy = dim[0], sz = dim[0] * dim[1];
#pragma unroll
    for (int rw = 0; rw < 4; ++rw) {
        const int dj = rw & 1, dk = rw >> 1;
        const bool any = T.ok[2 * rw] || T.ok[2 * rw + 1];
        // A mask rather than a select, so the offset stays straight-line code.
        uint32_t off = ((base + sy * (uint32_t)dj + sz * (uint32_t)dk) * 4u) & (0u - (uint32_t)any);
        // Opaque, so that each load takes it as a 32-bit offset on the array's
        // scalar base (the compiler otherwise forms 64-bit addresses here).
        asm volatile("" : "+v"(off));
        const f32x2_a4 va = *reinterpret_cast<const f32x2_a4*>(reinterpret_cast<const char*>(vel) + off);
        const f32x2_a4 vb = *reinterpret_cast<const f32x2_a4*>(reinterpret_cast<const char*>(sav) + off);
        fv[2 * rw] = va.x;
        fv[2 * rw + 1] = va.y;
        sw[2 * rw] = vb.x;
        sw[2 * rw + 1] = vb.y;
    }
}

__device__ __forceinline__ void sample_axis(const StepParams& P, int a, const float Pp[3],
                                            const float* __restrict__ vel, const float* __restrict__ sav,
                                            float& ofin, float& osav) {
    AxisTaps T;
    axis_taps(P, a, Pp, T);
    float fv[8], sw[8];
    axis_load_global(P, a, T, vel, sav, fv, sw);
    axis_blend(T, fv, sw, ofin, osav);
}

// What G2P reads for one particle: its position and velocity and the
// sampled (final, saved) grid velocity per axis.
struct G2PIn {
    float pos[3], vel[3], pic[3], sav[3];
};

// The loads and samples of G2P for the particle in slot q of rd.
template <int BF16>
__device__ __forceinline__ void g2p_load(const StepParams& P, const ParticleSoA& rd, uint32_t q,
                                         const float* __restrict__ u, const float* __restrict__ v,
                                         const float* __restrict__ w, const float* __restrict__ us,
                                         const float* __restrict__ vs, const float* __restrict__ ws, G2PIn& in) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        in.pos[a] = ld_at(rd.pos[a], q);
        in.vel[a] = load_vel<BF16>(rd.vel[a], q);
    }
    sample_axis(P, 0, in.pos, u, us, in.pic[0], in.sav[0]);
    sample_axis(P, 1, in.pos, v, vs, in.pic[1], in.sav[1]);
    sample_axis(P, 2, in.pos, w, ws, in.pic[2], in.sav[2]);
}

// FLIP/PIC blend + CFL clamp + wall friction + advection + wall clamp
// (sh/g2p.wgsl:131-198) of a loaded particle, written to slot p of sp.
template <int BF16>
__device__ __forceinline__ void g2p_finish(const StepParams& P, const G2PIn& in, const ParticleSoA& sp,
                                           uint32_t p) {
    const float h = P.h, dt = P.dt, alpha = P.flip;
    const float* Pp = in.pos;
    float nv[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float flip = in.vel[a] + (in.pic[a] - in.sav[a]);
        nv[a] = in.pic[a] * (1.0f - alpha) + flip * alpha;  // WGSL mix(pic, flip, alpha)
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
        st_at(sp.pos[a], p, np[a]);
        store_vel<BF16>(sp.vel[a], p, nv[a]);
    }
}

// G2P of one particle, read from slot q of rd and written to slot p of sp.
template <int BF16>
__device__ __forceinline__ void g2p_particle_from(const StepParams& P, const ParticleSoA& rd, uint32_t q,
                                                  const ParticleSoA& sp, uint32_t p, const float* __restrict__ u,
                                                  const float* __restrict__ v, const float* __restrict__ w,
                                                  const float* __restrict__ us, const float* __restrict__ vs,
                                                  const float* __restrict__ ws) {
    G2PIn in;
    g2p_load<BF16>(P, rd, q, u, v, w, us, vs, ws, in);
    g2p_finish<BF16>(P, in, sp, p);
}

template <int BF16>
__device__ __forceinline__ void g2p_particle(const StepParams& P, const ParticleSoA& sp, uint32_t p,
                                             const float* __restrict__ u, const float* __restrict__ v,
                                             const float* __restrict__ w, const float* __restrict__ us,
                                             const float* __restrict__ vs, const float* __restrict__ ws) {
    g2p_particle_from<BF16>(P, sp, p, sp, p, u, v, w, us, vs, ws);
}

// One particle of sh/scatter.wgsl:78-118 with global i32 atomics.
template <int BF16>
__device__ __forceinline__ void p2g_atomic_particle(const StepParams& P, const ParticleSoA& sp, uint32_t p,
                                                    int32_t* const nums[3], int32_t* const dens[3]) {
    const float pos[3] = {(sp.pos[0][p] - P.origin[0]) * P.inv_h, (sp.pos[1][p] - P.origin[1]) * P.inv_h,
                          (sp.pos[2][p] - P.origin[2]) * P.inv_h};
    const float pv[3] = {load_vel<BF16>(sp.vel[0], p), load_vel<BF16>(sp.vel[1], p), load_vel<BF16>(sp.vel[2], p)};
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

// Sort key of particle p plus whether its P2G cell differs from it.
__device__ __forceinline__ uint32_t particle_key(const StepParams& P, float x, float y, float z, bool& misfit) {
    uint32_t ki, kj, kk;
    const uint32_t key = cell_ijk(P, x, y, z, ki, kj, kk);
    misfit = floorf((x - P.origin[0]) * P.inv_h) != (float)ki || floorf((y - P.origin[1]) * P.inv_h) != (float)kj ||
             floorf((z - P.origin[2]) * P.inv_h) != (float)kk;
    return key;
}

// Wave-level list append of the lanes with mis set: every such lane's index p
// goes to list[slot] for slot < cap; *count (and *tank_count, nullable) grow
// by the number of such lanes.  All lanes of the wave must call it.
__device__ __forceinline__ void list_misfits(bool mis, uint32_t p, uint32_t* count, uint32_t* list, uint32_t cap,
                                             uint32_t* tank_count) {
    const uint64_t mm = __ballot(mis);
    if (!mm) return;
    const uint32_t lane = lane_id();
    uint32_t base = 0;
    if (lane == 0) {
        base = atomicAdd(count, (uint32_t)__popcll(mm));
        if (tank_count) atomicAdd(tank_count, (uint32_t)__popcll(mm));
    }
    base = __shfl(base, 0, 64);
    if (mis && list) {
        const uint32_t slot = base + (uint32_t)__popcll(mm & lanemask_lt());
        if (slot < cap) list[slot] = p;
    }
}

// occ[key] += 1 for every valid lane, one atomic per run of equal keys among
// the wave's lanes (sorted particles come in runs); the counts are integers,
// so the result equals one atomic per particle.  All lanes must call it.
__device__ __forceinline__ void hist_key_runs(uint32_t key, bool valid, uint32_t* occ) {
    const uint32_t lane = lane_id();
    const uint32_t prev = __shfl_up(key, 1, 64);
    const bool head = valid && (lane == 0 || prev != key);
    const uint64_t heads = __ballot(head);
    const uint64_t act = __ballot(valid);
    if (head) {
        const uint64_t later = lane == 63 ? 0ull : (heads & (~0ull << (lane + 1)));
        const uint32_t end = later ? (uint32_t)__builtin_ctzll(later) : (uint32_t)(64 - __builtin_clzll(act));
        atomicAdd(&occ[key], end - lane);
    }
}

}  // namespace fs
