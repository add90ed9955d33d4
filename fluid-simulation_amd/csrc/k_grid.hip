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
#include <algorithm>
#include <cstdlib>
#include "fs_cell.h"
#include "fs_gather.h"

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
        const uint8_t t = classify_cell(P, occ, c);
        ct[c] = t;
        liquid = t == kLiquid;
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
    FS_LAUNCH(k_classify, dim3(grid_for(P.C, 1024)), dim3(1024), 0, s, P, occ, ct, stats);
    return hipGetLastError();
}

// sh/scatter.wgsl:78-118: per particle, three axes, global i32 atomics.
template <int BF16>
__global__ void k_p2g_atomic(StepParams P, ParticleSoA sp, int32_t* __restrict__ un, int32_t* __restrict__ ud,
                             int32_t* __restrict__ vn, int32_t* __restrict__ vd, int32_t* __restrict__ wn,
                             int32_t* __restrict__ wd, const uint32_t* __restrict__ guard, uint32_t above) {
    if (guard && *guard <= above) return;
    int32_t* const nums[3] = {un, vn, wn};
    int32_t* const dens[3] = {ud, vd, wd};
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < P.N; p += stride)
        p2g_atomic_particle<BF16>(P, sp, p, nums, dens);
}

// A guarded launch is the fused engine's fallback and is usually a no-op, so
// it uses a small grid-stride grid; unguarded launches take one particle per
// thread like sh/scatter.wgsl.
hipError_t launch_p2g_atomic(const StepParams& P, const ParticleSoA& src, int bf16, int32_t* const num[3],
                             int32_t* const den[3], const uint32_t* guard, uint32_t above, hipStream_t s) {
    if (P.N == 0) return hipSuccess;
    unsigned g = grid_for(P.N, 64);
    if (guard && g > 4096) g = 4096;
    if (bf16)
        FS_LAUNCH(k_p2g_atomic<1>, dim3(g), dim3(64), 0, s, P, src, num[0], den[0], num[1], den[1], num[2],
                           den[2], guard, above);
    else
        FS_LAUNCH(k_p2g_atomic<0>, dim3(g), dim3(64), 0, s, P, src, num[0], den[0], num[1], den[1], num[2],
                           den[2], guard, above);
    return hipGetLastError();
}

// sh/scatter_local.wgsl:149-198: the reference's P2G on every substep when
// the particle sort is enabled (src/gpu/fluid.rs:519-530, 1572-1584).  A
// 64-lane workgroup accumulates its particles' taps in a workgroup-local
// open-addressed table (kLocalCap slots, linear probing, key = buffer <<
// shift | face index, stored +1 so that 0 marks an empty slot), then adds each
// occupied slot to global memory with one atomic; a tap that finds the table
// full goes to global memory directly.  All sums are wrapping i32 adds, so
// num/den equal the per-tap atomic form's bit for bit.
constexpr uint32_t kLocalCap = 1024;  // CAP of sh/scatter_local.wgsl:49

__device__ __forceinline__ void add_global6(uint32_t buf, uint32_t idx, int32_t v, int32_t* un, int32_t* ud,
                                            int32_t* vn, int32_t* vd, int32_t* wn, int32_t* wd) {
    switch (buf) {
        case 0: atomicAdd(&un[idx], v); break;
        case 1: atomicAdd(&ud[idx], v); break;
        case 2: atomicAdd(&vn[idx], v); break;
        case 3: atomicAdd(&vd[idx], v); break;
        case 4: atomicAdd(&wn[idx], v); break;
        default: atomicAdd(&wd[idx], v); break;
    }
}

template <int BF16>
__global__ __launch_bounds__(64) void k_p2g_local(StepParams P, ParticleSoA sp, int32_t* __restrict__ un,
                                                  int32_t* __restrict__ ud, int32_t* __restrict__ vn,
                                                  int32_t* __restrict__ vd, int32_t* __restrict__ wn,
                                                  int32_t* __restrict__ wd, uint32_t shift) {
    __shared__ uint32_t s_key[kLocalCap];
    __shared__ int32_t s_val[kLocalCap];
    for (uint32_t i = threadIdx.x; i < kLocalCap; i += 64) {
        s_key[i] = 0u;
        s_val[i] = 0;
    }
    __syncthreads();
    auto local_add = [&](uint32_t buf, uint32_t idx, int32_t v) {
        const uint32_t key = (buf << shift) | idx;
        const uint32_t k1 = key + 1u;
        uint32_t h = key * 2654435761u;  // Knuth's multiplicative hash, then xor-shift
        h = (h ^ (h >> 15)) & (kLocalCap - 1u);
        for (uint32_t probe = 0; probe < kLocalCap; ++probe) {
            const uint32_t cur = __hip_atomic_load(&s_key[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (cur == k1) {
                atomicAdd(&s_val[h], v);
                return;
            }
            if (cur == 0u) {
                const uint32_t old = atomicCAS(&s_key[h], 0u, k1);
                if (old == 0u || old == k1) {
                    atomicAdd(&s_val[h], v);
                    return;
                }
            }
            h = (h + 1u) & (kLocalCap - 1u);
        }
        add_global6(buf, idx, v, un, ud, vn, vd, wn, wd);
    };
    const uint32_t p = blockIdx.x * 64u + threadIdx.x;
    if (p < P.N) {
        const float pos[3] = {(sp.pos[0][p] - P.origin[0]) * P.inv_h, (sp.pos[1][p] - P.origin[1]) * P.inv_h,
                              (sp.pos[2][p] - P.origin[2]) * P.inv_h};
        const float pv[3] = {load_vel<BF16>(sp.vel[0], p), load_vel<BF16>(sp.vel[1], p),
                             load_vel<BF16>(sp.vel[2], p)};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            uint32_t dim[3];
            face_dims(P, a, dim);
            const float g[3] = {pos[0] + stagger(a, 0), pos[1] + stagger(a, 1), pos[2] + stagger(a, 2)};
            p2g_taps(g, dim, pv[a], P.scale, [&](int32_t ii, int32_t jj, int32_t kk, int32_t tn, int32_t td) {
                const uint32_t idx = (uint32_t)ii + dim[0] * ((uint32_t)jj + dim[1] * (uint32_t)kk);
                local_add(2u * a, idx, tn);
                local_add(2u * a + 1u, idx, td);
            });
        }
    }
    __syncthreads();
    const uint32_t mask = (1u << shift) - 1u;
    for (uint32_t i = threadIdx.x; i < kLocalCap; i += 64) {
        const uint32_t k1 = s_key[i];
        if (k1 == 0u) continue;
        const int32_t v = s_val[i];
        if (v == 0) continue;
        const uint32_t key = k1 - 1u;
        add_global6(key >> shift, key & mask, v, un, ud, vn, vd, wn, wd);
    }
}

bool p2g_local_supported(const StepParams& P) {
    uint32_t fmax = P.F[0] > P.F[1] ? P.F[0] : P.F[1];
    fmax = fmax > P.F[2] ? fmax : P.F[2];
    uint32_t shift = 0;
    while ((1ull << shift) < fmax) ++shift;
    // six buffers << shift, plus one, must stay below 2^32
    return shift <= 29u;
}

hipError_t launch_p2g_local(const StepParams& P, const ParticleSoA& src, int bf16, int32_t* const num[3],
                            int32_t* const den[3], hipStream_t s) {
    if (P.N == 0) return hipSuccess;
    uint32_t fmax = P.F[0] > P.F[1] ? P.F[0] : P.F[1];
    fmax = fmax > P.F[2] ? fmax : P.F[2];
    uint32_t shift = 0;
    while ((1ull << shift) < fmax) ++shift;
    if (shift > 29u) return hipErrorInvalidValue;
    const unsigned g = grid_for(P.N, 64);
    if (bf16)
        FS_LAUNCH(k_p2g_local<1>, dim3(g), dim3(64), 0, s, P, src, num[0], den[0], num[1], den[1], num[2], den[2],
                  shift);
    else
        FS_LAUNCH(k_p2g_local<0>, dim3(g), dim3(64), 0, s, P, src, num[0], den[0], num[1], den[1], num[2], den[2],
                  shift);
    return hipGetLastError();
}

// Clears num and den of the three axes when *guard > above (the fused
// engine's fallback to the atomic P2G), or always when guard is null: one
// launch for the six arrays.
__global__ void k_clear_faces_guarded(StepParams P, int32_t* __restrict__ n0, int32_t* __restrict__ n1,
                                      int32_t* __restrict__ n2, int32_t* __restrict__ d0, int32_t* __restrict__ d1,
                                      int32_t* __restrict__ d2, const uint32_t* __restrict__ guard, uint32_t above) {
    if (guard && *guard <= above) return;
    const uint32_t tot = P.F[0] + P.F[1] + P.F[2];
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t gi = blockIdx.x * blockDim.x + threadIdx.x; gi < tot; gi += stride) {
        int a;
        uint32_t f;
        split_face(P, gi, a, f);
        (a == 0 ? n0 : (a == 1 ? n1 : n2))[f] = 0;
        (a == 0 ? d0 : (a == 1 ? d1 : d2))[f] = 0;
    }
}

hipError_t launch_clear_faces_guarded(const StepParams& P, int32_t* const num[3], int32_t* const den[3],
                                      const uint32_t* guard, uint32_t above, hipStream_t s) {
    const size_t tot = (size_t)P.F[0] + P.F[1] + P.F[2];
    // Guarded (usually a no-op): a grid small enough to be resident at once.
    const unsigned need = grid_for(tot, 256), g = guard && need > 1024u ? 1024u : need;
    FS_LAUNCH(k_clear_faces_guarded, dim3(g), dim3(256), 0, s, P, num[0], num[1], num[2], den[0], den[1], den[2],
              guard, above);
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
    if (f >= faces_axis(P, a)) return;
    uint32_t d[3], i, j, k;
    face_dims(P, a, d);
    face_ijk(f, d, i, j, k);
    if (face_wet(P, a, ct, i, j, k)) vel[f] = vel[f] + grav_axis(P, a) * P.dt;
}

__global__ void k_enforce(StepParams P, int a, const uint8_t* __restrict__ ct, float* __restrict__ vel) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= faces_axis(P, a)) return;
    uint32_t d[3], i, j, k;
    face_dims(P, a, d);
    face_ijk(f, d, i, j, k);
    if (face_solid(P, a, ct, i, j, k)) vel[f] = 0.0f;
}

__global__ void k_gradient(StepParams P, int a, const float* __restrict__ pr, const uint8_t* __restrict__ ct,
                           float* __restrict__ vel) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= faces_axis(P, a)) return;
    uint32_t d[3], i, j, k;
    face_dims(P, a, d);
    face_ijk(f, d, i, j, k);
    vel[f] = face_gradient(P, a, pr, ct, i, j, k, vel[f]);
}

hipError_t launch_normalize(const StepParams& P, int a, const int32_t* num, const int32_t* den, float* vel,
                            hipStream_t s) {
    FS_LAUNCH(k_normalize, dim3(grid_for(P.F[a], 64)), dim3(64), 0, s, P.F[a], num, den, vel);
    return hipGetLastError();
}
hipError_t launch_copy_f32(const float* src, float* dst, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    FS_LAUNCH(k_copy, dim3(grid_for(n, 64)), dim3(64), 0, s, src, dst, n);
    return hipGetLastError();
}
hipError_t launch_forces(const StepParams& P, int a, const uint8_t* ct, float* vel, hipStream_t s) {
    FS_LAUNCH(k_forces, dim3(grid_for(P.F[a], 64)), dim3(64), 0, s, P, a, ct, vel);
    return hipGetLastError();
}
hipError_t launch_enforce(const StepParams& P, int a, const uint8_t* ct, float* vel, hipStream_t s) {
    FS_LAUNCH(k_enforce, dim3(grid_for(P.F[a], 64)), dim3(64), 0, s, P, a, ct, vel);
    return hipGetLastError();
}
hipError_t launch_gradient(const StepParams& P, int a, const float* p, const uint8_t* ct, float* vel,
                           hipStream_t s) {
    FS_LAUNCH(k_gradient, dim3(grid_for(P.F[a], 64)), dim3(64), 0, s, P, a, p, ct, vel);
    return hipGetLastError();
}

__global__ void k_divergence(StepParams P, const float* __restrict__ u, const float* __restrict__ v,
                             const float* __restrict__ w, const uint8_t* __restrict__ ct,
                             const uint32_t* __restrict__ occ, float* __restrict__ div) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= P.C) return;
    div[c] = ct[c] != kLiquid ? 0.0f : cell_div(P, u, v, w, occ, c);
}

hipError_t launch_divergence(const StepParams& P, float* const vel[3], const uint8_t* ct, const uint32_t* occ,
                             float* div, hipStream_t s) {
    FS_LAUNCH(k_divergence, dim3(grid_for(P.C, 64)), dim3(64), 0, s, P, vel[0], vel[1], vel[2], ct,
                       occ, div);
    return hipGetLastError();
}

// ---- fused face passes over the concatenated u|v|w face index space ----
struct FaceIntsC {
    const int32_t* num[3];
    const int32_t* den[3];
};
struct FaceFloats {
    float* p[3];
};

__global__ void k_face_epilogue(StepParams P, const int32_t* __restrict__ n0, const int32_t* __restrict__ n1,
                                const int32_t* __restrict__ n2, const int32_t* __restrict__ d0,
                                const int32_t* __restrict__ d1, const int32_t* __restrict__ d2,
                                const uint8_t* __restrict__ ct, float* __restrict__ v0, float* __restrict__ v1,
                                float* __restrict__ v2, float* __restrict__ s0, float* __restrict__ s1,
                                float* __restrict__ s2, const uint32_t* __restrict__ guard, uint32_t above) {
    if (guard && *guard <= above) return;
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
                                uint32_t above, hipStream_t s) {
    const size_t tot = (size_t)P.F[0] + P.F[1] + P.F[2];
    unsigned g = grid_for(tot, 256);
    // Usually a no-op fallback: a grid-stride grid small enough to be resident
    // at once, so the idle launch is one round of waves reading the guard.
    if (guard && g > 1024) g = 1024;
    FS_LAUNCH(k_face_epilogue, dim3(g), dim3(256), 0, s, P, num[0], num[1], num[2], den[0], den[1], den[2],
                       ct, vel[0], vel[1], vel[2], saved[0], saved[1], saved[2], guard, above);
    return hipGetLastError();
}

// The face epilogue of every face, one row of faces (fixed axis, j, k) per
// workgroup: the face's (i, j, k) come from the launch shape (one division
// per workgroup) instead of per-face divisions.
constexpr int kRowThreads = 128;
__global__ __launch_bounds__(kRowThreads) void k_face_epilogue_rows(StepParams P, FaceIntsC fi,
                                                                     const uint8_t* __restrict__ ct, FaceFloats vel,
                                                                     FaceFloats sav) {
    // The axis and its face grid by selects, not by indexing the argument
    // arrays with a run-time value (which would place them in scratch).
    const uint32_t rows0 = P.ny * P.nz, rows1 = (P.ny + 1u) * P.nz;
    uint32_t row = blockIdx.x;
    const int a = row < rows0 ? 0 : (row < rows0 + rows1 ? 1 : 2);
    row -= a == 0 ? 0u : (a == 1 ? rows0 : rows0 + rows1);
    const uint32_t d0 = a == 0 ? P.nx + 1u : P.nx, d1 = a == 1 ? P.ny + 1u : P.ny;
    const uint32_t j = row % d1, k = row / d1;
    const int32_t* __restrict__ num = pick3(a, fi.num[0], fi.num[1], fi.num[2]);
    const int32_t* __restrict__ den = pick3(a, fi.den[0], fi.den[1], fi.den[2]);
    float* __restrict__ v = pick3(a, vel.p[0], vel.p[1], vel.p[2]);
    float* __restrict__ sv = pick3(a, sav.p[0], sav.p[1], sav.p[2]);
    const float g = grav_axis(P, a);
    for (uint32_t i = threadIdx.x; i < d0; i += kRowThreads) {
        const uint32_t f = i + d0 * row;
        const int32_t dd = den[f];
        float vv = dd > 0 ? (float)num[f] / (float)dd : 0.0f;
        sv[f] = vv;
        if (face_wet(P, a, ct, i, j, k)) vv = vv + g * P.dt;
        if (face_solid(P, a, ct, i, j, k)) vv = 0.0f;
        v[f] = vv;
    }
}

hipError_t launch_face_epilogue_rows(const StepParams& P, const int32_t* const num[3], const int32_t* const den[3],
                                     const uint8_t* ct, float* const vel[3], float* const saved[3], hipStream_t s) {
    uint32_t rows = 0;
    for (int a = 0; a < 3; ++a) {
        const uint32_t d1 = a == 1 ? P.ny + 1u : P.ny, d2 = a == 2 ? P.nz + 1u : P.nz;
        rows += d1 * d2;
    }
    const FaceIntsC fi = {{num[0], num[1], num[2]}, {den[0], den[1], den[2]}};
    const FaceFloats fv = {{vel[0], vel[1], vel[2]}}, fs = {{saved[0], saved[1], saved[2]}};
    FS_LAUNCH(k_face_epilogue_rows, dim3(rows), dim3(kRowThreads), 0, s, P, fi, ct, fv, fs);
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
    FS_LAUNCH(k_gradient_enforce, dim3(grid_for(tot, 256)), dim3(256), 0, s, P, p, ct, vel[0], vel[1],
                       vel[2]);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Sorted-segment P2G gather (fs_gather.h), one workgroup per brick.  When more
// particles were listed as misfits than the gather walks, it exits and the
// atomic form runs instead.
template <int BF16, int V, int THREADS>
__global__ __launch_bounds__(THREADS) void k_p2g_gather(StepParams P, ParticleSoA sp, ParticleSoA src,
                                                           const uint32_t* __restrict__ occ,
                                                           const uint32_t* __restrict__ off,
                                                           const uint8_t* __restrict__ ct, GatherOut out,
                                                           uint32_t nbx, uint32_t nby, uint32_t nbz,
                                                           const uint32_t* __restrict__ mis_count,
                                                           const uint32_t* __restrict__ mis_list,
                                                           const uint32_t* __restrict__ perm,
                                                           const uint32_t* __restrict__ border,
                                                           uint32_t* __restrict__ weight) {
    const uint32_t mis_n = *mis_count;
    if (mis_n > kMisfitCap) return;
    const uint32_t G = nbx * nby * nbz;
    const uint32_t pos = xcd_remap(blockIdx.x, G);
    const uint32_t b = border ? border[pos] : pos;
    const uint32_t bxi = b % nbx, byi = (b / nbx) % nby, bzi = b / (nbx * nby);
    p2g_gather_brick<BF16, V, THREADS>(P, sp, src, occ, off, ct, out, (int32_t)bxi * Brick<V>::X, (int32_t)byi * Brick<V>::Y,
                              (int32_t)bzi * Brick<V>::Z, mis_list, mis_n, 0u, P.N, perm,
                              weight ? weight + b : nullptr);
}

// Dispatch order of the gather's bricks.  A brick's work is the particles of
// its halo, and where the liquid lies only part of the bricks have any: in
// brick order the loaded ones come late and the empty ones hold workgroup
// slots meanwhile.  Within the contiguous share of bricks that each XCD takes
// (xcd_remap), the bricks are put in order of decreasing halo particle count,
// in kOrderBuckets classes relative to the share's largest, empty ones last.
// The counts are those the previous gather found (each workgroup records its
// brick's; the liquid moves by less than a cell per substep), so the order
// costs one small kernel.  It only changes which workgroup runs a brick when:
// every brick writes its own faces, and its sums are integers, so stale or
// unset counts cost time, never results.
// Measured at 128^3 (4096 bricks, about 1536 resident at a time): 238 ->
// 222 us.  At 256^3 (32768 bricks) the brick order measured faster (1.15
// against 1.19 ms): with many rounds of workgroups the tail matters less
// than neighbouring bricks sharing their halo particles in L2 while they
// run together, so the order is used up to kOrderShareMax bricks per share.
constexpr int kOrderThreads = 1024;
constexpr uint32_t kOrderBuckets = 32;
constexpr uint32_t kOrderShareMax = 1024;  // bricks per XCD share (8192 bricks in all)

// One workgroup per XCD share of `share` bricks: a counting sort of the
// share's bricks into classes of decreasing weight (order within a class is
// whatever the LDS atomics give).
__global__ __launch_bounds__(kOrderThreads) void k_brick_order(const uint32_t* __restrict__ weight, uint32_t share,
                                                              uint32_t* __restrict__ order) {
    __shared__ uint32_t s_max;
    __shared__ uint32_t s_cnt[kOrderBuckets];
    const uint32_t c0 = blockIdx.x * share;
    if (threadIdx.x == 0) s_max = 0;
    if (threadIdx.x < kOrderBuckets) s_cnt[threadIdx.x] = 0;
    __syncthreads();
    uint32_t m = 0;
    for (uint32_t i = threadIdx.x; i < share; i += kOrderThreads) m = max(m, weight[c0 + i]);
    for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor(m, o, 64));
    if ((threadIdx.x & 63u) == 0) atomicMax(&s_max, m);
    __syncthreads();
    const uint64_t span = (uint64_t)s_max + 1u;
    // Heavier bricks in lower classes; empty ones in the last.
    auto bucket = [&](uint32_t w) -> uint32_t {
        if (w == 0) return kOrderBuckets - 1u;
        return (kOrderBuckets - 2u) - (uint32_t)((uint64_t)w * (kOrderBuckets - 1u) / span);
    };
    for (uint32_t i = threadIdx.x; i < share; i += kOrderThreads) atomicAdd(&s_cnt[bucket(weight[c0 + i])], 1u);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (uint32_t k = 0; k < kOrderBuckets; ++k) {
            const uint32_t c = s_cnt[k];
            s_cnt[k] = run;
            run += c;
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < share; i += kOrderThreads)
        order[c0 + atomicAdd(&s_cnt[bucket(weight[c0 + i])], 1u)] = c0 + i;
}

namespace {
template <int V>
hipError_t launch_gather_v(const StepParams& P, const ParticleSoA& sorted, const ParticleSoA& src, int bf16,
                           const uint32_t* occ, const uint32_t* cell_offset, const uint8_t* ct, const GatherOut& out,
                           const uint32_t* mis_count, const uint32_t* mis_list, const uint32_t* perm,
                           uint32_t* brick_ws, hipStream_t s) {
    const uint32_t nbx = (P.nx + Brick<V>::X - 1) / Brick<V>::X, nby = (P.ny + Brick<V>::Y - 1) / Brick<V>::Y,
                   nbz = (P.nz + Brick<V>::Z - 1) / Brick<V>::Z;
    const uint32_t G = nbx * nby * nbz;
    if (G <= kGSmallGrid) {
        // Every workgroup of a grid this small is resident at once.
        if (bf16)
            FS_LAUNCH((k_p2g_gather<1, V, kGThreadsSmall>), dim3(G), dim3(kGThreadsSmall), 0, s, P, sorted, src,
                               occ, cell_offset, ct, out, nbx, nby, nbz, mis_count, mis_list, perm, nullptr, nullptr);
        else
            FS_LAUNCH((k_p2g_gather<0, V, kGThreadsSmall>), dim3(G), dim3(kGThreadsSmall), 0, s, P, sorted, src,
                               occ, cell_offset, ct, out, nbx, nby, nbz, mis_count, mis_list, perm, nullptr, nullptr);
        return hipGetLastError();
    }
    const uint32_t* border = nullptr;
    uint32_t* weight = nullptr;
    if (brick_ws && G % 8u == 0 && G / 8u <= kOrderShareMax) {
        weight = brick_ws;
        uint32_t* order = brick_ws + G;
        FS_LAUNCH(k_brick_order, dim3(8), dim3(kOrderThreads), 0, s, weight, G / 8u, order);
        border = order;
    }
    if (bf16) {
        FS_LAUNCH((k_p2g_gather<1, V, kGThreadsLarge>), dim3(G), dim3(kGThreadsLarge), 0, s, P, sorted, src,
                           occ, cell_offset, ct, out, nbx, nby, nbz, mis_count, mis_list, perm, border, weight);
    } else {
        FS_LAUNCH((k_p2g_gather<0, V, kGThreadsLarge>), dim3(G), dim3(kGThreadsLarge), 0, s, P, sorted, src,
                           occ, cell_offset, ct, out, nbx, nby, nbz, mis_count, mis_list, perm, border, weight);
    }
    return hipGetLastError();
}
}  // namespace

size_t gather_brick_count_max(const StepParams& P) {
    auto nb = [&](uint32_t X, uint32_t Y, uint32_t Z) {
        return (size_t)((P.nx + X - 1) / X) * ((P.ny + Y - 1) / Y) * ((P.nz + Z - 1) / Z);
    };
    return std::max(nb(Brick<0>::X, Brick<0>::Y, Brick<0>::Z),
                    std::max(nb(Brick<1>::X, Brick<1>::Y, Brick<1>::Z), nb(Brick<2>::X, Brick<2>::Y, Brick<2>::Z)));
}

// Particle-owned brick scatter (fs_gather.h), one workgroup per 8^3 brick;
// num/den must be cleared before it and the face epilogue run after it.
struct FaceInts {
    int32_t* num[3];
    int32_t* den[3];
};

template <int BF16, int THREADS>
__global__ __launch_bounds__(THREADS) void k_p2g_bricks(StepParams P, ParticleSoA sp, const uint32_t* __restrict__ occ,
                                                         const uint32_t* __restrict__ off, FaceInts fi, uint32_t nbx,
                                                         uint32_t nby, uint32_t nbz,
                                                         const uint32_t* __restrict__ perm) {
    const uint32_t G = nbx * nby * nbz;
    const uint32_t b = xcd_remap(blockIdx.x, G);
    const uint32_t bxi = b % nbx, byi = (b / nbx) % nby, bzi = b / (nbx * nby);
    p2g_scatter_brick<BF16, THREADS>(P, sp, occ, off, fi.num, fi.den, (int32_t)bxi * kSB, (int32_t)byi * kSB,
                                     (int32_t)bzi * kSB, perm);
}

namespace {
template <int THREADS>
hipError_t launch_bricks_t(const StepParams& P, const ParticleSoA& sorted, int bf16, const uint32_t* occ,
                           const uint32_t* cell_offset, const FaceInts& fi, uint32_t nbx, uint32_t nby, uint32_t nbz,
                           const uint32_t* perm, hipStream_t s) {
    const uint32_t G = nbx * nby * nbz;
    if (bf16)
        FS_LAUNCH((k_p2g_bricks<1, THREADS>), dim3(G), dim3(THREADS), 0, s, P, sorted, occ, cell_offset, fi, nbx, nby,
                  nbz, perm);
    else
        FS_LAUNCH((k_p2g_bricks<0, THREADS>), dim3(G), dim3(THREADS), 0, s, P, sorted, occ, cell_offset, fi, nbx, nby,
                  nbz, perm);
    return hipGetLastError();
}
}  // namespace

hipError_t launch_p2g_bricks(const StepParams& P, const ParticleSoA& sorted, int bf16, const uint32_t* occ,
                             const uint32_t* cell_offset, int32_t* const num[3], int32_t* const den[3],
                             const uint32_t* perm, hipStream_t s) {
    const uint32_t nbx = (P.nx + kSB - 1) / kSB, nby = (P.ny + kSB - 1) / kSB, nbz = (P.nz + kSB - 1) / kSB;
    const FaceInts fi = {{num[0], num[1], num[2]}, {den[0], den[1], den[2]}};
    int threads = nbx * nby * nbz <= kGSmallGrid ? kSThreadsSmall : kSThreads;
    // FLUIDSTEP_P2G_THREADS=256|512|1024 fixes the workgroup size (read at
    // every launch).
    if (const char* e = std::getenv("FLUIDSTEP_P2G_THREADS")) {
        const int v = std::atoi(e);
        if (v == 256 || v == 512 || v == 1024) threads = v;
    }
    if (threads == 1024) return launch_bricks_t<1024>(P, sorted, bf16, occ, cell_offset, fi, nbx, nby, nbz, perm, s);
    if (threads == 512) return launch_bricks_t<512>(P, sorted, bf16, occ, cell_offset, fi, nbx, nby, nbz, perm, s);
    return launch_bricks_t<256>(P, sorted, bf16, occ, cell_offset, fi, nbx, nby, nbz, perm, s);
}

hipError_t launch_p2g_gather(const StepParams& P, const ParticleSoA& sorted, const ParticleSoA& src, int bf16,
                             const uint32_t* occ, const uint32_t* cell_offset, const uint8_t* ct, float* const vel[3],
                             float* const saved[3], const uint32_t* mis_count, const uint32_t* mis_list,
                             hipStream_t s, int brick_variant, const uint32_t* perm,
                             uint32_t* brick_ws) {
    const GatherOut out = {{vel[0], vel[1], vel[2]}, {saved[0], saved[1], saved[2]}};
    switch (brick_variant) {
        case 0: return launch_gather_v<0>(P, sorted, src, bf16, occ, cell_offset, ct, out, mis_count, mis_list, perm,
                                              brick_ws, s);
        case 1: return launch_gather_v<1>(P, sorted, src, bf16, occ, cell_offset, ct, out, mis_count, mis_list, perm,
                                              brick_ws, s);
        case 2: return launch_gather_v<2>(P, sorted, src, bf16, occ, cell_offset, ct, out, mis_count, mis_list, perm,
                                              brick_ws, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace fs
