// k_g2p.hip — grid-to-particle transfer with the FLIP/PIC blend, CFL clamp,
// wall friction, advection and wall clamp (sh/g2p.wgsl:131-198), and the
// uniform velocity impulse (sh/impulse.wgsl:26-36).
//
// Particles are read and written in place on the live ping-pong side (SoA).
// After the sort, consecutive lanes hold particles of the same or adjacent
// cells, so the 48 face loads per particle hit the same cache lines.
#include "fs_particle.h"

namespace fs {

namespace {
inline unsigned grid_for(size_t n, int per_block) { return (unsigned)((n + per_block - 1) / per_block); }
}

// With perm, particle p of sp is read from slot perm[p] of rd (a sort whose
// permutation was not applied yet): G2P moves the particles as it updates
// them.
template <int BF16>
__global__ __launch_bounds__(256) void k_g2p(StepParams P, ParticleSoA sp, ParticleSoA rd,
                                              const uint32_t* __restrict__ perm, const float* __restrict__ u,
                                              const float* __restrict__ v, const float* __restrict__ w,
                                              const float* __restrict__ us, const float* __restrict__ vs,
                                              const float* __restrict__ ws) {
    // Each XCD takes a contiguous eighth of the sorted particles, so the face
    // lines a region of particles reads are fetched into one XCD's L2.
    const uint32_t p = xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (p >= P.N) return;
    g2p_particle_from<BF16>(P, rd, perm ? ld_at(perm, p) : p, sp, p, u, v, w, us, vs, ws);
}

hipError_t launch_g2p(const StepParams& P, const ParticleSoA& live, int bf16, const float* const vel[3],
                      const float* const saved[3], hipStream_t s, const ParticleSoA* from, const uint32_t* perm) {
    if (P.N == 0) return hipSuccess;
    if ((from == nullptr) != (perm == nullptr)) return hipErrorInvalidValue;
    const ParticleSoA rd = from ? *from : live;
    const unsigned G = (grid_for(P.N, 256) + 7u) / 8u * 8u;  // whole XCD groups for xcd_remap
    if (bf16)
        FS_LAUNCH(k_g2p<1>, dim3(G), dim3(256), 0, s, P, live, rd, perm, vel[0], vel[1], vel[2], saved[0],
                  saved[1], saved[2]);
    else
        FS_LAUNCH(k_g2p<0>, dim3(G), dim3(256), 0, s, P, live, rd, perm, vel[0], vel[1], vel[2], saved[0],
                  saved[1], saved[2]);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Row-chunk G2P: one workgroup per run of up to kRowCells cells of one (y, z)
// row of the grid.  The particles whose cell lies in the chunk are one
// contiguous range [beg, end) of a cell order (the sort's permutation, or the
// unsorted substep's cell order) with bin starts `starts`.  Every face their
// stencils read lies in 21 face rows per field (u: y-1..y+1 x z-1..z+1; v:
// y..y+1 x z-1..z+1; w: y-1..y+1 x z..z+1, over x0-1..x1+1), so the workgroup
// copies those rows of vel and saved into LDS once and samples from there:
// about 5 global loads per particle instead of 24 face-pair gathers.  A wave
// in which some lane's stencil leaves the staged rows (a particle whose cell
// from inv_h differs from its sort cell, or one outside the tank) reads that
// axis from global memory instead.  The values and the arithmetic are those of
// sample_axis, so the results are the same bits.
namespace {
constexpr int kRowCells = 128;
constexpr int kRowPitch = kRowCells + 2;  // staged x-range: the chunk's cells + 1 (u) or + 2 (v, w)
constexpr int kRowRows = 21;              // staged rows per field
constexpr int kRowThreads = 256;
constexpr uint32_t kRowStageMin = 256;  // fewer particles: sample from global memory
}  // namespace

// Staged row r of a field: its axis and (y, z) offsets from that axis's first row.
__device__ __forceinline__ void row_of_stage(int r, int& a, int& ry, int& rz) {
    if (r < 9) { a = 0; ry = r % 3; rz = r / 3; }
    else if (r < 15) { a = 1; ry = (r - 9) & 1; rz = (r - 9) >> 1; }
    else { a = 2; ry = (r - 15) % 3; rz = (r - 15) / 3; }
}

template <int BF16>
__global__ __launch_bounds__(kRowThreads) void k_g2p_rows(StepParams P, ParticleSoA wr, ParticleSoA rd,
                                                          const uint32_t* __restrict__ order,
                                                          const uint32_t* __restrict__ starts,
                                                          const uint32_t* __restrict__ occ, int in_place,
                                                          const float* __restrict__ u, const float* __restrict__ v,
                                                          const float* __restrict__ w, const float* __restrict__ us,
                                                          const float* __restrict__ vs,
                                                          const float* __restrict__ ws) {
    __shared__ float s_f[2 * kRowRows * kRowPitch];
    const uint32_t nch = (P.nx + kRowCells - 1) / kRowCells;
    const uint32_t id = xcd_remap(blockIdx.x, gridDim.x);
    if (id >= nch * P.ny * P.nz) return;
    const int32_t x0 = (int32_t)(id % nch) * kRowCells;
    const int32_t cy = (int32_t)((id / nch) % P.ny), cz = (int32_t)(id / (nch * P.ny));
    const int32_t x1 = min(x0 + kRowCells, (int32_t)P.nx) - 1;
    const uint32_t rowc = P.nx * ((uint32_t)cy + P.ny * (uint32_t)cz);
    const uint32_t beg = starts[rowc + (uint32_t)x0];
    const uint32_t end = starts[rowc + (uint32_t)x1] + occ[rowc + (uint32_t)x1];
    if (beg >= end) return;
    const bool stage = end - beg >= kRowStageMin;
    const float* const fin[3] = {u, v, w};
    const float* const fsv[3] = {us, vs, ws};
    const int32_t ncell = x1 - x0 + 1;
    if (stage) {
        // The copy, by (field, axis): a fixed number of elements per thread,
        // each loaded unconditionally (an element outside the grid or the
        // chunk's x-range reads the array's first face instead and is stored
        // as 0), so every load is issued before the first LDS store waits.
        constexpr int kIt[3] = {(9 * kRowPitch + kRowThreads - 1) / kRowThreads,
                                (6 * kRowPitch + kRowThreads - 1) / kRowThreads,
                                (6 * kRowPitch + kRowThreads - 1) / kRowThreads};
        constexpr int kAll = kIt[0] + kIt[1] + kIt[2];
        float raw[2 * kAll];
        uint64_t keep = 0;
#pragma unroll
        for (int f = 0; f < 2; ++f) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                uint32_t dim[3];
                face_dims(P, a, dim);
                const int ry_n = a == 1 ? 2 : 3, rows = a == 0 ? 9 : 6;
                const float* src = f ? fsv[a] : fin[a];
#pragma unroll
                for (int j = 0; j < kIt[a]; ++j) {
                    const int k = f * kAll + (a == 0 ? 0 : a == 1 ? kIt[0] : kIt[0] + kIt[1]) + j;
                    const int el = (int)threadIdx.x + j * kRowThreads;
                    const int r = el / kRowPitch, ix = el - r * kRowPitch;
                    const int32_t gx = x0 - (a == 0 ? 0 : 1) + ix;
                    const int32_t gy = cy - (a == 1 ? 0 : 1) + r % ry_n;
                    const int32_t gz = cz - (a == 2 ? 0 : 1) + r / ry_n;
                    const bool ok = el < rows * kRowPitch && ix < ncell + (a == 0 ? 1 : 2) && gx >= 0 &&
                                    gx < (int32_t)dim[0] && gy >= 0 && gy < (int32_t)dim[1] && gz >= 0 &&
                                    gz < (int32_t)dim[2];
                    const uint32_t off =
                        ok ? (uint32_t)gx + dim[0] * ((uint32_t)gy + dim[1] * (uint32_t)gz) : 0u;
                    raw[k] = ld_at(src, off);
                    keep |= (uint64_t)ok << k;
                }
            }
        }
#pragma unroll
        for (int f = 0; f < 2; ++f) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const int rows = a == 0 ? 9 : 6, row0 = a == 0 ? 0 : a == 1 ? 9 : 15;
#pragma unroll
                for (int j = 0; j < kIt[a]; ++j) {
                    const int k = f * kAll + (a == 0 ? 0 : a == 1 ? kIt[0] : kIt[0] + kIt[1]) + j;
                    const int el = (int)threadIdx.x + j * kRowThreads;
                    if (el < rows * kRowPitch)
                        s_f[f * kRowRows * kRowPitch + row0 * kRowPitch + el] = ((keep >> k) & 1u) ? raw[k] : 0.0f;
                }
            }
        }
        __syncthreads();
    }
    // Software pipeline: the order entry two iterations ahead and the
    // particle one iteration ahead are loaded while this one is computed.
    uint32_t d = beg + threadIdx.x;
    uint32_t q = d < end ? ld_at(order, d) : 0u;
    uint32_t q_next = d + kRowThreads < end ? ld_at(order, d + kRowThreads) : 0u;
    float pp[3] = {0.0f, 0.0f, 0.0f}, pv[3] = {0.0f, 0.0f, 0.0f};
    if (d < end) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            pp[a] = ld_at(rd.pos[a], q);
            pv[a] = load_vel<BF16>(rd.vel[a], q);
        }
    }
    for (; d < end; d += kRowThreads) {
        G2PIn in;
        const uint32_t qc = q;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            in.pos[a] = pp[a];
            in.vel[a] = pv[a];
        }
        if (d + kRowThreads < end) {
            q = q_next;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                pp[a] = ld_at(rd.pos[a], q);
                pv[a] = load_vel<BF16>(rd.vel[a], q);
            }
        }
        if (d + 2 * kRowThreads < end) q_next = ld_at(order, d + 2 * kRowThreads);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            AxisTaps T;
            axis_taps(P, a, in.pos, T);
            // Stencil offsets in the staged rows (unsigned: below the first
            // row or column wraps to a large value).
            const int32_t ry_n = a == 1 ? 2 : 3, rz_n = a == 2 ? 2 : 3;
            const uint32_t ux = (uint32_t)T.bx - (uint32_t)(x0 - (a == 0 ? 0 : 1));
            const uint32_t uy = (uint32_t)T.by - (uint32_t)(cy - (a == 1 ? 0 : 1));
            const uint32_t uz = (uint32_t)T.bz - (uint32_t)(cz - (a == 2 ? 0 : 1));
            const bool inside = ux < (uint32_t)(ncell + (a == 0 ? 0 : 1)) && uy < (uint32_t)(ry_n - 1) &&
                                uz < (uint32_t)(rz_n - 1);
            float fv[8], sw[8];
            if (stage && __ballot(!inside) == 0ull) {
                const int row0 = (a == 0 ? 0 : a == 1 ? 9 : 15) + (int)uz * ry_n + (int)uy;
#pragma unroll
                for (int rw = 0; rw < 4; ++rw) {
                    const int i = (row0 + (rw >> 1) * ry_n + (rw & 1)) * kRowPitch + (int)ux;
                    fv[2 * rw] = s_f[i];
                    fv[2 * rw + 1] = s_f[i + 1];
                    sw[2 * rw] = s_f[kRowRows * kRowPitch + i];
                    sw[2 * rw + 1] = s_f[kRowRows * kRowPitch + i + 1];
                }
            } else {
                axis_load_global(P, a, T, fin[a], fsv[a], fv, sw);
            }
            axis_blend(T, fv, sw, in.pic[a], in.sav[a]);
        }
        g2p_finish<BF16>(P, in, wr, in_place ? qc : d);
    }
}

uint64_t g2p_row_chunks(const StepParams& P) {
    return (uint64_t)((P.nx + kRowCells - 1) / kRowCells) * P.ny * P.nz;
}

hipError_t launch_g2p_rows(const StepParams& P, const ParticleSoA& wr, const ParticleSoA& rd, const uint32_t* order,
                           const uint32_t* starts, const uint32_t* occ, int in_place, int bf16,
                           const float* const vel[3], const float* const saved[3], hipStream_t s) {
    if (P.N == 0) return hipSuccess;
    if (!order || !starts || !occ) return hipErrorInvalidValue;
    const uint64_t chunks = g2p_row_chunks(P);
    const uint64_t G = (chunks + 7u) / 8u * 8u;  // whole XCD groups for xcd_remap
    if (G > 0x7fffffffull) return hipErrorInvalidValue;
    if (bf16)
        FS_LAUNCH(k_g2p_rows<1>, dim3((unsigned)G), dim3(kRowThreads), 0, s, P, wr, rd, order, starts, occ,
                  in_place, vel[0], vel[1], vel[2], saved[0], saved[1], saved[2]);
    else
        FS_LAUNCH(k_g2p_rows<0>, dim3((unsigned)G), dim3(kRowThreads), 0, s, P, wr, rd, order, starts, occ,
                  in_place, vel[0], vel[1], vel[2], saved[0], saved[1], saved[2]);
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
    if (bf16) FS_LAUNCH(k_impulse<1>, dim3(grid_for(P.N, 256)), dim3(256), 0, s, P.N, live, dx, dy, dz);
    else FS_LAUNCH(k_impulse<0>, dim3(grid_for(P.N, 256)), dim3(256), 0, s, P.N, live, dx, dy, dz);
    return hipGetLastError();
}

}  // namespace fs
