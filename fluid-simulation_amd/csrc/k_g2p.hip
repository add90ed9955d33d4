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
