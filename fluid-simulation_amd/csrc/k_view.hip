// k_view.hip — the reference's buffer layouts, packed from this library's
// for a consumer that expects them: the 32-byte particle record
// {pos.xyz, 0, vel.xyz, 0} (src/gpu/fluid.rs:276-280) and u32 cell types.
// Both are bandwidth kernels: one pass over the source, coalesced stores.
#include "fs_face.h"
#include "fs_kernels.h"

namespace fs {

namespace {
constexpr int kViewBlock = 256;
}

// Lane i of a wave writes record i as two 16-byte halves, so a wave's stores
// cover 2 KB contiguously.
template <int BF16>
__global__ __launch_bounds__(kViewBlock) void k_pack_particles_vec4(ParticleSoA sp, uint32_t n, float4* __restrict__ dst) {
    const uint32_t i = blockIdx.x * kViewBlock + threadIdx.x;
    if (i >= n) return;
    const float4 lo = make_float4(sp.pos[0][i], sp.pos[1][i], sp.pos[2][i], 0.0f);
    const float4 hi = make_float4(load_vel<BF16>(sp.vel[0], i), load_vel<BF16>(sp.vel[1], i),
                                  load_vel<BF16>(sp.vel[2], i), 0.0f);
    dst[2u * i] = lo;
    dst[2u * i + 1u] = hi;
}

__global__ __launch_bounds__(kViewBlock) void k_widen_u8(const uint8_t* __restrict__ src, uint32_t n,
                                                        uint32_t* __restrict__ dst) {
    const uint32_t i = blockIdx.x * kViewBlock + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

hipError_t launch_pack_particles_vec4(const ParticleSoA& sp, uint32_t n, bool bf16, float* dst, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const dim3 grid((n + kViewBlock - 1) / kViewBlock);
    if (bf16)
        hipLaunchKernelGGL(k_pack_particles_vec4<1>, grid, dim3(kViewBlock), 0, s, sp, n, reinterpret_cast<float4*>(dst));
    else
        hipLaunchKernelGGL(k_pack_particles_vec4<0>, grid, dim3(kViewBlock), 0, s, sp, n, reinterpret_cast<float4*>(dst));
    return hipGetLastError();
}

hipError_t launch_widen_u8(const uint8_t* src, uint32_t n, uint32_t* dst, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_widen_u8, dim3((n + kViewBlock - 1) / kViewBlock), dim3(kViewBlock), 0, s, src, n, dst);
    return hipGetLastError();
}

}  // namespace fs
