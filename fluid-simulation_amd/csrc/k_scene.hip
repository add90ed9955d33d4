// k_scene.hip — initial particles of a scene on the GPU.
//
// generate_particles (src/gpu/fluid.rs:1833-1913) walks every block's
// lattice in order (ix, iy, iz innermost) and draws three numbers per
// particle from one linear congruential generator,
//     st <- 1664525 st + 1013904223 (mod 2^32),  rand01 = (st >> 8) / 2^24,
// seeded with 0x12345678, for the x, y and z jitter.  Particle g therefore
// uses draws 3g+1 .. 3g+3.  Here one thread per particle jumps the generator
// straight to state 3g (the affine map st -> a st + c composed with itself
// by squaring, as in F. Brown, "Random number generation with arbitrary
// strides", 1994) and takes its three draws, so every thread reproduces the
// sequential stream bit for bit.  The lattice point, jitter and clamp are
// the reference's f32 expressions in its order (no contraction).
#include "fs_scene.h"

namespace fs {

namespace {
constexpr uint32_t kLcgA = 1664525u, kLcgC = 1013904223u, kLcgSeed = 0x12345678u;

// State after n steps from st.
__device__ __forceinline__ uint32_t lcg_jump(uint32_t st, uint32_t n) {
    uint32_t A = 1u, C = 0u, a = kLcgA, c = kLcgC;
    while (n) {
        if (n & 1u) {
            A = a * A;
            C = a * C + c;
        }
        c = (a + 1u) * c;
        a = a * a;
        n >>= 1;
    }
    return A * st + C;
}

__device__ __forceinline__ float clamp_ref(float x, float lo, float hi) {
    // f32::clamp: min first, then max.
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
}
}  // namespace

__global__ __launch_bounds__(256) void k_seed_lattice(SeedLayout L, float* __restrict__ out) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= L.total) return;
    uint32_t b = 0;
    while (b + 1u < L.nblocks && g >= L.blocks[b + 1u].first) ++b;
    const SeedBlock& B = L.blocks[b];
    const uint32_t l = g - B.first;
    const uint32_t cz = B.counts[2], cyz = B.counts[1] * cz;
    const uint32_t idx[3] = {l / cyz, (l / cz) % B.counts[1], l % cz};
    uint32_t st = lcg_jump(kLcgSeed, 3u * g);
    float p[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        st = kLcgA * st + kLcgC;
        const float r = (float)(st >> 8) / 16777216.0f;
        const float j = ((r - 0.5f) * B.spacing) * 0.5f;
        p[a] = clamp_ref((B.wmin[a] + ((float)idx[a] + 0.5f) * B.spacing) + j, L.lo[a], L.hi[a]);
    }
    out[3u * g + 0u] = p[0];
    out[3u * g + 1u] = p[1];
    out[3u * g + 2u] = p[2];
}

hipError_t launch_seed_lattice(const SeedLayout& layout, float* pos_xyz, hipStream_t s) {
    if (layout.total == 0) return hipSuccess;
    hipLaunchKernelGGL(k_seed_lattice, dim3((layout.total + 255u) / 256u), dim3(256), 0, s, layout, pos_xyz);
    return hipGetLastError();
}

}  // namespace fs
