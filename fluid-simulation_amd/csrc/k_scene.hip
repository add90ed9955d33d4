// k_scene.hip — initial particles of a scene on the GPU.
//
// generate_particles (src/gpu/fluid.rs:1833-1913) walks every block's
// lattice in order (ix, iy, iz innermost) and draws three numbers per
// particle from one linear congruential generator,
//     st <- 1664525 st + 1013904223 (mod 2^32),  rand01 = (st >> 8) / 2^24,
// seeded with 0x12345678, for the x, y and z jitter.  Particle g therefore
// uses draws 3g+1 .. 3g+3.  Here each thread jumps the generator straight
// to state 3g of its first particle (the affine map st -> a st + c composed
// with itself by squaring, as in F. Brown, "Random number generation with
// arbitrary strides", 1994), takes its three draws, and jumps by 3 S to its
// next particle g + S, so every thread reproduces the sequential stream bit
// for bit.  The lattice point, jitter and clamp are the reference's f32
// expressions in its order (no contraction).
#include <algorithm>

#include "fs_scene.h"

namespace fs {

namespace {
constexpr uint32_t kLcgA = 1664525u, kLcgC = 1013904223u, kLcgSeed = 0x12345678u;

__device__ __forceinline__ float clamp_ref(float x, float lo, float hi) {
    // f32::clamp: min first, then max.
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
}
}  // namespace

__device__ __forceinline__ uint32_t div_magic(uint32_t n, uint32_t m, uint32_t shift) {
    return (uint32_t)(((uint64_t)n * m) >> shift);
}

constexpr uint32_t kSeedBlock = 256;

// Affine map of n generator steps: st -> A st + C.
struct LcgMap {
    uint32_t A, C;
};
__device__ __forceinline__ LcgMap lcg_map(uint32_t n) {
    LcgMap m{1u, 0u};
    uint32_t a = kLcgA, c = kLcgC;
    while (n) {
        if (n & 1u) {
            m.A = a * m.A;
            m.C = a * m.C + c;
        }
        c = (a + 1u) * c;
        a = a * a;
        n >>= 1;
    }
    return m;
}

// A grid-stride loop over particles g = blockIdx.x * kSeedBlock + tid + k S
// (S = the launched thread count): few workgroups, each long-lived, so the
// launch is not bound by the rate at which workgroups are dispatched.  A
// thread's next particle starts 3 S draws after its current one: one
// uniform affine map.  The loop runs per wave (its bound is the wave's first
// particle), so the lanes of a wave iterate together: a full wave's 64
// particles are consecutive, their 192 floats one contiguous run, staged in
// the wave's own LDS slice and stored as three 256-byte rows; the last,
// partial wave stores its live lanes directly.  Waves never wait on each
// other.
__global__ __launch_bounds__(kSeedBlock) void k_seed_lattice(SeedLayout L, float* __restrict__ out) {
    __shared__ float s_xyz[kSeedBlock / 64][192];
    float* wbuf = s_xyz[threadIdx.x >> 6];
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t S = gridDim.x * kSeedBlock;
    // The jump to the workgroup's first particle depends on blockIdx only
    // (uniform); each thread then jumps 3 * threadIdx.x further.
    const LcgMap to_block = lcg_map(3u * (blockIdx.x * kSeedBlock));
    const LcgMap to_thread = lcg_map(3u * threadIdx.x);
    const LcgMap next = lcg_map(3u * S);
    uint32_t pre = to_thread.A * (to_block.A * kLcgSeed + to_block.C) + to_thread.C;
    uint32_t b = 0;
    for (uint32_t gw = blockIdx.x * kSeedBlock + (threadIdx.x & ~63u); gw < L.total;
         gw += S, pre = next.A * pre + next.C) {
        const uint32_t g = gw + lane;
        float p[3] = {0.0f, 0.0f, 0.0f};
        if (g < L.total) {
            while (b + 1u < L.nblocks && g >= L.blocks[b + 1u].first) ++b;
            const SeedBlock& B = L.blocks[b];
            const uint32_t l = g - B.first;
            const uint32_t t = div_magic(l, B.mz, B.sz);   // l / counts[2]
            const uint32_t ix = div_magic(t, B.my, B.sy);  // t / counts[1]
            const uint32_t idx[3] = {ix, t - ix * B.counts[1], l - t * B.counts[2]};
            uint32_t st = pre;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                st = kLcgA * st + kLcgC;
                const float r = (float)(st >> 8) / 16777216.0f;
                const float j = ((r - 0.5f) * B.spacing) * 0.5f;
                p[a] = clamp_ref((B.wmin[a] + ((float)idx[a] + 0.5f) * B.spacing) + j, L.lo[a], L.hi[a]);
            }
        }
        if (gw + 64u <= L.total) {
            wbuf[3u * lane + 0u] = p[0];
            wbuf[3u * lane + 1u] = p[1];
            wbuf[3u * lane + 2u] = p[2];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            float* row = out + 3u * gw;
            row[lane] = wbuf[lane];
            row[64u + lane] = wbuf[64u + lane];
            row[128u + lane] = wbuf[128u + lane];
            // The next round's staging writes follow these reads.
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        } else if (g < L.total) {
            out[3u * g + 0u] = p[0];
            out[3u * g + 1u] = p[1];
            out[3u * g + 2u] = p[2];
        }
    }
}

hipError_t launch_seed_lattice(const SeedLayout& layout, float* pos_xyz, hipStream_t s) {
    if (layout.total == 0) return hipSuccess;
    // Eight workgroups per CU at most; small lattices take one thread each.
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const uint32_t need = (layout.total + kSeedBlock - 1u) / kSeedBlock;
    const uint32_t grid = std::min(need, (uint32_t)(8 * cus));
    hipLaunchKernelGGL(k_seed_lattice, dim3(grid), dim3(kSeedBlock), 0, s, layout, pos_xyz);
    return hipGetLastError();
}

}  // namespace fs
