// fs_scene.h — the lattice layout that generate_particles
// (src/gpu/fluid.rs:1833-1913) lays into the liquid blocks, as the seeding
// kernel takes it.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "fluidstep.h"

namespace fs {

// One block's lattice: counts[0] x counts[1] x counts[2] points at
// wmin + (i + 0.5) * spacing, particles first .. first + count - 1 of the
// whole layout (ix outermost, iz innermost).
struct SeedBlock {
    float wmin[3];
    float spacing;
    uint32_t counts[3];
    uint32_t first;
    // floor(n / counts[2]) and floor(n / counts[1]) for n < 2^31 as
    // (n * m) >> shift (magic_u31).
    uint32_t mz, sz, my, sy;
};

// Multiplier and shift with floor(n / d) == (n * m) >> shift for every
// n < 2^31 (d >= 1): l = ceil(log2 d), m = floor(2^(31+l) / d) + 1, shift =
// 31 + l (T. Granlund, P. Montgomery, "Division by invariant integers using
// multiplication", 1994, Theorem 4.2 with N = 31); m < 2^32.
inline void magic_u31(uint32_t d, uint32_t& m, uint32_t& shift) {
    uint32_t l = 0;
    while ((1ull << l) < d) ++l;
    m = (uint32_t)(((uint64_t)1 << (31 + l)) / d + 1u);
    shift = 31 + l;
}

struct SeedLayout {
    SeedBlock blocks[FS_SCENE_MAX_BLOCKS];
    uint32_t nblocks;
    uint32_t total;
    float lo[3], hi[3];  // clamp box: origin + 1.05 h .. origin + extent - 1.05 h
};

// Writes layout.total xyz triplets to pos_xyz (device).
hipError_t launch_seed_lattice(const SeedLayout& layout, float* pos_xyz, hipStream_t s);

}  // namespace fs
