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
};

struct SeedLayout {
    SeedBlock blocks[FS_SCENE_MAX_BLOCKS];
    uint32_t nblocks;
    uint32_t total;
    float lo[3], hi[3];  // clamp box: origin + 1.05 h .. origin + extent - 1.05 h
};

// Writes layout.total xyz triplets to pos_xyz (device).
hipError_t launch_seed_lattice(const SeedLayout& layout, float* pos_xyz, hipStream_t s);

}  // namespace fs
