// fs_scene.cpp — scene presets, spawn counts and the seeded lattice
// (src/scene/mod.rs, src/gpu/fluid.rs:1792-1913), host side.  Every float
// expression keeps the reference's f32 (or f64) type and operation order, so
// the blocks, counts and lattice parameters equal the reference's bit for
// bit; the particles themselves are laid on the GPU (k_scene.hip).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "fluidstep.h"
#include "fs_host_common.h"
#include "fs_scene.h"

using namespace fs;

namespace {
constexpr float kDefaultDropHeight = 0.72f;  // DEFAULT_DROP_HEIGHT (src/scene/mod.rs:11)
constexpr float kDefaultFillLevel = 0.2f;    // DEFAULT_FILL_LEVEL (:22)
constexpr float kReferenceDensity = 8.0f;    // REFERENCE_DENSITY (:15)
constexpr float kH = 2.0f / 64.0f;           // sim::H (src/sim/mod.rs:26)

float clampf_ref(float x, float lo, float hi) {  // f32::clamp
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
}

struct V3 {
    float x, y, z;
};
V3 vmax(V3 a, V3 b) { return {std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z)}; }
V3 vmin(V3 a, V3 b) { return {std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z)}; }
V3 vclamp(V3 v, V3 lo, V3 hi) { return vmin(vmax(v, lo), hi); }  // glam Vec3::clamp
V3 splat(float s) { return {s, s, s}; }
V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 mul(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
V3 scale(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }

struct Block {
    V3 min, max;
};
Block make_block(V3 lo, V3 hi) { return {lo, hi}; }

// scale_block_axes (src/scene/mod.rs:253-264)
Block scale_block_axes(Block b, V3 factor) {
    const V3 center = scale(0.5f, add(b.min, b.max));
    const V3 half = scale(0.5f, sub(b.max, b.min));
    const V3 scaled = mul(half, vmax(factor, splat(1.0e-3f)));
    b.min = vclamp(sub(center, scaled), splat(0.0f), splat(1.0f));
    b.max = vclamp(add(center, scaled), add(b.min, splat(1.0e-3f)), splat(1.0f));
    b.min = vmax(vmin(b.min, sub(b.max, splat(1.0e-3f))), splat(0.0f));
    return b;
}

// shift_block_y (:242-247)
Block shift_block_y(Block b, float delta) {
    const float shift = clampf_ref(delta, -b.min.y, 1.0f - b.max.y);
    b.min.y += shift;
    b.max.y += shift;
    return b;
}

// preset_blocks (:208-238)
std::vector<Block> preset_blocks(uint32_t preset, float drop_height, float fill_level) {
    const float fill = clampf_ref(fill_level, 0.0f, 1.0f);
    std::vector<Block> out;
    if (preset == FS_SCENE_DAM_BREAK) {
        const float ceiling = 0.98f;
        const float top = clampf_ref(fill * ceiling, 1.0e-3f, ceiling);
        out.push_back(make_block({0.05f, 0.0f, 0.05f}, {0.42f, top, 0.95f}));
    } else if (preset == FS_SCENE_DOUBLE_SPLASH) {
        const Block drops[2] = {make_block({0.1f, 0.45f, 0.3f}, {0.34f, 0.92f, 0.7f}),
                                make_block({0.66f, 0.45f, 0.3f}, {0.9f, 0.92f, 0.7f})};
        for (const Block& d : drops) {
            const Block s = shift_block_y(d, clampf_ref(drop_height, 0.0f, 1.0f) - kDefaultDropHeight);
            // scale_suspended_drop (:249-251)
            out.push_back(scale_block_axes(s, {1.0f, fill / kDefaultFillLevel, 1.0f}));
        }
    } else {  // FallingBlob, and ScenePreset::from_u32's default
        const float delta = clampf_ref(drop_height, 0.0f, 1.0f) - kDefaultDropHeight;
        const Block blob = make_block({0.19f, 0.2f, 0.19f}, {0.81f, 0.72f, 0.81f});
        // scale_block_about_center (:237-240)
        const float factor = std::pow(std::fmax(fill, 1.0e-4f) / kDefaultFillLevel, 1.0f / 3.0f);
        out.push_back(scale_block_axes(shift_block_y(blob, delta), splat(factor)));
    }
    return out;
}

// seeded_volume_fraction (:134-146): f32 sum of the blocks' normalized volumes.
float seeded_volume_fraction(const fs_liquid_block* b, uint32_t n) {
    float s = 0.0f;
    for (uint32_t i = 0; i < n; ++i) {
        const float ex = b[i].max[0] - b[i].min[0], ey = b[i].max[1] - b[i].min[1], ez = b[i].max[2] - b[i].min[2];
        s += std::fmax(std::fmax(ex, 0.0f) * std::fmax(ey, 0.0f) * std::fmax(ez, 0.0f), 0.0f);
    }
    return clampf_ref(s, 0.0f, 1.0f);
}

// seeded_cell_count (:148-151)
double seeded_cell_count(const uint32_t res[3], const fs_liquid_block* b, uint32_t n) {
    const double total = (double)res[0] * (double)res[1] * (double)res[2];
    return total * (double)seeded_volume_fraction(b, n);
}

// resolved_particle_count (:179-193)
uint32_t resolved_particle_count(uint32_t override_count, float density, const uint32_t res[3],
                                 const fs_liquid_block* b, uint32_t n) {
    if (override_count > 0) return override_count;
    const double cells = seeded_cell_count(res, b, n);
    if (cells <= 0.0) return 0;
    const double count = std::round(cells * (double)std::fmax(density, 0.0f));  // f64::round
    const double sat = count < 0.0 ? 0.0 : (count > 4294967295.0 ? 4294967295.0 : count);  // `as u32`
    return std::max((uint32_t)sat, 1024u);
}

// effective_particle_density_for_count (:278-289)
float effective_density_for_count(const uint32_t res[3], const fs_liquid_block* b, uint32_t n, uint32_t count) {
    const double cells = seeded_cell_count(res, b, n);
    if (cells <= 0.0) return kReferenceDensity;
    return (float)std::fmax((double)count / cells, 1.0e-3);
}

int32_t floor_to_i32(float v) {  // `(v).floor() as i32`, saturating, NaN -> 0
    const float f = std::floor(v);
    if (std::isnan(f)) return 0;
    if (f >= 2147483647.0f) return 2147483647;
    if (f <= -2147483648.0f) return (-2147483647 - 1);
    return (int32_t)f;
}

// The per-block lattice of generate_particles (src/gpu/fluid.rs:1838-1913) and
// estimated_particle_count (:1792-1831): budget by volume, spacing
// cbrt(vol / target), floor(extent / spacing) points per axis.
int build_layout(const fs_scene& sc, float h, const float origin[3], SeedLayout& L) {
    std::memset(&L, 0, sizeof L);
    if (sc.nblocks == 0 || sc.nblocks > FS_SCENE_MAX_BLOCKS)
        return fail(FS_ERR_INVALID_ARGUMENT, "scene needs 1 to FS_SCENE_MAX_BLOCKS liquid blocks");
    float extent[3];
    for (int a = 0; a < 3; ++a) extent[a] = (float)sc.res[a] * h;
    for (int a = 0; a < 3; ++a) {
        L.lo[a] = origin[a] + h * 1.05f;
        L.hi[a] = origin[a] + extent[a] - h * 1.05f;
    }
    float wmin[FS_SCENE_MAX_BLOCKS][3], ext[FS_SCENE_MAX_BLOCKS][3], vol[FS_SCENE_MAX_BLOCKS];
    float total_vol = 0.0f;
    for (uint32_t i = 0; i < sc.nblocks; ++i) {
        for (int a = 0; a < 3; ++a) {
            const float lo = origin[a] + sc.blocks[i].min[a] * extent[a];
            const float hi = origin[a] + sc.blocks[i].max[a] * extent[a];
            wmin[i][a] = lo;
            ext[i][a] = hi - lo;
        }
        vol[i] = std::fmax(ext[i][0] * ext[i][1] * ext[i][2], 1e-6f);
        total_vol += vol[i];
    }
    total_vol = std::fmax(total_vol, 1e-6f);
    const float total_target = (float)std::max(sc.particle_count, 1u);
    uint64_t first = 0;
    for (uint32_t i = 0; i < sc.nblocks; ++i) {
        const float target = std::fmax(total_target * (vol[i] / total_vol), 1.0f);
        const float spacing = std::fmax(std::cbrt(vol[i] / target), 1e-4f);
        SeedBlock& B = L.blocks[i];
        uint64_t cnt = 1;
        for (int a = 0; a < 3; ++a) {
            B.wmin[a] = wmin[i][a];
            B.counts[a] = (uint32_t)std::max(floor_to_i32(ext[i][a] / spacing), 1);
            cnt *= B.counts[a];
        }
        B.spacing = spacing;
        B.first = (uint32_t)first;
        magic_u31(B.counts[2], B.mz, B.sz);
        magic_u31(B.counts[1], B.my, B.sy);
        first += cnt;
        // Particle g draws the generator's states 3g+1 .. 3g+3 (32-bit).
        if (first > 0x55555555ull) return fail(FS_ERR_INVALID_ARGUMENT, "scene lattice too large");
    }
    L.nblocks = sc.nblocks;
    L.total = (uint32_t)first;
    return FS_OK;
}

double reg(const fs_registry* r, const char* id) {
    double v = 0.0;
    fs_registry_get(r, id, &v);
    return v;
}
}  // namespace

FS_API int fs_scene_preset_blocks(uint32_t preset, float drop_height, float fill_level, fs_liquid_block* blocks,
                                  uint32_t* nblocks) {
    if (!blocks || !nblocks) return fail(FS_ERR_INVALID_ARGUMENT, "null argument");
    const std::vector<Block> b = preset_blocks(preset, drop_height, fill_level);
    for (size_t i = 0; i < b.size(); ++i) {
        blocks[i].min[0] = b[i].min.x; blocks[i].min[1] = b[i].min.y; blocks[i].min[2] = b[i].min.z;
        blocks[i].max[0] = b[i].max.x; blocks[i].max[1] = b[i].max.y; blocks[i].max[2] = b[i].max.z;
    }
    *nblocks = (uint32_t)b.size();
    return FS_OK;
}

FS_API int fs_scene_from_registry(const fs_registry* r, fs_scene* out) {
    if (!r || !out) return fail(FS_ERR_INVALID_ARGUMENT, "null argument");
    fs_scene sc;
    std::memset(&sc, 0, sizeof sc);
    const uint32_t p = (uint32_t)reg(r, "scene.preset");
    sc.preset = p == 1u ? FS_SCENE_DAM_BREAK : (p == 2u ? FS_SCENE_DOUBLE_SPLASH : FS_SCENE_FALLING_BLOB);
    sc.res[0] = (uint32_t)reg(r, "grid.res_x");
    sc.res[1] = (uint32_t)reg(r, "grid.res_y");
    sc.res[2] = (uint32_t)reg(r, "grid.res_z");
    const float drop = (float)reg(r, "scene.drop_height");
    // Registry::fill_level: the 0-100 percentage as a fraction.
    const float fill = clampf_ref((float)reg(r, "scene.fill_level") / 100.0f, 0.0f, 1.0f);
    int rc = fs_scene_preset_blocks(sc.preset, drop, fill, sc.blocks, &sc.nblocks);
    if (rc != FS_OK) return rc;
    sc.particle_count = resolved_particle_count((uint32_t)reg(r, "particles.count"),
                                                (float)reg(r, "particles.density"), sc.res, sc.blocks, sc.nblocks);
    sc.effective_density = effective_density_for_count(sc.res, sc.blocks, sc.nblocks, sc.particle_count);
    // estimated_particle_count: the lattice of the reference tank (h = H,
    // centred origin).
    const float origin[3] = {-((float)sc.res[0] * kH) / 2.0f, -((float)sc.res[1] * kH) / 2.0f,
                             -((float)sc.res[2] * kH) / 2.0f};
    SeedLayout L;
    rc = build_layout(sc, kH, origin, L);
    if (rc != FS_OK) return rc;
    sc.generated_count = L.total;
    *out = sc;
    return FS_OK;
}

FS_API int fs_scene_seed(const fs_scene* scene, float h, const float origin[3], float* pos_xyz, uint32_t capacity,
                         uint32_t* count) {
    if (!scene || !origin || !count) return fail(FS_ERR_INVALID_ARGUMENT, "null argument");
    if (!(h > 0.0f) || !std::isfinite(h)) return fail(FS_ERR_INVALID_ARGUMENT, "h must be positive and finite");
    SeedLayout L;
    if (int rc = build_layout(*scene, h, origin, L); rc != FS_OK) return rc;
    *count = L.total;
    if (!pos_xyz) return FS_OK;
    if (capacity < L.total) return fail(FS_ERR_INVALID_ARGUMENT, "pos_xyz holds fewer triplets than the lattice");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(FS_ERR_DEVICE, "no HIP device available");
    hipStream_t s = nullptr;
    float* d = nullptr;
    auto done = [&](int rc) {
        if (d) (void)hipFree(d);
        if (s) (void)hipStreamDestroy(s);
        return rc;
    };
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
        return done(fail(FS_ERR_DEVICE, "hipStreamCreate failed"));
    const size_t bytes = (size_t)L.total * 3u * sizeof(float);
    if (hipMalloc(&d, bytes) != hipSuccess) return done(fail(FS_ERR_OUT_OF_MEMORY, "hipMalloc failed"));
    if (launch_seed_lattice(L, d, s) != hipSuccess) return done(fail(FS_ERR_DEVICE, "seed kernel launch failed"));
    if (hipMemcpyAsync(pos_xyz, d, bytes, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return done(fail(FS_ERR_DEVICE, "seed copy failed"));
    return done(FS_OK);
}

// ---- fs_scene_seed_noalloc: the lattice staged in a fixed device array ----
// The staging array of fs_scene_seed_noalloc: one per GPU, for the life of
// the process.  Host code never passes `scratch` itself to a kernel or a copy
// (on the host it names a stand-in, not device memory); it uses the device
// address that hipGetSymbolAddress returns on each device.
static __device__ __align__(256) float scratch[FS_SEED_SCRATCH_BYTES / sizeof(float)];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_ptr[kMaxDevices] = {};

// The current device's address of `scratch`, looked up on the first call on
// that device (with it current) and kept.
int scratch_on_current_device(float** out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail(FS_ERR_DEVICE, "hipGetDevice failed");
    if (dev < 0 || dev >= kMaxDevices) return fail(FS_ERR_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(FS_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<float*>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return FS_OK;
}
}  // namespace

FS_API int fs_scene_seed_noalloc(const fs_scene* scene, float h, const float origin[3], float* pos_xyz,
                                 uint32_t capacity, uint32_t* count) {
    if (!scene || !origin || !count) return fail(FS_ERR_INVALID_ARGUMENT, "null argument");
    if (!(h > 0.0f) || !std::isfinite(h)) return fail(FS_ERR_INVALID_ARGUMENT, "h must be positive and finite");
    SeedLayout L;
    if (int rc = build_layout(*scene, h, origin, L); rc != FS_OK) return rc;
    *count = L.total;
    if (!pos_xyz) return FS_OK;
    if (capacity < L.total) return fail(FS_ERR_INVALID_ARGUMENT, "pos_xyz holds fewer triplets than the lattice");
    const size_t bytes = (size_t)L.total * 3u * sizeof(float);
    if (bytes > FS_SEED_SCRATCH_BYTES)
        return fail(FS_ERR_INVALID_ARGUMENT, "lattice of " + std::to_string(L.total) +
                                                 " particles exceeds the staging array (" +
                                                 std::to_string(FS_SEED_SCRATCH_PARTICLES) + " particles)");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(FS_ERR_DEVICE, "no HIP device available");
    float* stage = nullptr;
    if (int rc = scratch_on_current_device(&stage)) return rc;
    // The kernel writes every staged float it covers, so nothing needs clearing.
    if (launch_seed_lattice(L, stage, nullptr) != hipSuccess) return fail(FS_ERR_DEVICE, "seed kernel launch failed");
    if (hipMemcpyAsync(pos_xyz, stage, bytes, hipMemcpyDeviceToHost, nullptr) != hipSuccess ||
        hipStreamSynchronize(nullptr) != hipSuccess)
        return fail(FS_ERR_DEVICE, "seed copy failed");
    return FS_OK;
}

FS_API int fs_create_from_settings(const fs_registry* r, const fs_scene* scene, fs_tank** out) {
    if (!r || !scene || !out) return fail(FS_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    // GpuFluid::new (src/gpu/fluid.rs:258-271): h = H, centred tank.
    const float nx = (float)(uint32_t)reg(r, "grid.res_x"), ny = (float)(uint32_t)reg(r, "grid.res_y"),
                nz = (float)(uint32_t)reg(r, "grid.res_z");
    const float origin[3] = {-nx * kH / 2.0f, -ny * kH / 2.0f, -nz * kH / 2.0f};
    fs_scene sc = *scene;
    sc.res[0] = (uint32_t)nx;
    sc.res[1] = (uint32_t)ny;
    sc.res[2] = (uint32_t)nz;
    uint32_t n = 0;
    int rc = fs_scene_seed(&sc, kH, origin, nullptr, 0, &n);
    if (rc != FS_OK) return rc;
    std::vector<float> pos((size_t)n * 3u);
    rc = fs_scene_seed(&sc, kH, origin, pos.data(), n, &n);
    if (rc != FS_OK) return rc;
    fs_params p;
    rc = fs_registry_to_params(r, kH, origin, scene->effective_density, &p);
    if (rc != FS_OK) return rc;
    return fs_create(&p, pos.data(), nullptr, n, out);
}
