/*
 * flip_oracle_scene.c — CPU restatement of fluid-lab's scene presets, spawn
 * counts and initial particle lattice.
 *
 * TEST INFRASTRUCTURE ONLY (see flip_oracle.h): the checker for
 * fs_scene_preset_blocks / fs_scene_from_registry / fs_scene_seed; the
 * product library never links or calls it.
 *
 * Restated from src/scene/mod.rs:11-324 and src/gpu/fluid.rs:1792-1913.
 * The generator is run sequentially, three draws per particle in lattice
 * order, exactly as generate_particles loops.  Every float expression keeps
 * the reference's f32 (or f64) type and operation order; this file is built
 * with -ffp-contract=off.  cbrtf and powf are the C library's (glibc here),
 * which is what the reference's f32::cbrt / f32::powf call on a native
 * build.
 *
 * Pins (tests/test_oracle_scene.py): the reference's scene tests
 * (src/scene/mod.rs:470-700: seeded fractions 0.199888 / 0.499720 /
 * 0.163170 / 0.173280, the ~512k default and ~1.31M wide-grid counts, the
 * 1024 floor, the override, drop-height and fill-level behaviour), the
 * generator's first draws computed with exact integer arithmetic, and
 * lattice invariants (counts are the floors of extent / spacing, every
 * point within a quarter spacing of its lattice site or on the clamp box).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#define ORC_SCENE_MAX_BLOCKS 8

static float clampf_r(float x, float lo, float hi) { /* f32::clamp */
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
}

typedef struct { float v[3]; } v3;

static v3 mk(float x, float y, float z) { v3 r = {{x, y, z}}; return r; }

/* scale_block_axes (src/scene/mod.rs:253-264); b = min[3], max[3] */
static void scale_axes(float *b, v3 f) {
    float center[3], half[3], sc[3], mn[3], mx[3];
    for (int a = 0; a < 3; ++a) {
        center[a] = 0.5f * (b[a] + b[3 + a]);
        half[a] = 0.5f * (b[3 + a] - b[a]);
        sc[a] = half[a] * fmaxf(f.v[a], 1.0e-3f);
    }
    for (int a = 0; a < 3; ++a) mn[a] = fminf(fmaxf(center[a] - sc[a], 0.0f), 1.0f);
    for (int a = 0; a < 3; ++a) mx[a] = fminf(fmaxf(center[a] + sc[a], mn[a] + 1.0e-3f), 1.0f);
    for (int a = 0; a < 3; ++a) mn[a] = fmaxf(fminf(mn[a], mx[a] - 1.0e-3f), 0.0f);
    for (int a = 0; a < 3; ++a) { b[a] = mn[a]; b[3 + a] = mx[a]; }
}

/* shift_block_y (:242-247) */
static void shift_y(float *b, float delta) {
    float s = clampf_r(delta, -b[1], 1.0f - b[4]);
    b[1] += s;
    b[4] += s;
}

/* preset_blocks (:208-238) with ScenePreset::from_u32 (:43-49).  Writes
 * 6 floats per block (min xyz, max xyz); returns the block count. */
uint32_t orc_scene_preset_blocks(uint32_t preset, float drop_height, float fill_level, float *out) {
    const float fill = clampf_r(fill_level, 0.0f, 1.0f);
    if (preset == 1u) { /* DamBreak */
        const float top = clampf_r(fill * 0.98f, 1.0e-3f, 0.98f);
        const float b[6] = {0.05f, 0.0f, 0.05f, 0.42f, top, 0.95f};
        memcpy(out, b, sizeof b);
        return 1;
    }
    if (preset == 2u) { /* DoubleSplash */
        const float src[2][6] = {{0.1f, 0.45f, 0.3f, 0.34f, 0.92f, 0.7f}, {0.66f, 0.45f, 0.3f, 0.9f, 0.92f, 0.7f}};
        for (int k = 0; k < 2; ++k) {
            float b[6];
            memcpy(b, src[k], sizeof b);
            shift_y(b, clampf_r(drop_height, 0.0f, 1.0f) - 0.72f);
            scale_axes(b, mk(1.0f, fill / 0.2f, 1.0f)); /* scale_suspended_drop */
            memcpy(out + 6 * k, b, sizeof b);
        }
        return 2;
    }
    /* FallingBlob */
    {
        float b[6] = {0.19f, 0.2f, 0.19f, 0.81f, 0.72f, 0.81f};
        const float delta = clampf_r(drop_height, 0.0f, 1.0f) - 0.72f;
        shift_y(b, delta);
        const float factor = powf(fmaxf(fill, 1.0e-4f) / 0.2f, 1.0f / 3.0f); /* scale_block_about_center */
        scale_axes(b, mk(factor, factor, factor));
        memcpy(out, b, sizeof b);
        return 1;
    }
}

/* seeded_volume_fraction (:134-146) */
float orc_scene_seeded_fraction(const float *blocks, uint32_t n) {
    float s = 0.0f;
    for (uint32_t i = 0; i < n; ++i) {
        const float *b = blocks + 6 * i;
        const float ex = b[3] - b[0], ey = b[4] - b[1], ez = b[5] - b[2];
        s += fmaxf(fmaxf(ex, 0.0f) * fmaxf(ey, 0.0f) * fmaxf(ez, 0.0f), 0.0f);
    }
    return clampf_r(s, 0.0f, 1.0f);
}

static double seeded_cells(const uint32_t res[3], const float *blocks, uint32_t n) { /* :148-151 */
    return (double)res[0] * (double)res[1] * (double)res[2] * (double)orc_scene_seeded_fraction(blocks, n);
}

/* resolved_particle_count (:179-193) */
uint32_t orc_scene_resolved_count(const uint32_t res[3], const float *blocks, uint32_t n, float density,
                                  uint32_t override_count) {
    if (override_count > 0) return override_count;
    const double cells = seeded_cells(res, blocks, n);
    if (cells <= 0.0) return 0;
    double c = round(cells * (double)fmaxf(density, 0.0f));
    if (c < 0.0) c = 0.0;
    if (c > 4294967295.0) c = 4294967295.0;
    const uint32_t u = (uint32_t)c;
    return u < 1024u ? 1024u : u;
}

/* effective_particle_density_for_count (:278-289) */
float orc_scene_effective_density(const uint32_t res[3], const float *blocks, uint32_t n, uint32_t count) {
    const double cells = seeded_cells(res, blocks, n);
    if (cells <= 0.0) return 8.0f;
    const double d = (double)count / cells;
    return (float)(d > 1.0e-3 ? d : 1.0e-3);
}

static int32_t floor_i32(float v) { /* `.floor() as i32` */
    const float f = floorf(v);
    if (isnan(f)) return 0;
    if (f >= 2147483647.0f) return 2147483647;
    if (f <= -2147483648.0f) return -2147483647 - 1;
    return (int32_t)f;
}

/* generate_particles (src/gpu/fluid.rs:1838-1913) for a tank of cell size h
 * at `origin` with extent res * h.  Writes up to `cap` xyz triplets to out
 * (NULL: count only) and returns the number of particles the lattice holds
 * (estimated_particle_count, :1792-1831, when h = H and origin = -extent/2).
 * info (nullable) receives 7 floats per block: wmin xyz, spacing, counts xyz. */
uint64_t orc_scene_generate(const float *blocks, uint32_t n, uint32_t particle_count, float h,
                            const float origin[3], const uint32_t res[3], float *out, uint64_t cap, float *info) {
    float extent[3], lo[3], hi[3];
    for (int a = 0; a < 3; ++a) extent[a] = (float)res[a] * h;
    for (int a = 0; a < 3; ++a) {
        lo[a] = origin[a] + h * 1.05f;
        hi[a] = origin[a] + extent[a] - h * 1.05f;
    }
    float wmin[ORC_SCENE_MAX_BLOCKS][3], ext[ORC_SCENE_MAX_BLOCKS][3], vol[ORC_SCENE_MAX_BLOCKS];
    float total_vol = 0.0f;
    if (n > ORC_SCENE_MAX_BLOCKS) n = ORC_SCENE_MAX_BLOCKS;
    for (uint32_t i = 0; i < n; ++i) {
        for (int a = 0; a < 3; ++a) {
            const float w0 = origin[a] + blocks[6 * i + a] * extent[a];
            const float w1 = origin[a] + blocks[6 * i + 3 + a] * extent[a];
            wmin[i][a] = w0;
            ext[i][a] = w1 - w0;
        }
        vol[i] = fmaxf(ext[i][0] * ext[i][1] * ext[i][2], 1e-6f);
        total_vol += vol[i];
    }
    total_vol = fmaxf(total_vol, 1e-6f);
    const float total_target = (float)(particle_count > 1u ? particle_count : 1u);
    uint32_t st = 0x12345678u;
    uint64_t k = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const float target = fmaxf(total_target * (vol[i] / total_vol), 1.0f);
        const float spacing = fmaxf(cbrtf(vol[i] / target), 1e-4f);
        int32_t c[3];
        for (int a = 0; a < 3; ++a) {
            c[a] = floor_i32(ext[i][a] / spacing);
            if (c[a] < 1) c[a] = 1;
        }
        if (info) {
            for (int a = 0; a < 3; ++a) info[7 * i + a] = wmin[i][a];
            info[7 * i + 3] = spacing;
            for (int a = 0; a < 3; ++a) info[7 * i + 4 + a] = (float)c[a];
        }
        for (int32_t ix = 0; ix < c[0]; ++ix)
            for (int32_t iy = 0; iy < c[1]; ++iy)
                for (int32_t iz = 0; iz < c[2]; ++iz) {
                    const int32_t id[3] = {ix, iy, iz};
                    float j[3];
                    for (int a = 0; a < 3; ++a) { /* rand01 */
                        st = st * 1664525u + 1013904223u;
                        const float r = (float)(st >> 8) / (float)(1u << 24);
                        j[a] = (r - 0.5f) * spacing * 0.5f;
                    }
                    if (out && k < cap)
                        for (int a = 0; a < 3; ++a)
                            out[3 * k + a] = clampf_r(wmin[i][a] + ((float)id[a] + 0.5f) * spacing + j[a], lo[a], hi[a]);
                    ++k;
                }
    }
    return k;
}
