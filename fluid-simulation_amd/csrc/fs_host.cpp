// fs_host.cpp — the C-ABI of include/fluidstep.h: settings registry, the tank
// (device buffers + stream), and the per-substep launch sequence of both
// engines.
//
// Launch order per substep follows GpuContext::record_substep_coarse
// (src/gpu/mod.rs:993-1033): clear -> mark -> [sort] -> classify -> P2G ->
// normalize/save/forces/enforce -> divergence -> CG -> gradient/enforce -> G2P.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <iterator>
#include <string>
#include <vector>

#include "fluidstep.h"
#include "fs_kernels.h"
#include "fs_host_common.h"
#include "fs_tank.h"

using namespace fs;

// ---------------------------------------------------------------------------
// Error reporting
namespace fs {
thread_local std::string g_last_error;
thread_local uint64_t g_dispatches = 0;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}
}  // namespace fs

namespace {
// ---------------------------------------------------------------------------
// Registry rows of the step (src/settings/mod.rs:407-732, 1349-1391): id,
// type, default, inclusive bounds, apply class.
struct Row {
    const char* id;
    bool is_u32;
    double def, lo, hi;
    int apply;
};
const Row kRows[] = {
    {"scene.preset", true, 0, 0, 2, FS_APPLY_RESET},
    {"scene.fill_level", false, 20.0, 0.0, 100.0, FS_APPLY_RESET},
    {"scene.drop_height", false, 1.0, 0.0, 1.0, FS_APPLY_RESET},
    {"grid.res_x", true, 80, 16, 128, FS_APPLY_RESET},
    {"grid.res_y", true, 40, 16, 128, FS_APPLY_RESET},
    {"grid.res_z", true, 80, 16, 128, FS_APPLY_RESET},
    {"particles.density", false, 10.0, 1.0, 32.0, FS_APPLY_RESET},
    {"particles.count", true, 0, 0, 134217728, FS_APPLY_RESET},
    {"physics.gravity", false, -9.81, -40.0, 40.0, FS_APPLY_LIVE},
    {"physics.fixed_dt", false, 1.0 / 120.0, 1.0 / 480.0, 1.0 / 30.0, FS_APPLY_RESET},
    {"physics.sim_speed", false, 1.0, 0.05, 1.0, FS_APPLY_LIVE},
    {"physics.flip_blend", false, 0.9, 0.0, 1.0, FS_APPLY_LIVE},
    {"physics.max_substeps", true, 2, 1, 16, FS_APPLY_RESET},
    {"physics.wall_friction", false, 0.0, 0.0, 1.0, FS_APPLY_LIVE},
    {"physics.rest_density", false, 0.0, 0.0, 32.0, FS_APPLY_LIVE},
    {"physics.volume_stiffness", false, 0.45, 0.0, 4.0, FS_APPLY_LIVE},
    {"physics.drift_clamp", false, 0.5, 0.05, 2.0, FS_APPLY_LIVE},
    {"physics.cfl", false, 2.0, 1.0, 6.0, FS_APPLY_LIVE},
    {"interaction.auto_roll_enabled", true, 0, 0, 1, FS_APPLY_LIVE},
    {"interaction.auto_roll_strength", false, 0.222, 0.0, 1.2, FS_APPLY_LIVE},
    {"interaction.auto_roll_cadence", false, 0.8375, 0.5, 8.0, FS_APPLY_LIVE},
    {"interaction.wave_enabled", true, 0, 0, 1, FS_APPLY_LIVE},
    {"interaction.wave_strength", false, 2.145, 0.0, 3.0, FS_APPLY_LIVE},
    {"interaction.wave_frequency", false, 0.22775, 0.05, 4.0, FS_APPLY_LIVE},
    {"classify.liquid_threshold", true, 1, 1, 8, FS_APPLY_LIVE},
    {"classify.surface_dilation", true, 0, 0, 1, FS_APPLY_LIVE},
    {"solver.pressure_iterations", true, 30, 1, 200, FS_APPLY_LIVE},
    {"solver.pressure_warm_start", true, 1, 0, 1, FS_APPLY_LIVE},
    {"solver.pressure_residual_tolerance", false, 0.0, 0.0, 0.1, FS_APPLY_LIVE},
    {"dev.particle_sort", true, 1, 0, 1, FS_APPLY_RESET},
    {"dev.particle_sort_period", true, 4, 1, 16, FS_APPLY_RESET},
};
constexpr int kNumRows = (int)(sizeof(kRows) / sizeof(kRows[0]));

int row_index(const char* id) {
    if (!id) return -1;
    for (int i = 0; i < kNumRows; ++i)
        if (std::strcmp(kRows[i].id, id) == 0) return i;
    return -1;
}

// Stored value as the registry keeps it: U32 or the f64 widening of an f32.
double f32_store(double v) { return (double)(float)v; }

// Registry::set_value_f64_result (src/settings/mod.rs:1733-1821): non-finite
// rejected; U32 rows round half away from zero then clamp; F32 rows cast to
// f32 then clamp.  `clamped` reports the pre-clamp value was out of range.
void mutate(std::vector<double>& values, const char* id, double v, fs_mutation* m) {
    fs_mutation r;
    std::memset(&r, 0, sizeof r);
    r.requested_value = v;
    r.stored_value = NAN;
    r.apply_class = -1;
    r.status = FS_STATUS_UNKNOWN_ID;
    if (!std::isfinite(v)) {
        r.status = FS_STATUS_NON_FINITE_REJECTED;
    } else {
        int i = row_index(id);
        if (i >= 0) {
            const Row& row = kRows[i];
            double stored;
            bool clamped;
            if (row.is_u32) {
                double rounded = std::round(v);  // f64::round: half away from zero
                // `rounded as u32` saturates in Rust (NaN excluded above).
                double sat = rounded < 0.0 ? 0.0 : (rounded > 4294967295.0 ? 4294967295.0 : rounded);
                stored = sat < row.lo ? row.lo : (sat > row.hi ? row.hi : sat);
                clamped = rounded < row.lo || rounded > row.hi;
            } else {
                float in = (float)v;
                float lo = (float)row.lo, hi = (float)row.hi;
                float st = in < lo ? lo : (in > hi ? hi : in);
                stored = (double)st;
                clamped = in < lo || in > hi;
            }
            values[i] = stored;
            r.stored_value = stored;
            r.clamped = clamped ? 1 : 0;
            r.apply_class = row.apply;
            r.status = row.apply == FS_APPLY_LIVE ? FS_STATUS_APPLIED : FS_STATUS_STORED;
        }
    }
    r.accepted = (r.status == FS_STATUS_APPLIED || r.status == FS_STATUS_STORED) ? 1 : 0;
    r.needs_reset = (r.accepted && r.apply_class == FS_APPLY_RESET) ? 1 : 0;
    if (m) *m = r;
}

std::vector<double> default_values() {
    std::vector<double> v(kNumRows);
    for (int i = 0; i < kNumRows; ++i) v[i] = kRows[i].is_u32 ? kRows[i].def : f32_store(kRows[i].def);
    return v;
}

double value_of(const std::vector<double>& v, const char* id) {
    int i = row_index(id);
    return i >= 0 ? v[i] : NAN;
}

}  // namespace

std::vector<double> fs_default_settings() { return default_values(); }

struct fs_registry {
    std::vector<double> values = default_values();
};

namespace fs {

// Params uniform fill (src/gpu/fluid.rs:289-318) + live clamps of the
// setters (src/gpu/fluid.rs:1295-1371).
void build_step_params(const fs_params& q, uint32_t n, StepParams& P) {
    std::memset(&P, 0, sizeof P);
    P.nx = q.nx; P.ny = q.ny; P.nz = q.nz;
    P.C = q.nx * q.ny * q.nz;
    P.F[0] = (q.nx + 1) * q.ny * q.nz;
    P.F[1] = q.nx * (q.ny + 1) * q.nz;
    P.F[2] = q.nx * q.ny * (q.nz + 1);
    P.N = n;
    P.h = q.h;
    P.inv_h = 1.0f / q.h;
    P.dt = q.dt;
    P.scale = 65536.0f;  // FIXED_SCALE (src/gpu/fluid.rs:41)
    P.rho = q.rho;
    P.flip = std::fmin(std::fmax(q.flip_blend, 0.0f), 1.0f);
    P.friction = std::fmin(std::fmax(q.wall_friction, 0.0f), 1.0f);
    P.cfl = std::fmax(q.cfl, 1.0f);
    for (int a = 0; a < 3; ++a) {
        P.origin[a] = q.origin[a];
        P.grav[a] = q.gravity[a];
    }
    P.rest = std::fmax(q.rest_density, 0.1f);
    P.stiff = std::fmax(q.volume_stiffness, 0.0f);
    P.drift = std::fmax(q.drift_clamp, 0.0f);
    uint32_t th = q.liquid_threshold < 1u ? 1u : q.liquid_threshold;
    P.thresh = (uint32_t)(float)th;
    P.dilation = q.surface_dilation >= 1u ? 1u : 0u;
    P.warm = q.pressure_warm_start ? 1u : 0u;
    P.iters = q.pressure_iterations < 1u ? 1u : q.pressure_iterations;
    float tol = std::fmin(std::fmax(q.pressure_residual_tolerance, 0.0f), 0.1f);
    P.tol = tol;
    P.tol_sq = tol * tol;
}

bool all_finite(const float* a, size_t n) {
    if (!a) return true;
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(a[i])) return false;
    return true;
}

// Creation-time checks shared by fs_create and fs_batch_create.
int check_params(const fs_params& q, uint32_t n, const float* pos_xyz, const float* vel_xyz) {
    if (q.nx < 3 || q.ny < 3 || q.nz < 3) return fail(FS_ERR_INVALID_ARGUMENT, "grid must be at least 3 cells per axis");
    // Kernels address cell and face arrays with 32-bit byte offsets.
    if ((uint64_t)(q.nx + 1) * (q.ny + 1) * (q.nz + 1) >= (1ull << 30))
        return fail(FS_ERR_INVALID_ARGUMENT, "grid too large for 32-bit byte offsets");
    if (!(q.h > 0.0f) || !std::isfinite(q.h) || !(q.dt > 0.0f) || !std::isfinite(q.dt) || !(q.rho > 0.0f))
        return fail(FS_ERR_INVALID_ARGUMENT, "h, dt and rho must be positive and finite");
    const float scalars[] = {q.origin[0], q.origin[1], q.origin[2], q.gravity[0], q.gravity[1], q.gravity[2],
                             q.flip_blend, q.wall_friction, q.rest_density, q.volume_stiffness, q.drift_clamp,
                             q.cfl, q.pressure_residual_tolerance, q.rho};
    if (!all_finite(scalars, sizeof scalars / sizeof scalars[0]))
        return fail(FS_ERR_NOT_FINITE, "non-finite parameter");
    if (q.engine > FS_ENGINE_FUSED || q.precision > FS_PRECISION_MIXED || q.velocity_storage > FS_VELOCITY_BF16)
        return fail(FS_ERR_INVALID_ARGUMENT, "unknown engine/precision/velocity_storage");
    if (q.engine == FS_ENGINE_REFERENCE && q.precision != FS_PRECISION_F32)
        return fail(FS_ERR_UNSUPPORTED, "the reference engine runs FS_PRECISION_F32 only");
    if (q.engine == FS_ENGINE_FUSED && q.precision != FS_PRECISION_MIXED)
        return fail(FS_ERR_UNSUPPORTED, "the fused engine runs FS_PRECISION_MIXED only");
    if (n > 0 && (!all_finite(pos_xyz, 3 * (size_t)n) || !all_finite(vel_xyz, vel_xyz ? 3 * (size_t)n : 0)))
        return fail(FS_ERR_NOT_FINITE, "non-finite particle position or velocity");
    if ((uint64_t)n > (1ull << 31)) return fail(FS_ERR_INVALID_ARGUMENT, "too many particles");
    return FS_OK;
}

// Grid of a persistent (cooperative) kernel whose co-resident figure
// (occupancy x CUs) is max_blocks.  A grid barrier's latency grows with the
// number of workgroups (on MI355X about 7 us at 256 and 11 us at 512), so a
// kernel resident at two or more workgroups per CU runs half its figure; one
// resident at one workgroup per CU runs one per CU.  Whole XCD groups
// (multiples of 8 blocks) either way.
int half_resident_blocks(int max_blocks) {
    int b = (max_blocks / 2 / 8) * 8;
    return b > 0 ? b : 1;
}

int persistent_grid_blocks(int max_blocks, int cus) {
    if (max_blocks <= cus) {
        const int b = (max_blocks / 8) * 8;
        return b > 0 ? b : 1;
    }
    return half_resident_blocks(max_blocks);
}

}  // namespace fs

namespace {

int upload_particles(fs_tank* t, const float* pos, const float* vel) {
    const uint32_t n = t->N;
    if (n == 0) return FS_OK;
    std::vector<float> col(n);
    std::vector<uint16_t> colb(t->bf16 ? n : 0);
    for (int a = 0; a < 3; ++a) {
        for (uint32_t i = 0; i < n; ++i) col[i] = pos[3 * (size_t)i + a];
        FS_HIP(hipMemcpyAsync(t->side[0].pos[a], col.data(), n * sizeof(float), hipMemcpyHostToDevice, t->stream));
        FS_HIP(hipStreamSynchronize(t->stream));
    }
    for (int a = 0; a < 3; ++a) {
        for (uint32_t i = 0; i < n; ++i) col[i] = vel ? vel[3 * (size_t)i + a] : 0.0f;
        if (t->bf16) {
            for (uint32_t i = 0; i < n; ++i) {
                uint32_t u;
                std::memcpy(&u, &col[i], 4);
                uint32_t rr = 0x7fffu + ((u >> 16) & 1u);
                colb[i] = (uint16_t)((u + rr) >> 16);
            }
            FS_HIP(hipMemcpyAsync(t->side[0].vel[a], colb.data(), n * 2, hipMemcpyHostToDevice, t->stream));
        } else {
            FS_HIP(hipMemcpyAsync(t->side[0].vel[a], col.data(), n * 4, hipMemcpyHostToDevice, t->stream));
        }
        FS_HIP(hipStreamSynchronize(t->stream));
    }
    t->cur = 0;
    return FS_OK;
}

// Detailed timing: records a boundary event closing the span since the
// previous one under `tag` (a section index, kTagCg + bucket, kTagCgSolve).
constexpr int kTagCg = 100, kTagCgSolve = 200, kTagStart = -1;
// FINE_SECTIONS indices (src/gpu/timing.rs:32-59).
enum : int {
    kSecClear, kSecMark, kSecSort, kSecClassify, kSecScatter, kSecNormalize, kSecSaveVel = 8, kSecForces = 11,
    kSecBoundPre = 14, kSecDivergence = 17, kSecCgInit, kSecGradient, kSecBoundPost = 22, kSecG2p = 25
};
constexpr int kFineSections = 26;
// CG_CATS indices (src/gpu/timing.rs:105).
enum : int { kCatSpmv, kCatReduce, kCatUpdate, kCatScalars, kCgCategories };

int mark(fs_tank* t, int tag) {
    if (!t->dev_armed) return FS_OK;
    if (t->dev_n == t->dev_ev.size()) {
        hipEvent_t e = nullptr;
        FS_HIP(hipEventCreate(&e));
        t->dev_ev.push_back(e);
        t->dev_tag.push_back(0);
    }
    FS_HIP(hipEventRecord(t->dev_ev[t->dev_n], t->stream));
    t->dev_tag[t->dev_n++] = tag;
    return FS_OK;
}
#define FS_MARK(t, tag)                         \
    do {                                        \
        if (int rc_ = mark((t), (tag))) return rc_; \
    } while (0)

bool env_is(const char* name, const char* value) {
    const char* e = std::getenv(name);
    return e && std::strcmp(e, value) == 0;
}

// The sorted particles are now on the other side: it becomes the live one.
void flip_live_side(fs_tank* t) {
    t->cur ^= 1;
    t->sort_swapped = true;  // src/gpu/fluid.rs:1570
    t->perm_valid = 1;
}

// Stable sort of the live side into the other side (both engines).  With
// move = false the permutation is computed but not applied: the particles
// stay where they are (t->cur unchanged), and the caller fills the other
// side through t->perm and only then calls flip_live_side, so an error
// return in between leaves the tank on its complete, unsorted side.
int sort_particles(fs_tank* t, bool move = true) {
    const StepParams& P = t->P;
    hipStream_t s = t->stream;
    FS_HIP(launch_exclusive_scan(t->occ, t->cell_offset, P.C, t->scan_ws, s));
    FS_HIP(launch_stable_permutation(t->keys, P.N, P.C, t->ktmp0, t->ktmp1, t->vtmp, t->perm, t->radix_ws, s));
    if (move) {
        FS_HIP(launch_permute(P.N, t->perm, t->side[t->cur], t->side[t->cur ^ 1], t->bf16, s));
        flip_live_side(t);
    }
    return FS_OK;
}

// ---- one substep, reference-structured engine: one kernel per WGSL pass ----
int substep_reference(fs_tank* t, bool sort_now) {
    const StepParams& P = t->P;
    hipStream_t s = t->stream;
    if (t->timing) FS_HIP(hipEventRecord(t->ev[0], s));
    FS_MARK(t, kTagStart);
    for (int a = 0; a < 3; ++a) {
        FS_HIP(launch_fill_u32((uint32_t*)t->num[a], 0u, P.F[a], s));
        FS_HIP(launch_fill_u32((uint32_t*)t->den[a], 0u, P.F[a], s));
    }
    FS_HIP(launch_fill_u32(t->occ, 0u, P.C, s));
    FS_HIP(launch_fill_u32(t->counters, 0u, 4, s));
    if (!P.warm) FS_HIP(launch_fill_f32(t->p, 0.0f, P.C, s));
    FS_MARK(t, kSecClear);
    FS_HIP(launch_mark(P, t->side[t->cur].pos, t->occ, s));
    FS_MARK(t, kSecMark);
    if (sort_now) {
        // Keys for the sort: same expression as mark; the histogram it also
        // produces goes to scratch (cell_offset is rewritten by the scan).
        FS_HIP(launch_key_hist(P, t->side[t->cur].pos, t->keys, t->cell_offset, t->counters + 1, nullptr, 0u, s));
        int rc = sort_particles(t);
        if (rc != FS_OK) return rc;
    }
    FS_MARK(t, kSecSort);  // an empty span on substeps without a sort
    ParticleSoA live = t->side[t->cur];
    FS_HIP(launch_classify(P, t->occ, t->ct, t->counters, s));
    FS_MARK(t, kSecClassify);
    // With the sort enabled the reference scatters through the workgroup-local
    // table on every substep (src/gpu/fluid.rs:519-530, 1572-1584); without
    // it, per-tap global atomics (sh/scatter.wgsl).
    if (t->prm.particle_sort && p2g_local_supported(P) && !env_is("FLUIDSTEP_REF_P2G", "atomic"))
        FS_HIP(launch_p2g_local(P, live, t->bf16, t->num, t->den, s));
    else
        FS_HIP(launch_p2g_atomic(P, live, t->bf16, t->num, t->den, nullptr, 0u, s));
    FS_MARK(t, kSecScatter);
    for (int a = 0; a < 3; ++a) {
        FS_HIP(launch_normalize(P, a, t->num[a], t->den[a], t->vel[a], s));
        FS_MARK(t, kSecNormalize + a);
    }
    for (int a = 0; a < 3; ++a) {
        FS_HIP(launch_copy_f32(t->vel[a], t->saved[a], P.F[a], s));
        FS_MARK(t, kSecSaveVel + a);
    }
    for (int a = 0; a < 3; ++a) {
        FS_HIP(launch_forces(P, a, t->ct, t->vel[a], s));
        FS_MARK(t, kSecForces + a);
    }
    for (int a = 0; a < 3; ++a) {
        FS_HIP(launch_enforce(P, a, t->ct, t->vel[a], s));
        FS_MARK(t, kSecBoundPre + a);
    }
    if (t->timing) FS_HIP(hipEventRecord(t->ev[1], s));
    FS_HIP(launch_divergence(P, t->vel, t->ct, t->occ, t->div, s));
    FS_MARK(t, kSecDivergence);
    FS_HIP(launch_cg_init_ref(P, t->div, t->ct, t->p, t->r, t->d0, t->scal, s));
    FS_HIP(launch_cg_reduce_ref(P, t->r, t->r, t->partials, t->scal, s));
    FS_HIP(launch_cg_reduce_final_ref(P, t->partials, t->scal, s));
    FS_HIP(launch_cg_set_rsold_ref(t->scal, s));
    FS_MARK(t, kSecCgInit);
    if (t->timing) FS_HIP(hipEventRecord(t->ev[4], s));
    // Six timed passes per iteration in solver order, bucketed as the
    // reference's CG_BUCKET (src/gpu/timing.rs:110-116).
    for (uint32_t it = 0; it < P.iters; ++it) {
        FS_HIP(launch_cg_iter_count_ref(t->scal, t->counters + 2, s));
        FS_HIP(launch_cg_spmv_ref(P, t->ct, t->d0, t->q, t->scal, s));
        FS_MARK(t, kTagCg + kCatSpmv);
        FS_HIP(launch_cg_reduce_ref(P, t->d0, t->q, t->partials, t->scal, s));
        FS_HIP(launch_cg_reduce_final_ref(P, t->partials, t->scal, s));
        FS_MARK(t, kTagCg + kCatReduce);
        FS_HIP(launch_cg_alpha_ref(t->scal, s));
        FS_MARK(t, kTagCg + kCatScalars);
        FS_HIP(launch_cg_update_ref(P, t->scal, t->d0, t->q, t->p, t->r, s));
        FS_MARK(t, kTagCg + kCatUpdate);
        FS_HIP(launch_cg_reduce_ref(P, t->r, t->r, t->partials, t->scal, s));
        FS_HIP(launch_cg_reduce_final_ref(P, t->partials, t->scal, s));
        FS_MARK(t, kTagCg + kCatReduce);
        FS_HIP(launch_cg_beta_ref(t->scal, s));
        FS_HIP(launch_cg_dir_ref(P, t->scal, t->r, t->d0, s));
        FS_MARK(t, kTagCg + kCatScalars);
    }
    if (t->timing) FS_HIP(hipEventRecord(t->ev[2], s));
    for (int a = 0; a < 3; ++a) {
        FS_HIP(launch_gradient(P, a, t->p, t->ct, t->vel[a], s));
        FS_MARK(t, kSecGradient + a);
    }
    for (int a = 0; a < 3; ++a) {
        FS_HIP(launch_enforce(P, a, t->ct, t->vel[a], s));
        FS_MARK(t, kSecBoundPost + a);
    }
    FS_HIP(launch_g2p(P, live, t->bf16, t->vel, t->saved, s));
    FS_MARK(t, kSecG2p);
    if (t->timing) FS_HIP(hipEventRecord(t->ev[3], s));
    t->gather_last = 0;
    return FS_OK;
}

// ---- one substep, fused engine ----
}  // namespace

namespace fs {

int substep_fused(fs_tank* t, bool sort_now, bool stop_before_g2p) {
    const StepParams& P = t->P;
    hipStream_t s = t->stream;
    if (t->timing) FS_HIP(hipEventRecord(t->ev[0], s));
    // Detailed timing: each fused kernel is timed under the first section
    // whose work it does (the rest report as fused, zero span).
    FS_MARK(t, kTagStart);
    FS_HIP(launch_fill_u32(t->occ, 0u, P.C, s));
    FS_HIP(launch_fill_u32(t->counters, 0u, 4, s));
    FS_MARK(t, kSecClear);
    FS_HIP(launch_key_hist(P, t->side[t->cur].pos, t->keys, t->occ, t->counters + 1, t->mis_list, kMisfitCap, s));
    FS_MARK(t, kSecMark);
    if (sort_now) {
        // The permutation is applied by G2P, which rewrites every particle
        // anyway; until then the particles lie on the pre-sort side, which
        // stays the live side until G2P is enqueued.
        int rc = sort_particles(t, false);
        if (rc != FS_OK) return rc;
    }
    FS_MARK(t, kSecSort);
    // live: the side G2P writes (the sorted side on a sort substep);
    // pre: the side the particles lie on now.
    ParticleSoA live = t->side[sort_now ? t->cur ^ 1 : t->cur];
    const ParticleSoA pre = t->side[t->cur];
    FS_HIP(launch_classify(P, t->occ, t->ct, t->counters, s));
    FS_MARK(t, kSecClassify);
    if (sort_now || t->gperm) {
        // The gather walks the particles by cell through an order of them.
        // On a sort substep they still lie on the pre-sort side, in the
        // sort's order (perm), which G2P applies as it writes.  Otherwise they
        // lie on the live side and stay there; a cell order of them (gperm,
        // not stable) serves, since the gather's integer sums are the atomic
        // form's terms in any order.
        const ParticleSoA& from = sort_now ? pre : live;
        const uint32_t* order = t->perm;
        // Bin starts of this order: the sort's (cell_offset) or, without a
        // sort, the cell order's own (goff), so that cell_offset keeps the
        // bin starts of the last sort as the reference's buffer does.
        const uint32_t* starts = t->cell_offset;
        if (!sort_now) {
            FS_HIP(launch_exclusive_scan(t->occ, t->goff, P.C, t->scan_ws, s));
            FS_HIP(launch_cell_order(t->keys, P.N, t->goff, t->cursor, P.C, t->gperm, s));
            order = t->gperm;
            starts = t->goff;
        }
        if (t->p2g_particle) {
            // Particle-owned: each particle's taps added once, by the brick
            // of its sort cell, into cleared num/den; then the face pass.
            FS_HIP(launch_clear_faces_guarded(P, t->num, t->den, nullptr, 0u, s));
            FS_HIP(launch_p2g_bricks(P, from, t->bf16, t->occ, starts, t->num, t->den, order, s));
            FS_MARK(t, kSecScatter);
            FS_HIP(launch_face_epilogue_rows(P, t->num, t->den, t->ct, t->vel, t->saved, s));
        } else {
            FS_HIP(launch_p2g_gather(P, from, from, t->bf16, t->occ, starts, t->ct, t->vel, t->saved,
                                     t->counters + 1, t->mis_list, s, t->gather_brick, order, t->brick_ws));
            // Fallback when more particles' P2G cells differ from their sort
            // cells than the gather walks (any particle order will do).
            FS_HIP(launch_clear_faces_guarded(P, t->num, t->den, t->counters + 1, kMisfitCap, s));
            FS_HIP(launch_p2g_atomic(P, from, t->bf16, t->num, t->den, t->counters + 1, kMisfitCap, s));
            FS_MARK(t, kSecScatter);
            FS_HIP(launch_face_epilogue(P, t->num, t->den, t->ct, t->vel, t->saved, t->counters + 1, kMisfitCap,
                                        s));
        }
        FS_MARK(t, kSecNormalize);
        t->gather_last = 1;
    } else {
        for (int a = 0; a < 3; ++a) {
            FS_HIP(launch_fill_u32((uint32_t*)t->num[a], 0u, P.F[a], s));
            FS_HIP(launch_fill_u32((uint32_t*)t->den[a], 0u, P.F[a], s));
        }
        FS_HIP(launch_p2g_atomic(P, live, t->bf16, t->num, t->den, nullptr, 0u, s));
        FS_MARK(t, kSecScatter);
        FS_HIP(launch_face_epilogue(P, t->num, t->den, t->ct, t->vel, t->saved, nullptr, 0u, s));
        FS_MARK(t, kSecNormalize);
        t->gather_last = 0;
    }
    if (t->timing) FS_HIP(hipEventRecord(t->ev[1], s));
    FS_HIP(launch_div_cginit(P, t->vel, t->ct, t->occ, t->div, t->b, t->p, t->r, t->d0, t->cg, t->scan_ws, s));
    FS_HIP(launch_cg_reset(t->cg, s));
    FS_MARK(t, kSecDivergence);
    if (t->timing) FS_HIP(hipEventRecord(t->ev[4], s));
    if (t->cg_mode == 0) {
        FS_HIP(launch_cg_persistent(P, t->cg_vectors(), t->cg, t->cg_blocks, s));
    } else if (!t->cg_use_graph) {
        FS_HIP(launch_cg_multi(P, t->ct, t->b, t->p, t->r, t->d0, t->d1, t->q, t->cg, t->cg_multi_blocks, s));
    } else {
        if (!t->cg_exec) {
            // Capture the whole solve (2 x cap + 2 kernels) once; the graph is
            // dropped whenever the step constants change.
            hipGraph_t g = nullptr;
            FS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            const uint64_t before = g_dispatches;
            hipError_t le =
                launch_cg_multi(P, t->ct, t->b, t->p, t->r, t->d0, t->d1, t->q, t->cg, t->cg_multi_blocks, s);
            t->cg_graph_dispatches = g_dispatches - before;  // issued by each replay, not by the capture
            g_dispatches = before;
            hipError_t ce = hipStreamEndCapture(s, &g);
            if (le != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return fail(FS_ERR_DEVICE, std::string("CG capture: ") + hipGetErrorString(le));
            }
            FS_HIP(ce);
            hipError_t ie = hipGraphInstantiate(&t->cg_exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            FS_HIP(ie);
        }
        FS_HIP(hipGraphLaunch(t->cg_exec, s));
        g_dispatches += t->cg_graph_dispatches;
    }
    FS_MARK(t, kTagCgSolve);
    if (t->timing) FS_HIP(hipEventRecord(t->ev[2], s));
    FS_HIP(launch_gradient_enforce(P, t->p, t->ct, t->vel, s));
    FS_MARK(t, kSecGradient);
    if (stop_before_g2p) {
        if (sort_now) {
            FS_HIP(launch_permute(P.N, t->perm, pre, live, t->bf16, s));
            flip_live_side(t);
        }
        return FS_OK;
    }
    // The gather's cell order (the sort's permutation and bin starts, or the
    // unsorted substep's gperm and goff) also serves the row-chunk G2P.  By
    // default it runs on sort substeps of 2^21 particles or more when the row
    // chunks hold 128 or more on average: with fewer particles, or sparser
    // chunks, its workgroups are too few or their lanes mostly idle, and one
    // particle per thread measured faster.  On a substep without a sort it
    // would read and write the particles in place through the cell order,
    // both scattered, where one particle per thread in storage order reads
    // and writes them contiguously and measured faster.
    const bool rows_forced = t->g2p_rows == 1;
    const bool rows = rows_forced || (t->g2p_rows < 0 && P.N >= (1u << 21) &&
                                      (uint64_t)P.N >= 128u * g2p_row_chunks(P));
    if (sort_now) {
        if (rows)
            FS_HIP(launch_g2p_rows(P, live, pre, t->perm, t->cell_offset, t->occ, 0, t->bf16, t->vel, t->saved, s));
        else
            FS_HIP(launch_g2p(P, live, t->bf16, t->vel, t->saved, s, &pre, t->perm));
        flip_live_side(t);
    } else if (rows_forced && t->gather_last) {
        FS_HIP(launch_g2p_rows(P, live, live, t->gperm, t->goff, t->occ, 1, t->bf16, t->vel, t->saved, s));
    } else {
        FS_HIP(launch_g2p(P, live, t->bf16, t->vel, t->saved, s));
    }
    FS_MARK(t, kSecG2p);
    if (t->timing) FS_HIP(hipEventRecord(t->ev[3], s));
    return FS_OK;
}

// Adds the spans of the substep's boundary events to the detailed totals.
int collect_detailed(fs_tank* t) {
    if (t->dev_n == 0) return FS_OK;
    FS_HIP(hipEventSynchronize(t->dev_ev[t->dev_n - 1]));
    for (size_t i = 1; i < t->dev_n; ++i) {
        float ms = 0.0f;
        FS_HIP(hipEventElapsedTime(&ms, t->dev_ev[i - 1], t->dev_ev[i]));
        const int tag = t->dev_tag[i];
        if (tag >= 0 && tag < kFineSections) t->sec_ms[tag] += ms;
        else if (tag >= kTagCg && tag < kTagCg + kCgCategories) t->cg_cat_ms[tag - kTagCg] += ms;
        if (tag >= kTagCg) t->cg_total_ms += ms;
    }
    t->dev_n = 0;
    t->substeps_timed += 1;
    return FS_OK;
}

int do_step(fs_tank* t, uint32_t substeps) {
    double acc[4] = {0, 0, 0, 0};
    if (t->detailed) {
        // Totals over this call's substeps (GpuTimers' frame totals).
        std::fill(std::begin(t->sec_ms), std::end(t->sec_ms), 0.0);
        std::fill(std::begin(t->cg_cat_ms), std::end(t->cg_cat_ms), 0.0);
        t->cg_total_ms = 0.0;
        t->substeps_timed = 0;
        t->cg_iters_timed = t->P.iters;
    }
    t->call_substeps = substeps;
    for (uint32_t i = 0; i < substeps; ++i) {
        bool sort_now = false;
        if (t->prm.particle_sort) {
            const uint64_t period = t->prm.particle_sort_period ? t->prm.particle_sort_period : 1;
            sort_now = (t->tick % period) == 0;
            t->tick += 1;
        }
        const uint64_t d0 = g_dispatches;
        t->dev_armed = t->detailed != 0;
        t->dev_n = 0;
        int rc = t->prm.engine == FS_ENGINE_FUSED ? substep_fused(t, sort_now) : substep_reference(t, sort_now);
        t->dev_armed = false;
        if (rc != FS_OK) return rc;
        t->dispatches_last = (uint32_t)(g_dispatches - d0);
        t->sorted_last = sort_now ? 1 : 0;
        t->substeps += 1;
        if (t->detailed)
            if (int rc2 = collect_detailed(t)) return rc2;
        if (t->timing) {
            FS_HIP(hipEventSynchronize(t->ev[3]));
            float a = 0, b = 0, c = 0, g = 0;
            FS_HIP(hipEventElapsedTime(&a, t->ev[0], t->ev[1]));
            FS_HIP(hipEventElapsedTime(&b, t->ev[1], t->ev[2]));
            FS_HIP(hipEventElapsedTime(&c, t->ev[2], t->ev[3]));
            FS_HIP(hipEventElapsedTime(&g, t->ev[4], t->ev[2]));
            acc[0] += a; acc[1] += b; acc[2] += c; acc[3] += g;
        }
    }
    if (t->timing) {
        t->ms[0] = acc[0]; t->ms[1] = acc[1]; t->ms[2] = acc[2];
        t->ms[3] = acc[0] + acc[1] + acc[2];
        t->ms[4] = acc[3];
    }
    return FS_OK;
}

}  // namespace fs

namespace {

int check_tank(const fs_tank* t) {
    if (!t) return fail(FS_ERR_INVALID_ARGUMENT, "null tank handle");
    return FS_OK;
}

// Clears the solver's sticky failure word (after it was reported, or on reset).
hipError_t clear_solver_failure(fs_tank* t) {
    const uint32_t zero = 0u;
    return hipMemcpy(&t->cg.state->failed, &zero, sizeof zero, hipMemcpyHostToDevice);
}

}  // namespace

// ===========================================================================
extern "C" {

FS_API const char* fs_last_error(void) { return g_last_error.c_str(); }
FS_API const char* fs_version(void) { return "fluidstep 0.1 (gfx950)"; }

FS_API fs_registry* fs_registry_create(void) { return new (std::nothrow) fs_registry(); }
FS_API void fs_registry_destroy(fs_registry* r) { delete r; }

FS_API int fs_registry_set(fs_registry* r, const char* id, double value, fs_mutation* out) {
    if (!r || !id) return fail(FS_ERR_INVALID_ARGUMENT, "null registry or id");
    mutate(r->values, id, value, out);
    return FS_OK;
}

FS_API int fs_registry_get(const fs_registry* r, const char* id, double* value) {
    if (!r || !id || !value) return fail(FS_ERR_INVALID_ARGUMENT, "null argument");
    int i = row_index(id);
    if (i < 0) return fail(FS_ERR_INVALID_ARGUMENT, std::string("unknown setting id ") + id);
    *value = r->values[i];
    return FS_OK;
}

FS_API int fs_registry_count(const fs_registry* r) { return r ? kNumRows : 0; }
FS_API const char* fs_registry_id(const fs_registry* r, int index) {
    return (r && index >= 0 && index < kNumRows) ? kRows[index].id : nullptr;
}

FS_API int fs_params_default(fs_params* p) {
    if (!p) return fail(FS_ERR_INVALID_ARGUMENT, "null params");
    fs_registry r;
    const float H = 2.0f / 64.0f;  // sim::H (src/sim/mod.rs:26)
    float origin[3] = {-(80.0f) * H / 2.0f, -(40.0f) * H / 2.0f, -(80.0f) * H / 2.0f};
    return fs_registry_to_params(&r, H, origin, 8.0f, p);
}

FS_API int fs_registry_to_params(const fs_registry* r, float h, const float origin[3], float scene_density,
                                 fs_params* out) {
    if (!r || !out || !origin) return fail(FS_ERR_INVALID_ARGUMENT, "null argument");
    const auto& v = r->values;
    fs_params p;
    std::memset(&p, 0, sizeof p);
    p.nx = (uint32_t)value_of(v, "grid.res_x");
    p.ny = (uint32_t)value_of(v, "grid.res_y");
    p.nz = (uint32_t)value_of(v, "grid.res_z");
    p.h = h;
    for (int a = 0; a < 3; ++a) p.origin[a] = origin[a];
    p.dt = (float)value_of(v, "physics.fixed_dt");
    p.gravity[0] = 0.0f;
    p.gravity[1] = (float)value_of(v, "physics.gravity");
    p.gravity[2] = 0.0f;
    p.rho = 1000.0f;
    p.flip_blend = (float)value_of(v, "physics.flip_blend");
    p.wall_friction = (float)value_of(v, "physics.wall_friction");
    // effective_rest_density / effective_surface_dilation (src/scene/mod.rs:302-324)
    float rest = (float)value_of(v, "physics.rest_density");
    p.rest_density = rest > 0.0f ? rest : std::fmax(scene_density, 1.0e-3f);
    p.volume_stiffness = (float)value_of(v, "physics.volume_stiffness");
    p.drift_clamp = (float)value_of(v, "physics.drift_clamp");
    p.liquid_threshold = (uint32_t)value_of(v, "classify.liquid_threshold");
    uint32_t dil = (uint32_t)value_of(v, "classify.surface_dilation");
    uint32_t autod = scene_density < 8.0f ? 1u : 0u;
    p.surface_dilation = dil > autod ? dil : autod;
    p.cfl = (float)value_of(v, "physics.cfl");
    p.pressure_iterations = (uint32_t)value_of(v, "solver.pressure_iterations");
    p.pressure_warm_start = (uint32_t)value_of(v, "solver.pressure_warm_start");
    p.pressure_residual_tolerance = (float)value_of(v, "solver.pressure_residual_tolerance");
    p.particle_sort = (uint32_t)value_of(v, "dev.particle_sort");
    p.particle_sort_period = (uint32_t)value_of(v, "dev.particle_sort_period");
    p.engine = FS_ENGINE_FUSED;
    p.precision = FS_PRECISION_MIXED;
    p.velocity_storage = FS_VELOCITY_F32;
    p.scene_density = scene_density;
    *out = p;
    return FS_OK;
}

FS_API int fs_create(const fs_params* prm, const float* pos_xyz, const float* vel_xyz, uint32_t n, fs_tank** out) {
    if (!prm || !out || (n > 0 && !pos_xyz)) return fail(FS_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    const fs_params& q = *prm;
    const uint64_t C = (uint64_t)q.nx * q.ny * q.nz;
    if (int rc0 = check_params(q, n, pos_xyz, vel_xyz); rc0 != FS_OK) return rc0;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(FS_ERR_DEVICE, "no HIP device available");

    fs_tank* t = new (std::nothrow) fs_tank();
    if (!t) return fail(FS_ERR_OUT_OF_MEMORY, "host allocation failed");
    t->prm = q;
    t->N = n;
    t->bf16 = q.velocity_storage == FS_VELOCITY_BF16;
    build_step_params(q, n, t->P);
    int rc = [&]() -> int {
        FS_HIP(hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking));
        for (int sd = 0; sd < 2; ++sd)
            for (int a = 0; a < 3; ++a) {
                FS_HIP(t->alloc(t->side[sd].pos[a], n));
                if (t->bf16) {
                    uint16_t* vp;
                    FS_HIP(t->alloc(vp, n));
                    t->side[sd].vel[a] = vp;
                } else {
                    float* vp;
                    FS_HIP(t->alloc(vp, n));
                    t->side[sd].vel[a] = vp;
                }
            }
        FS_HIP(t->alloc(t->keys, n));
        FS_HIP(t->alloc(t->ktmp0, n));
        FS_HIP(t->alloc(t->ktmp1, n));
        FS_HIP(t->alloc(t->vtmp, n));
        FS_HIP(t->alloc(t->perm, n));
        const bool unsorted_substeps = !q.particle_sort || q.particle_sort_period > 1;
        if (q.engine == FS_ENGINE_FUSED && unsorted_substeps && !env_is("FLUIDSTEP_UNSORTED_P2G", "atomic")) {
            FS_HIP(t->alloc(t->gperm, n));
            FS_HIP(t->alloc(t->goff, C));
            FS_HIP(t->alloc(t->cursor, C));
        }
        FS_HIP(t->alloc(t->radix_ws, radix_workspace(n)));
        FS_HIP(t->alloc(t->scan_ws, scan_workspace(C) + scan_workspace(C / 64 + 1)));
        FS_HIP(t->alloc(t->occ, C));
        FS_HIP(t->alloc(t->cell_offset, C));
        FS_HIP(t->alloc(t->ct, C));
        for (int a = 0; a < 3; ++a) {
            FS_HIP(t->alloc(t->num[a], t->P.F[a]));
            FS_HIP(t->alloc(t->den[a], t->P.F[a]));
            FS_HIP(t->alloc(t->vel[a], t->P.F[a]));
            FS_HIP(t->alloc(t->saved[a], t->P.F[a]));
        }
        FS_HIP(t->alloc(t->div, C));
        FS_HIP(t->alloc(t->b, C));
        FS_HIP(t->alloc(t->p, C));
        FS_HIP(t->alloc(t->r, C));
        FS_HIP(t->alloc(t->d0, C));
        FS_HIP(t->alloc(t->d1, C));
        FS_HIP(t->alloc(t->q, C));
        FS_HIP(t->alloc(t->r1, C));
        for (int k = 0; k < 2; ++k) {
            FS_HIP(t->alloc(t->cg_w[k], C));
            FS_HIP(t->alloc(t->cg_s[k], C));
        }
        FS_HIP(t->alloc(t->scal, 8));
        FS_HIP(t->alloc(t->partials, (C + 255) / 256));
        FS_HIP(t->alloc(t->counters, 4));
        FS_HIP(t->alloc(t->mis_list, kMisfitCap));
        // FLUIDSTEP_GATHER_ORDER=0 runs the gather's bricks in brick order.
        const char* gorder = std::getenv("FLUIDSTEP_GATHER_ORDER");
        if (!(gorder && std::atoi(gorder) == 0)) {
            const size_t nb = 2 * gather_brick_count_max(t->P);
            FS_HIP(t->alloc(t->brick_ws, nb));
            FS_HIP(hipMemset(t->brick_ws, 0, nb * sizeof(uint32_t)));
        }
        const size_t ntiles = C / 64 + 1;
        int dev = 0, cus = 0;
        FS_HIP(hipGetDevice(&dev));
        FS_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        t->cg_blocks = cg_persistent_max_blocks();
        if (t->cg_blocks <= 0) return fail(FS_ERR_DEVICE, "cannot size the persistent CG grid");
        const int max_blocks = t->cg_blocks;
        t->cg_blocks = persistent_grid_blocks(max_blocks, cus);
        // The override may only shrink the grid below the default sizing.
        if (const char* env = std::getenv("FLUIDSTEP_CG_BLOCKS")) {
            int v = std::atoi(env);
            if (v > 0 && v <= t->cg_blocks) t->cg_blocks = v;
        }
        if (t->cg_blocks <= 0) t->cg_blocks = 1;
        t->cg_multi_blocks = 4 * cus;
        if (const char* env = std::getenv("FLUIDSTEP_CG_MULTI_BLOCKS")) {
            int v = std::atoi(env);
            if (v > 0 && v <= 65536) t->cg_multi_blocks = v;
        }
        if (const char* env = std::getenv("FLUIDSTEP_CG_MODE")) t->cg_mode = std::strcmp(env, "persistent") == 0 ? 0 : 1;
        if (const char* env = std::getenv("FLUIDSTEP_CG_GRAPH")) t->cg_use_graph = std::atoi(env) != 0;
        if (const char* env = std::getenv("FLUIDSTEP_P2G_FORM")) t->p2g_particle = std::strcmp(env, "particle") == 0;
        if (const char* env = std::getenv("FLUIDSTEP_G2P_ROWS")) t->g2p_rows = std::atoi(env) != 0 ? 1 : 0;
        if (const char* env = std::getenv("FLUIDSTEP_GATHER_BRICK")) {
            const int v = std::atoi(env);
            if (v >= 0 && v <= 2) t->gather_brick = v;
        }
        if (std::getenv("FLUIDSTEP_VERBOSE"))
            std::fprintf(stderr, "fluidstep: CG %s, persistent grid %d (co-resident max %d), multi grid %d, graph %d\n",
                         t->cg_mode == 0 ? "persistent" : "multi-kernel", t->cg_blocks, max_blocks,
                         t->cg_multi_blocks, t->cg_use_graph);
        uint32_t* sync = nullptr;
        FS_HIP(t->alloc(sync, cg_sync_words(max_blocks)));
        cg_bind_sync(t->cg, sync, max_blocks);
        FS_HIP(t->alloc(t->cg.nmask, C));
        // Two halves of up to two sums per block: the persistent form alternates
        // the halves between reductions.
        FS_HIP(t->alloc(t->cg.partials, 2 * 2 * (size_t)std::max(max_blocks, t->cg_multi_blocks)));
        FS_HIP(t->alloc(t->cg.result, 4));
        FS_HIP(t->alloc(t->cg.tile_flags, ntiles));
        FS_HIP(t->alloc(t->cg.tile_scan, ntiles));
        FS_HIP(t->alloc(t->cg.tiles, ntiles));
        FS_HIP(t->alloc(t->cg.n_tiles, 1));
        FS_HIP(t->alloc(t->cg.state, 1));
        for (auto& e : t->ev) FS_HIP(hipEventCreate(&e));
        // Zero everything once so read-outs before the first step are defined.
        for (int a = 0; a < 3; ++a) {
            FS_HIP(launch_fill_f32(t->vel[a], 0.0f, t->P.F[a], t->stream));
            FS_HIP(launch_fill_f32(t->saved[a], 0.0f, t->P.F[a], t->stream));
        }
        FS_HIP(launch_fill_f32(t->p, 0.0f, C, t->stream));
        FS_HIP(launch_fill_f32(t->div, 0.0f, C, t->stream));
        FS_HIP(launch_fill_u32(t->occ, 0u, C, t->stream));
        FS_HIP(launch_fill_u32(t->cell_offset, 0u, C, t->stream));
        FS_HIP(launch_fill_u32((uint32_t*)t->scal, 0u, 8, t->stream));
        FS_HIP(launch_fill_u32(t->counters, 0u, 4, t->stream));
        FS_HIP(launch_cg_reset(t->cg, t->stream));
        FS_HIP(launch_fill_u32((uint32_t*)t->cg.state, 0u, sizeof(CgState) / 4, t->stream));
        // cg_scalars[6] = tol_sq (set_pressure_residual_tolerance)
        FS_HIP(hipMemcpyAsync(t->scal + 6, &t->P.tol_sq, sizeof(float), hipMemcpyHostToDevice, t->stream));
        FS_HIP(hipStreamSynchronize(t->stream));
        return FS_OK;
    }();
    if (rc != FS_OK) {
        delete t;
        return rc;
    }
    if (n > 0) {
        t->init_pos.assign(pos_xyz, pos_xyz + 3 * (size_t)n);
        if (vel_xyz) t->init_vel.assign(vel_xyz, vel_xyz + 3 * (size_t)n);
        else t->init_vel.assign(3 * (size_t)n, 0.0f);
        rc = upload_particles(t, t->init_pos.data(), t->init_vel.data());
        if (rc != FS_OK) {
            delete t;
            return rc;
        }
    }
    // Registry view of the creation parameters (for fs_set_param results).
    *out = t;
    return FS_OK;
}

FS_API int fs_destroy(fs_tank* t) {
    if (!t) return FS_OK;
    if (t->stream) (void)hipStreamSynchronize(t->stream);
    delete t;
    return FS_OK;
}

FS_API int fs_reset(fs_tank* t) {
    if (int rc = check_tank(t)) return rc;
    FS_HIP(hipStreamSynchronize(t->stream));
    if (t->N > 0) {
        int rc = upload_particles(t, t->init_pos.data(), t->init_vel.data());
        if (rc != FS_OK) return rc;
    }
    FS_HIP(launch_fill_f32(t->p, 0.0f, t->P.C, t->stream));
    FS_HIP(hipStreamSynchronize(t->stream));
    FS_HIP(clear_solver_failure(t));
    t->tick = 0;
    t->substeps = 0;
    t->perm_valid = 0;
    t->sort_swapped = true;  // src/gpu/fluid.rs:1415
    return FS_OK;
}

FS_API int fs_load_state(fs_tank* t, const float* pos_xyz, const float* vel_xyz, const float* pressure) {
    if (int rc = check_tank(t)) return rc;
    if (t->N > 0 && !pos_xyz) return fail(FS_ERR_INVALID_ARGUMENT, "null positions");
    if (!all_finite(pos_xyz, 3 * (size_t)t->N) || !all_finite(vel_xyz, vel_xyz ? 3 * (size_t)t->N : 0) ||
        !all_finite(pressure, pressure ? t->P.C : 0))
        return fail(FS_ERR_NOT_FINITE, "non-finite state");
    FS_HIP(hipStreamSynchronize(t->stream));
    if (t->N > 0) {
        int rc = upload_particles(t, pos_xyz, vel_xyz);
        if (rc != FS_OK) return rc;
        t->sort_swapped = true;  // the live side is side 0 again
    }
    if (pressure) {
        FS_HIP(hipMemcpyAsync(t->p, pressure, sizeof(float) * t->P.C, hipMemcpyHostToDevice, t->stream));
        FS_HIP(hipStreamSynchronize(t->stream));
    }
    return FS_OK;
}

FS_API int fs_step(fs_tank* t, uint32_t substeps) {
    if (int rc = check_tank(t)) return rc;
    return do_step(t, substeps);
}

FS_API int fs_get_device_view(fs_tank* t, fs_device_view* out) {
    if (int rc = check_tank(t)) return rc;
    if (!out) return fail(FS_ERR_INVALID_ARGUMENT, "null view");
    fs_device_view v;
    std::memset(&v, 0, sizeof v);
    const ParticleSoA& sd = t->side[t->cur];
    for (int a = 0; a < 3; ++a) {
        v.pos[a] = sd.pos[a];
        v.vel[a] = sd.vel[a];
        v.face_vel[a] = t->vel[a];
    }
    v.n = t->N;
    v.velocity_storage = t->bf16 ? FS_VELOCITY_BF16 : FS_VELOCITY_F32;
    v.live_side = (uint32_t)t->cur;
    v.dims[0] = t->P.nx;
    v.dims[1] = t->P.ny;
    v.dims[2] = t->P.nz;
    v.cell_type = t->ct;
    v.pressure = t->p;
    v.occupancy = t->occ;
    v.stats = t->counters;
    v.stream = (void*)t->stream;
    *out = v;
    return FS_OK;
}

FS_API int fs_take_sort_swapped(fs_tank* t) {
    if (check_tank(t) != FS_OK) return 0;
    const bool s = t->sort_swapped;
    t->sort_swapped = false;
    return s ? 1 : 0;
}

FS_API int fs_pack_particles_vec4(fs_tank* t, float* dst_device) {
    if (int rc = check_tank(t)) return rc;
    if (!dst_device && t->N > 0) return fail(FS_ERR_INVALID_ARGUMENT, "null destination");
    FS_HIP(launch_pack_particles_vec4(t->side[t->cur], t->N, t->bf16, dst_device, t->stream));
    return FS_OK;
}

FS_API int fs_pack_cell_types_u32(fs_tank* t, uint32_t* dst_device) {
    if (int rc = check_tank(t)) return rc;
    if (!dst_device) return fail(FS_ERR_INVALID_ARGUMENT, "null destination");
    FS_HIP(launch_widen_u8(t->ct, t->P.C, dst_device, t->stream));
    return FS_OK;
}

FS_API int fs_synchronize(fs_tank* t) {
    if (int rc = check_tank(t)) return rc;
    FS_HIP(hipStreamSynchronize(t->stream));
    if (t->prm.engine == FS_ENGINE_FUSED) {
        // A solve that timed out in any substep since the last check leaves
        // the sticky word set (a later successful solve does not clear it):
        // report it once and clear it.
        CgState st;
        FS_HIP(hipMemcpy(&st, t->cg.state, sizeof st, hipMemcpyDeviceToHost));
        if (st.failed) {
            FS_HIP(clear_solver_failure(t));
            return fail(FS_ERR_SOLVER, "persistent CG grid barrier timed out");
        }
    }
    return FS_OK;
}

FS_API int fs_set_timing(fs_tank* t, int enabled) {
    if (int rc = check_tank(t)) return rc;
    t->timing = enabled ? 1 : 0;
    return FS_OK;
}

FS_API int fs_set_param(fs_tank* t, const char* id, double value, fs_mutation* out) {
    if (int rc = check_tank(t)) return rc;
    if (!id) return fail(FS_ERR_INVALID_ARGUMENT, "null id");
    fs_mutation m;
    mutate(t->settings, id, value, &m);
    m.applied_live = 0;
    if (m.status == FS_STATUS_APPLIED) {
        const double v = m.stored_value;
        const std::string s(id);
        fs_params& q = t->prm;
        bool changed = true;
        if (s == "physics.gravity") {
            q.gravity[0] = 0.0f; q.gravity[1] = (float)v; q.gravity[2] = 0.0f;  // push_gravity, identity box
        } else if (s == "physics.flip_blend") q.flip_blend = (float)v;
        else if (s == "physics.wall_friction") q.wall_friction = (float)v;
        else if (s == "physics.rest_density") q.rest_density = v > 0.0 ? (float)v : std::fmax(q.scene_density, 1.0e-3f);
        else if (s == "physics.volume_stiffness") q.volume_stiffness = (float)v;
        else if (s == "physics.drift_clamp") q.drift_clamp = (float)v;
        else if (s == "physics.cfl") q.cfl = (float)v;
        else if (s == "classify.liquid_threshold") q.liquid_threshold = (uint32_t)v;
        else if (s == "classify.surface_dilation") {
            uint32_t autod = q.scene_density < 8.0f ? 1u : 0u;
            q.surface_dilation = (uint32_t)v > autod ? (uint32_t)v : autod;
        } else if (s == "solver.pressure_iterations") q.pressure_iterations = (uint32_t)v;
        else if (s == "solver.pressure_warm_start") q.pressure_warm_start = v >= 0.5 ? 1u : 0u;
        else if (s == "solver.pressure_residual_tolerance") q.pressure_residual_tolerance = (float)v;
        else changed = false;  // physics.sim_speed: frame-driver knob
        if (changed) {
            FS_HIP(hipStreamSynchronize(t->stream));
            build_step_params(q, t->N, t->P);
            t->drop_graph();
            FS_HIP(hipMemcpyAsync(t->scal + 6, &t->P.tol_sq, sizeof(float), hipMemcpyHostToDevice, t->stream));
            FS_HIP(hipStreamSynchronize(t->stream));
        }
        m.applied_live = 1;
    }
    if (out) *out = m;
    return FS_OK;
}

FS_API int fs_set_gravity(fs_tank* t, float gx, float gy, float gz) {
    if (int rc = check_tank(t)) return rc;
    if (!std::isfinite(gx) || !std::isfinite(gy) || !std::isfinite(gz)) return fail(FS_ERR_NOT_FINITE, "non-finite gravity");
    t->prm.gravity[0] = gx; t->prm.gravity[1] = gy; t->prm.gravity[2] = gz;
    FS_HIP(hipStreamSynchronize(t->stream));
    build_step_params(t->prm, t->N, t->P);
    t->drop_graph();
    return FS_OK;
}

FS_API int fs_apply_impulse(fs_tank* t, float dx, float dy, float dz) {
    if (int rc = check_tank(t)) return rc;
    if (!std::isfinite(dx) || !std::isfinite(dy) || !std::isfinite(dz)) return fail(FS_ERR_NOT_FINITE, "non-finite impulse");
    FS_HIP(launch_impulse(t->P, t->side[t->cur], t->bf16, dx, dy, dz, t->stream));
    return FS_OK;
}

FS_API int fs_read_particles(fs_tank* t, float* pos_xyz, float* vel_xyz) {
    if (int rc = check_tank(t)) return rc;
    FS_HIP(hipStreamSynchronize(t->stream));
    const uint32_t n = t->N;
    if (n == 0) return FS_OK;
    std::vector<float> col(n);
    std::vector<uint16_t> colb(n);
    for (int a = 0; a < 3; ++a) {
        if (pos_xyz) {
            FS_HIP(hipMemcpy(col.data(), t->side[t->cur].pos[a], n * 4, hipMemcpyDeviceToHost));
            for (uint32_t i = 0; i < n; ++i) pos_xyz[3 * (size_t)i + a] = col[i];
        }
        if (vel_xyz) {
            if (t->bf16) {
                FS_HIP(hipMemcpy(colb.data(), t->side[t->cur].vel[a], n * 2, hipMemcpyDeviceToHost));
                for (uint32_t i = 0; i < n; ++i) {
                    uint32_t u = ((uint32_t)colb[i]) << 16;
                    std::memcpy(&vel_xyz[3 * (size_t)i + a], &u, 4);
                }
            } else {
                FS_HIP(hipMemcpy(col.data(), t->side[t->cur].vel[a], n * 4, hipMemcpyDeviceToHost));
                for (uint32_t i = 0; i < n; ++i) vel_xyz[3 * (size_t)i + a] = col[i];
            }
        }
    }
    return FS_OK;
}

FS_API int fs_read_faces(fs_tank* t, int axis, int which, float* out) {
    if (int rc = check_tank(t)) return rc;
    if (axis < 0 || axis > 2 || !out || (which != FS_FACE_VELOCITY && which != FS_FACE_SAVED))
        return fail(FS_ERR_INVALID_ARGUMENT, "bad axis/which/out");
    FS_HIP(hipStreamSynchronize(t->stream));
    const float* src = which == FS_FACE_VELOCITY ? t->vel[axis] : t->saved[axis];
    FS_HIP(hipMemcpy(out, src, sizeof(float) * t->P.F[axis], hipMemcpyDeviceToHost));
    return FS_OK;
}

FS_API int fs_read_cells_f32(fs_tank* t, int which, float* out) {
    if (int rc = check_tank(t)) return rc;
    if (!out || (which != FS_CELL_PRESSURE && which != FS_CELL_DIVERGENCE))
        return fail(FS_ERR_INVALID_ARGUMENT, "bad which/out");
    FS_HIP(hipStreamSynchronize(t->stream));
    FS_HIP(hipMemcpy(out, which == FS_CELL_PRESSURE ? t->p : t->div, sizeof(float) * t->P.C, hipMemcpyDeviceToHost));
    return FS_OK;
}

FS_API int fs_read_cells_u32(fs_tank* t, int which, uint32_t* out) {
    if (int rc = check_tank(t)) return rc;
    if (!out) return fail(FS_ERR_INVALID_ARGUMENT, "null out");
    FS_HIP(hipStreamSynchronize(t->stream));
    const uint32_t C = t->P.C;
    if (which == FS_CELL_TYPE) {
        std::vector<uint8_t> tmp(C);
        FS_HIP(hipMemcpy(tmp.data(), t->ct, C, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < C; ++i) out[i] = tmp[i];
    } else if (which == FS_CELL_OCCUPANCY) {
        FS_HIP(hipMemcpy(out, t->occ, sizeof(uint32_t) * C, hipMemcpyDeviceToHost));
    } else if (which == FS_CELL_OFFSET) {
        FS_HIP(hipMemcpy(out, t->cell_offset, sizeof(uint32_t) * C, hipMemcpyDeviceToHost));
    } else {
        return fail(FS_ERR_INVALID_ARGUMENT, "bad which");
    }
    return FS_OK;
}

FS_API int fs_read_permutation(fs_tank* t, uint32_t* perm) {
    if (int rc = check_tank(t)) return rc;
    if (!perm) return fail(FS_ERR_INVALID_ARGUMENT, "null out");
    if (!t->perm_valid) return fail(FS_ERR_INVALID_ARGUMENT, "no sort has run since create/reset");
    FS_HIP(hipStreamSynchronize(t->stream));
    if (t->N) FS_HIP(hipMemcpy(perm, t->perm, sizeof(uint32_t) * t->N, hipMemcpyDeviceToHost));
    return FS_OK;
}

FS_API int fs_get_stats(fs_tank* t, fs_stats* out) {
    if (int rc = check_tank(t)) return rc;
    if (!out) return fail(FS_ERR_INVALID_ARGUMENT, "null out");
    FS_HIP(hipStreamSynchronize(t->stream));
    fs_stats s;
    std::memset(&s, 0, sizeof s);
    uint32_t cnt[4];
    FS_HIP(hipMemcpy(cnt, t->counters, sizeof cnt, hipMemcpyDeviceToHost));
    s.liquid_cells = cnt[0];
    s.sorted = (uint32_t)t->sorted_last;
    s.p2g_gather = (uint32_t)(t->gather_last && (t->p2g_particle || cnt[1] <= kMisfitCap));
    if (t->prm.engine == FS_ENGINE_FUSED) {
        CgState st;
        FS_HIP(hipMemcpy(&st, t->cg.state, sizeof st, hipMemcpyDeviceToHost));
        if (st.error || st.failed) return fail(FS_ERR_SOLVER, "persistent CG grid barrier timed out");
        s.cg_iterations = st.iterations;
        s.rs_initial = st.rs_init;
        s.rs_final = st.rs_old;
        s.true_residual_sq = st.true_res_sq;
    } else {
        float sc[8];
        FS_HIP(hipMemcpy(sc, t->scal, sizeof sc, hipMemcpyDeviceToHost));
        s.cg_iterations = cnt[2];
        s.rs_initial = sc[4];
        s.rs_final = sc[0];
        s.true_residual_sq = NAN;
    }
    s.ms_prep = t->ms[0];
    s.ms_pressure = t->ms[1];
    s.ms_finish = t->ms[2];
    s.ms_total = t->ms[3];
    s.ms_cg = t->ms[4];
    s.substeps = t->substeps;
    s.misfits = cnt[1];
    if (t->prm.engine == FS_ENGINE_FUSED) {
        uint32_t nt = 0;
        FS_HIP(hipMemcpy(&nt, t->cg.n_tiles, sizeof nt, hipMemcpyDeviceToHost));
        s.cg_active_cells = nt * 64u;
    }
    *out = s;
    return FS_OK;
}

FS_API int fs_get_params(const fs_tank* t, fs_params* out) {
    if (!t || !out) return fail(FS_ERR_INVALID_ARGUMENT, "null argument");
    *out = t->prm;
    return FS_OK;
}

FS_API uint32_t fs_particle_count(const fs_tank* t) { return t ? t->N : 0u; }
FS_API uint64_t fs_buffer_bytes(const fs_tank* t) { return t ? t->bytes : 0u; }

}  // extern "C"
