// fs_frame.cpp — the host loop around the step: the fixed-timestep
// accumulator (src/timestep.rs), the seeded interaction schedules and the
// run state of FluidApp (src/lib.rs:55-264, 352-398, 506-601, 907-937).
//
// The accumulator keeps the reference's f32 expressions in its order, so its
// counts and stats equal the reference's bit for bit.  The schedules follow
// the reference's steps and draw order; their quaternion arithmetic is
// written here from glam's definitions (slerp, axis-angle, rotation of a
// vector), with the C library's acos/sin/cos.
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>

#include "fluidstep.h"
#include "fs_host_common.h"

using namespace fs;

namespace {

constexpr float kMaxRenderDt = 1.0f / 30.0f;              // MAX_RENDER_DT_S (src/timestep.rs:14)
constexpr float kInteractionMaxDt = 1.0f / 30.0f;         // INTERACTION_MAX_DT_S (src/lib.rs:53)
constexpr uint64_t kInteractionSeed = 0x464c5549445f5631ull;  // INTERACTION_SEED (src/lib.rs:52)

float finite_nonnegative(float v) { return std::isfinite(v) ? std::fmax(v, 0.0f) : 0.0f; }
float real_time_factor(float sim_advanced, float wall_dt) { return wall_dt > 0.0f ? sim_advanced / wall_dt : 0.0f; }

}  // namespace

// ---- TimestepController (src/timestep.rs:63-247) ----
struct fs_timestep {
    float accumulator = 0.0f;
    float fixed_dt = 0.0f;
    uint32_t max_substeps = 0;
    float dropped_time = 0.0f;
    fs_timestep_stats last{};

    fs_timestep(float dt, uint32_t cap) : fixed_dt(dt), max_substeps(cap) { std::memset(&last, 0, sizeof last); }

    fs_timestep_stats idle(float wall_dt, float scale) const {
        fs_timestep_stats s;
        std::memset(&s, 0, sizeof s);
        s.fixed_dt = fixed_dt;
        s.max_substeps = max_substeps;
        s.wall_dt = wall_dt;
        s.sim_time_scale = scale;
        s.policy = FS_TIMESTEP_FIXED_DROP_CAP;
        return s;
    }

    uint32_t steps_for_frame(float render_dt_s, float sim_time_scale) {  // steps_for_frame_scaled (:104-159)
        const float wall = finite_nonnegative(render_dt_s);
        const float scale = finite_nonnegative(sim_time_scale);
        const float clamped = std::fmin(wall * scale, kMaxRenderDt);
        accumulator += clamped;
        const float before = accumulator;
        uint32_t natural = 0;
        if (fixed_dt > 0.0f) {
            const float q = std::floor(accumulator / fixed_dt);  // `as u32` saturates
            natural = !(q > 0.0f) ? 0u : (q >= 4294967295.0f ? 4294967295u : (uint32_t)q);
        }
        const uint32_t n = natural < max_substeps ? natural : max_substeps;
        accumulator -= (float)n * fixed_dt;
        float dropped = 0.0f;
        if (natural > max_substeps) {
            dropped = accumulator;
            dropped_time += dropped;
            accumulator = 0.0f;
        }
        const float advanced = (float)n * fixed_dt;
        fs_timestep_stats s;
        s.substeps = n;
        s.natural_substeps = natural;
        s.max_substeps = max_substeps;
        s.substep_cap_hit = natural > max_substeps ? 1u : 0u;
        s.fixed_dt = fixed_dt;
        s.accumulated_before = before;
        s.accumulated_after = accumulator;
        s.dropped_this_frame = dropped;
        s.sim_advanced = advanced;
        s.wall_dt = wall;
        s.sim_time_scale = scale;
        s.real_time_factor = real_time_factor(advanced, wall);
        s.policy = FS_TIMESTEP_FIXED_DROP_CAP;
        last = s;
        return n;
    }

    void record_manual(float render_dt_s, uint32_t substeps, float sim_time_scale) {  // :163-183
        fs_timestep_stats s = idle(finite_nonnegative(render_dt_s), finite_nonnegative(sim_time_scale));
        s.substeps = substeps;
        s.sim_advanced = (float)substeps * fixed_dt;
        s.real_time_factor = real_time_factor(s.sim_advanced, s.wall_dt);
        last = s;
    }

    void record_export(float sim_dt_s, uint32_t substeps) {  // :187-205
        fs_timestep_stats s = idle(finite_nonnegative(sim_dt_s), 1.0f);
        s.substeps = substeps;
        s.max_substeps = substeps;
        s.natural_substeps = substeps;
        s.sim_advanced = (float)substeps * fixed_dt;
        s.real_time_factor = real_time_factor(s.sim_advanced, s.wall_dt);
        s.policy = FS_TIMESTEP_FIXED_EXPLICIT_EXPORT;
        last = s;
    }

    void record_idle(float render_dt_s, float sim_time_scale) {  // :209-216
        last = idle(finite_nonnegative(render_dt_s), finite_nonnegative(sim_time_scale));
    }

    void reset() {  // :237-244; the cumulative dropped time is kept
        accumulator = 0.0f;
        last = idle(0.0f, 1.0f);
    }
};

// ---- quaternion and vector arithmetic (glam's definitions, f32) ----
namespace {
struct V3 {
    float x, y, z;
};
struct Q {
    float x, y, z, w;
};
float dot(Q a, Q b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
Q scale(Q a, float s) { return {a.x * s, a.y * s, a.z * s, a.w * s}; }
Q add(Q a, Q b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
Q sub(Q a, Q b) { return {a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w}; }
Q normalize(Q a) { return scale(a, 1.0f / std::sqrt(dot(a, a))); }
float dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 scale3(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
V3 normalize3(V3 a) { return scale3(a, 1.0f / std::sqrt(dot3(a, a))); }
float acos_clamped(float v) { return std::acos(std::fmin(std::fmax(v, -1.0f), 1.0f)); }

Q from_axis_angle(V3 axis, float angle) {
    const float s = std::sin(angle * 0.5f), c = std::cos(angle * 0.5f);
    return {axis.x * s, axis.y * s, axis.z * s, c};
}
void to_axis_angle(Q q, V3& axis, float& angle) {
    angle = acos_clamped(q.w) * 2.0f;
    const float scale_sq = std::fmax(1.0f - q.w * q.w, 0.0f);
    if (scale_sq >= 1.0e-8f * 1.0e-8f)
        axis = scale3({q.x, q.y, q.z}, 1.0f / std::sqrt(scale_sq));
    else
        axis = {1.0f, 0.0f, 0.0f};
}
Q lerp(Q a, Q b, float s) {
    const float bias = dot(a, b) >= 0.0f ? 1.0f : -1.0f;
    return normalize(add(a, scale(sub(scale(b, bias), a), s)));
}
// Shortest-path spherical interpolation; linear near identical ends.
Q slerp(Q a, Q b, float s) {
    float d = dot(a, b);
    if (d < 0.0f) {
        b = scale(b, -1.0f);
        d = -d;
    }
    if (d > 1.0f - FLT_EPSILON) return lerp(a, b, s);
    const float theta = acos_clamped(d);
    const float s1 = std::sin(theta * (1.0f - s)), s2 = std::sin(theta * s), st = std::sin(theta);
    return scale(add(scale(a, s1), scale(b, s2)), 1.0f / st);
}
Q conjugate(Q q) { return {-q.x, -q.y, -q.z, q.w}; }
V3 rotate(Q q, V3 v) {  // q v q^-1 for a unit q
    const V3 b{q.x, q.y, q.z};
    const float b2 = dot3(b, b);
    const V3 t1 = scale3(v, q.w * q.w - b2), t2 = scale3(b, dot3(v, b) * 2.0f), t3 = scale3(cross(b, v), q.w * 2.0f);
    return {t1.x + t2.x + t3.x, t1.y + t2.y + t3.y, t1.z + t2.z + t3.z};
}
// clamp_rotation (src/lib.rs:246-255)
Q clamp_rotation(Q q, float max_angle) {
    max_angle = std::fmax(max_angle, 0.0f);
    V3 axis;
    float angle;
    to_axis_angle(q, axis, angle);
    if (angle > max_angle && dot3(axis, axis) > 0.0f) return normalize(from_axis_angle(axis, max_angle));
    return normalize(q);
}
constexpr Q kIdentity{0.0f, 0.0f, 0.0f, 1.0f};
}  // namespace

// ---- InteractionState (src/lib.rs:55-236) ----
struct fs_interactions {
    uint64_t state = 0;
    // AutoRollSchedule
    bool roll_enabled_last = false;
    float roll_elapsed = 0.0f, roll_duration = 0.0f;
    Q roll_from = kIdentity, roll_target = kIdentity;
    // WaveSchedule
    bool wave_enabled_last = false;
    float wave_time_until_next = 0.0f;
    uint32_t wave_phase = 0;

    explicit fs_interactions(uint64_t seed) { reset(seed); }
    void reset(uint64_t seed) {
        *this = fs_interactions();
        state = seed == 0 ? 0x9e3779b97f4a7c15ull : seed;  // InteractionRng::new
    }
    fs_interactions() = default;

    uint32_t next_u32() {  // :70-76
        state = state * 6364136223846793005ull + 1442695040888963407ull;
        return (uint32_t)(state >> 32);
    }
    float next_f32() {  // :78-81: (u32::MAX as f32 + 1) is 2^32
        return (float)next_u32() * (1.0f / 4294967296.0f);
    }
    float signed_f32() { return next_f32() * 2.0f - 1.0f; }

    void schedule_auto_roll(Q from, float strength, float cadence) {  // :188-203
        cadence = std::fmax(cadence, 0.1f);
        const float jitter = 0.75f + next_f32() * 0.5f;
        const float hx = signed_f32();
        const float hz = signed_f32();
        const V3 horizontal{hx, 0.0f, hz};
        const V3 axis = dot3(horizontal, horizontal) > 1.0e-6f ? normalize3(horizontal) : V3{1.0f, 0.0f, 0.0f};
        const float angle = strength * (0.35f + next_f32() * 0.65f);
        roll_from = clamp_rotation(normalize(from), strength);
        roll_target = normalize(from_axis_angle(axis, angle));
        roll_duration = cadence * jitter;
        roll_elapsed = 0.0f;
    }

    bool auto_roll(float dt, Q current, bool enabled, float strength, float cadence, Q& out) {  // :146-186
        strength = std::fmax(strength, 0.0f);
        if (!enabled || strength <= FLT_EPSILON) {
            roll_enabled_last = false;
            return false;
        }
        dt = finite_nonnegative(dt);
        if (!roll_enabled_last || roll_duration <= FLT_EPSILON) {
            schedule_auto_roll(current, strength, cadence);
            roll_enabled_last = true;
        }
        roll_elapsed += dt;
        while (roll_duration > FLT_EPSILON && roll_elapsed >= roll_duration) {
            const float overflow = roll_elapsed - roll_duration;
            schedule_auto_roll(roll_target, strength, cadence);
            roll_enabled_last = true;
            roll_elapsed = std::fmin(overflow, roll_duration);
        }
        float t = roll_elapsed / roll_duration;
        t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
        const float eased = t * t * (3.0f - 2.0f * t);
        out = clamp_rotation(normalize(slerp(roll_from, roll_target, eased)), strength);
        return true;
    }

    bool wave(float dt, bool enabled, float strength, float frequency, float dv[3]) {  // :205-236
        strength = std::fmax(strength, 0.0f);
        frequency = std::fmax(frequency, 0.0f);
        if (!enabled || strength <= FLT_EPSILON || frequency <= FLT_EPSILON) {
            wave_enabled_last = false;
            return false;
        }
        if (!wave_enabled_last) {
            wave_enabled_last = true;
            wave_time_until_next = 0.0f;
        }
        wave_time_until_next -= finite_nonnegative(dt);
        if (wave_time_until_next > 0.0f) return false;
        // wave_direction (:257-264): +X, -X, +Z, -Z (negation flips the zeros' signs)
        static const V3 dirs[4] = {{1.0f, 0.0f, 0.0f}, {-1.0f, -0.0f, -0.0f}, {0.0f, 0.0f, 1.0f}, {-0.0f, -0.0f, -1.0f}};
        const V3 d = dirs[wave_phase % 4u];
        wave_phase += 1u;
        wave_time_until_next += 1.0f / frequency;
        if (wave_time_until_next < 0.0f) wave_time_until_next = 1.0f / frequency;
        dv[0] = d.x * strength;
        dv[1] = 0.0f;
        dv[2] = d.z * strength;
        return true;
    }
};

// ---- FluidApp's frame loop ----
struct fs_frame {
    fs_tank* tank = nullptr;
    const fs_registry* settings = nullptr;
    fs_timestep timestep{0.0f, 0};
    fs_interactions interactions{kInteractionSeed};
    bool paused = false;
    uint32_t pending_steps = 0;
    uint64_t tick = 0;
    uint32_t reset_count = 0;
    Q box_orient = kIdentity;
};

namespace {
double setting(const fs_registry* r, const char* id) {
    double v = 0.0;
    fs_registry_get(r, id, &v);
    return v;
}

fs_timestep timestep_from(const fs_registry* r) {
    return fs_timestep((float)setting(r, "physics.fixed_dt"), (uint32_t)setting(r, "physics.max_substeps"));
}

// push_gravity (src/lib.rs:907-912): the world gravity in the box's frame.
int push_gravity(fs_frame* f) {
    const float mag = (float)setting(f->settings, "physics.gravity");
    const V3 g = rotate(conjugate(f->box_orient), {0.0f, mag, 0.0f});
    return fs_set_gravity(f->tank, g.x, g.y, g.z);
}

// update_interactions (src/lib.rs:914-937)
int update_interactions(fs_frame* f, float dt) {
    const fs_registry* r = f->settings;
    Q next;
    if (f->interactions.auto_roll(dt, f->box_orient, setting(r, "interaction.auto_roll_enabled") != 0.0,
                                  (float)setting(r, "interaction.auto_roll_strength"),
                                  (float)setting(r, "interaction.auto_roll_cadence"), next)) {
        if (std::fabs(dot(f->box_orient, next)) < 0.999999f) {
            f->box_orient = next;
            if (int rc = push_gravity(f)) return rc;
        }
    }
    float dv[3];
    if (f->interactions.wave(dt, setting(r, "interaction.wave_enabled") != 0.0,
                             (float)setting(r, "interaction.wave_strength"),
                             (float)setting(r, "interaction.wave_frequency"), dv))
        return fs_apply_impulse(f->tank, dv[0], dv[1], dv[2]);
    return FS_OK;
}

double finite_or_zero(double v) { return std::isfinite(v) ? v : 0.0; }
}  // namespace

// ---- C-ABI ----
FS_API fs_timestep* fs_timestep_create(float fixed_dt, uint32_t max_substeps) {
    return new (std::nothrow) fs_timestep(fixed_dt, max_substeps);
}
FS_API void fs_timestep_destroy(fs_timestep* c) { delete c; }
FS_API uint32_t fs_timestep_steps_for_frame(fs_timestep* c, float render_dt_s, float sim_time_scale) {
    return c ? c->steps_for_frame(render_dt_s, sim_time_scale) : 0u;
}
FS_API void fs_timestep_record_manual(fs_timestep* c, float render_dt_s, uint32_t substeps, float sim_time_scale) {
    if (c) c->record_manual(render_dt_s, substeps, sim_time_scale);
}
FS_API void fs_timestep_record_export(fs_timestep* c, float sim_dt_s, uint32_t substeps) {
    if (c) c->record_export(sim_dt_s, substeps);
}
FS_API void fs_timestep_record_idle(fs_timestep* c, float render_dt_s, float sim_time_scale) {
    if (c) c->record_idle(render_dt_s, sim_time_scale);
}
FS_API void fs_timestep_reset(fs_timestep* c) {
    if (c) c->reset();
}
FS_API int fs_timestep_last_stats(const fs_timestep* c, fs_timestep_stats* out) {
    if (!c || !out) return fail(FS_ERR_INVALID_ARGUMENT, "null argument");
    *out = c->last;
    return FS_OK;
}
FS_API float fs_timestep_total_dropped(const fs_timestep* c) { return c ? c->dropped_time : 0.0f; }

FS_API fs_interactions* fs_interactions_create(uint64_t seed) { return new (std::nothrow) fs_interactions(seed); }
FS_API void fs_interactions_destroy(fs_interactions* s) { delete s; }
FS_API void fs_interactions_reset(fs_interactions* s, uint64_t seed) {
    if (s) s->reset(seed);
}
FS_API int fs_interactions_auto_roll(fs_interactions* s, float dt_s, const float q_current[4], int enabled,
                                     float strength, float cadence_s, float q_out[4]) {
    if (!s || !q_current || !q_out) return 0;
    Q out;
    const Q cur{q_current[0], q_current[1], q_current[2], q_current[3]};
    if (!s->auto_roll(dt_s, cur, enabled != 0, strength, cadence_s, out)) return 0;
    q_out[0] = out.x;
    q_out[1] = out.y;
    q_out[2] = out.z;
    q_out[3] = out.w;
    return 1;
}
FS_API int fs_interactions_wave(fs_interactions* s, float dt_s, int enabled, float strength, float frequency_hz,
                                float dv_out[3]) {
    if (!s || !dv_out) return 0;
    return s->wave(dt_s, enabled != 0, strength, frequency_hz, dv_out) ? 1 : 0;
}
FS_API uint32_t fs_interactions_next_u32(fs_interactions* s) { return s ? s->next_u32() : 0u; }

FS_API int fs_frame_create(fs_tank* t, const fs_registry* settings, fs_frame** out) {
    if (!t || !settings || !out) return fail(FS_ERR_INVALID_ARGUMENT, "null argument");
    fs_frame* f = new (std::nothrow) fs_frame();
    if (!f) return fail(FS_ERR_OUT_OF_MEMORY, "host allocation failed");
    f->tank = t;
    f->settings = settings;
    f->timestep = timestep_from(settings);
    if (int rc = push_gravity(f)) {
        delete f;
        return rc;
    }
    *out = f;
    return FS_OK;
}
FS_API void fs_frame_destroy(fs_frame* f) { delete f; }

FS_API int fs_frame_run(fs_frame* f, double render_dt_ms, fs_timestep_stats* out) {
    if (!f) return fail(FS_ERR_INVALID_ARGUMENT, "null frame driver");
    const double ms = std::fmax(finite_or_zero(render_dt_ms), 0.0);
    const float dt_s = (float)ms / 1000.0f;
    const float sim_speed = (float)setting(f->settings, "physics.sim_speed");
    if (!f->paused) {
        float idt = dt_s * sim_speed;
        idt = idt < 0.0f ? 0.0f : (idt > kInteractionMaxDt ? kInteractionMaxDt : idt);
        if (int rc = update_interactions(f, idt)) return rc;
        const uint32_t n = f->timestep.steps_for_frame(dt_s, sim_speed);
        if (n > 0) {
            if (int rc = fs_step(f->tank, n)) return rc;
            f->tick += n;
        }
    } else {
        f->timestep.reset();  // no build-up while paused
        if (f->pending_steps > 0) {
            f->pending_steps -= 1;
            if (int rc = fs_step(f->tank, 1)) return rc;
            f->tick += 1;
            f->timestep.record_manual(dt_s, 1, sim_speed);
        } else {
            f->timestep.record_idle(dt_s, sim_speed);
        }
    }
    if (out) *out = f->timestep.last;
    return FS_OK;
}

FS_API int fs_frame_export(fs_frame* f, uint32_t substeps, double sim_dt_s, fs_timestep_stats* out) {
    if (!f) return fail(FS_ERR_INVALID_ARGUMENT, "null frame driver");
    const double dt = std::fmax(finite_or_zero(sim_dt_s), 0.0);
    if (int rc = update_interactions(f, (float)dt)) return rc;
    if (substeps > 0) {
        if (int rc = fs_step(f->tank, substeps)) return rc;
        f->tick += substeps;
    }
    f->timestep.record_export((float)dt, substeps);
    if (out) *out = f->timestep.last;
    return FS_OK;
}

FS_API int fs_frame_set_paused(fs_frame* f, int paused) {
    if (!f) return fail(FS_ERR_INVALID_ARGUMENT, "null frame driver");
    f->paused = paused != 0;
    return FS_OK;
}
FS_API int fs_frame_queue_step(fs_frame* f) {
    if (!f) return fail(FS_ERR_INVALID_ARGUMENT, "null frame driver");
    f->pending_steps += 1;
    return FS_OK;
}

FS_API int fs_frame_reset(fs_frame* f) {
    if (!f) return fail(FS_ERR_INVALID_ARGUMENT, "null frame driver");
    if (int rc = fs_reset(f->tank)) return rc;
    f->tick = 0;
    f->pending_steps = 0;
    f->reset_count += 1;
    f->timestep = timestep_from(f->settings);
    f->box_orient = kIdentity;
    // interaction_seed_for_reset (src/lib.rs:992-994)
    f->interactions.reset(kInteractionSeed ^ ((uint64_t)f->reset_count * 0x9e3779b97f4a7c15ull));
    return push_gravity(f);
}

FS_API uint64_t fs_frame_tick(const fs_frame* f) { return f ? f->tick : 0u; }
FS_API int fs_frame_timestep_stats(const fs_frame* f, fs_timestep_stats* out, float* total_dropped_s) {
    if (!f || !out) return fail(FS_ERR_INVALID_ARGUMENT, "null argument");
    *out = f->timestep.last;
    if (total_dropped_s) *total_dropped_s = f->timestep.dropped_time;
    return FS_OK;
}
FS_API int fs_frame_orientation(const fs_frame* f, float q_out[4]) {
    if (!f || !q_out) return fail(FS_ERR_INVALID_ARGUMENT, "null argument");
    q_out[0] = f->box_orient.x;
    q_out[1] = f->box_orient.y;
    q_out[2] = f->box_orient.z;
    q_out[3] = f->box_orient.w;
    return FS_OK;
}
