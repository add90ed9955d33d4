/*
 * fluidstep.h — C-ABI of the MI355X FLIP/PIC simulation step.
 *
 * The drop-in boundary is fluid-lab's `GpuFluid` (app/crates/fluid-lab/src/
 * gpu/fluid.rs:100-1716) together with the live-setting path of its registry
 * (src/settings/mod.rs:1733-1821) and `GpuContext::step`
 * (src/gpu/mod.rs:950-1033).  Each entry point below names the reference
 * interface it replaces.  Conventions: plain pointers and sizes; `int` status
 * codes (FS_OK == 0, negative on error, message in fs_last_error()); no
 * exceptions or panics cross the ABI; the caller owns every host array; a
 * handle owns its device buffers and HIP stream and is not thread-safe;
 * read-outs reflect the state after the last completed fs_step/fs_reset.
 * The product library needs a gfx950 GPU: without one, fs_create fails with
 * FS_ERR_DEVICE (there is no CPU fallback).
 */
#ifndef FLUIDSTEP_H
#define FLUIDSTEP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(_WIN32)
#define FS_API __declspec(dllexport)
#else
#define FS_API __attribute__((visibility("default")))
#endif

/* ---- status codes ---- */
#define FS_OK 0
#define FS_ERR_INVALID_ARGUMENT (-1)
#define FS_ERR_UNSUPPORTED (-2)
#define FS_ERR_DEVICE (-3)
#define FS_ERR_OUT_OF_MEMORY (-4)
#define FS_ERR_NOT_FINITE (-5)
#define FS_ERR_SOLVER (-6)

/* ---- step configuration ---- */
#define FS_ENGINE_REFERENCE 0 /* one kernel per WGSL pass, host-driven CG loop */
#define FS_ENGINE_FUSED 1     /* fused passes, sorted-segment P2G, persistent CG */
#define FS_PRECISION_F32 0    /* f32 everywhere, the WGSL f32 reduction trees */
#define FS_PRECISION_MIXED 1  /* f32 vectors, f64 dots / scalars / true residual */
#define FS_VELOCITY_F32 0
#define FS_VELOCITY_BF16 1    /* particle velocity stored as bf16, math in f32 */

/* Creation parameters: the reference Params uniform (src/gpu/fluid.rs:25-39,
 * filled at :289-318) with every field explicit. */
typedef struct fs_params {
    uint32_t nx, ny, nz;          /* grid.res_x/y/z */
    float h;                      /* uniform cell size (sim::H in the reference) */
    float origin[3];              /* world position of cell (0,0,0)'s low corner */
    float dt;                     /* physics.fixed_dt */
    float gravity[3];             /* grav (set_gravity_vec) */
    float rho;                    /* 1000 in the reference */
    float flip_blend;             /* physics.flip_blend */
    float wall_friction;          /* physics.wall_friction */
    float rest_density;           /* effective rest packing (particles per cell) */
    float volume_stiffness;       /* physics.volume_stiffness */
    float drift_clamp;            /* physics.drift_clamp */
    uint32_t liquid_threshold;    /* classify.liquid_threshold */
    uint32_t surface_dilation;    /* effective classify.surface_dilation */
    float cfl;                    /* physics.cfl */
    uint32_t pressure_iterations; /* solver.pressure_iterations (CG cap) */
    uint32_t pressure_warm_start; /* solver.pressure_warm_start */
    float pressure_residual_tolerance; /* solver.pressure_residual_tolerance */
    uint32_t particle_sort;       /* dev.particle_sort */
    uint32_t particle_sort_period;/* dev.particle_sort_period */
    uint32_t engine;              /* FS_ENGINE_* */
    uint32_t precision;           /* FS_PRECISION_* */
    uint32_t velocity_storage;    /* FS_VELOCITY_* */
    float scene_density;          /* particles per seeded cell, for Auto rest/dilation */
    uint32_t reserved[4];
} fs_params;

/* Result of a setting mutation (SettingMutationResult, src/settings/mod.rs:43-127). */
#define FS_STATUS_APPLIED 0
#define FS_STATUS_STORED 1
#define FS_STATUS_UNKNOWN_ID 2
#define FS_STATUS_NON_FINITE_REJECTED 3
#define FS_APPLY_LIVE 0
#define FS_APPLY_RESET 1
#define FS_APPLY_RELOAD 2
typedef struct fs_mutation {
    int32_t status;        /* FS_STATUS_* */
    int32_t accepted;      /* status is neither unknown_id nor non_finite_rejected */
    int32_t clamped;
    int32_t apply_class;   /* FS_APPLY_*, -1 when unknown */
    int32_t applied_live;  /* the step's parameters changed now */
    int32_t needs_reset;
    double requested_value;
    double stored_value;   /* NaN when nothing was stored */
} fs_mutation;

/* Per-step diagnostics (stats buffer, cg_scalars, timing sections). */
typedef struct fs_stats {
    uint32_t liquid_cells;     /* stats[0], sh/classify.wgsl:55-57 */
    uint32_t cg_iterations;    /* iterations run while the solve was active */
    uint32_t sorted;           /* the last substep re-sorted the particles */
    uint32_t p2g_gather;       /* the last substep used the sorted-segment gather */
    double rs_initial;         /* cg_scalars[4] */
    double rs_final;           /* cg_scalars[0] after the loop */
    double true_residual_sq;   /* |b - A p|^2 in f64 (mixed precision), else NaN */
    double ms_total;           /* device time of the last fs_step call (timing on) */
    double ms_prep;            /* clear, mark, sort, classify, P2G epilogue */
    double ms_pressure;        /* divergence + CG */
    double ms_finish;          /* gradient + enforce + G2P */
    uint64_t substeps;         /* substeps run since create/reset */
    double ms_cg;              /* the CG iterations alone (fused: the persistent kernel) */
    uint32_t cg_active_cells;  /* cells in 64-cell tiles holding liquid (fused engine) */
    uint32_t misfits;          /* particles whose P2G cell differs from their sort cell */
} fs_stats;

typedef struct fs_tank fs_tank;
typedef struct fs_registry fs_registry;

/* ---- settings registry (host only; no GPU needed) ---- */
/* Registry::default (src/settings/mod.rs:404) restricted to the step's rows. */
FS_API fs_registry* fs_registry_create(void);
FS_API void fs_registry_destroy(fs_registry* r);
/* Registry::set_value_f64_result (src/settings/mod.rs:1733-1821). */
FS_API int fs_registry_set(fs_registry* r, const char* id, double value, fs_mutation* out);
FS_API int fs_registry_get(const fs_registry* r, const char* id, double* value);
FS_API int fs_registry_count(const fs_registry* r);
FS_API const char* fs_registry_id(const fs_registry* r, int index);
/* GpuFluid::new's Params fill (src/gpu/fluid.rs:258-318) from registry values:
 * h, origin and scene_density come from the caller. */
FS_API int fs_registry_to_params(const fs_registry* r, float h, const float origin[3], float scene_density,
                                 fs_params* out);

/* ---- step lifecycle ---- */
/* Defaults of the reference registry (80x40x80, H = 2/64, centred origin). */
FS_API int fs_params_default(fs_params* p);
/* GpuFluid::new (src/gpu/fluid.rs:247-1197): allocate every buffer and upload
 * the initial particles.  pos_xyz / vel_xyz hold n xyz triplets (vel may be
 * NULL = zero).  Non-finite inputs are rejected, and so are grids of 2^30 or
 * more faces ((nx+1)(ny+1)(nz+1)): kernels address their arrays with 32-bit
 * byte offsets. */
FS_API int fs_create(const fs_params* p, const float* pos_xyz, const float* vel_xyz, uint32_t n, fs_tank** out);
FS_API int fs_destroy(fs_tank* t);
/* GpuFluid::reset (src/gpu/fluid.rs:1408-1418): initial particles back,
 * pressure zeroed, sort tick 0. */
FS_API int fs_reset(fs_tank* t);
/* Overwrite the live state (particles, and the warm-start pressure when not
 * NULL); used to start a step from a given state.  No reference counterpart. */
FS_API int fs_load_state(fs_tank* t, const float* pos_xyz, const float* vel_xyz, const float* pressure);
/* GpuContext::step(n) (src/gpu/mod.rs:950-989): n substeps in the order of
 * record_substep_coarse (src/gpu/mod.rs:993-1033).  Returns after enqueueing;
 * read-outs and fs_synchronize wait. */
FS_API int fs_step(fs_tank* t, uint32_t substeps);
FS_API int fs_synchronize(fs_tank* t);
/* Record per-section device times into fs_stats (GpuTimers, src/gpu/timing.rs). */
FS_API int fs_set_timing(fs_tank* t, int enabled);

/* ---- live setters ---- */
/* FluidApp::mutate_setting + apply_live_setting (src/lib.rs:680-867) ->
 * GpuFluid::set_* (src/gpu/fluid.rs:1295-1371). */
FS_API int fs_set_param(fs_tank* t, const char* id, double value, fs_mutation* out);
/* GpuFluid::set_gravity_vec (src/gpu/fluid.rs:1347-1350). */
FS_API int fs_set_gravity(fs_tank* t, float gx, float gy, float gz);
/* GpuFluid::apply_impulse (src/gpu/fluid.rs:1374-1395). */
FS_API int fs_apply_impulse(fs_tank* t, float dx, float dy, float dz);

/* ---- read-outs (GpuFluid::*_buffer accessors, src/gpu/fluid.rs:1199-1292) ---- */
#define FS_FACE_VELOCITY 0 /* u/v/w_vel after the step */
#define FS_FACE_SAVED 1    /* u/v/w_saved: post-P2G, pre-force */
#define FS_CELL_PRESSURE 0
#define FS_CELL_DIVERGENCE 1
#define FS_CELL_TYPE 0      /* 0 Solid, 1 Liquid, 2 Air */
#define FS_CELL_OCCUPANCY 1 /* particles per cell (mark) */
#define FS_CELL_OFFSET 2    /* bin starts of the last sort */
FS_API int fs_read_particles(fs_tank* t, float* pos_xyz, float* vel_xyz);
FS_API int fs_read_faces(fs_tank* t, int axis, int which, float* out);
FS_API int fs_read_cells_f32(fs_tank* t, int which, float* out);
FS_API int fs_read_cells_u32(fs_tank* t, int which, uint32_t* out);
/* perm[d] = the source index of sorted slot d in the last sort. */
FS_API int fs_read_permutation(fs_tank* t, uint32_t* perm);
FS_API int fs_get_stats(fs_tank* t, fs_stats* out);
FS_API int fs_get_params(const fs_tank* t, fs_params* out);
FS_API uint32_t fs_particle_count(const fs_tank* t);
FS_API uint64_t fs_buffer_bytes(const fs_tank* t);
/* Thread-local message of the last failing call on this thread. */
FS_API const char* fs_last_error(void);
FS_API const char* fs_version(void);

/* ---- batched tanks ----
 * Many independent tanks of uneven sizes stepped together on one GPU.  Each
 * kernel phase of a substep is one launch over every tank (a tank table and
 * per-phase work tables on the device); the pressure solve is one persistent
 * launch in which each tank's blocks run that tank's CG with their own
 * reductions; the whole substep is replayed from a captured hipGraph.  Per
 * tank the result is the fused engine's step (FS_ENGINE_FUSED,
 * FS_PRECISION_MIXED, particles sorted every substep); velocity storage must
 * be the same for every tank.  The reference steps one GpuFluid per
 * GpuContext (src/gpu/mod.rs:950-1033); a batch replaces a loop of
 * GpuFluid::new / GpuContext::step over tanks. */
typedef struct fs_batch fs_batch;

/* Tile shapes the autotuner picks from (each an index into a fixed menu). */
typedef struct fs_batch_tuning {
    uint32_t gather_brick; /* P2G gather brick: 0 = 8x8x8, 1 = 16x8x8, 2 = 32x4x4 cells */
    uint32_t sort_tile;    /* radix-sort keys per workgroup: 0 = 4096, 1 = 2048, 2 = 8192 */
    uint32_t cg_unroll;    /* CG tiles in flight per wave: 0 = 4, 1 = 2, 2 = 1 */
    uint32_t cg_blocks;    /* persistent CG workgroups (0 = default) */
    double ms_gather;      /* best measured time of each tuned phase (ms) */
    double ms_sort;
    double ms_cg;
} fs_batch_tuning;

/* params[i], pos_xyz[i], vel_xyz[i] (NULL = zero), counts[i] describe tank i. */
FS_API int fs_batch_create(const fs_params* params, const float* const* pos_xyz, const float* const* vel_xyz,
                           const uint32_t* counts, uint32_t ntanks, fs_batch** out);
FS_API int fs_batch_destroy(fs_batch* b);
/* n substeps of every tank (replays the captured substep graph). */
FS_API int fs_batch_step(fs_batch* b, uint32_t substeps);
FS_API int fs_batch_synchronize(fs_batch* b);
FS_API int fs_batch_set_timing(fs_batch* b, int enabled);
/* Time each menu entry of the tuned phases on the batch's current state
 * (restored afterwards) and keep the fastest; out (nullable) receives the
 * choice.  Without a call the first menu entries are used. */
FS_API int fs_batch_autotune(fs_batch* b, fs_batch_tuning* out);
FS_API int fs_batch_set_tuning(fs_batch* b, const fs_batch_tuning* tuning);
FS_API uint32_t fs_batch_tank_count(const fs_batch* b);
FS_API uint64_t fs_batch_buffer_bytes(const fs_batch* b);
/* Read-outs of tank `tank`, laid out as the single-tank read-outs; the
 * permutation holds tank-local indices.  Stats: per-tank solver figures, the
 * batch's phase times. */
FS_API int fs_batch_read_particles(fs_batch* b, uint32_t tank, float* pos_xyz, float* vel_xyz);
FS_API int fs_batch_read_faces(fs_batch* b, uint32_t tank, int axis, int which, float* out);
FS_API int fs_batch_read_cells_f32(fs_batch* b, uint32_t tank, int which, float* out);
FS_API int fs_batch_read_cells_u32(fs_batch* b, uint32_t tank, int which, uint32_t* out);
FS_API int fs_batch_read_permutation(fs_batch* b, uint32_t tank, uint32_t* perm);
FS_API int fs_batch_get_stats(fs_batch* b, uint32_t tank, fs_stats* out);

/* ---- differentiation ----
 * Reverse-mode gradients of a loss of the final particle state with respect
 * to the initial particle state and the gravity vector.  The reference has no
 * differentiation; what is differentiated is its step, GpuContext::step
 * (src/gpu/mod.rs:950-1033) as run by the fused engine.  Conventions: the
 * sort order, cell types, stencils, wet/solid faces and the CFL and wall-clamp
 * branches of the forward are held fixed; fixed-point rounding of P2G counts
 * as the identity; the pressure solve counts as exact (its adjoint is a CG
 * solve with the same operator, to a 1e-6 relative residual).  Needs
 * FS_ENGINE_FUSED, FS_PRECISION_MIXED, FS_VELOCITY_F32, particles sorted
 * every substep and wall_friction 0 (FS_ERR_UNSUPPORTED otherwise). */
/* From the tank's live state, run `steps` substeps keeping a snapshot
 * (particles + warm-start pressure) every `every` substeps. */
FS_API int fs_forward_checkpointed(fs_tank* t, uint32_t steps, uint32_t every);
/* Reverse pass over the last fs_forward_checkpointed.  gpos_final /
 * gvel_final: dL/d(final positions / velocities), xyz per particle in the
 * order fs_read_particles gives after the forward (NULL = zero).  Outputs
 * (nullable): dL/d(initial positions / velocities) in the order of the state
 * the forward started from, and dL/d(gravity) (3 doubles).  Recomputes each
 * checkpoint interval; the tank is left in the forward's final state. */
FS_API int fs_step_adjoint(fs_tank* t, const float* gpos_final, const float* gvel_final, float* gpos_init,
                           float* gvel_init, double* gravity_grad);

/* ---- live device buffers (zero-copy hand-off to a consumer on the GPU) ----
 * The reference lends its buffers to the renderer and slice inspector:
 * particle_buffer() (the live ping-pong side), take_sort_swapped(),
 * cell_type_buffer(), pressure_buffer(), u/v/w_vel_buffer(), stats_buffer()
 * (src/gpu/fluid.rs:1205-1245); the consumer rebinds when the live side
 * flips (src/gpu/mod.rs:983-988).  Here the same state is lent as device
 * pointers in this library's layouts.  All of the tank's work is ordered on
 * its HIP stream (`stream`): a consumer enqueues its reads on that stream or
 * waits for it (fs_synchronize).  Pointers stay valid until fs_destroy; the
 * particle pointers name the live side until the next fs_step or fs_reset,
 * which may flip it (then fs_take_sort_swapped reports 1). */
typedef struct {
    const float* pos[3];        /* live particle side, SoA: n floats each */
    const void* vel[3];         /* n floats, or n bf16 as uint16 (FS_VELOCITY_BF16) */
    uint32_t n;
    uint32_t velocity_storage;  /* FS_VELOCITY_* */
    uint32_t live_side;         /* 0 or 1 */
    uint32_t dims[3];           /* nx, ny, nz */
    const uint8_t* cell_type;   /* nx*ny*nz: 0 solid, 1 liquid, 2 air (the reference stores u32) */
    const float* pressure;      /* nx*ny*nz */
    const float* face_vel[3];   /* (nx+1)ny nz, nx(ny+1)nz, nx ny(nz+1) */
    const uint32_t* occupancy;  /* nx*ny*nz particles per cell */
    const uint32_t* stats;      /* stats[0] = liquid cell count (sh/classify.wgsl:55-57) */
    void* stream;               /* the tank's hipStream_t */
} fs_device_view;

FS_API int fs_get_device_view(fs_tank* t, fs_device_view* out);
/* take_sort_swapped (src/gpu/fluid.rs:1219-1226): 1 if the live particle
 * side flipped (a sort, a reset or a state load) since the last call, then 0. */
FS_API int fs_take_sort_swapped(fs_tank* t);
/* The reference's particle record {pos.xyz, 0, vel.xyz, 0} (32 B,
 * src/gpu/fluid.rs:276-280, read by sh/particles.wgsl:22) for the live side,
 * written to dst (device, 8n floats), enqueued on the tank's stream. */
FS_API int fs_pack_particles_vec4(fs_tank* t, float* dst_device);
/* Cell types as u32, the reference's cell_type buffer, written to dst
 * (device, nx*ny*nz), enqueued on the tank's stream. */
FS_API int fs_pack_cell_types_u32(fs_tank* t, uint32_t* dst_device);

/* ---- frame driver: the host loop around fs_step ----
 * TimestepController (src/timestep.rs:63-247): a fixed-dt accumulator fed by
 * render deltas, each clamped to 1/30 s, running at most max_substeps per
 * frame and dropping the excess when capped.  Host only. */
#define FS_TIMESTEP_FIXED_DROP_CAP 0        /* "fixed-drop-cap" (TIMESTEP_POLICY_LABEL) */
#define FS_TIMESTEP_FIXED_EXPLICIT_EXPORT 1 /* "fixed-explicit-export" */
/* TimestepFrameStats (src/timestep.rs:20-49). */
typedef struct {
    uint32_t substeps, natural_substeps, max_substeps, substep_cap_hit;
    float fixed_dt, accumulated_before, accumulated_after, dropped_this_frame, sim_advanced, wall_dt,
        sim_time_scale, real_time_factor;
    uint32_t policy; /* FS_TIMESTEP_* */
} fs_timestep_stats;
typedef struct fs_timestep fs_timestep;
FS_API fs_timestep* fs_timestep_create(float fixed_dt, uint32_t max_substeps); /* TimestepController::new */
FS_API void fs_timestep_destroy(fs_timestep* c);
/* steps_for_frame_scaled (:104-159): substeps to run for this frame. */
FS_API uint32_t fs_timestep_steps_for_frame(fs_timestep* c, float render_dt_s, float sim_time_scale);
FS_API void fs_timestep_record_manual(fs_timestep* c, float render_dt_s, uint32_t substeps, float sim_time_scale);
FS_API void fs_timestep_record_export(fs_timestep* c, float sim_dt_s, uint32_t substeps);
FS_API void fs_timestep_record_idle(fs_timestep* c, float render_dt_s, float sim_time_scale);
FS_API void fs_timestep_reset(fs_timestep* c);
FS_API int fs_timestep_last_stats(const fs_timestep* c, fs_timestep_stats* out);
FS_API float fs_timestep_total_dropped(const fs_timestep* c);

/* Interaction schedules (InteractionState, src/lib.rs:55-264): the seeded
 * auto-roll of the tank (a rotation that eases to a new random target every
 * ~cadence seconds, at most `strength` radians from upright) and the periodic
 * wave impulse (+x, -x, +z, -z in turn).  Host only.  Quaternions are
 * x, y, z, w. */
typedef struct fs_interactions fs_interactions;
FS_API fs_interactions* fs_interactions_create(uint64_t seed); /* InteractionState::new */
FS_API void fs_interactions_destroy(fs_interactions* s);
FS_API void fs_interactions_reset(fs_interactions* s, uint64_t seed);
/* update_auto_roll (:146-186): 1 and the new pose in q_out, or 0 (disabled). */
FS_API int fs_interactions_auto_roll(fs_interactions* s, float dt_s, const float q_current[4], int enabled,
                                     float strength, float cadence_s, float q_out[4]);
/* update_wave (:205-236): 1 and the impulse (velocity change) in dv_out, or 0. */
FS_API int fs_interactions_wave(fs_interactions* s, float dt_s, int enabled, float strength, float frequency_hz,
                                float dv_out[3]);
/* InteractionRng::next_u32 (:70-76), for checking the generator. */
FS_API uint32_t fs_interactions_next_u32(fs_interactions* s);

/* FluidApp's frame loop around one tank (src/lib.rs:352-398, 570-601, 506-560,
 * 907-937): run state, the timestep controller, the interactions and the
 * tank's gravity in the box frame (box_orient^-1 * (0, physics.gravity, 0)).
 * Settings are read from `settings` at every call (the caller's registry,
 * which must outlive the driver): physics.fixed_dt and physics.max_substeps
 * at create and reset, physics.sim_speed and interaction.* per frame. */
typedef struct fs_frame fs_frame;
FS_API int fs_frame_create(fs_tank* t, const fs_registry* settings, fs_frame** out);
FS_API void fs_frame_destroy(fs_frame* f);
/* FluidApp::frame: Running runs the accumulator's substeps (after the
 * interactions, at min(render_dt * sim_speed, 1/30 s)); Paused runs one
 * queued step or none.  render_dt_ms: the browser rAF delta (non-finite or
 * negative reads as 0). */
FS_API int fs_frame_run(fs_frame* f, double render_dt_ms, fs_timestep_stats* out);
/* FluidApp::export_frame (:570-601): exactly `substeps` substeps for
 * sim_dt_s seconds of simulation. */
FS_API int fs_frame_export(fs_frame* f, uint32_t substeps, double sim_dt_s, fs_timestep_stats* out);
FS_API int fs_frame_set_paused(fs_frame* f, int paused);
FS_API int fs_frame_queue_step(fs_frame* f); /* the step button while paused */
/* FluidApp::reset (:520-560) for the tank: fs_reset, tick and queue cleared,
 * a new timestep controller from the settings, upright box, interactions
 * reseeded (interaction_seed_for_reset, :992-994), gravity pushed. */
FS_API int fs_frame_reset(fs_frame* f);
FS_API uint64_t fs_frame_tick(const fs_frame* f);
FS_API int fs_frame_orientation(const fs_frame* f, float q_out[4]); /* box_orient */
/* The timestep stats of the last frame and the total dropped simulation time
 * (s), the pair FluidApp hands to Profiler::set_timestep_stats
 * (src/profiler/mod.rs:172-193). */
FS_API int fs_frame_timestep_stats(const fs_frame* f, fs_timestep_stats* out, float* total_dropped_s);

/* ---- timing and stats contract ----
 * GpuTimers' detailed mode (src/gpu/timing.rs:32-116, recorded by
 * record_substep_detailed, src/gpu/mod.rs:1039-1212): device time per fine
 * section (FINE_SECTIONS) and per CG category (CG_CATS, the six passes of an
 * iteration bucketed as CG_BUCKET), summed over the substeps of the last
 * fs_step call, from HIP events on the tank's stream.  Consecutive events
 * share a boundary, so the sections tile the substep.  The reference engine
 * times every section and the four categories.  The fused engine times each
 * of its kernels under the first section whose work it does (the others are
 * flagged fused and read 0) and the CG as one persistent solve
 * (cg_split = 0, the categories NaN).  Detailed timing adds one event per
 * section and six per CG iteration and waits for each substep. */
#define FS_FINE_SECTIONS 26
#define FS_CG_CATEGORIES 4
typedef struct {
    uint32_t detailed;    /* detailed timing was on for the last fs_step */
    uint32_t substeps;    /* substeps summed (Readout.substeps) */
    uint32_t cg_iters;    /* CG iterations per substep (Readout.cg_iters) */
    uint32_t cg_split;    /* 1: cg_ms holds the categories */
    double section_ms[FS_FINE_SECTIONS];
    double cg_ms[FS_CG_CATEGORIES];
    double cg_total_ms;
    /* Readout.prep_ms / pressure_ms / finish_ms: the sections rolled up at
     * PREP_END ("divergence") and FINISH_START ("gradient_u"), the CG in
     * pressure (src/gpu/timing.rs:62-70, 472-481). */
    double prep_ms, pressure_ms, finish_ms;
    uint32_t fused[FS_FINE_SECTIONS]; /* 1: done inside an earlier section's kernel */
} fs_detailed_timing;
FS_API int fs_set_detailed_timing(fs_tank* t, int enabled);
FS_API int fs_get_detailed_timing(fs_tank* t, fs_detailed_timing* out);
FS_API const char* fs_fine_section_name(int i); /* FINE_SECTIONS[i]; NULL out of range */
FS_API const char* fs_cg_category_name(int i);  /* CG_CATS[i]; NULL out of range */
/* GpuFluid::dispatches_per_substep (src/gpu/fluid.rs:1273-1276): kernel
 * dispatches the last substep issued, counted as they were enqueued. */
FS_API uint32_t fs_dispatches_per_substep(const fs_tank* t);
/* Profiler::stats_json (src/profiler/mod.rs:416-594): the simulation's part
 * of the object, with the reference's field names and number formatting.
 * Timestep fields come from f (NULL: the profiler's defaults, zeros).  "gpu"
 * is the GpuTimers readout: detailed when detailed timing is on, else coarse
 * (fs_set_timing), else null.  Browser and render fields are not emitted.
 * Writes a NUL-terminated string when cap > *len_out; with buf NULL only
 * *len_out is set. */
FS_API int fs_stats_json(fs_tank* t, const fs_frame* f, char* buf, size_t cap, size_t* len_out);

/* ---- scene presets and particle seeding ----
 * The reference builds its initial particles from the registry on the host:
 * SceneConfig::from_settings (src/scene/mod.rs:96-116) picks the preset's
 * liquid blocks and resolves the spawn count, and generate_particles
 * (src/gpu/fluid.rs:1833-1913) lays a jittered lattice into each block. */
#define FS_SCENE_FALLING_BLOB 0
#define FS_SCENE_DAM_BREAK 1
#define FS_SCENE_DOUBLE_SPLASH 2
#define FS_SCENE_MAX_BLOCKS 8

/* LiquidBlock (src/scene/mod.rs:61-75): an AABB in normalized [0,1]^3 tank
 * space, y up. */
typedef struct {
    float min[3];
    float max[3];
} fs_liquid_block;

/* SceneConfig (src/scene/mod.rs:84-91) plus the counts derived from it. */
typedef struct {
    uint32_t preset;           /* FS_SCENE_* */
    uint32_t res[3];           /* grid resolution */
    uint32_t nblocks;
    fs_liquid_block blocks[FS_SCENE_MAX_BLOCKS];
    uint32_t particle_count;   /* resolved spawn target (resolved_particle_count) */
    float effective_density;   /* effective_particle_density: particles per seeded cell */
    uint32_t generated_count;  /* particles the lattice holds (estimated_particle_count, src/gpu/fluid.rs:1792-1831) */
} fs_scene;

/* preset_blocks (src/scene/mod.rs:208-238): the blocks of `preset` for a
 * drop height and a fill fraction in [0,1].  *nblocks <= FS_SCENE_MAX_BLOCKS. */
FS_API int fs_scene_preset_blocks(uint32_t preset, float drop_height, float fill_level, fs_liquid_block* blocks,
                                  uint32_t* nblocks);
/* SceneConfig::from_settings (src/scene/mod.rs:96-116) from the registry's
 * scene.preset, scene.drop_height, scene.fill_level, grid.res_*,
 * particles.density and particles.count rows. */
FS_API int fs_scene_from_registry(const fs_registry* r, fs_scene* out);
/* generate_particles (src/gpu/fluid.rs:1833-1913) for a tank of cell size h
 * at `origin` (extent res * h), computed on the GPU.  *count receives the
 * number of particles the lattice holds; pos_xyz (host, `capacity` xyz
 * triplets) receives them, or is NULL to query the count only
 * (FS_ERR_INVALID_ARGUMENT if capacity < *count).  Bit-identical to the
 * reference's sequential generator. */
FS_API int fs_scene_seed(const fs_scene* scene, float h, const float origin[3], float* pos_xyz, uint32_t capacity,
                         uint32_t* count);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc): the same
 * seeding as fs_scene_seed, allocating nothing (no hipMalloc, no stream).
 * The lattice is staged in a fixed device array of FS_SEED_SCRATCH_BYTES
 * (one per GPU, held until the process ends) and copied to pos_xyz on the
 * null stream; the call returns when the copy is done.  Lattices of more
 * than FS_SEED_SCRATCH_PARTICLES particles fail with FS_ERR_INVALID_ARGUMENT.
 * Every call on a GPU shares that GPU's array: calls that overlap on one GPU
 * (from several threads) race.  Same results as fs_scene_seed. */
#define FS_SEED_SCRATCH_BYTES (128u * 1024u * 1024u)
#define FS_SEED_SCRATCH_PARTICLES (FS_SEED_SCRATCH_BYTES / 12u)
FS_API int fs_scene_seed_noalloc(const fs_scene* scene, float h, const float origin[3], float* pos_xyz,
                                 uint32_t capacity, uint32_t* count);
/* GpuFluid::new (src/gpu/fluid.rs:247-318) from a registry and a scene:
 * h = 2/64, the tank centred on the origin, the scene's particles seeded on
 * the GPU at rest, and Params from the registry with the scene's effective
 * density (auto rest density and surface dilation, src/scene/mod.rs:283-324).
 * engine, precision and velocity storage as in fs_params_default. */
FS_API int fs_create_from_settings(const fs_registry* r, const fs_scene* scene, fs_tank** out);

#ifdef __cplusplus
}
#endif
#endif
