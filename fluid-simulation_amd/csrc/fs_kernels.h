// fs_kernels.h — host-side launchers of the step's HIP kernels.  Every
// launcher enqueues on the given stream and returns the launch status.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "fs_device.h"

namespace fs {

// Particle SoA on one ping-pong side.  Velocity is f32 or bf16 storage.
struct ParticleSoA {
    float* pos[3];
    void* vel[3];
};

// Kernel dispatches issued by this thread's launchers (each launcher counts
// what it enqueues); the host reads the difference across a substep.
extern thread_local uint64_t g_dispatches;
#define FS_LAUNCH(...)               \
    do {                             \
        ++::fs::g_dispatches;        \
        hipLaunchKernelGGL(__VA_ARGS__); \
    } while (0)

// Listed misfit particles (P2G cell != sort cell) the brick gather walks;
// beyond this many the step takes the atomic P2G form instead.
constexpr uint32_t kMisfitCap = 4096;

// ---- k_bin.hip: clears, keys, histogram, scans, stable radix permutation ----
hipError_t launch_fill_u32(uint32_t* p, uint32_t value, size_t count, hipStream_t s);
hipError_t launch_fill_f32(float* p, float value, size_t count, hipStream_t s);
// sh/mark.wgsl as written: one global atomic per particle.
hipError_t launch_mark(const StepParams& P, const float* const pos[3], uint32_t* occ, hipStream_t s);
// Keys + histogram with wave-aggregated atomics; keys[p] = cell_key(p).
// mis_count counts particles whose P2G cell (floor((pos-origin)*inv_h))
// differs from their sort cell; the first mis_cap of them are listed in
// mis_list (nullable) for the brick gather.
hipError_t launch_key_hist(const StepParams& P, const float* const pos[3], uint32_t* keys, uint32_t* occ,
                           uint32_t* mis_count, uint32_t* mis_list, uint32_t mis_cap, hipStream_t s);
// Fill that runs only when *guard > above.
hipError_t launch_fill_u32_guarded(uint32_t* p, uint32_t value, size_t count, const uint32_t* guard,
                                   uint32_t above, hipStream_t s);
// Exclusive scan workspace size (u32 elements) for `len` inputs.
size_t scan_workspace(size_t len);
hipError_t launch_exclusive_scan(const uint32_t* in, uint32_t* out, size_t len, uint32_t* ws,
                                 hipStream_t s);
size_t radix_workspace(size_t n);  // u32 elements (histogram table + scan ws)
// Stable LSD radix sort of keys[0..n) (keys < key_limit); writes perm[d] = the
// source index of sorted slot d.  keys_tmp/vals_tmp are n-sized scratch.
// tile_variant picks the keys per workgroup: 0 = 4096, 1 = 2048, 2 = 8192.
// order: the particles grouped by cell in key order (not stable), from the
// keys, their bin starts and a scratch cursor of ncells words.
hipError_t launch_cell_order(const uint32_t* keys, uint32_t n, const uint32_t* cell_offset, uint32_t* cursor,
                             uint32_t ncells, uint32_t* order, hipStream_t s);
hipError_t launch_stable_permutation(const uint32_t* keys, uint32_t n, uint32_t key_limit,
                                     uint32_t* keys_tmp0, uint32_t* keys_tmp1, uint32_t* vals_tmp,
                                     uint32_t* perm, uint32_t* ws, hipStream_t s, int tile_variant = 0);
// dst[d] = src[perm[d]] for the six particle columns.
hipError_t launch_permute(uint32_t n, const uint32_t* perm, const ParticleSoA& src,
                          const ParticleSoA& dst, int bf16, hipStream_t s);

// ---- k_grid.hip: classify, P2G, face passes, divergence, gradient ----
hipError_t launch_classify(const StepParams& P, const uint32_t* occ, uint8_t* ct, uint32_t* stats,
                           hipStream_t s);
// sh/scatter.wgsl: per-particle i32 atomics into num/den.
// guard (nullable): run only when *guard > above.
// The six num/den arrays cleared when *guard > above (fused fallback).
hipError_t launch_clear_faces_guarded(const StepParams& P, int32_t* const num[3], int32_t* const den[3],
                                      const uint32_t* guard, uint32_t above, hipStream_t s);
hipError_t launch_p2g_atomic(const StepParams& P, const ParticleSoA& src, int bf16,
                             int32_t* const num[3], int32_t* const den[3], const uint32_t* guard,
                             uint32_t above, hipStream_t s);
// sh/scatter_local.wgsl: workgroup-local hash pre-accumulation, one global
// atomic per touched face slot (same num/den bits as launch_p2g_atomic).
// Supported when every face index fits the key below 2^29 (else use the
// per-tap form).
bool p2g_local_supported(const StepParams& P);
hipError_t launch_p2g_local(const StepParams& P, const ParticleSoA& src, int bf16, int32_t* const num[3],
                            int32_t* const den[3], hipStream_t s);
// Reference-structured per-axis passes (one kernel each, sh/*.wgsl).
hipError_t launch_normalize(const StepParams& P, int a, const int32_t* num, const int32_t* den,
                            float* vel, hipStream_t s);
hipError_t launch_copy_f32(const float* src, float* dst, size_t n, hipStream_t s);
hipError_t launch_forces(const StepParams& P, int a, const uint8_t* ct, float* vel, hipStream_t s);
hipError_t launch_enforce(const StepParams& P, int a, const uint8_t* ct, float* vel, hipStream_t s);
hipError_t launch_gradient(const StepParams& P, int a, const float* p, const uint8_t* ct, float* vel,
                           hipStream_t s);
hipError_t launch_divergence(const StepParams& P, float* const vel[3], const uint8_t* ct,
                             const uint32_t* occ, float* div, hipStream_t s);
// Fused: normalize -> save -> forces -> enforce for all three axes.
hipError_t launch_face_epilogue(const StepParams& P, const int32_t* const num[3],
                                const int32_t* const den[3], const uint8_t* ct, float* const vel[3],
                                float* const saved[3], const uint32_t* guard, uint32_t above,
                                hipStream_t s);
// Fused: gradient -> enforce for all three axes.
hipError_t launch_gradient_enforce(const StepParams& P, const float* p, const uint8_t* ct,
                                   float* const vel[3], hipStream_t s);
// Sorted-segment P2G gather per brick of cells, with the face epilogue fused.
// brick_variant: 0 = 8x8x8, 1 = 16x8x8, 2 = 32x4x4 cells per workgroup.
// src is the pre-sort side, which the listed misfit indices refer to; the
// gather does nothing when *mis_count exceeds kMisfitCap (fs_gather.h).  With
// perm (the stable cell order of `sorted`, whose particles were not moved),
// sorted slot d is particle perm[d].  With brick_ws (gather_brick_count_max
// * 2 uint32 of workspace, zeroed once) a grid of more bricks than are
// resident at once runs its bricks in order of decreasing halo particle
// count, as the previous call counted them.
// The face epilogue (normalize, save, forces, enforce) of every face from
// num/den, one row of faces per workgroup.
hipError_t launch_face_epilogue_rows(const StepParams& P, const int32_t* const num[3], const int32_t* const den[3],
                                     const uint8_t* ct, float* const vel[3], float* const saved[3], hipStream_t s);
// Particle-owned P2G (fs_gather.h p2g_scatter_brick): every particle's taps
// added once, by the brick of its sort cell; num/den cleared before, the face
// epilogue after.  perm as for launch_p2g_gather.
hipError_t launch_p2g_bricks(const StepParams& P, const ParticleSoA& sorted, int bf16, const uint32_t* occ,
                             const uint32_t* cell_offset, int32_t* const num[3], int32_t* const den[3],
                             const uint32_t* perm, hipStream_t s);
hipError_t launch_p2g_gather(const StepParams& P, const ParticleSoA& sorted, const ParticleSoA& src, int bf16,
                             const uint32_t* occ, const uint32_t* cell_offset, const uint8_t* ct,
                             float* const vel[3], float* const saved[3], const uint32_t* mis_count,
                             const uint32_t* mis_list, hipStream_t s, int brick_variant = 0,
                             const uint32_t* perm = nullptr, uint32_t* brick_ws = nullptr);
// Workspace of launch_p2g_gather's brick order: 2 * this many uint32 (the
// largest brick count over the brick shapes); without it the bricks run in
// brick order.
size_t gather_brick_count_max(const StepParams& P);

// ---- k_cg.hip: pressure solve ----
// Reference-structured CG passes (sh/cg_*.wgsl), f32 scalars in `scal`.
hipError_t launch_cg_init_ref(const StepParams& P, const float* div, const uint8_t* ct, float* p,
                              float* r, float* d, float* scal, hipStream_t s);
hipError_t launch_cg_reduce_ref(const StepParams& P, const float* a, const float* b, float* partials,
                                const float* scal, hipStream_t s);
hipError_t launch_cg_reduce_final_ref(const StepParams& P, const float* partials, float* scal,
                                      hipStream_t s);
hipError_t launch_cg_set_rsold_ref(float* scal, hipStream_t s);
hipError_t launch_cg_spmv_ref(const StepParams& P, const uint8_t* ct, const float* d, float* q,
                              const float* scal, hipStream_t s);
hipError_t launch_cg_alpha_ref(float* scal, hipStream_t s);
hipError_t launch_cg_update_ref(const StepParams& P, const float* scal, const float* d, const float* q,
                                float* p, float* r, hipStream_t s);
hipError_t launch_cg_beta_ref(float* scal, hipStream_t s);
hipError_t launch_cg_dir_ref(const StepParams& P, const float* scal, const float* r, float* d,
                             hipStream_t s);
hipError_t launch_cg_iter_count_ref(const float* scal, uint32_t* count, hipStream_t s);

// Mixed-precision solve state (device), f64 scalars.
struct CgState {
    double rs_old, rs_init, dq, rr;
    double true_res_sq;
    float alpha, beta;       // f32 step lengths of the current iteration
    uint32_t iterations;
    uint32_t active;
    uint32_t error;          // nonzero: the last solve's grid barrier timed out
    uint32_t failed;         // sticky: some solve timed out since the host last
                             // cleared it (kernels only ever set it)
};
struct CgWork {
    // grid-barrier + reduction scratch; every word below up to `slots_data`
    // (inclusive) lives in one area of cg_sync_words(max_blocks) u32 words
    // that launch_cg_reset clears before each solve
    uint32_t* sync;          // base of the area
    uint32_t sync_words;
    uint32_t* bar_count;     // [1]
    uint32_t* group_count;   // [8 * 32]: one arrival counter per 128-byte line
    uint32_t* bar_gen;       // [1]
    uint32_t* abort_flag;    // [1]
    uint64_t* slots;         // [2][max_blocks][2] epoch-tagged partials (exchanges)
    uint64_t* slots_data;    // [2][max_blocks][2] epoch-tagged partials (tagged data barriers)
    uint16_t* nmask;         // [C] stencil mask of liquid cells
    double* partials;        // [max_blocks]
    double* result;          // [4]
    uint32_t* tile_flags;    // [C/64 + 1]
    uint32_t* tile_scan;     // [C/64 + 1]
    uint32_t* tiles;         // [C/64] active tile list
    uint32_t* n_tiles;       // [1]
    CgState* state;          // [1]
    long long spin_ticks;    // limit of every bounded wait (wall_clock64 ticks)
    uint32_t tiles_per_block;  // persistent solve: most tiles a workgroup takes (0: one batch per wave)
};
// tiles[scan[t]] = t for flagged tiles; *n_active = the flagged count.
hipError_t launch_compact_tiles(uint32_t ntiles, const uint32_t* flags, const uint32_t* scan, uint32_t* tiles,
                                uint32_t* n_active, hipStream_t s);
hipError_t launch_div_cginit(const StepParams& P, float* const vel[3], const uint8_t* ct,
                             const uint32_t* occ, float* div, float* b, float* p, float* r, float* d,
                             const CgWork& w, uint32_t* scan_ws, hipStream_t s);
// Vectors of the persistent solve.  On entry r[0] holds r_0 = b - A x_0 and
// pd holds d_0 = r_0 (k_div_cginit / the adjoint's init write both); x is
// updated in place.  The single-reduction form uses all of them; the
// two-phase form uses b, x, r[0] (residual), pd and s[0] (directions) and
// w[0] (A d).
struct CgVectors {
    const float* b;
    float* x;
    float* r[2];
    float* w[2];
    float* s[2];
    float* pd;
};

// Size of the sync area for a grid of up to max_blocks, and the binding of
// the counters, flags and slots to an allocation of that many words.
size_t cg_sync_words(int max_blocks);
void cg_bind_sync(CgWork& w, uint32_t* area, int max_blocks);
// Clears the sync area (one fill kernel); enqueue before every solve.
hipError_t launch_cg_reset(const CgWork& w, hipStream_t s);
// The whole solve in one cooperative launch; partials must hold
// 2 x 2 x grid_blocks doubles.
hipError_t launch_cg_persistent(const StepParams& P, const CgVectors& v, const CgWork& w, int grid_blocks,
                                hipStream_t s);
// Co-resident blocks of the persistent kernel (occupancy x CUs) and its
// workgroup size.
int cg_persistent_max_blocks();
int cg_persistent_threads();
// Same solve as two kernels per iteration (SpMV+d.q, update+r.r) with
// last-block-done reductions; P.iters iterations are enqueued (converged
// iterations return at once).  Suitable for hipGraph capture.
hipError_t launch_cg_multi(const StepParams& P, const uint8_t* ct, const float* b, float* p, float* r,
                           float* d0, float* d1, float* q, const CgWork& w, int grid_blocks, hipStream_t s);

// ---- k_g2p.hip ----
// With from and perm (both or neither), particle d of live is read from
// slot perm[d] of *from, i.e. the sort's permutation is applied as G2P writes.
hipError_t launch_g2p(const StepParams& P, const ParticleSoA& live, int bf16,
                      const float* const vel[3], const float* const saved[3], hipStream_t s,
                      const ParticleSoA* from = nullptr, const uint32_t* perm = nullptr);
// G2P by row chunks of a cell order: the particles of sorted slots d (order[d]
// on rd) are updated and written to slot d of wr, or back to order[d] when
// in_place.  starts/occ: the order's bin starts and the cells' counts.
uint64_t g2p_row_chunks(const StepParams& P);  // workgroups (row chunks) of launch_g2p_rows
hipError_t launch_g2p_rows(const StepParams& P, const ParticleSoA& wr, const ParticleSoA& rd, const uint32_t* order,
                           const uint32_t* starts, const uint32_t* occ, int in_place, int bf16,
                           const float* const vel[3], const float* const saved[3], hipStream_t s);
// The reference's 32-byte particle record {pos.xyz, 0, vel.xyz, 0} of the
// side `sp` into dst (8n floats), and u8 -> u32 widening (k_view.hip).
hipError_t launch_pack_particles_vec4(const ParticleSoA& sp, uint32_t n, bool bf16, float* dst, hipStream_t s);
hipError_t launch_widen_u8(const uint8_t* src, uint32_t n, uint32_t* dst, hipStream_t s);
hipError_t launch_impulse(const StepParams& P, const ParticleSoA& live, int bf16, float dx, float dy,
                          float dz, hipStream_t s);

// ---- k_adjoint.hip: reverse pass of one fused substep ----
struct AdjFaces {
    const float* vel[3];    // final face velocities of the forward substep
    const float* saved[3];  // saved (normalized P2G) velocities
    const int32_t* den[3];  // P2G weight sums (fixed point)
    float* vbar[3];         // adjoint of the final face velocities (accumulated by k_adj_g2p)
    float* sbar[3];         // adjoint of the saved velocities
    float* numbar[3];       // adjoint of sum w v per face
    float* denbar[3];       // adjoint of sum w per face
};
struct AdjParticles {
    float* gx[3];
    float* gv[3];
};
struct AdjParticlesIn {
    const float* gx[3];
    const float* gv[3];
};
hipError_t launch_adj_g2p(const StepParams& P, const ParticleSoA& sorted, const AdjFaces& faces,
                          const AdjParticles& out, const AdjParticlesIn& in, hipStream_t s);
hipError_t launch_adj_pbar(const StepParams& P, const uint8_t* ct, const AdjFaces& faces, float* pbar,
                           hipStream_t s);
hipError_t launch_adj_cg_init(const StepParams& P, const uint16_t* nmask, const float* pbar, float* rhs, float* z,
                              float* r, float* d, hipStream_t s);
hipError_t launch_adj_faces(const StepParams& P, const uint8_t* ct, const float* z, const AdjFaces& faces,
                            double* grav, hipStream_t s);
hipError_t launch_adj_p2g(const StepParams& P, const ParticleSoA& sorted, const AdjFaces& faces,
                          const AdjParticles& out, hipStream_t s);
hipError_t launch_adj_unsort(uint32_t n, const uint32_t* perm, const AdjParticles& sorted,
                             const AdjParticles& in_order, hipStream_t s);

}  // namespace fs
