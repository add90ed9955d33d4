// fs_adjoint.cpp — checkpointed forward and reverse pass of the C-ABI
// (fs_forward_checkpointed, fs_step_adjoint).
//
// The forward keeps a snapshot (live particles + warm-start pressure) every
// `every` substeps.  The reverse pass walks the checkpoint intervals from the
// last to the first: it restores the interval's snapshot, re-runs the
// interval forward recording each substep's input state, then for each
// substep of the interval, last first, restores that input, replays the
// substep up to (not including) G2P so the sorted particles and every grid
// intermediate are on the device, and runs the adjoint kernels of
// k_adjoint.hip.  The replay reproduces the forward bit for bit (the sort,
// the integer P2G and the fixed-order CG reductions are deterministic).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "fluidstep.h"
#include "fs_host_common.h"
#include "fs_tank.h"

using namespace fs;

namespace {

// The adjoint solve runs to the forward's relative tolerance (1e-6 when the
// forward runs a fixed iteration count), capped at the larger of the
// forward's cap and kAdjIters.  Its error enters the gradient at that
// relative level.
constexpr float kAdjTolDefault = 1e-6f;
constexpr uint32_t kAdjIters = 400;
// The adjoint counts P2G's fixed-point rounding as the identity: each
// substep is replayed for the reverse pass with P2G at 2^22 instead of the
// step's 2^16, so rounding moves the linearization by ~4e-5 relative
// instead of ~3e-3.  A face's i32 sums stay below 2^31 up to about
// 500 particle-weights times m/s (a weight sum of 30 at the 3.75 m/s CFL
// cap of the 64^3 tank uses a quarter of the range).
constexpr float kAdjFixedScale = 4194304.0f;

int alloc_snapshot(fs_tank* t, fs_tank::Snapshot& s) {
    for (int a = 0; a < 3; ++a) {
        FS_HIP(t->adj_alloc(s.pos[a], t->N));
        FS_HIP(t->adj_alloc(s.vel[a], t->N));
    }
    FS_HIP(t->adj_alloc(s.p, t->P.C));
    return FS_OK;
}

int snapshot(fs_tank* t, const fs_tank::Snapshot& s) {
    const size_t nb = sizeof(float) * t->N;
    const ParticleSoA& live = t->side[t->cur];
    for (int a = 0; a < 3; ++a) {
        if (nb) FS_HIP(hipMemcpyAsync(s.pos[a], live.pos[a], nb, hipMemcpyDeviceToDevice, t->stream));
        if (nb) FS_HIP(hipMemcpyAsync(s.vel[a], live.vel[a], nb, hipMemcpyDeviceToDevice, t->stream));
    }
    FS_HIP(hipMemcpyAsync(s.p, t->p, sizeof(float) * t->P.C, hipMemcpyDeviceToDevice, t->stream));
    return FS_OK;
}

int restore(fs_tank* t, const fs_tank::Snapshot& s) {
    const size_t nb = sizeof(float) * t->N;
    const ParticleSoA& live = t->side[t->cur];
    for (int a = 0; a < 3; ++a) {
        if (nb) FS_HIP(hipMemcpyAsync(live.pos[a], s.pos[a], nb, hipMemcpyDeviceToDevice, t->stream));
        if (nb) FS_HIP(hipMemcpyAsync(live.vel[a], s.vel[a], nb, hipMemcpyDeviceToDevice, t->stream));
    }
    FS_HIP(hipMemcpyAsync(t->p, s.p, sizeof(float) * t->P.C, hipMemcpyDeviceToDevice, t->stream));
    return FS_OK;
}

int check_adjoint_params(const fs_tank* t) {
    const fs_params& q = t->prm;
    if (q.engine != FS_ENGINE_FUSED || q.precision != FS_PRECISION_MIXED)
        return fail(FS_ERR_UNSUPPORTED, "the adjoint runs the fused engine in mixed precision");
    if (q.velocity_storage != FS_VELOCITY_F32)
        return fail(FS_ERR_UNSUPPORTED, "the adjoint needs f32 particle velocity storage");
    if (!q.particle_sort || q.particle_sort_period > 1)
        return fail(FS_ERR_UNSUPPORTED, "the adjoint needs the particles sorted every substep");
    if (t->P.friction > 0.0f) return fail(FS_ERR_UNSUPPORTED, "the adjoint is implemented for wall_friction 0");
    return FS_OK;
}

// Adjoint of one substep whose input state is the tank's live state: on
// entry adj_gx/adj_gv hold dL/d(output) in the substep's output order, on
// exit dL/d(input) in the input order; adj_grav accumulates dL/d(gravity).
int adjoint_substep_at_scale(fs_tank* t);

int adjoint_substep(fs_tank* t) {
    const float scale0 = t->P.scale;
    t->P.scale = kAdjFixedScale;
    const int rc = adjoint_substep_at_scale(t);
    t->P.scale = scale0;
    return rc;
}

int adjoint_substep_at_scale(fs_tank* t) {
    hipStream_t s = t->stream;
    const StepParams& P = t->P;
    if (int rc = substep_fused(t, true, true)) return rc;
    const ParticleSoA sorted = t->side[t->cur];
    // P2G weight sums of the sorted particles (the forward's den).
    for (int a = 0; a < 3; ++a) {
        FS_HIP(launch_fill_u32((uint32_t*)t->num[a], 0u, P.F[a], s));
        FS_HIP(launch_fill_u32((uint32_t*)t->den[a], 0u, P.F[a], s));
        FS_HIP(launch_fill_f32(t->adj_vbar[a], 0.0f, P.F[a], s));
        FS_HIP(launch_fill_f32(t->adj_sbar[a], 0.0f, P.F[a], s));
    }
    FS_HIP(launch_p2g_atomic(P, sorted, 0, t->num, t->den, nullptr, 0u, s));
    AdjFaces faces;
    for (int a = 0; a < 3; ++a) {
        faces.vel[a] = t->vel[a];
        faces.saved[a] = t->saved[a];
        faces.den[a] = t->den[a];
        faces.vbar[a] = t->adj_vbar[a];
        faces.sbar[a] = t->adj_sbar[a];
        faces.numbar[a] = t->adj_numbar[a];
        faces.denbar[a] = t->adj_denbar[a];
    }
    const AdjParticles part_sorted = {{t->adj_sgx[0], t->adj_sgx[1], t->adj_sgx[2]},
                                      {t->adj_sgv[0], t->adj_sgv[1], t->adj_sgv[2]}};
    const AdjParticles part_in = {{t->adj_gx[0], t->adj_gx[1], t->adj_gx[2]},
                                  {t->adj_gv[0], t->adj_gv[1], t->adj_gv[2]}};
    const AdjParticlesIn out_bar = {{t->adj_gx[0], t->adj_gx[1], t->adj_gx[2]},
                                    {t->adj_gv[0], t->adj_gv[1], t->adj_gv[2]}};
    FS_HIP(launch_adj_g2p(P, sorted, faces, part_sorted, out_bar, s));
    FS_HIP(launch_adj_pbar(P, t->ct, faces, t->adj_pbar, s));
    FS_HIP(launch_adj_cg_init(P, t->cg.nmask, t->adj_pbar, t->adj_rhs, t->adj_z, t->r, t->d0, s));
    StepParams Pa = P;
    Pa.tol = P.tol > 0.0f ? P.tol : kAdjTolDefault;
    Pa.tol_sq = Pa.tol * Pa.tol;
    Pa.iters = kAdjIters > P.iters ? kAdjIters : P.iters;
    FS_HIP(launch_cg_reset(t->cg, s));
    CgVectors av = t->cg_vectors();
    av.b = t->adj_rhs;
    av.x = t->adj_z;
    FS_HIP(launch_cg_persistent(Pa, av, t->cg, t->cg_blocks, s));
    FS_HIP(launch_adj_faces(P, t->ct, t->adj_z, faces, t->adj_grav, s));
    FS_HIP(launch_adj_p2g(P, sorted, faces, part_sorted, s));
    FS_HIP(launch_adj_unsort(P.N, t->perm, part_sorted, part_in, s));
    CgState st;
    FS_HIP(hipMemcpyAsync(&st, t->cg.state, sizeof st, hipMemcpyDeviceToHost, s));
    FS_HIP(hipStreamSynchronize(s));
    if (st.error) return fail(FS_ERR_SOLVER, "adjoint CG grid barrier timed out");
    return FS_OK;
}

// AoS xyz (host, nullable = zero) <-> three device columns.
int upload_xyz(fs_tank* t, const float* xyz, float* const col[3]) {
    const uint32_t n = t->N;
    if (!n) return FS_OK;
    std::vector<float> c(n);
    for (int a = 0; a < 3; ++a) {
        for (uint32_t i = 0; i < n; ++i) c[i] = xyz ? xyz[3 * (size_t)i + a] : 0.0f;
        FS_HIP(hipMemcpy(col[a], c.data(), sizeof(float) * n, hipMemcpyHostToDevice));
    }
    return FS_OK;
}

int download_xyz(fs_tank* t, float* const col[3], float* xyz) {
    const uint32_t n = t->N;
    if (!n || !xyz) return FS_OK;
    std::vector<float> c(n);
    for (int a = 0; a < 3; ++a) {
        FS_HIP(hipMemcpy(c.data(), col[a], sizeof(float) * n, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < n; ++i) xyz[3 * (size_t)i + a] = c[i];
    }
    return FS_OK;
}

}  // namespace

extern "C" {

FS_API int fs_forward_checkpointed(fs_tank* t, uint32_t steps, uint32_t every) {
    if (!t) return fail(FS_ERR_INVALID_ARGUMENT, "null tank handle");
    if (steps == 0 || every == 0) return fail(FS_ERR_INVALID_ARGUMENT, "steps and every must be positive");
    if (int rc = check_adjoint_params(t)) return rc;
    FS_HIP(hipStreamSynchronize(t->stream));
    t->drop_adjoint();
    const uint32_t K = (steps + every - 1) / every;
    const uint32_t tape_len = every < steps ? every : steps;
    int rc = [&]() -> int {
        t->ckpt.resize(K);
        for (auto& s : t->ckpt)
            if (int e = alloc_snapshot(t, s)) return e;
        t->tape.resize(tape_len);
        for (auto& s : t->tape)
            if (int e = alloc_snapshot(t, s)) return e;
        if (int e = alloc_snapshot(t, t->adj_final)) return e;
        for (int a = 0; a < 3; ++a) {
            FS_HIP(t->adj_alloc(t->adj_gx[a], t->N));
            FS_HIP(t->adj_alloc(t->adj_gv[a], t->N));
            FS_HIP(t->adj_alloc(t->adj_sgx[a], t->N));
            FS_HIP(t->adj_alloc(t->adj_sgv[a], t->N));
            FS_HIP(t->adj_alloc(t->adj_vbar[a], t->P.F[a]));
            FS_HIP(t->adj_alloc(t->adj_sbar[a], t->P.F[a]));
            FS_HIP(t->adj_alloc(t->adj_numbar[a], t->P.F[a]));
            FS_HIP(t->adj_alloc(t->adj_denbar[a], t->P.F[a]));
        }
        FS_HIP(t->adj_alloc(t->adj_pbar, t->P.C));
        FS_HIP(t->adj_alloc(t->adj_rhs, t->P.C));
        FS_HIP(t->adj_alloc(t->adj_z, t->P.C));
        FS_HIP(t->adj_alloc(t->adj_grav, 4));
        return FS_OK;
    }();
    if (rc != FS_OK) {
        t->drop_adjoint();
        return rc;
    }
    for (uint32_t s = 0; s < steps; ++s) {
        if (s % every == 0)
            if (int e = snapshot(t, t->ckpt[s / every])) return e;
        if (int e = do_step(t, 1)) return e;
    }
    if (int e = snapshot(t, t->adj_final)) return e;
    FS_HIP(hipStreamSynchronize(t->stream));
    t->adj_steps = steps;
    t->adj_every = every;
    t->adj_valid = 1;
    return FS_OK;
}

FS_API int fs_step_adjoint(fs_tank* t, const float* gpos_final, const float* gvel_final, float* gpos_init,
                           float* gvel_init, double* gravity_grad) {
    if (!t) return fail(FS_ERR_INVALID_ARGUMENT, "null tank handle");
    if (!t->adj_valid) return fail(FS_ERR_INVALID_ARGUMENT, "no checkpointed forward to differentiate");
    if (int rc = upload_xyz(t, gpos_final, t->adj_gx)) return rc;
    if (int rc = upload_xyz(t, gvel_final, t->adj_gv)) return rc;
    FS_HIP(hipMemset(t->adj_grav, 0, 4 * sizeof(double)));
    const uint32_t steps = t->adj_steps, every = t->adj_every;
    const uint32_t K = (steps + every - 1) / every;
    for (int seg = (int)K - 1; seg >= 0; --seg) {
        const uint32_t s0 = (uint32_t)seg * every;
        const uint32_t s1 = s0 + every < steps ? s0 + every : steps;
        if (int rc = restore(t, t->ckpt[seg])) return rc;
        // Recompute the interval, recording each substep's input.
        for (uint32_t s = s0; s < s1; ++s) {
            if (int rc = snapshot(t, t->tape[s - s0])) return rc;
            if (s + 1 < s1)
                if (int rc = do_step(t, 1)) return rc;
        }
        for (int s = (int)s1 - 1; s >= (int)s0; --s) {
            if (int rc = restore(t, t->tape[(uint32_t)s - s0])) return rc;
            if (int rc = adjoint_substep(t)) return rc;
        }
    }
    if (int rc = download_xyz(t, t->adj_gx, gpos_init)) return rc;
    if (int rc = download_xyz(t, t->adj_gv, gvel_init)) return rc;
    if (gravity_grad) FS_HIP(hipMemcpy(gravity_grad, t->adj_grav, 3 * sizeof(double), hipMemcpyDeviceToHost));
    // Leave the tank in the state the forward ended in.
    if (int rc = restore(t, t->adj_final)) return rc;
    FS_HIP(hipStreamSynchronize(t->stream));
    return FS_OK;
}

}  // extern "C"
