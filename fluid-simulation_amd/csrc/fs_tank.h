// fs_tank.h — the single-tank handle behind fs_tank* (device buffers, stream,
// solver settings) and the host step functions shared by the C-ABI files.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "fluidstep.h"
#include "fs_kernels.h"

using fs::CgWork;
using fs::ParticleSoA;
using fs::StepParams;

std::vector<double> fs_default_settings();

// ---------------------------------------------------------------------------
// Tank
struct fs_tank {
    fs_params prm{};
    StepParams P{};
    hipStream_t stream = nullptr;
    uint32_t N = 0;
    int bf16 = 0;
    std::vector<double> settings = fs_default_settings();

    std::vector<float> init_pos, init_vel;  // GpuFluid::initial (reset source)

    ParticleSoA side[2]{};
    int cur = 0;
    bool sort_swapped = false;  // take_sort_swapped: the live side flipped since the last query
    uint32_t *keys = nullptr, *ktmp0 = nullptr, *ktmp1 = nullptr, *vtmp = nullptr, *perm = nullptr;
    // Fused engine, substeps without a sort: the live side's stable cell order,
    // through which the gather reads the particles (they stay in place).
    uint32_t* gperm = nullptr;
    uint32_t* goff = nullptr;    // bin starts of gperm
    uint32_t* cursor = nullptr;  // per-cell slot counters while gperm is built
    uint32_t *radix_ws = nullptr, *scan_ws = nullptr;
    uint32_t *occ = nullptr, *cell_offset = nullptr;
    uint8_t* ct = nullptr;
    int32_t *num[3] = {}, *den[3] = {};
    float *vel[3] = {}, *saved[3] = {};
    float *div = nullptr, *b = nullptr, *p = nullptr, *r = nullptr, *d0 = nullptr, *d1 = nullptr, *q = nullptr;
    float *r1 = nullptr, *cg_w[2] = {}, *cg_s[2] = {};  // single-reduction CG: r, w = A r, s = A pd
    float* scal = nullptr;      // reference CG scalars [8]
    float* partials = nullptr;  // reference CG partials
    uint32_t* counters = nullptr;  // [0] liquid cells, [1] misfits, [2] CG iterations (reference engine)
    uint32_t* mis_list = nullptr;  // indices (pre-sort side) of the first kMisfitCap misfit particles
    uint32_t* brick_ws = nullptr;  // gather brick weights and dispatch order (null: brick order)
    CgWork cg{};
    // Persistent-solve vectors of the forward step (b, p, r|r1, cg_w, cg_s, d0).
    fs::CgVectors cg_vectors() const {
        return fs::CgVectors{b, p, {r, r1}, {cg_w[0], cg_w[1]}, {cg_s[0], cg_s[1]}, d0};
    }
    int cg_blocks = 0;        // persistent form
    int cg_mode = 0;          // 0 persistent kernel, 1 two kernels per iteration
    int cg_multi_blocks = 0;  // multi-kernel form
    int cg_use_graph = 1;     // replay the multi-kernel solve from a captured hipGraph
    int gather_brick = 0;     // P2G gather brick: 0 = 8x8x8, 1 = 16x8x8, 2 = 32x4x4 cells
    int p2g_particle = 0;     // 1: the particle-owned brick scatter instead of the face-owning gather
    int g2p_rows = -1;        // G2P by row chunks with the faces staged in LDS: 1 on, 0 off, -1 for large dense tanks
    hipGraphExec_t cg_exec = nullptr;

    uint64_t tick = 0;  // sort tick (advance_sort_tick, src/gpu/fluid.rs:1491-1498)
    uint64_t substeps = 0;
    int sorted_last = 0, gather_last = 0, perm_valid = 0;
    int timing = 0;
    hipEvent_t ev[5] = {};
    // Detailed timing (the reference's FINE_SECTIONS and CG_CATS): boundary
    // events of the substep in flight, each closing a span with a tag
    // (0..25 a section, 100 + c a CG category, 200 the whole fused solve).
    int detailed = 0;
    bool dev_armed = false;
    std::vector<hipEvent_t> dev_ev;
    std::vector<int> dev_tag;
    size_t dev_n = 0;
    double sec_ms[26] = {}, cg_cat_ms[4] = {}, cg_total_ms = 0.0;
    uint32_t cg_iters_timed = 0, substeps_timed = 0;
    // Kernel dispatches of the last substep, and those one replay of the
    // captured CG graph issues.
    uint32_t dispatches_last = 0;
    uint32_t call_substeps = 0;  // substeps of the last fs_step call
    uint64_t cg_graph_dispatches = 0;
    double ms[5] = {0, 0, 0, 0, 0};
    uint64_t bytes = 0;

    std::vector<void*> allocs;

    // Checkpointed forward and reverse pass (fs_forward_checkpointed /
    // fs_step_adjoint).  A snapshot is the live particle side and the
    // warm-start pressure.
    struct Snapshot {
        float* pos[3];
        float* vel[3];
        float* p;
    };
    std::vector<Snapshot> ckpt;  // one per `adj_every` substeps of the forward
    std::vector<Snapshot> tape;  // the inputs of the substeps of one checkpoint interval
    Snapshot adj_final{};        // the state the forward ended in
    uint32_t adj_steps = 0, adj_every = 0;
    int adj_valid = 0;
    float *adj_gx[3] = {}, *adj_gv[3] = {};    // adjoint of the current step's input state
    float *adj_sgx[3] = {}, *adj_sgv[3] = {};  // the same in the step's sorted order
    float *adj_vbar[3] = {}, *adj_sbar[3] = {}, *adj_numbar[3] = {}, *adj_denbar[3] = {};
    float *adj_pbar = nullptr, *adj_rhs = nullptr, *adj_z = nullptr;
    double* adj_grav = nullptr;
    std::vector<void*> adj_allocs;

    template <typename T>
    hipError_t alloc(T*& ptr, size_t count) {
        void* p_ = nullptr;
        size_t nb = count * sizeof(T);
        if (nb == 0) nb = 16;
        hipError_t e = hipMalloc(&p_, nb);
        if (e != hipSuccess) return e;
        allocs.push_back(p_);
        bytes += nb;
        ptr = (T*)p_;
        return hipSuccess;
    }
    template <typename T>
    hipError_t adj_alloc(T*& ptr, size_t count) {
        void* p_ = nullptr;
        size_t nb = count * sizeof(T);
        if (nb == 0) nb = 16;
        hipError_t e = hipMalloc(&p_, nb);
        if (e != hipSuccess) return e;
        adj_allocs.push_back(p_);
        ptr = (T*)p_;
        return hipSuccess;
    }
    void drop_adjoint() {
        for (void* a : adj_allocs) (void)hipFree(a);
        adj_allocs.clear();
        ckpt.clear();
        tape.clear();
        adj_valid = 0;
    }
    void drop_graph() {
        if (cg_exec) (void)hipGraphExecDestroy(cg_exec);
        cg_exec = nullptr;
    }
    ~fs_tank() {
        drop_graph();
        drop_adjoint();
        for (void* a : allocs) (void)hipFree(a);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        for (auto& e : dev_ev) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};


namespace fs {
// One substep of the fused engine; with stop_before_g2p the particles are
// left sorted on the live side and G2P is not run (adjoint replay).
int substep_fused(fs_tank* t, bool sort_now, bool stop_before_g2p = false);
// n substeps of the tank's engine (sort schedule, timing, counters).
int do_step(fs_tank* t, uint32_t substeps);
}  // namespace fs
