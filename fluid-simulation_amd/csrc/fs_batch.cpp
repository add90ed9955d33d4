// fs_batch.cpp — the batched-tank C-ABI of include/fluidstep.h: one set of
// concatenated device arrays for all tanks, the per-phase work tables, the
// substep launch sequence (one launch per phase over every tank), its capture
// into a hipGraph per ping-pong side, and the autotuner of the tile shapes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "fluidstep.h"
#include "fs_batch.h"
#include "fs_host_common.h"

using namespace fs;

struct fs_batch {
    uint32_t T = 0;
    std::vector<fs_params> prm;
    std::vector<BatchTank> tanks;  // host copy of the device table
    uint32_t N_tot = 0, C_tot = 0, tiles_tot = 0, F_tot[3] = {0, 0, 0};
    int bf16 = 0;
    hipStream_t stream = nullptr;
    std::vector<void*> allocs;
    uint64_t bytes = 0;

    BatchTank* d_tanks = nullptr;
    WorkItem *wl_part = nullptr, *wl_cell = nullptr, *wl_face = nullptr, *wl_brick[3] = {}, *wl_cg = nullptr;
    uint32_t n_part = 0, n_cell = 0, n_face = 0, n_brick[3] = {0, 0, 0}, n_cg = 0;

    ParticleSoA side[2]{};
    int cur = 0;
    uint32_t *keys = nullptr, *ktmp0 = nullptr, *ktmp1 = nullptr, *vtmp = nullptr, *perm = nullptr;
    uint32_t *radix_ws = nullptr, *scan_ws = nullptr;
    uint32_t *occ = nullptr, *cell_offset = nullptr;
    uint8_t* ct = nullptr;
    BatchGrid grid{};
    float *div = nullptr, *b = nullptr, *p = nullptr, *r = nullptr, *d0 = nullptr, *d1 = nullptr, *q = nullptr;
    uint32_t* counters = nullptr;  // [4 T] per tank + [4 T] global misfit count
    uint32_t* mis_list = nullptr;
    BatchCgWork cg{};
    int cg_max_blocks = 0;      // co-resident figure of the CG kernel
    int cg_default_blocks = 0;  // persistent_grid_blocks of it (or a smaller override)
    int cg_limit_blocks = 0;    // the largest grid an override may ask for
    int cg_grid = 0;            // workgroups of the current CG plan (padded to 8)

    fs_batch_tuning tuning{};
    int use_graph = 1;
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    int perm_valid = 0;
    uint64_t substeps = 0;
    int timing = 0;
    hipEvent_t ev[2] = {};
    double ms_total = 0.0;

    uint32_t* mis_count() const { return counters + 4 * T; }

    template <typename X>
    hipError_t alloc(X*& ptr, size_t count) {
        void* p_ = nullptr;
        size_t nb = count * sizeof(X);
        if (nb == 0) nb = 16;
        hipError_t e = hipMalloc(&p_, nb);
        if (e != hipSuccess) return e;
        allocs.push_back(p_);
        bytes += nb;
        ptr = (X*)p_;
        return hipSuccess;
    }
    void drop_graphs() {
        for (auto& g : graph) {
            if (g) (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    }
    ~fs_batch() {
        drop_graphs();
        for (void* a : allocs) (void)hipFree(a);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// Consecutive items of a table go to one XCD: hardware workgroup w runs on
// XCD w % 8, so w takes item (w % 8) * (G / 8) + w / 8 (G padded to 8).
std::vector<WorkItem> xcd_order(std::vector<WorkItem> items) {
    while (items.size() % 8) items.push_back({kNoTank, 0});
    const size_t G = items.size(), per = G / 8;
    std::vector<WorkItem> out(G);
    for (size_t w = 0; w < G; ++w) out[w] = items[(w % 8) * per + w / 8];
    return out;
}

int upload_table(fs_batch* bt, const std::vector<WorkItem>& items, WorkItem*& dev, uint32_t& n) {
    const std::vector<WorkItem> ord = xcd_order(items);
    n = (uint32_t)ord.size();
    if (!dev) FS_HIP(bt->alloc(dev, ord.size() ? ord.size() : 1));
    if (n) FS_HIP(hipMemcpy(dev, ord.data(), ord.size() * sizeof(WorkItem), hipMemcpyHostToDevice));
    return FS_OK;
}

// Expected liquid tiles of a tank: its particles over the rest density.
double expected_tiles(const fs_batch* bt, uint32_t t) {
    const BatchTank& k = bt->tanks[t];
    const double cells = (double)k.P.N / std::max(1.0, (double)k.P.rest);
    return std::min((double)k.ntiles, std::max(1.0, std::ceil(cells / 64.0)));
}

// Split G persistent workgroups into one group per tank (at least one each,
// the rest by expected liquid tiles, at most one workgroup per 16 tiles), and
// place each group's workgroups on as few XCDs as possible.
int plan_cg(fs_batch* bt, int G) {
    const uint32_t T = bt->T;
    if (G < 8) G = 8;
    if ((uint32_t)G < T)
        return fail(FS_ERR_UNSUPPORTED, "more tanks than co-resident CG workgroups (" + std::to_string(G) + ")");
    std::vector<double> est(T);
    std::vector<uint32_t> nb(T, 1), cap(T);
    for (uint32_t t = 0; t < T; ++t) {
        est[t] = expected_tiles(bt, t);
        cap[t] = std::max(1u, (uint32_t)std::ceil(est[t] / 16.0));
    }
    int extra = G - (int)T;
    while (extra > 0) {
        int best = -1;
        double best_load = 0.0;
        for (uint32_t t = 0; t < T; ++t) {
            if (nb[t] >= cap[t]) continue;
            const double load = est[t] / nb[t];
            if (best < 0 || load > best_load) {
                best = (int)t;
                best_load = load;
            }
        }
        if (best < 0) break;
        nb[best] += 1;
        --extra;
    }
    uint32_t used = 0;
    for (uint32_t t = 0; t < T; ++t) {
        bt->tanks[t].cg_base = used;
        bt->tanks[t].cg_nb = nb[t];
        used += nb[t];
    }
    const uint32_t Ghw = (used + 7u) / 8u * 8u;
    std::vector<WorkItem> hw(Ghw, WorkItem{kNoTank, 0});
    std::vector<std::vector<uint32_t>> freeslots(8);
    for (uint32_t w = 0; w < Ghw; ++w) freeslots[w % 8].push_back(w);
    for (auto& f : freeslots) std::reverse(f.begin(), f.end());  // pop_back takes the lowest slot
    std::vector<uint32_t> order(T);
    for (uint32_t t = 0; t < T; ++t) order[t] = t;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t c) { return nb[a] > nb[c]; });
    for (uint32_t t : order) {
        uint32_t need = nb[t], rank = 0;
        while (need) {
            int x = 0;
            for (int i = 1; i < 8; ++i)
                if (freeslots[i].size() > freeslots[x].size()) x = i;
            while (need && !freeslots[x].empty()) {
                hw[freeslots[x].back()] = WorkItem{t, rank++};
                freeslots[x].pop_back();
                --need;
            }
        }
    }
    // Padding workgroups meet the grid barriers too, each with a slot of its own.
    uint32_t pad_slot = used;
    for (auto& w : hw)
        if (w.tank == kNoTank) w.start = pad_slot++;
    if (!bt->wl_cg) FS_HIP(bt->alloc(bt->wl_cg, (size_t)bt->cg_max_blocks + 8));
    if (Ghw > (uint32_t)bt->cg_max_blocks + 8) return fail(FS_ERR_UNSUPPORTED, "CG plan exceeds its table");
    FS_HIP(hipMemcpy(bt->wl_cg, hw.data(), hw.size() * sizeof(WorkItem), hipMemcpyHostToDevice));
    FS_HIP(hipMemcpy(bt->d_tanks, bt->tanks.data(), bt->tanks.size() * sizeof(BatchTank), hipMemcpyHostToDevice));
    bt->n_cg = Ghw;
    bt->cg_grid = (int)Ghw;
    return FS_OK;
}

int enqueue_substep(fs_batch* bt, int cur, hipStream_t s) {
    const ParticleSoA src = bt->side[cur];
    const ParticleSoA dst = bt->side[cur ^ 1];
    FS_HIP(launch_fill_u32(bt->occ, 0u, bt->C_tot, s));
    FS_HIP(launch_fill_u32(bt->counters, 0u, 8 * (size_t)bt->T, s));
    FS_HIP(launch_batch_key_hist(bt->d_tanks, bt->wl_part, bt->n_part, src, bt->keys, bt->occ, bt->counters,
                                 bt->mis_count(), bt->mis_list, s));
    FS_HIP(launch_exclusive_scan(bt->occ, bt->cell_offset, bt->C_tot, bt->scan_ws, s));
    FS_HIP(launch_stable_permutation(bt->keys, bt->N_tot, bt->C_tot, bt->ktmp0, bt->ktmp1, bt->vtmp, bt->perm,
                                     bt->radix_ws, s, (int)bt->tuning.sort_tile));
    // The permutation is applied by G2P, which rewrites every particle; until
    // then the gather reads the particles on src through it.
    FS_HIP(launch_batch_classify(bt->d_tanks, bt->wl_cell, bt->n_cell, bt->occ, bt->ct, bt->counters, s));
    const int v = (int)bt->tuning.gather_brick;
    FS_HIP(launch_batch_gather(bt->d_tanks, bt->wl_brick[v], bt->n_brick[v], v, bt->bf16, src, src, bt->occ,
                               bt->cell_offset, bt->ct, bt->grid, bt->mis_count(), bt->mis_list, s, bt->perm));
    // Atomic form when more particles' P2G cells differ from their sort cells
    // than the gather walks (every launch below is a no-op otherwise).
    for (int a = 0; a < 3; ++a) {
        FS_HIP(launch_fill_u32_guarded((uint32_t*)bt->grid.num[a], 0u, bt->F_tot[a], bt->mis_count(), kMisfitCap, s));
        FS_HIP(launch_fill_u32_guarded((uint32_t*)bt->grid.den[a], 0u, bt->F_tot[a], bt->mis_count(), kMisfitCap, s));
    }
    FS_HIP(launch_batch_p2g_fallback(bt->d_tanks, bt->wl_part, bt->n_part, bt->wl_face, bt->n_face, bt->bf16, src,
                                     bt->ct, bt->grid, bt->mis_count(), s));
    FS_HIP(launch_batch_div_cginit(bt->d_tanks, bt->wl_cell, bt->n_cell, bt->grid, bt->ct, bt->occ, bt->div, bt->b,
                                   bt->p, bt->r, bt->d0, bt->cg, s));
    FS_HIP(launch_exclusive_scan(bt->cg.tile_flags, bt->cg.tile_scan, bt->tiles_tot + 1, bt->scan_ws, s));
    FS_HIP(launch_compact_tiles(bt->tiles_tot + 1, bt->cg.tile_flags, bt->cg.tile_scan, bt->cg.tiles, bt->cg.n_tiles, s));
    FS_HIP(launch_fill_u32(bt->cg.arrive, 0u, 9 * 32, s));
    FS_HIP(launch_fill_u32(bt->cg.abort_flag, 0u, 1, s));
    FS_HIP(launch_fill_u32(reinterpret_cast<uint32_t*>(bt->cg.slots), 0u, bt->cg.slot_words, s));
    FS_HIP(launch_batch_cg(bt->d_tanks, bt->wl_cg, bt->n_cg, (int)bt->tuning.cg_unroll, bt->b, bt->p, bt->r, bt->d0,
                           bt->d1, bt->q, bt->cg, s));
    FS_HIP(launch_batch_gradient_enforce(bt->d_tanks, bt->wl_face, bt->n_face, bt->p, bt->ct, bt->grid, s));
    FS_HIP(launch_batch_g2p(bt->d_tanks, bt->wl_part, bt->n_part, bt->bf16, dst, bt->grid, s, &src, bt->perm));
    return FS_OK;
}

// Capture one substep (for the current side) into a graph; on any capture
// failure the batch runs its launches directly from then on.
int ensure_graph(fs_batch* bt, int cur) {
    if (!bt->use_graph || bt->graph[cur]) return FS_OK;
    hipGraph_t g = nullptr;
    FS_HIP(hipStreamBeginCapture(bt->stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_substep(bt, cur, bt->stream);
    const hipError_t ce = hipStreamEndCapture(bt->stream, &g);
    hipError_t ie = hipErrorUnknown;
    if (rc == FS_OK && ce == hipSuccess && g) ie = hipGraphInstantiate(&bt->graph[cur], g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    if (ie != hipSuccess) {
        bt->graph[cur] = nullptr;
        bt->drop_graphs();
        bt->use_graph = 0;
        (void)hipGetLastError();
        if (std::getenv("FLUIDSTEP_VERBOSE"))
            std::fprintf(stderr, "fluidstep: batch substep capture failed (%s); launching directly\n",
                         g_last_error.c_str());
    }
    return FS_OK;
}

int batch_step(fs_batch* bt, uint32_t substeps) {
    double acc = 0.0;
    for (uint32_t i = 0; i < substeps; ++i) {
        if (bt->timing) FS_HIP(hipEventRecord(bt->ev[0], bt->stream));
        if (int rc = ensure_graph(bt, bt->cur)) return rc;
        if (bt->use_graph) {
            FS_HIP(hipGraphLaunch(bt->graph[bt->cur], bt->stream));
        } else {
            if (int rc = enqueue_substep(bt, bt->cur, bt->stream)) return rc;
        }
        if (bt->timing) {
            FS_HIP(hipEventRecord(bt->ev[1], bt->stream));
            FS_HIP(hipEventSynchronize(bt->ev[1]));
            float ms = 0.0f;
            FS_HIP(hipEventElapsedTime(&ms, bt->ev[0], bt->ev[1]));
            acc += ms;
        }
        bt->cur ^= 1;
        bt->perm_valid = 1;
        bt->substeps += 1;
    }
    if (bt->timing) bt->ms_total = acc;
    return FS_OK;
}

int check_batch(const fs_batch* bt, uint32_t tank) {
    if (!bt) return fail(FS_ERR_INVALID_ARGUMENT, "null batch handle");
    if (tank >= bt->T) return fail(FS_ERR_INVALID_ARGUMENT, "tank index out of range");
    return FS_OK;
}

int check_tuning(const fs_batch_tuning& t) {
    if (t.gather_brick > 2 || t.sort_tile > 2 || t.cg_unroll > 2)
        return fail(FS_ERR_INVALID_ARGUMENT, "tuning menu index out of range");
    return FS_OK;
}

// Time `reps` replays of fn on the batch stream (ms per replay).
template <typename Fn>
int time_launches(fs_batch* bt, int reps, double& ms, Fn&& fn) {
    FS_HIP(hipEventRecord(bt->ev[0], bt->stream));
    for (int i = 0; i < reps; ++i)
        if (int rc = fn()) return rc;
    FS_HIP(hipEventRecord(bt->ev[1], bt->stream));
    FS_HIP(hipEventSynchronize(bt->ev[1]));
    float f = 0.0f;
    FS_HIP(hipEventElapsedTime(&f, bt->ev[0], bt->ev[1]));
    ms = f / reps;
    return FS_OK;
}

}  // namespace

extern "C" {

FS_API int fs_batch_create(const fs_params* params, const float* const* pos_xyz, const float* const* vel_xyz,
                           const uint32_t* counts, uint32_t ntanks, fs_batch** out) {
    if (!params || !counts || !out || ntanks == 0 || (!pos_xyz))
        return fail(FS_ERR_INVALID_ARGUMENT, "null argument or empty batch");
    *out = nullptr;
    uint64_t N_tot = 0, C_tot = 0, F_tot[3] = {0, 0, 0};
    for (uint32_t t = 0; t < ntanks; ++t) {
        const fs_params& q = params[t];
        if (counts[t] > 0 && !pos_xyz[t]) return fail(FS_ERR_INVALID_ARGUMENT, "null particle array");
        if (int rc = check_params(q, counts[t], pos_xyz[t], vel_xyz ? vel_xyz[t] : nullptr)) return rc;
        if (q.engine != FS_ENGINE_FUSED || q.precision != FS_PRECISION_MIXED)
            return fail(FS_ERR_UNSUPPORTED, "batched tanks run FS_ENGINE_FUSED with FS_PRECISION_MIXED");
        if (!q.particle_sort || q.particle_sort_period > 1)
            return fail(FS_ERR_UNSUPPORTED, "batched tanks sort their particles every substep");
        if (q.velocity_storage != params[0].velocity_storage)
            return fail(FS_ERR_UNSUPPORTED, "every tank of a batch needs the same velocity storage");
        N_tot += counts[t];
        C_tot += ((uint64_t)q.nx * q.ny * q.nz + 63) / 64 * 64;
        F_tot[0] += (uint64_t)(q.nx + 1) * q.ny * q.nz;
        F_tot[1] += (uint64_t)q.nx * (q.ny + 1) * q.nz;
        F_tot[2] += (uint64_t)q.nx * q.ny * (q.nz + 1);
    }
    if (N_tot >= (1ull << 31) || C_tot >= (1ull << 31) || F_tot[0] >= (1ull << 31) || F_tot[1] >= (1ull << 31) ||
        F_tot[2] >= (1ull << 31))
        return fail(FS_ERR_INVALID_ARGUMENT, "batch too large for 32-bit indices");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(FS_ERR_DEVICE, "no HIP device available");

    fs_batch* bt = new (std::nothrow) fs_batch();
    if (!bt) return fail(FS_ERR_OUT_OF_MEMORY, "host allocation failed");
    bt->T = ntanks;
    bt->prm.assign(params, params + ntanks);
    bt->bf16 = params[0].velocity_storage == FS_VELOCITY_BF16;
    bt->N_tot = (uint32_t)N_tot;
    bt->C_tot = (uint32_t)C_tot;
    bt->tiles_tot = (uint32_t)(C_tot / 64);
    for (int a = 0; a < 3; ++a) bt->F_tot[a] = (uint32_t)F_tot[a];
    bt->tanks.resize(ntanks);
    {
        uint32_t po = 0, co = 0, fo[3] = {0, 0, 0};
        for (uint32_t t = 0; t < ntanks; ++t) {
            BatchTank& k = bt->tanks[t];
            std::memset(&k, 0, sizeof k);
            build_step_params(params[t], counts[t], k.P);
            k.part_off = po;
            k.cell_off = co;
            k.tile_off = co / 64;
            k.ntiles = (k.P.C + 63) / 64;
            for (int a = 0; a < 3; ++a) {
                k.face_off[a] = fo[a];
                fo[a] += k.P.F[a];
            }
            po += counts[t];
            co += k.ntiles * 64;
        }
    }
    int rc = [&]() -> int {
        FS_HIP(hipStreamCreateWithFlags(&bt->stream, hipStreamNonBlocking));
        const size_t N = bt->N_tot, C = bt->C_tot;
        for (int sd = 0; sd < 2; ++sd)
            for (int a = 0; a < 3; ++a) {
                FS_HIP(bt->alloc(bt->side[sd].pos[a], N));
                if (bt->bf16) {
                    uint16_t* vp;
                    FS_HIP(bt->alloc(vp, N));
                    bt->side[sd].vel[a] = vp;
                } else {
                    float* vp;
                    FS_HIP(bt->alloc(vp, N));
                    bt->side[sd].vel[a] = vp;
                }
            }
        FS_HIP(bt->alloc(bt->keys, N));
        FS_HIP(bt->alloc(bt->ktmp0, N));
        FS_HIP(bt->alloc(bt->ktmp1, N));
        FS_HIP(bt->alloc(bt->vtmp, N));
        FS_HIP(bt->alloc(bt->perm, N));
        FS_HIP(bt->alloc(bt->radix_ws, radix_workspace(N)));
        FS_HIP(bt->alloc(bt->scan_ws, scan_workspace(C) + scan_workspace(bt->tiles_tot + 1)));
        FS_HIP(bt->alloc(bt->occ, C));
        FS_HIP(bt->alloc(bt->cell_offset, C));
        FS_HIP(bt->alloc(bt->ct, C));
        for (int a = 0; a < 3; ++a) {
            FS_HIP(bt->alloc(bt->grid.vel[a], bt->F_tot[a]));
            FS_HIP(bt->alloc(bt->grid.saved[a], bt->F_tot[a]));
            FS_HIP(bt->alloc(bt->grid.num[a], bt->F_tot[a]));
            FS_HIP(bt->alloc(bt->grid.den[a], bt->F_tot[a]));
        }
        FS_HIP(bt->alloc(bt->div, C));
        FS_HIP(bt->alloc(bt->b, C));
        FS_HIP(bt->alloc(bt->p, C));
        FS_HIP(bt->alloc(bt->r, C));
        FS_HIP(bt->alloc(bt->d0, C));
        FS_HIP(bt->alloc(bt->d1, C));
        FS_HIP(bt->alloc(bt->q, C));
        FS_HIP(bt->alloc(bt->counters, 8 * (size_t)ntanks));
        FS_HIP(bt->alloc(bt->mis_list, kMisfitCap));
        FS_HIP(bt->alloc(bt->cg.nmask, C));
        FS_HIP(bt->alloc(bt->cg.tile_flags, bt->tiles_tot + 1));
        FS_HIP(bt->alloc(bt->cg.tile_scan, bt->tiles_tot + 1));
        FS_HIP(bt->alloc(bt->cg.tiles, bt->tiles_tot + 1));
        FS_HIP(bt->alloc(bt->cg.n_tiles, 1));
        FS_HIP(bt->alloc(bt->cg.arrive, 9 * 32));
        FS_HIP(bt->alloc(bt->cg.gen, 32));
        FS_HIP(bt->alloc(bt->cg.abort_flag, 1));
        FS_HIP(bt->alloc(bt->cg.state, ntanks));
        bt->cg_max_blocks = batch_cg_max_blocks();
        if (bt->cg_max_blocks <= 0) return fail(FS_ERR_DEVICE, "cannot size the persistent CG grid");
        int dev = 0, cus = 0;
        FS_HIP(hipGetDevice(&dev));
        FS_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        // Overrides may only shrink the grid below the default sizing: a grid
        // above it is not launched (see persistent_grid_blocks).
        bt->cg_default_blocks = persistent_grid_blocks(bt->cg_max_blocks, cus);
        bt->cg_limit_blocks = bt->cg_default_blocks;
        if (const char* env = std::getenv("FLUIDSTEP_CG_BLOCKS")) {
            const int v = std::atoi(env);
            if (v > 0 && v <= bt->cg_limit_blocks) bt->cg_default_blocks = v;
        }
        FS_HIP(bt->alloc(bt->cg.partials, 2 * ((size_t)bt->cg_max_blocks + 8)));
        FS_HIP(bt->alloc(bt->cg.flags, 2 * ((size_t)bt->cg_max_blocks + 8)));
        bt->cg.slot_words = (uint32_t)(16 * ((size_t)bt->cg_max_blocks + 8));
        FS_HIP(bt->alloc(bt->cg.slots, bt->cg.slot_words / 2));
        FS_HIP(bt->alloc(bt->d_tanks, ntanks));
        if (const char* env = std::getenv("FLUIDSTEP_BATCH_GRAPH")) bt->use_graph = std::atoi(env) != 0;
        // Work tables.
        std::vector<WorkItem> part, cell, face, brick[3];
        static const int bx[3] = {8, 16, 32}, by[3] = {8, 8, 4}, bz[3] = {8, 8, 4};
        for (uint32_t t = 0; t < ntanks; ++t) {
            const StepParams& P = bt->tanks[t].P;
            for (uint32_t s0 = 0; s0 < P.N; s0 += kBatchBlock) part.push_back({t, s0});
            for (uint32_t s0 = 0; s0 < P.C; s0 += kBatchBlock) cell.push_back({t, s0});
            const uint32_t nf = P.F[0] + P.F[1] + P.F[2];
            for (uint32_t s0 = 0; s0 < nf; s0 += kBatchBlock) face.push_back({t, s0});
            for (int v = 0; v < 3; ++v) {
                const uint32_t nb = ((P.nx + bx[v] - 1) / bx[v]) * ((P.ny + by[v] - 1) / by[v]) *
                                    ((P.nz + bz[v] - 1) / bz[v]);
                for (uint32_t k = 0; k < nb; ++k) brick[v].push_back({t, k});
            }
        }
        if (int e = upload_table(bt, part, bt->wl_part, bt->n_part)) return e;
        if (int e = upload_table(bt, cell, bt->wl_cell, bt->n_cell)) return e;
        if (int e = upload_table(bt, face, bt->wl_face, bt->n_face)) return e;
        for (int v = 0; v < 3; ++v)
            if (int e = upload_table(bt, brick[v], bt->wl_brick[v], bt->n_brick[v])) return e;
        if (int e = plan_cg(bt, bt->cg_default_blocks)) return e;
        for (auto& e : bt->ev) FS_HIP(hipEventCreate(&e));
        // Defined contents before the first step.
        for (int a = 0; a < 3; ++a) {
            FS_HIP(launch_fill_f32(bt->grid.vel[a], 0.0f, bt->F_tot[a], bt->stream));
            FS_HIP(launch_fill_f32(bt->grid.saved[a], 0.0f, bt->F_tot[a], bt->stream));
        }
        FS_HIP(launch_fill_f32(bt->p, 0.0f, C, bt->stream));
        FS_HIP(launch_fill_f32(bt->div, 0.0f, C, bt->stream));
        FS_HIP(launch_fill_u32(bt->occ, 0u, C, bt->stream));
        FS_HIP(launch_fill_u32(bt->cell_offset, 0u, C, bt->stream));
        FS_HIP(launch_fill_u32((uint32_t*)bt->ct, 0u, C / 4, bt->stream));
        FS_HIP(launch_fill_u32(bt->counters, 0u, 8 * (size_t)ntanks, bt->stream));
        FS_HIP(launch_fill_u32(bt->cg.tile_flags, 0u, bt->tiles_tot + 1, bt->stream));
        FS_HIP(launch_fill_u32(bt->cg.gen, 0u, 32, bt->stream));
        FS_HIP(launch_fill_u32((uint32_t*)bt->cg.state, 0u, sizeof(BatchCgState) / 4 * ntanks, bt->stream));
        // Particles (xyz interleaved on the host -> SoA columns).
        std::vector<float> col;
        std::vector<uint16_t> colb;
        for (uint32_t t = 0; t < ntanks; ++t) {
            const uint32_t n = counts[t];
            if (!n) continue;
            const uint32_t off = bt->tanks[t].part_off;
            col.resize(n);
            for (int a = 0; a < 3; ++a) {
                for (uint32_t i = 0; i < n; ++i) col[i] = pos_xyz[t][3 * (size_t)i + a];
                FS_HIP(hipMemcpy(bt->side[0].pos[a] + off, col.data(), n * 4, hipMemcpyHostToDevice));
            }
            for (int a = 0; a < 3; ++a) {
                const float* v = vel_xyz ? vel_xyz[t] : nullptr;
                for (uint32_t i = 0; i < n; ++i) col[i] = v ? v[3 * (size_t)i + a] : 0.0f;
                if (bt->bf16) {
                    colb.resize(n);
                    for (uint32_t i = 0; i < n; ++i) {
                        uint32_t u;
                        std::memcpy(&u, &col[i], 4);
                        colb[i] = (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
                    }
                    FS_HIP(hipMemcpy((uint16_t*)bt->side[0].vel[a] + off, colb.data(), n * 2, hipMemcpyHostToDevice));
                } else {
                    FS_HIP(hipMemcpy((float*)bt->side[0].vel[a] + off, col.data(), n * 4, hipMemcpyHostToDevice));
                }
            }
        }
        FS_HIP(hipStreamSynchronize(bt->stream));
        if (std::getenv("FLUIDSTEP_VERBOSE"))
            std::fprintf(stderr, "fluidstep: batch of %u tanks, %u particles, %u cells, CG grid %d (max %d)\n",
                         ntanks, bt->N_tot, bt->C_tot, bt->cg_grid, bt->cg_max_blocks);
        return FS_OK;
    }();
    if (rc != FS_OK) {
        delete bt;
        return rc;
    }
    *out = bt;
    return FS_OK;
}

FS_API int fs_batch_destroy(fs_batch* bt) {
    if (!bt) return FS_OK;
    if (bt->stream) (void)hipStreamSynchronize(bt->stream);
    delete bt;
    return FS_OK;
}

FS_API int fs_batch_step(fs_batch* bt, uint32_t substeps) {
    if (!bt) return fail(FS_ERR_INVALID_ARGUMENT, "null batch handle");
    return batch_step(bt, substeps);
}

FS_API int fs_batch_synchronize(fs_batch* bt) {
    if (!bt) return fail(FS_ERR_INVALID_ARGUMENT, "null batch handle");
    FS_HIP(hipStreamSynchronize(bt->stream));
    // A solve that timed out in any substep since the last check leaves its
    // tank's sticky word set; report it once and clear it.
    std::vector<BatchCgState> st(bt->T);
    FS_HIP(hipMemcpy(st.data(), bt->cg.state, st.size() * sizeof(BatchCgState), hipMemcpyDeviceToHost));
    int first = -1;
    for (uint32_t t = 0; t < bt->T; ++t)
        if (st[t].failed) {
            if (first < 0) first = (int)t;
            const uint32_t zero = 0u;
            FS_HIP(hipMemcpy(&bt->cg.state[t].failed, &zero, sizeof zero, hipMemcpyHostToDevice));
        }
    if (first >= 0)
        return fail(FS_ERR_SOLVER, "batched CG group barrier timed out (tank " + std::to_string(first) + ")");
    return FS_OK;
}

FS_API int fs_batch_set_timing(fs_batch* bt, int enabled) {
    if (!bt) return fail(FS_ERR_INVALID_ARGUMENT, "null batch handle");
    bt->timing = enabled ? 1 : 0;
    return FS_OK;
}

FS_API int fs_batch_set_tuning(fs_batch* bt, const fs_batch_tuning* tuning) {
    if (!bt || !tuning) return fail(FS_ERR_INVALID_ARGUMENT, "null argument");
    if (int rc = check_tuning(*tuning)) return rc;
    FS_HIP(hipStreamSynchronize(bt->stream));
    int G = tuning->cg_blocks ? (int)tuning->cg_blocks : bt->cg_default_blocks;
    if (G > bt->cg_limit_blocks)
        return fail(FS_ERR_INVALID_ARGUMENT, "cg_blocks above the persistent grid limit (" +
                                                 std::to_string(bt->cg_limit_blocks) + ")");
    if (int rc = plan_cg(bt, G)) return rc;
    bt->tuning = *tuning;
    bt->tuning.cg_blocks = (uint32_t)G;
    bt->drop_graphs();
    return FS_OK;
}

FS_API int fs_batch_autotune(fs_batch* bt, fs_batch_tuning* out) {
    if (!bt) return fail(FS_ERR_INVALID_ARGUMENT, "null batch handle");
    hipStream_t s = bt->stream;
    FS_HIP(hipStreamSynchronize(s));
    const int cur = bt->cur;
    const ParticleSoA src = bt->side[cur], dst = bt->side[cur ^ 1];
    const size_t N = bt->N_tot, C = bt->C_tot;
    // The trials overwrite the other particle side and the pressure; keep the
    // warm-start pressure to put back (the live side is only read).
    float* p_keep = nullptr;
    FS_HIP(hipMalloc(&p_keep, C * sizeof(float)));
    struct Free {
        float* p;
        ~Free() { if (p) (void)hipFree(p); }
    } guard{p_keep};
    FS_HIP(hipMemcpyAsync(p_keep, bt->p, C * sizeof(float), hipMemcpyDeviceToDevice, s));
    fs_batch_tuning best = bt->tuning;
    // Keys and bins of the current state.
    FS_HIP(launch_fill_u32(bt->occ, 0u, C, s));
    FS_HIP(launch_fill_u32(bt->counters, 0u, 8 * (size_t)bt->T, s));
    FS_HIP(launch_batch_key_hist(bt->d_tanks, bt->wl_part, bt->n_part, src, bt->keys, bt->occ, bt->counters,
                                 bt->mis_count(), bt->mis_list, s));
    FS_HIP(launch_exclusive_scan(bt->occ, bt->cell_offset, C, bt->scan_ws, s));
    // Sort tile: the permutation is a function of the keys alone.
    double best_ms = 1e30;
    for (int v = 0; v < 3; ++v) {
        double ms = 0.0;
        auto run = [&]() -> int {
            FS_HIP(launch_stable_permutation(bt->keys, (uint32_t)N, (uint32_t)C, bt->ktmp0, bt->ktmp1, bt->vtmp,
                                             bt->perm, bt->radix_ws, s, v));
            return FS_OK;
        };
        if (int rc = run()) return rc;
        if (int rc = time_launches(bt, 3, ms, run)) return rc;
        if (ms < best_ms) {
            best_ms = ms;
            best.sort_tile = (uint32_t)v;
        }
    }
    best.ms_sort = best_ms;
    FS_HIP(launch_stable_permutation(bt->keys, (uint32_t)N, (uint32_t)C, bt->ktmp0, bt->ktmp1, bt->vtmp, bt->perm,
                                     bt->radix_ws, s, (int)best.sort_tile));
    FS_HIP(launch_permute((uint32_t)N, bt->perm, src, dst, bt->bf16, s));
    FS_HIP(launch_batch_classify(bt->d_tanks, bt->wl_cell, bt->n_cell, bt->occ, bt->ct, bt->counters, s));
    // Gather brick: vel/saved are a function of the sorted particles.
    best_ms = 1e30;
    for (int v = 0; v < 3; ++v) {
        double ms = 0.0;
        auto run = [&]() -> int {
            FS_HIP(launch_batch_gather(bt->d_tanks, bt->wl_brick[v], bt->n_brick[v], v, bt->bf16, src, src, bt->occ,
                                       bt->cell_offset, bt->ct, bt->grid, bt->mis_count(), bt->mis_list, s, bt->perm));
            return FS_OK;
        };
        if (int rc = run()) return rc;
        if (int rc = time_launches(bt, 3, ms, run)) return rc;
        if (ms < best_ms) {
            best_ms = ms;
            best.gather_brick = (uint32_t)v;
        }
    }
    best.ms_gather = best_ms;
    // P2G result for the CG trials (also the fallback form when it applies).
    const int gv = (int)best.gather_brick;
    FS_HIP(launch_batch_gather(bt->d_tanks, bt->wl_brick[gv], bt->n_brick[gv], gv, bt->bf16, dst, src, bt->occ,
                               bt->cell_offset, bt->ct, bt->grid, bt->mis_count(), bt->mis_list, s));
    for (int a = 0; a < 3; ++a) {
        FS_HIP(launch_fill_u32_guarded((uint32_t*)bt->grid.num[a], 0u, bt->F_tot[a], bt->mis_count(), kMisfitCap, s));
        FS_HIP(launch_fill_u32_guarded((uint32_t*)bt->grid.den[a], 0u, bt->F_tot[a], bt->mis_count(), kMisfitCap, s));
    }
    FS_HIP(launch_batch_p2g_fallback(bt->d_tanks, bt->wl_part, bt->n_part, bt->wl_face, bt->n_face, bt->bf16, dst,
                                     bt->ct, bt->grid, bt->mis_count(), s));
    // CG unroll x workgroups: each trial solves the same system from the same
    // warm start.
    best_ms = 1e30;
    const int Gd = bt->cg_default_blocks;
    const int Gs[2] = {Gd, std::max(8, (Gd / 2) / 8 * 8)};
    for (int gi = 0; gi < 2; ++gi) {
        if (gi == 1 && Gs[1] == Gs[0]) break;
        if (int rc = plan_cg(bt, Gs[gi])) return rc;
        for (int u = 0; u < 3; ++u) {
            FS_HIP(hipMemcpyAsync(bt->p, p_keep, C * sizeof(float), hipMemcpyDeviceToDevice, s));
            FS_HIP(launch_batch_div_cginit(bt->d_tanks, bt->wl_cell, bt->n_cell, bt->grid, bt->ct, bt->occ, bt->div,
                                           bt->b, bt->p, bt->r, bt->d0, bt->cg, s));
            FS_HIP(launch_exclusive_scan(bt->cg.tile_flags, bt->cg.tile_scan, bt->tiles_tot + 1, bt->scan_ws, s));
            FS_HIP(launch_compact_tiles(bt->tiles_tot + 1, bt->cg.tile_flags, bt->cg.tile_scan, bt->cg.tiles,
                                        bt->cg.n_tiles, s));
            FS_HIP(launch_fill_u32(bt->cg.arrive, 0u, 9 * 32, s));
            FS_HIP(launch_fill_u32(bt->cg.abort_flag, 0u, 1, s));
            FS_HIP(launch_fill_u32(reinterpret_cast<uint32_t*>(bt->cg.slots), 0u, bt->cg.slot_words, s));
            double ms = 0.0;
            auto run = [&]() -> int {
                FS_HIP(launch_batch_cg(bt->d_tanks, bt->wl_cg, bt->n_cg, u, bt->b, bt->p, bt->r, bt->d0, bt->d1,
                                       bt->q, bt->cg, s));
                return FS_OK;
            };
            if (int rc = time_launches(bt, 1, ms, run)) return rc;
            std::vector<BatchCgState> st(bt->T);
            FS_HIP(hipMemcpy(st.data(), bt->cg.state, st.size() * sizeof(BatchCgState), hipMemcpyDeviceToHost));
            for (const auto& x : st)
                if (x.error || x.failed)
                    return fail(FS_ERR_SOLVER, "batched CG group barrier timed out during tuning");
            if (ms < best_ms) {
                best_ms = ms;
                best.cg_unroll = (uint32_t)u;
                best.cg_blocks = (uint32_t)Gs[gi];
            }
        }
    }
    best.ms_cg = best_ms;
    if (int rc = plan_cg(bt, (int)best.cg_blocks)) return rc;
    FS_HIP(hipMemcpyAsync(bt->p, p_keep, C * sizeof(float), hipMemcpyDeviceToDevice, s));
    FS_HIP(hipStreamSynchronize(s));
    bt->tuning = best;
    bt->drop_graphs();
    if (out) *out = best;
    return FS_OK;
}

FS_API uint32_t fs_batch_tank_count(const fs_batch* bt) { return bt ? bt->T : 0u; }
FS_API uint64_t fs_batch_buffer_bytes(const fs_batch* bt) { return bt ? bt->bytes : 0u; }

FS_API int fs_batch_read_particles(fs_batch* bt, uint32_t tank, float* pos_xyz, float* vel_xyz) {
    if (int rc = check_batch(bt, tank)) return rc;
    FS_HIP(hipStreamSynchronize(bt->stream));
    const BatchTank& k = bt->tanks[tank];
    const uint32_t n = k.P.N, off = k.part_off;
    if (!n) return FS_OK;
    std::vector<float> col(n);
    std::vector<uint16_t> colb(n);
    const ParticleSoA& sd = bt->side[bt->cur];
    for (int a = 0; a < 3; ++a) {
        if (pos_xyz) {
            FS_HIP(hipMemcpy(col.data(), sd.pos[a] + off, n * 4, hipMemcpyDeviceToHost));
            for (uint32_t i = 0; i < n; ++i) pos_xyz[3 * (size_t)i + a] = col[i];
        }
        if (vel_xyz) {
            if (bt->bf16) {
                FS_HIP(hipMemcpy(colb.data(), (const uint16_t*)sd.vel[a] + off, n * 2, hipMemcpyDeviceToHost));
                for (uint32_t i = 0; i < n; ++i) {
                    const uint32_t u = ((uint32_t)colb[i]) << 16;
                    std::memcpy(&vel_xyz[3 * (size_t)i + a], &u, 4);
                }
            } else {
                FS_HIP(hipMemcpy(col.data(), (const float*)sd.vel[a] + off, n * 4, hipMemcpyDeviceToHost));
                for (uint32_t i = 0; i < n; ++i) vel_xyz[3 * (size_t)i + a] = col[i];
            }
        }
    }
    return FS_OK;
}

FS_API int fs_batch_read_faces(fs_batch* bt, uint32_t tank, int axis, int which, float* out) {
    if (int rc = check_batch(bt, tank)) return rc;
    if (axis < 0 || axis > 2 || !out || (which != FS_FACE_VELOCITY && which != FS_FACE_SAVED))
        return fail(FS_ERR_INVALID_ARGUMENT, "bad axis/which/out");
    FS_HIP(hipStreamSynchronize(bt->stream));
    const BatchTank& k = bt->tanks[tank];
    const float* src = (which == FS_FACE_VELOCITY ? bt->grid.vel[axis] : bt->grid.saved[axis]) + k.face_off[axis];
    FS_HIP(hipMemcpy(out, src, sizeof(float) * k.P.F[axis], hipMemcpyDeviceToHost));
    return FS_OK;
}

FS_API int fs_batch_read_cells_f32(fs_batch* bt, uint32_t tank, int which, float* out) {
    if (int rc = check_batch(bt, tank)) return rc;
    if (!out || (which != FS_CELL_PRESSURE && which != FS_CELL_DIVERGENCE))
        return fail(FS_ERR_INVALID_ARGUMENT, "bad which/out");
    FS_HIP(hipStreamSynchronize(bt->stream));
    const BatchTank& k = bt->tanks[tank];
    FS_HIP(hipMemcpy(out, (which == FS_CELL_PRESSURE ? bt->p : bt->div) + k.cell_off, sizeof(float) * k.P.C,
                     hipMemcpyDeviceToHost));
    return FS_OK;
}

FS_API int fs_batch_read_cells_u32(fs_batch* bt, uint32_t tank, int which, uint32_t* out) {
    if (int rc = check_batch(bt, tank)) return rc;
    if (!out) return fail(FS_ERR_INVALID_ARGUMENT, "null out");
    FS_HIP(hipStreamSynchronize(bt->stream));
    const BatchTank& k = bt->tanks[tank];
    const uint32_t C = k.P.C;
    if (which == FS_CELL_TYPE) {
        std::vector<uint8_t> tmp(C);
        FS_HIP(hipMemcpy(tmp.data(), bt->ct + k.cell_off, C, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < C; ++i) out[i] = tmp[i];
    } else if (which == FS_CELL_OCCUPANCY) {
        FS_HIP(hipMemcpy(out, bt->occ + k.cell_off, sizeof(uint32_t) * C, hipMemcpyDeviceToHost));
    } else if (which == FS_CELL_OFFSET) {
        // Bin starts are batch-wide particle indices; report them per tank.
        FS_HIP(hipMemcpy(out, bt->cell_offset + k.cell_off, sizeof(uint32_t) * C, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < C; ++i) out[i] -= k.part_off;
    } else {
        return fail(FS_ERR_INVALID_ARGUMENT, "bad which");
    }
    return FS_OK;
}

FS_API int fs_batch_read_permutation(fs_batch* bt, uint32_t tank, uint32_t* perm) {
    if (int rc = check_batch(bt, tank)) return rc;
    if (!perm) return fail(FS_ERR_INVALID_ARGUMENT, "null out");
    if (!bt->perm_valid) return fail(FS_ERR_INVALID_ARGUMENT, "no sort has run since create");
    FS_HIP(hipStreamSynchronize(bt->stream));
    const BatchTank& k = bt->tanks[tank];
    if (!k.P.N) return FS_OK;
    FS_HIP(hipMemcpy(perm, bt->perm + k.part_off, sizeof(uint32_t) * k.P.N, hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < k.P.N; ++i) perm[i] -= k.part_off;
    return FS_OK;
}

FS_API int fs_batch_get_stats(fs_batch* bt, uint32_t tank, fs_stats* out) {
    if (int rc = check_batch(bt, tank)) return rc;
    if (!out) return fail(FS_ERR_INVALID_ARGUMENT, "null out");
    FS_HIP(hipStreamSynchronize(bt->stream));
    const BatchTank& k = bt->tanks[tank];
    fs_stats s;
    std::memset(&s, 0, sizeof s);
    uint32_t cnt[4], mis = 0, ts[2] = {0, 0};
    FS_HIP(hipMemcpy(cnt, bt->counters + 4 * tank, sizeof cnt, hipMemcpyDeviceToHost));
    FS_HIP(hipMemcpy(&mis, bt->mis_count(), sizeof mis, hipMemcpyDeviceToHost));
    FS_HIP(hipMemcpy(&ts[0], bt->cg.tile_scan + k.tile_off, 4, hipMemcpyDeviceToHost));
    FS_HIP(hipMemcpy(&ts[1], bt->cg.tile_scan + k.tile_off + k.ntiles, 4, hipMemcpyDeviceToHost));
    BatchCgState st;
    FS_HIP(hipMemcpy(&st, bt->cg.state + tank, sizeof st, hipMemcpyDeviceToHost));
    if (st.error || st.failed) return fail(FS_ERR_SOLVER, "batched CG group barrier timed out");
    s.liquid_cells = cnt[0];
    s.misfits = cnt[1];
    s.sorted = bt->substeps ? 1u : 0u;
    s.p2g_gather = (uint32_t)(bt->substeps && mis <= kMisfitCap);
    s.cg_iterations = st.iterations;
    s.rs_initial = st.rs_init;
    s.rs_final = st.rs_final;
    s.true_residual_sq = st.true_res_sq;
    s.cg_active_cells = (ts[1] - ts[0]) * 64u;
    s.ms_total = bt->ms_total;
    s.substeps = bt->substeps;
    *out = s;
    return FS_OK;
}

}  // extern "C"
