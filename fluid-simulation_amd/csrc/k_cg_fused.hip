// k_cg_fused.hip — the mixed-precision pressure solve of the fused engine.
//
//  * k_div_cginit: divergence (sh/divergence.wgsl) and CG init
//    (sh/cg_init.wgsl) per cell in one pass; it also packs each liquid cell's
//    stencil into a 16-bit mask (non-solid neighbour count, which neighbours
//    are liquid) and flags the 64-cell tiles that hold liquid.
//  * the liquid tiles are compacted into a list (exclusive scan), so the solve
//    never touches all-air or all-solid tiles.
//  * the solve itself, in one of three forms:
//      - k_cg_persistent (the default): one cooperative launch for the whole
//        solve on the fewest workgroups that keep every wave's tiles in one
//        batch (solve_blocks).  A workgroup keeps its cells in registers when
//        they fit (ResidentCells) and streams them otherwise.  Two grid-wide
//        reductions per iteration: a counter-free exchange after phase A
//        (grid_exchange) and a memory-ordering one after phase B
//        (grid_sum_tagged, grid_sum or the fence-free coherent_sum, by
//        workgroup count and residency);
//      - k_cg_single: the same CG in the single-reduction arrangement (one
//        sweep and one barrier per iteration), kept as a measured
//        alternative: its rounding path misses the per-step parity bar;
//      - k_cg_phase_a / k_cg_phase_b: two launches per iteration, each ending
//        in a last-block-done reduction (no block ever waits), enqueued for
//        the iteration cap and replayed from a hipGraph by the host.
//    Per iteration: phase A forms dn = r + beta d at every stencil point
//    (identical expression, so neighbours agree bit for bit with the
//    reference's separate sh/cg_dir.wgsl pass), q = A dn and dn.q; phase B
//    updates p += alpha dn, r -= alpha q and r.r.  Vectors are f32; dots are
//    f64 summed in a fixed order; alpha and beta are rounded to f32 as
//    src/sim/pressure.rs:282,296 do; the stop test rs_new <= tol^2 rs_init is
//    f64 (src/sim/pressure.rs:262-292).
//  * the f64 true residual |b - A p|^2 is reduced after the loop.
#include <cstdlib>
#include <cstring>
#include "fs_cell.h"
#include "fs_cg.h"

namespace fs {
// A solve whose grid barrier timed out: the last-solve word and the sticky
// word (cleared only by the host) are both set.
__device__ inline void cg_mark_failed(CgState* st) {
    st->error = 1u;
    st->failed = 1u;
}

namespace {
inline unsigned grid_for(size_t n, int per_block) { return (unsigned)((n + per_block - 1) / per_block); }
constexpr int kCgMultiThreads = 256;  // multi-kernel workgroup
constexpr int kLdsTiles = 2048;       // tile ids a persistent block keeps in LDS
// Sync area layout (u32 words): each counter or flag on its own 128-byte line.
constexpr size_t kSyncBarCount = 0, kSyncGroups = 32, kSyncAbort = 32 + 8 * 32, kSyncGen = kSyncAbort + 32,
                 kSyncSlots = kSyncGen + 32;
}  // namespace

// Two slot areas (exchange, data barrier) of two epoch halves of max_blocks
// slots of kSlotStride u64 each.
size_t cg_sync_words(int max_blocks) {
    return kSyncSlots + 8u * kSlotStride * (size_t)(max_blocks > 0 ? max_blocks : 1);
}

void cg_bind_sync(CgWork& w, uint32_t* area, int max_blocks) {
    w.sync = area;
    w.sync_words = (uint32_t)cg_sync_words(max_blocks);
    w.bar_count = area + kSyncBarCount;
    w.group_count = area + kSyncGroups;
    w.abort_flag = area + kSyncAbort;
    w.bar_gen = area + kSyncGen;
    w.slots = reinterpret_cast<uint64_t*>(area + kSyncSlots);
    w.slots_data =
        reinterpret_cast<uint64_t*>(area + kSyncSlots + 4u * kSlotStride * (size_t)(max_blocks > 0 ? max_blocks : 1));
    w.spin_ticks = kSpinTimeoutTicks;
}

hipError_t launch_cg_reset(const CgWork& w, hipStream_t s) { return launch_fill_u32(w.sync, 0u, w.sync_words, s); }

__global__ void k_div_cginit(StepParams P, const float* __restrict__ u, const float* __restrict__ v,
                             const float* __restrict__ w, const uint8_t* __restrict__ ct,
                             const uint32_t* __restrict__ occ, float* __restrict__ div, float* __restrict__ b,
                             float* __restrict__ p, float* __restrict__ r, float* __restrict__ d,
                             uint16_t* __restrict__ nmask, uint32_t* __restrict__ tile_flags) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    bool liquid = false;
    if (c < P.C) liquid = div_cginit_cell(P, u, v, w, ct, occ, c, div, b, p, r, d, nmask);
    const uint64_t m = __ballot(liquid);
    const uint32_t tile = c / kTile;
    if (lane_id() == 0 && (size_t)tile * kTile < P.C) tile_flags[tile] = m ? 1u : 0u;
}

__global__ void k_compact_tiles(uint32_t ntiles, const uint32_t* __restrict__ flags, const uint32_t* __restrict__ scan,
                                uint32_t* __restrict__ tiles, uint32_t* __restrict__ n_active) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    if (flags[t]) tiles[scan[t]] = t;
    if (t == ntiles - 1u) *n_active = scan[t] + flags[t];
}

hipError_t launch_compact_tiles(uint32_t ntiles, const uint32_t* flags, const uint32_t* scan, uint32_t* tiles,
                                uint32_t* n_active, hipStream_t s) {
    FS_LAUNCH(k_compact_tiles, dim3(grid_for(ntiles, 256)), dim3(256), 0, s, ntiles, flags, scan, tiles,
                       n_active);
    return hipGetLastError();
}

hipError_t launch_div_cginit(const StepParams& P, float* const vel[3], const uint8_t* ct, const uint32_t* occ,
                             float* div, float* b, float* p, float* r, float* d, const CgWork& w, uint32_t* scan_ws,
                             hipStream_t s) {
    // blockDim is a multiple of 64, so each wave is exactly one tile.
    FS_LAUNCH(k_div_cginit, dim3(grid_for(P.C, 256)), dim3(256), 0, s, P, vel[0], vel[1], vel[2], ct, occ,
                       div, b, p, r, d, w.nmask, w.tile_flags);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const uint32_t ntiles = (P.C + kTile - 1) / kTile;
    e = launch_exclusive_scan(w.tile_flags, w.tile_scan, ntiles, scan_ws, s);
    if (e != hipSuccess) return e;
    FS_LAUNCH(k_compact_tiles, dim3(grid_for(ntiles, 256)), dim3(256), 0, s, ntiles, w.tile_flags,
                       w.tile_scan, w.tiles, w.n_tiles);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Deterministic f64 grid reductions.
//
// Each block sums its waves in wave order and publishes one partial.  Arrival
// is counted in two levels (a counter per group of blocks with equal
// blockIdx % 8, then one counter over the groups), so no single address takes
// more than G/8 atomics.  The partials are summed in one fixed order (by the
// block arriving last in last_block_sum, by every block in grid_sum), so the
// total does not depend on timing.  Ordering:
// every wave's stores reach L2 before the workgroup barrier; thread 0's
// release on its counter writes the XCD's L2 back; the consumer's acquire
// invalidates its caches once.  Waiting blocks poll with relaxed loads (an
// acquire per poll would invalidate the shared L2 on every spin).
// Two-level arrival (thread 0 only); returns true for the block arriving last.
// Arrivals are release RMWs; only a block that turns out to be a group's or
// the grid's last arrival issues the acquire fence (fence-after-RMW
// synchronizes with every release earlier in the counter's RMW chain).
__device__ __forceinline__ bool arrive(const CgWork& w, uint32_t G) {
    const uint32_t grp = blockIdx.x & 7u;
    const uint32_t in_grp = (G - grp + 7u) / 8u;
    const uint32_t ngrp = G < 8u ? G : 8u;
    uint32_t* gc = w.group_count + grp * 32u;  // one 128-byte line per group
    const uint32_t prev = __hip_atomic_fetch_add(gc, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    if (prev != in_grp - 1u) return false;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    __hip_atomic_store(gc, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t prev2 = __hip_atomic_fetch_add(w.bar_count, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    if (prev2 != ngrp - 1u) return false;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    __hip_atomic_store(w.bar_count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return true;
}

// Grid barrier + sum for the persistent kernels.  Orders memory: every
// block's stores before the barrier are visible to every block after it.
// The last arrival only bumps the generation word; afterwards every block
// sums the G published partials itself in the same fixed order, so all
// blocks hold the identical total without another round trip.  Partials
// alternate between two halves of the array so a fast block's next
// publication cannot overwrite a value a slow block is still summing.
// Returns false on abort.
template <int THREADS, int NV>
__device__ bool grid_sum_nv(const double (&v)[NV], const CgWork& w, uint32_t G, double (&out)[NV]) {
    __shared__ int s_last, s_abort;
    __shared__ uint32_t s_gen;
    double bs[NV];
    block_sum_nv<THREADS, NV>(v, bs);
    if (threadIdx.x == 0) {
        const uint32_t g0 = ld_relaxed(w.bar_gen);
#pragma unroll
        for (int j = 0; j < NV; ++j) w.partials[((size_t)(g0 & 1u) * G + blockIdx.x) * NV + j] = bs[j];
        s_gen = g0;
        s_abort = 0;
        s_last = arrive(w, G) ? 1 : 0;
    }
    __syncthreads();
    if (s_last) {
        if (threadIdx.x == 0)
            __hip_atomic_store(w.bar_gen, s_gen + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else if (threadIdx.x == 0) {
        const long long t0 = wall_clock64();
        while (ld_relaxed(w.bar_gen) == s_gen) {
            if (ld_relaxed(w.abort_flag)) {
                s_abort = 1;
                break;
            }
            if (wall_clock64() - t0 > w.spin_ticks) {
                __hip_atomic_store(w.abort_flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                s_abort = 1;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
    if (s_abort) return false;
    const double* base = w.partials + (size_t)(s_gen & 1u) * G * NV;
    double acc[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) acc[j] = 0.0;
    for (uint32_t i = threadIdx.x; i < G; i += THREADS)
#pragma unroll
        for (int j = 0; j < NV; ++j)
            acc[j] += __hip_atomic_load(&base[(size_t)i * NV + j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    reduce_loaded<THREADS, NV>(acc, out);
    return true;
}

template <int THREADS>
__device__ __forceinline__ bool grid_sum(double v, const CgWork& w, uint32_t G, double& out) {
    const double in[1] = {v};
    double o[1];
    if (!grid_sum_nv<THREADS, 1>(in, w, G, o)) return false;
    out = o[0];
    return true;
}

// Reduction without counters or fences, for a total whose consumers read no
// data another block wrote since the previous grid_sum: every block publishes
// its partial in its own epoch-tagged slot (fs_cg.h) and reads all G slots
// until each carries the epoch; the total is then formed in grid_sum's order.
// No write-back or invalidate is issued.  Slots alternate between two halves
// by epoch parity: a block publishes epoch e+1 only after reading every
// epoch-e slot, and epoch e+2 only after every block has published e+1, i.e.
// finished reading epoch e.  The slots start cleared and epochs start at 1.
// Returns false on abort.
template <int THREADS>
__device__ __forceinline__ bool grid_exchange_all(double v, const CgWork& w, uint64_t* slots, uint32_t G, uint32_t epoch,
                                  double& out) {
    __shared__ int s_abort;
    uint64_t* half = slots + (size_t)(epoch & 1u) * G * kSlotStride;
    const double in[1] = {v};
    double bs[1];
    block_sum_nv<THREADS, 1>(in, bs);
    if (threadIdx.x == 0) {
        s_abort = 0;
        publish_tagged(half, blockIdx.x, epoch, bs[0], 0u, kSlotStride);
    }
    __syncthreads();
    double acc[1] = {0.0};
    for (uint32_t i = threadIdx.x; i < G; i += THREADS) {
        double x;
        uint32_t f;
        if (!poll_tagged(half, i, epoch, w.abort_flag, x, f, w.spin_ticks, kSlotStride)) {
            s_abort = 1;
            break;
        }
        acc[0] += x;
    }
    double o[1];
    reduce_loaded<THREADS, 1>(acc, o);
    if (s_abort) return false;
    out = o[0];
    return true;
}

// Sum over tagged slots read by wave 0 alone, for G <= kPollSlots * 64.
// Thread 0 publishes the block's partial in its epoch-tagged slot; lane l
// of wave 0 polls slots l, l + 64, ... all at once (their tags are loaded
// together each round) until each carries the epoch.  The total is formed
// in reduce_loaded's order, as if thread i held slot i: the wave sums of
// slots [64k, 64k + 64), added in k order from 0.0.  So every block, and
// every reduction of this file, holds the identical total.  Slots alternate by epoch parity as in grid_exchange_all.
// FENCED makes it a grid barrier as well (grid_sum_tagged): every block's
// stores before it are visible to every block after it.  Thread 0 issues a
// release fence (writes the XCD's L2 back) before publishing; the polling
// lanes issue the acquire fence after their reads (each lane's fence follows
// its own reads of the tags it waited for), and the workgroup barrier orders
// the other waves after it.  The fenced form has its own slot area and epoch
// sequence.
constexpr uint32_t kPollSlots = 4;
template <int THREADS, bool FENCED>
__device__ __forceinline__ bool tagged_sum(double v, const CgWork& w, uint64_t* slots, uint32_t G, uint32_t epoch,
                                           double& out) {
    __shared__ double s_out;
    __shared__ int s_abort;
    uint64_t* half = slots + (size_t)(epoch & 1u) * G * kSlotStride;
    const double in[1] = {v};
    double bs[1];
    block_sum_nv<THREADS, 1>(in, bs);
    if (threadIdx.x == 0) {
        s_abort = 0;
        if (FENCED) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        publish_tagged(half, blockIdx.x, epoch, bs[0], 0u, kSlotStride);
    }
    if (threadIdx.x < 64) {
        const uint32_t lane = threadIdx.x;
        double x[kPollSlots];
        uint32_t pending = 0;  // bit k: slot lane + 64 k not yet seen at this epoch
#pragma unroll
        for (uint32_t k = 0; k < kPollSlots; ++k) {
            x[k] = 0.0;
            if (lane + 64u * k < G) pending |= 1u << k;
        }
        const long long t0 = wall_clock64();
        bool ok = true;
        for (uint32_t spin = 0; pending; ++spin) {
            // Rows k of slots that exist (64 k < G, the same for all lanes)
            // are loaded by every lane, at an index clamped to the last slot
            // (in bounds); only pending slots take the value.
            uint64_t hi[kPollSlots], lo[kPollSlots];
#pragma unroll
            for (uint32_t k = 0; k < kPollSlots; ++k) {
                hi[k] = lo[k] = 0;
                if (64u * k < G) {
                    const uint32_t i = lane + 64u * k < G ? lane + 64u * k : G - 1u;
                    hi[k] = __hip_atomic_load(half + kSlotStride * i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    lo[k] = __hip_atomic_load(half + kSlotStride * i + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
#pragma unroll
            for (uint32_t k = 0; k < kPollSlots; ++k) {
                if ((pending >> k) & 1u && (uint32_t)(hi[k] >> 32) == epoch && (uint32_t)(lo[k] >> 33) == epoch) {
                    x[k] = 0.0 + __longlong_as_double((long long)(((hi[k] & 0xffffffffull) << 32) |
                                                                  (lo[k] & 0xffffffffull)));
                    pending &= ~(1u << k);
                }
            }
            if (pending && (spin & 31u) == 31u &&
                (ld_relaxed(w.abort_flag) || wall_clock64() - t0 > w.spin_ticks)) {
                __hip_atomic_store(w.abort_flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = false;
                break;
            }
        }
        if (!ok) s_abort = 1;
        if (FENCED) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        // Rows past G hold +0.0 in every lane and would add +0.0; a sum that
        // starts at +0.0 is never -0.0, so leaving those terms out changes
        // no bit of reduce_loaded's total.
        double t = 0.0;
#pragma unroll
        for (uint32_t k = 0; k < kPollSlots; ++k)
            if (64u * k < G) t += wave_sum(x[k]);
        if (lane == 0) s_out = t;
    }
    __syncthreads();
    if (s_abort) return false;
    out = s_out;
    return true;
}

template <int THREADS>
__device__ __forceinline__ bool grid_sum_tagged(double v, const CgWork& w, uint32_t G, uint32_t epoch, double& out) {
    // tagged_sum reads at most kPollSlots * 64 slots; larger grids (512-thread
    // workgroups) take the counter barrier.
    return G <= kPollSlots * 64u ? tagged_sum<THREADS, true>(v, w, w.slots_data, G, epoch, out)
                                 : grid_sum<THREADS>(v, w, G, out);
}

// The phase-A exchange: up to kPollSlots * 64 workgroups, wave 0 alone
// reads the slots (fewer workgroup barriers); above, every thread reads one.
template <int THREADS>
__device__ __forceinline__ bool grid_exchange(double v, const CgWork& w, uint32_t G, uint32_t epoch, double& out) {
    return G <= kPollSlots * 64u ? tagged_sum<THREADS, false>(v, w, w.slots, G, epoch, out)
                                 : grid_exchange_all<THREADS>(v, w, w.slots, G, epoch, out);
}

// Fence-free grid barrier + sum for the two-phase solve, in which the only
// data that crosses workgroups are r and d (neighbour values): written with
// store_through (agent-scope sc1 stores, written through the L2) and read
// with ld_coherent_at (agent-scope sc1 loads, which bypass the L1 and read
// at the coherence point).  Each wave first waits until its own stores are
// acknowledged (s_waitcnt vmcnt(0): an acknowledged sc1 store is visible at
// agent scope); the workgroup barrier inside the sum then orders those
// waits before thread 0 publishes the tagged partial.  A reader issues its
// loads only after its workgroup has seen every block's tag, so each load
// returns the value stored before that tag.  These are the rules by which
// agent-scope atomics are ordered on this target; here they replace the L2
// write-back and invalidate of a fenced barrier.  It uses the tagged slots
// and epochs of grid_sum_tagged.
template <int THREADS>
__device__ __forceinline__ bool coherent_sum(double v, const CgWork& w, uint32_t G, uint32_t epoch, double& out) {
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
    return G <= kPollSlots * 64u ? tagged_sum<THREADS, false>(v, w, w.slots_data, G, epoch, out)
                                 : grid_exchange_all<THREADS>(v, w, w.slots_data, G, epoch, out);
}

// Last-block-done reduction: no block waits; the last to arrive gets the
// total (valid in thread 0) and returns true.
template <int THREADS>
__device__ bool last_block_sum(double v, const CgWork& w, double& tot) {
    __shared__ double s_w[THREADS / 64];
    __shared__ double s_tree[THREADS];
    __shared__ int s_last;
    const double bs = block_sum<THREADS>(v, s_w);
    if (threadIdx.x == 0) {
        w.partials[blockIdx.x] = bs;
        s_last = arrive(w, gridDim.x) ? 1 : 0;
    }
    __syncthreads();
    if (!s_last) return false;
    tot = sum_partials<THREADS>(w.partials, gridDim.x, s_tree);
    return true;
}

// ---- persistent form ----
// Workgroups that take part in a persistent solve over nt liquid tiles: the
// fewest whole XCD groups (multiples of 8) that leave each workgroup at most
// per_block tiles, capped at the launched grid.  A grid barrier's latency
// grows with the number of arrivals, so a small solve runs on few
// workgroups; the others exit at once and never meet a barrier.  Every
// workgroup reads the same nt, so all agree on the count.
__device__ __forceinline__ uint32_t solve_blocks(uint32_t nt, uint32_t per_block) {
    const uint32_t launched = gridDim.x;
    uint32_t g = (nt + per_block - 1u) / per_block;
    g = (g + 7u) / 8u * 8u;
    if (g < 8u) g = 8u;
    return g < launched ? g : launched;
}

// Form of the reductions that order memory: kDbCounter (grid_sum),
// kDbTagged (grid_sum_tagged), kDbCoherent (coherent_sum in the loop) or
// kDbAuto: tagged slots up to kTaggedMaxBlocks participating workgroups,
// coherent_sum above.
// Measured per iteration on MI355X, tagged vs counter: 8 workgroups 7.2 vs
// 9.5 us, 16 about 9.3 vs 11.0, 208 equal, 256 (streaming 256^3) 51.0 vs
// 50.4; coherent vs fenced: 8 workgroups 6.97 vs 6.13 (tagged), 16 8.44 vs
// 8.05 (tagged), 208 resident 11.4 vs 14.1 (counter), 256 streaming 41.0 vs
// 45.6 (counter).
constexpr int kDbCounter = 0, kDbTagged = 1, kDbAuto = 2, kDbCoherent = 3;
constexpr uint32_t kTaggedMaxBlocks = 128;
template <int DB>
__device__ __forceinline__ bool use_tagged(uint32_t G) {
    return DB == kDbTagged || ((DB == kDbAuto || DB == kDbCoherent) && G <= kTaggedMaxBlocks);
}

// The iteration loop of the two-phase solve over one block's tiles; RES
// keeps the cells in registers (every wave has at most one batch of kU
// tiles); COH uses coherent_sum, write-through r and d stores and coherent
// neighbour loads.  Returns false on abort.  All forms run the same
// reductions in the same order and the same f32 expressions.
template <int THREADS, bool EXCH, int DB, bool RES, bool COH, typename TileId>
__device__ __forceinline__ bool cg_two_phase(const StepParams& P, const float* __restrict__ b, float* __restrict__ p,
                                             float* __restrict__ r, float* __restrict__ d0, float* __restrict__ d1,
                                             float* __restrict__ q, const CgWork& w, uint32_t G, const TileWalk& tw,
                                             TileId&& tile_id, double& rs_init_out, double& rs_final,
                                             uint32_t& iters) {
    const uint32_t step = tw.stride * kU;
    // The reductions that order memory: counters (grid_sum) or tagged slots
    // (grid_sum_tagged, epochs 1 = r.r of r_0, k + 2 = iteration k).
    auto data_sum = [&](double v, uint32_t epoch, double& out) -> bool {
        if (COH) return coherent_sum<THREADS>(v, w, G, epoch, out);
        return use_tagged<DB>(G) ? grid_sum_tagged<THREADS>(v, w, G, epoch, out) : grid_sum<THREADS>(v, w, G, out);
    };
    ResidentCells<kU> rc;
    double acc = 0.0;
    if (RES)
        acc = resident_load<kU>(P, w.nmask, r, d0, p, tw, tile_id, rc);
    else
        for (uint32_t ti = tw.begin; ti < tw.end; ti += step) acc += rr_batch<kU>(P, r, tw, ti, tile_id);
    double rs0;
    if (!data_sum(acc, 1u, rs0)) return false;
    double rs_old = rs0;
    const double rs_init = rs0;
    const double tol = (double)P.tol;
    const double tol_sq = tol * tol;
    float beta = 0.0f;
    float* dprev = d0;
    float* dnext = d1;
    // Streaming form: p takes the updates of iterations k-1 and k together at
    // odd k (phase_b_batch kPTwo); an even last iteration leaves one pending.
    float alpha_prev = 0.0f, last_alpha = 0.0f;
    const float* last_dn = d0;
    iters = 0;
    for (uint32_t it = 0; it < P.iters; ++it) {
        ++iters;
        const bool first = it == 0;
        float* dn_buf = first ? dprev : dnext;
        acc = 0.0;
        if (RES)
            acc = resident_phase_a<kU, COH>(P, r, dprev, dnext, beta, first, rc);
        else
            for (uint32_t ti = tw.begin; ti < tw.end; ti += step)
                acc += phase_a_batch<kU, COH>(P, w.nmask, r, dprev, dnext, q, beta, first, tw, ti, tile_id);
        double dq;
        const bool ok = EXCH ? grid_exchange<THREADS>(acc, w, G, it + 1u, dq) : grid_sum<THREADS>(acc, w, G, dq);
        if (!ok) return false;
        const float alpha = fabs(dq) > 1e-30 ? (float)(rs_old / dq) : 0.0f;
        acc = 0.0;
        if (RES)
            acc = resident_phase_b<kU>(P, r, alpha, rc);
        else if (it & 1u)
            for (uint32_t ti = tw.begin; ti < tw.end; ti += step)
                acc += phase_b_batch<kU, kPTwo, COH>(P, dn_buf, q, p, r, alpha, dprev, alpha_prev, tw, ti,
                                                             tile_id);
        else
            for (uint32_t ti = tw.begin; ti < tw.end; ti += step)
                acc += phase_b_batch<kU, kPNone, COH>(P, dn_buf, q, p, r, alpha, nullptr, 0.0f, tw, ti,
                                                              tile_id);
        alpha_prev = alpha;
        last_alpha = alpha;
        last_dn = dn_buf;
        double rr;
        if (!data_sum(acc, it + 2u, rr)) return false;
        beta = rs_old > 1e-30 ? (float)(rr / rs_old) : 0.0f;
        rs_old = rr;
        if (tol > 0.0 && rr <= tol_sq * rs_init) break;
        if (!first) {
            float* t = dprev;
            dprev = dnext;
            dnext = t;
        }
    }
    if (RES) {
#pragma unroll
        for (int u = 0; u < kU; ++u)
            if (rc.c[u] < P.C) p[rc.c[u]] = rc.p[u];
    } else if (iters & 1u) {
        for (uint32_t ti = tw.begin; ti < tw.end; ti += step) p_update_batch<kU>(P, last_dn, p, last_alpha, tw, ti, tile_id);
    }
    rs_init_out = rs_init;
    rs_final = rs_old;
    return true;
}

// THREADS: workgroup size; EXCH: the reduction after phase A (dn.q, which
// phase B consumes with data only its own lanes wrote) uses grid_exchange;
// DB: the form of the reductions that order memory (use_tagged).
template <int THREADS, bool EXCH, int DB>
__global__ __launch_bounds__(THREADS) void k_cg_persistent(StepParams P, const uint8_t* __restrict__ ct,
                                                               const float* __restrict__ b, float* __restrict__ p,
                                                               float* __restrict__ r, float* __restrict__ d0,
                                                               float* __restrict__ d1, float* __restrict__ q,
                                                               CgWork w) {
    __shared__ uint32_t s_tiles[kLdsTiles];
    (void)ct;
    const uint32_t nt = *w.n_tiles;
    // At most 7/8 of one batch of kU tiles per wave: on the 128^3 tank 240
    // workgroups instead of 208, 9.35 instead of 9.67 us per iteration (less
    // phase-A work per CU outweighs the barrier's added arrivals).
    const uint32_t per_max = THREADS / 64 * kU;
    const uint32_t per = w.tiles_per_block && w.tiles_per_block < per_max ? w.tiles_per_block
                                                                          : per_max - THREADS / 128;
    const uint32_t G = solve_blocks(nt, per);
    if (blockIdx.x >= G) return;
    const uint32_t lb = xcd_remap(blockIdx.x, G);
    const TileWalk tw = tile_walk(nt, lb, G, THREADS / 64);
    // This block's slice of the tile list, kept in LDS when it fits.
    const uint32_t s0 = (uint32_t)(((uint64_t)nt * lb) / G);
    const uint32_t s1 = (uint32_t)(((uint64_t)nt * (lb + 1u)) / G);
    const bool in_lds = s1 - s0 <= (uint32_t)kLdsTiles;
    if (in_lds)
        for (uint32_t i = s0 + threadIdx.x; i < s1; i += THREADS) s_tiles[i - s0] = w.tiles[i];
    __syncthreads();
    const uint32_t* gt = w.tiles;
    auto tile_id = [&](uint32_t t) -> uint32_t { return in_lds ? s_tiles[t - s0] : gt[t]; };
    const uint32_t step = tw.stride * kU;
    // A block whose slice gives every wave at most one batch of kU tiles keeps
    // its cells in registers; the others stream the vectors.
    const bool resident = s1 - s0 <= (uint32_t)(THREADS / 64 * kU);
    // The barrier form depends on DB and G only, so every block takes the
    // same one; resident and streaming blocks share r and d the same way.
    const bool coherent = DB == kDbCoherent || (DB == kDbAuto && G > kTaggedMaxBlocks);
    double rs_init = 0.0, rs_old = 0.0;
    uint32_t iters = 0;
    bool ok;
    if (coherent && resident)
        ok = cg_two_phase<THREADS, EXCH, DB, true, true>(P, b, p, r, d0, d1, q, w, G, tw, tile_id, rs_init, rs_old,
                                                         iters);
    else if (coherent)
        ok = cg_two_phase<THREADS, EXCH, DB, false, true>(P, b, p, r, d0, d1, q, w, G, tw, tile_id, rs_init, rs_old,
                                                          iters);
    else if (resident)
        ok = cg_two_phase<THREADS, EXCH, DB, true, false>(P, b, p, r, d0, d1, q, w, G, tw, tile_id, rs_init, rs_old,
                                                          iters);
    else
        ok = cg_two_phase<THREADS, EXCH, DB, false, false>(P, b, p, r, d0, d1, q, w, G, tw, tile_id, rs_init,
                                                           rs_old, iters);
    if (!ok) {
        if (blockIdx.x == 0 && threadIdx.x == 0) cg_mark_failed(w.state);
        return;
    }
    // Resident blocks store p after the loop and streaming ones may finish a
    // pending update there: one more barrier publishes p before the true
    // residual reads neighbours' p.
    double unused;
    const bool okp = use_tagged<DB>(G) ? grid_sum_tagged<THREADS>(0.0, w, G, iters + 2u, unused)
                                       : grid_sum<THREADS>(0.0, w, G, unused);
    if (!okp) {
        if (blockIdx.x == 0 && threadIdx.x == 0) cg_mark_failed(w.state);
        return;
    }
    double acc = 0.0;
    for (uint32_t ti = tw.begin; ti < tw.end; ti += step) acc += true_res_batch<kU>(P, w.nmask, b, p, tw, ti, tile_id);
    double tr;
    const bool ok2 = EXCH ? grid_exchange<THREADS>(acc, w, G, iters + 1u, tr) : grid_sum<THREADS>(acc, w, G, tr);
    if (!ok2) {
        if (blockIdx.x == 0 && threadIdx.x == 0) cg_mark_failed(w.state);
        return;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        w.state->rs_init = rs_init;
        w.state->rs_old = rs_old;
        w.state->true_res_sq = tr;
        w.state->iterations = iters;
        w.state->active = 0u;
        w.state->error = 0u;
    }
}

// ---- single-reduction form ----
// The same CG as k_cg_persistent in the Chronopoulos-Gear arrangement (one
// sweep and one grid barrier per iteration, fs_cg.h).  In exact arithmetic
// its iterates are those of the two-phase form; with f32 vectors they differ
// by rounding.  Scalars per iteration k (f64, alpha/beta rounded to f32 as
// src/sim/pressure.rs:282,296 round theirs):
//     beta_k  = gamma_k / gamma_{k-1}                           (k > 0)
//     alpha_k = gamma_k / (delta_k - beta_k gamma_k / alpha_{k-1})
//     alpha_0 = gamma_0 / delta_0
// with gamma = r.r and delta = (A r).r; the stop test gamma_{k+1} <=
// tol^2 gamma_0 is that of src/sim/pressure.rs:262-292.
template <int THREADS, int U>
__global__ __launch_bounds__(THREADS) void k_cg_single(StepParams P, CgVectors v, CgWork w) {
    __shared__ uint32_t s_tiles[kLdsTiles];
    const uint32_t nt = *w.n_tiles;
    const uint32_t G = solve_blocks(nt, THREADS / 64 * kU);
    if (blockIdx.x >= G) return;
    const uint32_t lb = xcd_remap(blockIdx.x, G);
    const TileWalk tw = tile_walk(nt, lb, G, THREADS / 64);
    const uint32_t s0 = (uint32_t)(((uint64_t)nt * lb) / G);
    const uint32_t s1 = (uint32_t)(((uint64_t)nt * (lb + 1u)) / G);
    const bool in_lds = s1 - s0 <= (uint32_t)kLdsTiles;
    if (in_lds)
        for (uint32_t i = s0 + threadIdx.x; i < s1; i += THREADS) s_tiles[i - s0] = w.tiles[i];
    __syncthreads();
    const uint32_t* gt = w.tiles;
    auto tile_id = [&](uint32_t t) -> uint32_t { return in_lds ? s_tiles[t - s0] : gt[t]; };
    const uint32_t step = tw.stride * U;

    double part[2] = {0.0, 0.0}, tot[2];
    for (uint32_t ti = tw.begin; ti < tw.end; ti += step)
        cgs_init_batch<U>(P, w.nmask, v.r[0], v.w[0], tw, ti, tile_id, part[0], part[1]);
    if (!grid_sum_nv<THREADS, 2>(part, w, G, tot)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) cg_mark_failed(w.state);
        return;
    }
    const double rs_init = tot[0];
    const double tol = (double)P.tol;
    const double tol_sq = tol * tol;
    double gam = tot[0], del = tot[1], gam_prev = 0.0;
    float alpha_prev = 0.0f;
    uint32_t iters = 0;
    for (uint32_t it = 0; it < P.iters; ++it) {
        ++iters;
        const bool first = it == 0;
        float alpha, beta;
        if (first) {
            beta = 0.0f;
            alpha = fabs(del) > 1e-30 ? (float)(gam / del) : 0.0f;
        } else {
            beta = gam_prev > 1e-30 ? (float)(gam / gam_prev) : 0.0f;
            const double den = alpha_prev != 0.0f ? del - (double)beta * gam / (double)alpha_prev : del;
            alpha = fabs(den) > 1e-30 ? (float)(gam / den) : 0.0f;
        }
        const uint32_t cur = it & 1u, nxt = cur ^ 1u;
        const CgSweepIo io{v.r[cur], v.w[cur], v.s[nxt], v.r[nxt], v.w[nxt], v.s[cur], v.pd, v.x};
        part[0] = 0.0;
        part[1] = 0.0;
        for (uint32_t ti = tw.begin; ti < tw.end; ti += step)
            cgs_sweep_batch<U>(P, w.nmask, io, alpha, beta, first, tw, ti, tile_id, part[0], part[1]);
        if (!grid_sum_nv<THREADS, 2>(part, w, G, tot)) {
            if (blockIdx.x == 0 && threadIdx.x == 0) cg_mark_failed(w.state);
            return;
        }
        gam_prev = gam;
        alpha_prev = alpha;
        gam = tot[0];
        del = tot[1];
        if (tol > 0.0 && gam <= tol_sq * rs_init) break;
    }
    double acc = 0.0;
    for (uint32_t ti = tw.begin; ti < tw.end; ti += step) acc += true_res_batch<U>(P, w.nmask, v.b, v.x, tw, ti, tile_id);
    double tr;
    if (!grid_exchange<THREADS>(acc, w, G, 1u, tr)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) cg_mark_failed(w.state);
        return;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        w.state->rs_init = rs_init;
        w.state->rs_old = gam;
        w.state->true_res_sq = tr;
        w.state->iterations = iters;
        w.state->active = 0u;
        w.state->error = 0u;
    }
}

namespace {
// Persistent-kernel variant: FLUIDSTEP_CG_FORM=twophase (default) or
// single; workgroup size 1024 (default) or 512 (FLUIDSTEP_CG_THREADS);
// for the two-phase form FLUIDSTEP_CG_EXCHANGE=0 keeps grid_sum after
// phase A, and FLUIDSTEP_CG_DATA_BARRIER=counter|tagged|coherent fixes the
// form of the memory-ordering reductions (default kDbAuto: by workgroup
// count and residency).
struct PersistentVariant {
    int threads;
    bool single;
    bool exch;
    int db;
    const void* fn;
};

template <int T, bool X>
const void* two_phase_fn_x(int db) {
    if (db == kDbCounter) return (const void*)k_cg_persistent<T, X, kDbCounter>;
    if (db == kDbTagged) return (const void*)k_cg_persistent<T, X, kDbTagged>;
    if (db == kDbCoherent) return (const void*)k_cg_persistent<T, X, kDbCoherent>;
    return (const void*)k_cg_persistent<T, X, kDbAuto>;
}
template <int T>
const void* two_phase_fn(bool exch, int db) {
    return exch ? two_phase_fn_x<T, true>(db) : two_phase_fn_x<T, false>(db);
}
// Workgroup size, form and exchange are fixed per process (the sync area is
// sized from them); the data-barrier form is read at every launch, so one
// process can run the same solve under each form.
PersistentVariant persistent_variant() {
    static const PersistentVariant base = [] {
        PersistentVariant r{1024, false, true, kDbAuto, nullptr};
        if (const char* e = std::getenv("FLUIDSTEP_CG_THREADS")) r.threads = std::atoi(e) == 512 ? 512 : 1024;
        if (const char* e = std::getenv("FLUIDSTEP_CG_FORM")) r.single = std::strcmp(e, "single") == 0;
        if (const char* e = std::getenv("FLUIDSTEP_CG_EXCHANGE")) r.exch = std::atoi(e) != 0;
        return r;
    }();
    PersistentVariant r = base;
    if (const char* e = std::getenv("FLUIDSTEP_CG_DATA_BARRIER"))
        r.db = std::strcmp(e, "tagged") == 0     ? kDbTagged
               : std::strcmp(e, "counter") == 0  ? kDbCounter
               : std::strcmp(e, "coherent") == 0 ? kDbCoherent
                                                 : kDbAuto;
    if (r.single)
        r.fn = r.threads == 1024 ? (const void*)k_cg_single<1024, 2> : (const void*)k_cg_single<512, 2>;
    else
        r.fn = r.threads == 1024 ? two_phase_fn<1024>(r.exch, r.db) : two_phase_fn<512>(r.exch, r.db);
    return r;
}
}  // namespace

int cg_persistent_threads() { return persistent_variant().threads; }

int cg_persistent_max_blocks() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    const PersistentVariant v = persistent_variant();
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, v.fn, v.threads, 0) != hipSuccess) return 0;
    return cus * per_cu;
}

hipError_t launch_cg_persistent(const StepParams& P, const CgVectors& v, const CgWork& w_in, int grid_blocks,
                                hipStream_t s) {
    const PersistentVariant pv = persistent_variant();
    // FLUIDSTEP_CG_SPIN_TICKS shortens the bounded waits (read at every
    // launch): a test forces the solver-failure path with it, every block
    // still leaving the kernel through the abort flag.
    CgWork w = w_in;
    // FLUIDSTEP_CG_TILES_PER_BLOCK lowers the tiles a workgroup takes, so more
    // workgroups take part (read at every launch).
    if (const char* env = std::getenv("FLUIDSTEP_CG_TILES_PER_BLOCK")) w.tiles_per_block = (uint32_t)std::atoi(env);
    if (const char* env = std::getenv("FLUIDSTEP_CG_SPIN_TICKS")) {
        const long long v_ = std::atoll(env);
        if (v_ > 0 && v_ < w.spin_ticks) w.spin_ticks = v_;
    }
    if (pv.single) {
        void* args[] = {(void*)&P, (void*)&v, (void*)&w};
        ++g_dispatches;
        return hipLaunchCooperativeKernel(pv.fn, dim3(grid_blocks), dim3(pv.threads), args, 0, s);
    }
    const uint8_t* ct = nullptr;
    const float* b = v.b;
    float *p = v.x, *r = v.r[0], *d0 = v.pd, *d1 = v.s[0], *q = v.w[0];
    void* args[] = {(void*)&P, (void*)&ct, (void*)&b, (void*)&p, (void*)&r,
                    (void*)&d0, (void*)&d1, (void*)&q, (void*)&w};
    ++g_dispatches;
    return hipLaunchCooperativeKernel(pv.fn, dim3(grid_blocks), dim3(pv.threads), args, 0, s);
}

// ---- multi-kernel form ----
__global__ __launch_bounds__(kCgMultiThreads) void k_cg_rs0(StepParams P, const float* __restrict__ r, CgWork w) {
    const uint32_t G = gridDim.x;
    const TileWalk tw = tile_walk(*w.n_tiles, xcd_remap(blockIdx.x, G), G, kCgMultiThreads / 64);
    const uint32_t* gt = w.tiles;
    auto tile_id = [&](uint32_t t) -> uint32_t { return gt[t]; };
    double acc = 0.0;
    for (uint32_t ti = tw.begin; ti < tw.end; ti += tw.stride * kU) acc += rr_batch<kU>(P, r, tw, ti, tile_id);
    double tot;
    if (last_block_sum<kCgMultiThreads>(acc, w, tot) && threadIdx.x == 0) {
        CgState* st = w.state;
        st->rs_old = tot;
        st->rs_init = tot;
        st->alpha = 0.0f;
        st->beta = 0.0f;
        st->iterations = 0u;
        st->active = 1u;
        st->error = 0u;
    }
}

__global__ __launch_bounds__(kCgMultiThreads) void k_cg_phase_a(StepParams P, const float* __restrict__ r,
                                                                 const float* __restrict__ dprev,
                                                                 float* __restrict__ dnext, float* __restrict__ q,
                                                                 int first, CgWork w) {
    CgState* st = w.state;
    if (st->active == 0u) return;
    const float beta = st->beta;
    const uint32_t G = gridDim.x;
    const TileWalk tw = tile_walk(*w.n_tiles, xcd_remap(blockIdx.x, G), G, kCgMultiThreads / 64);
    const uint32_t* gt = w.tiles;
    auto tile_id = [&](uint32_t t) -> uint32_t { return gt[t]; };
    double acc = 0.0;
    for (uint32_t ti = tw.begin; ti < tw.end; ti += tw.stride * kU)
        acc += phase_a_batch<kU>(P, w.nmask, r, dprev, dnext, q, beta, first != 0, tw, ti, tile_id);
    double dq;
    if (last_block_sum<kCgMultiThreads>(acc, w, dq) && threadIdx.x == 0) {
        st->alpha = fabs(dq) > 1e-30 ? (float)(st->rs_old / dq) : 0.0f;
        st->iterations += 1u;
    }
}

__global__ __launch_bounds__(kCgMultiThreads) void k_cg_phase_b(StepParams P, const float* __restrict__ dn,
                                                                 const float* __restrict__ q, float* __restrict__ p,
                                                                 float* __restrict__ r, CgWork w) {
    CgState* st = w.state;
    if (st->active == 0u) return;
    const float alpha = st->alpha;
    const uint32_t G = gridDim.x;
    const TileWalk tw = tile_walk(*w.n_tiles, xcd_remap(blockIdx.x, G), G, kCgMultiThreads / 64);
    const uint32_t* gt = w.tiles;
    auto tile_id = [&](uint32_t t) -> uint32_t { return gt[t]; };
    double acc = 0.0;
    for (uint32_t ti = tw.begin; ti < tw.end; ti += tw.stride * kU)
        acc += phase_b_batch<kU, kPOne>(P, dn, q, p, r, alpha, nullptr, 0.0f, tw, ti, tile_id);
    double rr;
    if (last_block_sum<kCgMultiThreads>(acc, w, rr) && threadIdx.x == 0) {
        const double rs_old = st->rs_old;
        st->beta = rs_old > 1e-30 ? (float)(rr / rs_old) : 0.0f;
        st->rs_old = rr;
        const double tol = (double)P.tol;
        if (tol > 0.0 && rr <= tol * tol * st->rs_init) st->active = 0u;
    }
}

__global__ __launch_bounds__(kCgMultiThreads) void k_cg_true_res(StepParams P, const float* __restrict__ b,
                                                                  const float* __restrict__ p, CgWork w) {
    const uint32_t G = gridDim.x;
    const TileWalk tw = tile_walk(*w.n_tiles, xcd_remap(blockIdx.x, G), G, kCgMultiThreads / 64);
    const uint32_t* gt = w.tiles;
    auto tile_id = [&](uint32_t t) -> uint32_t { return gt[t]; };
    double acc = 0.0;
    for (uint32_t ti = tw.begin; ti < tw.end; ti += tw.stride * kU)
        acc += true_res_batch<kU>(P, w.nmask, b, p, tw, ti, tile_id);
    double tr;
    if (last_block_sum<kCgMultiThreads>(acc, w, tr) && threadIdx.x == 0) {
        w.state->true_res_sq = tr;
        w.state->active = 0u;
    }
}

hipError_t launch_cg_multi(const StepParams& P, const uint8_t* ct, const float* b, float* p, float* r, float* d0,
                           float* d1, float* q, const CgWork& w, int grid_blocks, hipStream_t s) {
    (void)ct;
    const dim3 g(grid_blocks), t(kCgMultiThreads);
    FS_LAUNCH(k_cg_rs0, g, t, 0, s, P, r, w);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    float* D[2] = {d0, d1};
    for (uint32_t it = 0; it < P.iters; ++it) {
        const int first = it == 0;
        const float* dprev = first ? d0 : D[(it + 1u) & 1u];
        float* dnext = first ? d1 : D[it & 1u];
        const float* dn = first ? d0 : D[it & 1u];
        FS_LAUNCH(k_cg_phase_a, g, t, 0, s, P, r, dprev, dnext, q, first, w);
        FS_LAUNCH(k_cg_phase_b, g, t, 0, s, P, dn, q, p, r, w);
    }
    FS_LAUNCH(k_cg_true_res, g, t, 0, s, P, b, p, w);
    return hipGetLastError();
}

}  // namespace fs
