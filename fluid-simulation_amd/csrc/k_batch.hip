// k_batch.hip — the batched-tank step: every phase of the fused substep as one
// launch over all tanks of a batch.  Each workgroup takes one WorkItem of its
// phase's table (fs_batch.h), loads that tank's constants, offsets the
// concatenated arrays to the tank and runs the same per-particle, per-cell,
// per-face and per-brick device functions as the single-tank kernels
// (fs_particle.h, fs_cell.h, fs_face.h, fs_gather.h, fs_cg.h), so a tank of a
// batch computes what a single tank computes.
//
// The pressure solve is one cooperative launch.  Its workgroups are split into
// one group per tank (sized on the host by the tank's expected liquid volume);
// a group runs that tank's mixed-precision CG over the tank's liquid tiles
// and reduces over its own partial slots, while the grid barriers are shared
// by all tanks (lock step); a tank that has converged idles at the barriers.
#include <cstdlib>
#include <cstring>
#include "fs_batch.h"
#include "fs_cell.h"
#include "fs_gather.h"

namespace fs {

namespace {
constexpr int kBCgThreads = 1024;  // CG workgroup (16 waves), one per CU
constexpr int kBLdsTiles = 2048;   // tile ids a CG workgroup keeps in LDS
}  // namespace

// ---------------------------------------------------------------------------
// Keys, histogram, misfit list (k_key_hist per tank; keys are global cells).
__global__ __launch_bounds__(kBatchBlock) void kb_key_hist(const BatchTank* __restrict__ T,
                                                           const WorkItem* __restrict__ wl, ParticleSoA sp,
                                                           uint32_t* __restrict__ keys, uint32_t* __restrict__ occ,
                                                           uint32_t* __restrict__ counters,
                                                           uint32_t* __restrict__ mis_count,
                                                           uint32_t* __restrict__ mis_list) {
    const WorkItem wi = wl[blockIdx.x];
    if (wi.tank == kNoTank) return;
    const StepParams P = T[wi.tank].P;
    const uint32_t part_off = T[wi.tank].part_off, cell_off = T[wi.tank].cell_off;
    const uint32_t local = wi.start + threadIdx.x;
    const bool valid = local < P.N;
    const uint32_t p = part_off + local;
    uint32_t key = 0xffffffffu;
    bool mis = false;
    if (valid) {
        key = cell_off + particle_key(P, sp.pos[0][p], sp.pos[1][p], sp.pos[2][p], mis);
        keys[p] = key;
    }
    list_misfits(mis, p, mis_count, mis_list, kMisfitCap, &counters[4 * wi.tank + 1]);
    hist_key_runs(key, valid, occ);
}

hipError_t launch_batch_key_hist(const BatchTank* T, const WorkItem* wl, uint32_t nwl, const ParticleSoA& sp,
                                 uint32_t* keys, uint32_t* occ, uint32_t* counters, uint32_t* mis_count,
                                 uint32_t* mis_list, hipStream_t s) {
    if (nwl == 0) return hipSuccess;
    hipLaunchKernelGGL(kb_key_hist, dim3(nwl), dim3(kBatchBlock), 0, s, T, wl, sp, keys, occ, counters, mis_count,
                       mis_list);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBatchBlock) void kb_classify(const BatchTank* __restrict__ T,
                                                           const WorkItem* __restrict__ wl,
                                                           const uint32_t* __restrict__ occ, uint8_t* __restrict__ ct,
                                                           uint32_t* __restrict__ counters) {
    const WorkItem wi = wl[blockIdx.x];
    if (wi.tank == kNoTank) return;
    const StepParams P = T[wi.tank].P;
    const uint32_t co = T[wi.tank].cell_off;
    const uint32_t c = wi.start + threadIdx.x;
    bool liquid = false;
    if (c < P.C) {
        const uint8_t t = classify_cell(P, occ + co, c);
        ct[co + c] = t;
        liquid = t == kLiquid;
    }
    __shared__ uint32_t s_cnt;
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    const uint64_t m = __ballot(liquid);
    if (lane_id() == 0 && m) atomicAdd(&s_cnt, (uint32_t)__popcll(m));
    __syncthreads();
    if (threadIdx.x == 0 && s_cnt) atomicAdd(&counters[4 * wi.tank], s_cnt);
}

hipError_t launch_batch_classify(const BatchTank* T, const WorkItem* wl, uint32_t nwl, const uint32_t* occ,
                                 uint8_t* ct, uint32_t* counters, hipStream_t s) {
    if (nwl == 0) return hipSuccess;
    hipLaunchKernelGGL(kb_classify, dim3(nwl), dim3(kBatchBlock), 0, s, T, wl, occ, ct, counters);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Gather: work item = (tank, brick index of the tank's brick grid).
template <int BF16, int V>
__global__ __launch_bounds__(kGThreads) void kb_gather(const BatchTank* __restrict__ T,
                                                        const WorkItem* __restrict__ wl, ParticleSoA sp,
                                                        ParticleSoA src, const uint32_t* __restrict__ occ,
                                                        const uint32_t* __restrict__ off,
                                                        const uint8_t* __restrict__ ct, BatchGrid g,
                                                        const uint32_t* __restrict__ mis_count,
                                                        const uint32_t* __restrict__ mis_list,
                                                        const uint32_t* __restrict__ perm) {
    const uint32_t mis_n = *mis_count;
    if (mis_n > kMisfitCap) return;  // the atomic form handles this substep
    const WorkItem wi = wl[blockIdx.x];
    if (wi.tank == kNoTank) return;
    const BatchTank& bt = T[wi.tank];
    const StepParams P = bt.P;
    const uint32_t co = bt.cell_off, part_off = bt.part_off;
    const uint32_t nbx = (P.nx + Brick<V>::X - 1) / Brick<V>::X, nby = (P.ny + Brick<V>::Y - 1) / Brick<V>::Y;
    const uint32_t b = wi.start;
    const uint32_t bxi = b % nbx, byi = (b / nbx) % nby, bzi = b / (nbx * nby);
    const GatherOut out = {{g.vel[0] + bt.face_off[0], g.vel[1] + bt.face_off[1], g.vel[2] + bt.face_off[2]},
                           {g.saved[0] + bt.face_off[0], g.saved[1] + bt.face_off[1], g.saved[2] + bt.face_off[2]}};
    p2g_gather_brick<BF16, V>(P, sp, src, occ + co, off + co, ct + co, out, (int32_t)bxi * Brick<V>::X,
                              (int32_t)byi * Brick<V>::Y, (int32_t)bzi * Brick<V>::Z, mis_list, mis_n, part_off,
                              part_off + P.N, perm);
}

namespace {
template <int V>
hipError_t launch_batch_gather_v(const BatchTank* T, const WorkItem* wl, uint32_t nwl, int bf16,
                                 const ParticleSoA& sorted, const ParticleSoA& src, const uint32_t* occ,
                                 const uint32_t* off, const uint8_t* ct, const BatchGrid& g,
                                 const uint32_t* mis_count, const uint32_t* mis_list, const uint32_t* perm,
                                 hipStream_t s) {
    if (bf16)
        hipLaunchKernelGGL((kb_gather<1, V>), dim3(nwl), dim3(kGThreads), 0, s, T, wl, sorted, src, occ, off, ct, g,
                           mis_count, mis_list, perm);
    else
        hipLaunchKernelGGL((kb_gather<0, V>), dim3(nwl), dim3(kGThreads), 0, s, T, wl, sorted, src, occ, off, ct, g,
                           mis_count, mis_list, perm);
    return hipGetLastError();
}
}  // namespace

hipError_t launch_batch_gather(const BatchTank* T, const WorkItem* wl, uint32_t nwl, int brick_variant, int bf16,
                               const ParticleSoA& sorted, const ParticleSoA& src, const uint32_t* occ,
                               const uint32_t* cell_offset, const uint8_t* ct, const BatchGrid& g,
                               const uint32_t* mis_count, const uint32_t* mis_list, hipStream_t s,
                               const uint32_t* perm) {
    if (nwl == 0) return hipSuccess;
    switch (brick_variant) {
        case 0: return launch_batch_gather_v<0>(T, wl, nwl, bf16, sorted, src, occ, cell_offset, ct, g, mis_count, mis_list,
                                                    perm, s);
        case 1: return launch_batch_gather_v<1>(T, wl, nwl, bf16, sorted, src, occ, cell_offset, ct, g, mis_count, mis_list,
                                                    perm, s);
        case 2: return launch_batch_gather_v<2>(T, wl, nwl, bf16, sorted, src, occ, cell_offset, ct, g, mis_count, mis_list,
                                                    perm, s);
        default: return hipErrorInvalidValue;
    }
}

// ---------------------------------------------------------------------------
// Fallback: atomic P2G and the face epilogue, when too many misfits.  Both
// are usually no-ops, so they walk their work tables with a small
// grid-stride grid.
constexpr uint32_t kFallbackGrid = 2048;

template <int BF16>
__global__ __launch_bounds__(kBatchBlock) void kb_p2g_atomic(const BatchTank* __restrict__ T,
                                                             const WorkItem* __restrict__ wl, uint32_t nwl,
                                                             ParticleSoA sp, BatchGrid g,
                                                             const uint32_t* __restrict__ mis_count) {
    if (*mis_count <= kMisfitCap) return;
    for (uint32_t w = blockIdx.x; w < nwl; w += gridDim.x) {
        const WorkItem wi = wl[w];
        if (wi.tank == kNoTank) continue;
        const BatchTank& bt = T[wi.tank];
        const StepParams P = bt.P;
        const uint32_t local = wi.start + threadIdx.x;
        if (local >= P.N) continue;
        int32_t* const nums[3] = {g.num[0] + bt.face_off[0], g.num[1] + bt.face_off[1], g.num[2] + bt.face_off[2]};
        int32_t* const dens[3] = {g.den[0] + bt.face_off[0], g.den[1] + bt.face_off[1], g.den[2] + bt.face_off[2]};
        p2g_atomic_particle<BF16>(P, sp, bt.part_off + local, nums, dens);
    }
}

__global__ __launch_bounds__(kBatchBlock) void kb_face_epilogue(const BatchTank* __restrict__ T,
                                                                const WorkItem* __restrict__ wl, uint32_t nwl,
                                                                const uint8_t* __restrict__ ct, BatchGrid g,
                                                                const uint32_t* __restrict__ mis_count) {
    if (*mis_count <= kMisfitCap) return;
    for (uint32_t w = blockIdx.x; w < nwl; w += gridDim.x) {
        const WorkItem wi = wl[w];
        if (wi.tank == kNoTank) continue;
        const BatchTank& bt = T[wi.tank];
        const StepParams P = bt.P;
        int a;
        uint32_t f;
        if (!split_face(P, wi.start + threadIdx.x, a, f)) continue;
        const uint32_t fo = bt.face_off[a];
        uint32_t d[3], i, j, k;
        face_dims(P, a, d);
        face_ijk(f, d, i, j, k);
        float vv, ss;
        face_epilogue(P, a, ct + bt.cell_off, i, j, k, g.num[a][fo + f], g.den[a][fo + f], vv, ss);
        g.vel[a][fo + f] = vv;
        g.saved[a][fo + f] = ss;
    }
}

hipError_t launch_batch_p2g_fallback(const BatchTank* T, const WorkItem* wl_part, uint32_t npart,
                                     const WorkItem* wl_face, uint32_t nface, int bf16, const ParticleSoA& sorted,
                                     const uint8_t* ct, const BatchGrid& g, const uint32_t* mis_count,
                                     hipStream_t s) {
    if (npart) {
        const uint32_t gp = npart < kFallbackGrid ? npart : kFallbackGrid;
        if (bf16)
            hipLaunchKernelGGL(kb_p2g_atomic<1>, dim3(gp), dim3(kBatchBlock), 0, s, T, wl_part, npart, sorted, g,
                               mis_count);
        else
            hipLaunchKernelGGL(kb_p2g_atomic<0>, dim3(gp), dim3(kBatchBlock), 0, s, T, wl_part, npart, sorted, g,
                               mis_count);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (nface) {
        const uint32_t gf = nface < kFallbackGrid ? nface : kFallbackGrid;
        hipLaunchKernelGGL(kb_face_epilogue, dim3(gf), dim3(kBatchBlock), 0, s, T, wl_face, nface, ct, g, mis_count);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBatchBlock) void kb_div_cginit(const BatchTank* __restrict__ T,
                                                             const WorkItem* __restrict__ wl, BatchGrid g,
                                                             const uint8_t* __restrict__ ct,
                                                             const uint32_t* __restrict__ occ, float* __restrict__ div,
                                                             float* __restrict__ b, float* __restrict__ p,
                                                             float* __restrict__ r, float* __restrict__ d,
                                                             BatchCgWork w) {
    const WorkItem wi = wl[blockIdx.x];
    if (wi.tank == kNoTank) return;
    const BatchTank& bt = T[wi.tank];
    const StepParams P = bt.P;
    const uint32_t co = bt.cell_off;
    const uint32_t c = wi.start + threadIdx.x;  // wi.start is a multiple of 256: waves are whole tiles
    bool liquid = false;
    if (c < P.C)
        liquid = div_cginit_cell(P, g.vel[0] + bt.face_off[0], g.vel[1] + bt.face_off[1], g.vel[2] + bt.face_off[2],
                                 ct + co, occ + co, c, div + co, b + co, p + co, r + co, d + co, w.nmask + co);
    const uint64_t m = __ballot(liquid);
    const uint32_t tile = c / kTile;
    if (lane_id() == 0 && tile * kTile < P.C) w.tile_flags[bt.tile_off + tile] = m ? 1u : 0u;
}

hipError_t launch_batch_div_cginit(const BatchTank* T, const WorkItem* wl, uint32_t nwl, const BatchGrid& g,
                                   const uint8_t* ct, const uint32_t* occ, float* div, float* b, float* p, float* r,
                                   float* d, const BatchCgWork& w, hipStream_t s) {
    if (nwl == 0) return hipSuccess;
    hipLaunchKernelGGL(kb_div_cginit, dim3(nwl), dim3(kBatchBlock), 0, s, T, wl, g, ct, occ, div, b, p, r, d, w);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Lock-step reductions of the batched solve.  Every workgroup of the launch
// meets every other at each reduction: the one after phase B (whose r the
// next phase A reads at neighbours) is a grid barrier (two-level arrival as
// in the single-tank solver: a counter per group of workgroups with equal
// blockIdx % 8, then one over the groups; release RMWs, one acquire for the
// last arrival, relaxed polling of the generation word by the others); the
// one after phase A is the counter-free exchange (lockstep_exchange).  Each
// workgroup publishes its f64 partial in its slot and whether its tank is
// still iterating; afterwards it sums the slots of its own tank,
// [base, base + nb), in slot order (so every workgroup of a tank holds the
// identical total) and ORs all flags.  The barriers of all tanks coincide, so
// the L2 invalidation that each acquire performs falls between phases rather
// than in the middle of other tanks' phases.  Partials and flags alternate
// between two halves by generation parity.
__device__ __forceinline__ bool arrive_grid(const BatchCgWork& w) {
    const uint32_t G = gridDim.x;
    const uint32_t grp = blockIdx.x & 7u;
    const uint32_t in_grp = (G - grp + 7u) / 8u;
    const uint32_t ngrp = G < 8u ? G : 8u;
    uint32_t* gc = w.arrive + grp * 32u;
    uint32_t* top = w.arrive + 8u * 32u;
    const uint32_t prev = __hip_atomic_fetch_add(gc, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    if (prev != in_grp - 1u) return false;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    __hip_atomic_store(gc, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t prev2 = __hip_atomic_fetch_add(top, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    if (prev2 != ngrp - 1u) return false;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    __hip_atomic_store(top, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return true;
}

// Per-tank total in one fixed order over the slots [base, base + nb): thread
// i adds slot i, i + THREADS, ... when it lies in the range, then the block
// reduces (reduce_loaded), so every workgroup of a tank holds the identical
// total.
template <int THREADS>
__device__ __forceinline__ double tank_total(double loaded) {
    const double in[1] = {loaded};
    double o[1];
    reduce_loaded<THREADS, 1>(in, o);
    return o[0];
}

template <int THREADS>
__device__ bool lockstep_sum(double v, uint32_t flag, const BatchCgWork& w, uint32_t slot, uint32_t base, uint32_t nb,
                             double& out, bool& any) {
    __shared__ int s_last, s_abort;
    __shared__ uint32_t s_gen;
    const double in[1] = {v};
    double bs[1];
    block_sum_nv<THREADS, 1>(in, bs);
    const uint32_t G = gridDim.x;
    if (threadIdx.x == 0) {
        const uint32_t g0 = ld_relaxed(w.gen);
        const size_t half = (size_t)(g0 & 1u) * G;
        w.partials[half + slot] = bs[0];
        w.flags[half + slot] = flag;
        s_gen = g0;
        s_abort = 0;
        s_last = arrive_grid(w) ? 1 : 0;
    }
    __syncthreads();
    if (s_last) {
        if (threadIdx.x == 0) __hip_atomic_store(w.gen, s_gen + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else if (threadIdx.x == 0) {
        const long long t0 = wall_clock64();
        while (ld_relaxed(w.gen) == s_gen) {
            if (ld_relaxed(w.abort_flag)) {
                s_abort = 1;
                break;
            }
            if (wall_clock64() - t0 > kSpinTimeoutTicks) {
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
    const size_t half = (size_t)(s_gen & 1u) * G;
    uint32_t f = 0;
    double acc = 0.0;
    for (uint32_t i = threadIdx.x; i < G; i += THREADS) {
        f |= __hip_atomic_load(&w.flags[half + i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (i >= base && i < base + nb)
            acc += __hip_atomic_load(&w.partials[half + i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    any = __syncthreads_or(f != 0u) != 0;
    out = tank_total<THREADS>(acc);
    return true;
}

// The reduction after phase A without counters or fences (phase B then reads
// only data its own lanes wrote): every workgroup publishes its partial and
// flag in its epoch-tagged slot (fs_cg.h) in `slots` (2 x G slots) and reads
// all G slots; the tank total and the OR of the flags are formed as in
// lockstep_sum.  Slots alternate between two halves by epoch parity (a
// workgroup publishes e+2 only after all have published e+1, i.e. finished
// reading e).
template <int THREADS>
__device__ bool lockstep_exchange(double v, uint32_t flag, const BatchCgWork& w, uint64_t* slots, uint32_t slot,
                                  uint32_t base, uint32_t nb, uint32_t epoch, double& out, bool& any) {
    __shared__ int s_abort;
    const uint32_t G = gridDim.x;
    uint64_t* half = slots + (size_t)(epoch & 1u) * G * 2u;
    const double in[1] = {v};
    double bs[1];
    block_sum_nv<THREADS, 1>(in, bs);
    if (threadIdx.x == 0) {
        s_abort = 0;
        publish_tagged(half, slot, epoch, bs[0], flag);
    }
    __syncthreads();
    uint32_t f = 0;
    double acc = 0.0;
    for (uint32_t i = threadIdx.x; i < G; i += THREADS) {
        double x;
        uint32_t fi;
        if (!poll_tagged(half, i, epoch, w.abort_flag, x, fi)) {
            s_abort = 1;
            break;
        }
        f |= fi;
        if (i >= base && i < base + nb) acc += x;
    }
    any = __syncthreads_or(f != 0u) != 0;
    const double tot = tank_total<THREADS>(acc);
    if (s_abort) return false;
    out = tot;
    return true;
}

// The fence-free grid barrier after phase B (coherent_sum of the single-tank
// solver): the only data that cross workgroups in the loop are r and d,
// written with store_through and read at neighbours with ld_coherent_at
// (phase_a_batch / phase_b_batch with COH).  Each wave waits until its
// stores are acknowledged (s_waitcnt vmcnt(0); an acknowledged sc1 store is
// visible at agent scope) before the workgroup barrier inside the exchange,
// after which thread 0 publishes; a workgroup reads neighbours only after it
// has seen every slot's tag.  Its slots follow the exchange's in `slots`,
// with their own epochs.
template <int THREADS>
__device__ __forceinline__ bool lockstep_coherent(double v, uint32_t flag, const BatchCgWork& w, uint32_t slot,
                                                  uint32_t base, uint32_t nb, uint32_t epoch, double& out,
                                                  bool& any) {
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
    return lockstep_exchange<THREADS>(v, flag, w, w.slots + 4u * gridDim.x, slot, base, nb, epoch, out, any);
}

// Workgroups of a tank take its slice of the liquid-tile list; workgroups
// whose tank has converged (or padding workgroups, which have none) keep
// meeting the barriers with zero partials until no tank iterates.  COH takes
// lockstep_coherent after phase B instead of lockstep_sum.
template <int U, bool COH>
__global__ __launch_bounds__(kBCgThreads) void kb_cg(const BatchTank* __restrict__ T, const WorkItem* __restrict__ wl,
                                                      const float* __restrict__ b, float* __restrict__ p,
                                                      float* __restrict__ r, float* __restrict__ d0,
                                                      float* __restrict__ d1, float* __restrict__ q, BatchCgWork w) {
    __shared__ uint32_t s_tiles[kBLdsTiles];
    const WorkItem wi = wl[blockIdx.x];
    const bool has_tank = wi.tank != kNoTank;
    const uint32_t tank = has_tank ? wi.tank : 0u;
    const BatchTank& bt = T[tank];
    const StepParams P = bt.P;
    const uint32_t co = bt.cell_off, tile_off = bt.tile_off;
    const uint32_t G = has_tank ? bt.cg_nb : 1u, rank = has_tank ? wi.start : 0u;
    const uint32_t base = has_tank ? bt.cg_base : 0u, nb = has_tank ? bt.cg_nb : 0u;
    const uint32_t slot = has_tank ? base + rank : wi.start;  // padding items carry their own slot
    // The tank's liquid tiles are the slice [tb, tb + nt) of the compacted list.
    uint32_t tb = 0, nt = 0;
    if (has_tank) {
        tb = w.tile_scan[tile_off];
        nt = w.tile_scan[tile_off + bt.ntiles] - tb;
    }
    const TileWalk tw = tile_walk(nt, rank, G, kBCgThreads / 64);
    const uint32_t s0 = (uint32_t)(((uint64_t)nt * rank) / G);
    const uint32_t s1 = (uint32_t)(((uint64_t)nt * (rank + 1u)) / G);
    const bool in_lds = s1 - s0 <= (uint32_t)kBLdsTiles;
    if (in_lds)
        for (uint32_t i = s0 + threadIdx.x; i < s1; i += kBCgThreads) s_tiles[i - s0] = w.tiles[tb + i] - tile_off;
    __syncthreads();
    const uint32_t* gt = w.tiles + tb;
    auto tile_id = [&](uint32_t t) -> uint32_t { return in_lds ? s_tiles[t - s0] : gt[t] - tile_off; };
    const uint32_t step = tw.stride * U;
    const uint16_t* nm = w.nmask + co;
    const float* bb = b + co;
    float* pp = p + co;
    float* rr_ = r + co;
    float* q_ = q + co;

    auto report = [&](double rs_init, double rs_final, double tr, uint32_t iters, uint32_t err) {
        if (has_tank && rank == 0 && threadIdx.x == 0) {
            BatchCgState* st = w.state + tank;
            st->rs_init = rs_init;
            st->rs_final = rs_final;
            st->true_res_sq = tr;
            st->iterations = iters;
            st->error = err;
            if (err) st->failed = 1u;
        }
    };

    double acc = 0.0;
    for (uint32_t ti = tw.begin; ti < tw.end; ti += step) acc += rr_batch<U>(P, rr_, tw, ti, tile_id);
    double rs0;
    bool any = false;
    if (!lockstep_sum<kBCgThreads>(acc, has_tank ? 1u : 0u, w, slot, base, nb, rs0, any)) {
        report(0.0, 0.0, 0.0, 0u, 1u);
        return;
    }
    double rs_old = rs0;
    const double rs_init = rs0;
    const double tol = (double)P.tol;
    const double tol_sq = tol * tol;
    float beta = 0.0f;
    float* dprev = d0 + co;
    float* dnext = d1 + co;
    uint32_t iters = 0;
    bool active = has_tank;
    // p takes the updates of iterations k-1 and k together at odd k
    // (phase_b_batch kPTwo); a tank stopping after an even k applies the
    // pending one then.  A tank's iteration index k is `it` while it is
    // active (all tanks start together).
    float alpha_prev = 0.0f;
    for (uint32_t it = 0;; ++it) {
        const bool first = it == 0;
        float* dn_buf = first ? dprev : dnext;
        acc = 0.0;
        if (active) {
            ++iters;
            for (uint32_t ti = tw.begin; ti < tw.end; ti += step)
                acc += phase_a_batch<U, COH>(P, nm, rr_, dprev, dnext, q_, beta, first, tw, ti, tile_id);
        }
        double dq;
        if (!lockstep_exchange<kBCgThreads>(acc, active ? 1u : 0u, w, w.slots, slot, base, nb, it + 1u, dq, any)) {
            report(rs_init, rs_old, 0.0, iters, 1u);
            return;
        }
        if (!any) break;  // every tank has finished (the same flags are seen by all)
        const float alpha = fabs(dq) > 1e-30 ? (float)(rs_old / dq) : 0.0f;
        acc = 0.0;
        if (active && (it & 1u))
            for (uint32_t ti = tw.begin; ti < tw.end; ti += step)
                acc += phase_b_batch<U, kPTwo, COH>(P, dn_buf, q_, pp, rr_, alpha, dprev, alpha_prev, tw, ti,
                                                    tile_id);
        else if (active)
            for (uint32_t ti = tw.begin; ti < tw.end; ti += step)
                acc += phase_b_batch<U, kPNone, COH>(P, dn_buf, q_, pp, rr_, alpha, nullptr, 0.0f, tw, ti, tile_id);
        double rr;
        const bool okb = COH ? lockstep_coherent<kBCgThreads>(acc, active ? 1u : 0u, w, slot, base, nb, it + 1u, rr, any)
                             : lockstep_sum<kBCgThreads>(acc, active ? 1u : 0u, w, slot, base, nb, rr, any);
        if (!okb) {
            report(rs_init, rs_old, 0.0, iters, 1u);
            return;
        }
        if (active) {
            beta = rs_old > 1e-30 ? (float)(rr / rs_old) : 0.0f;
            rs_old = rr;
            alpha_prev = alpha;
            if ((tol > 0.0 && rr <= tol_sq * rs_init) || iters >= P.iters) {
                active = false;
                if (!(it & 1u))
                    for (uint32_t ti = tw.begin; ti < tw.end; ti += step)
                        p_update_batch<U>(P, dn_buf, pp, alpha, tw, ti, tile_id);
            } else if (!first) {
                float* t = dprev;
                dprev = dnext;
                dnext = t;
            }
        }
    }
    // A pending p update of a stopping tank comes after its last barrier:
    // one more barrier publishes p before the true residual reads neighbours.
    double unused;
    if (!lockstep_sum<kBCgThreads>(0.0, 0u, w, slot, base, nb, unused, any)) {
        report(rs_init, rs_old, 0.0, iters, 1u);
        return;
    }
    acc = 0.0;
    for (uint32_t ti = tw.begin; ti < tw.end; ti += step) acc += true_res_batch<U>(P, nm, bb, pp, tw, ti, tile_id);
    double tr;
    if (!lockstep_sum<kBCgThreads>(acc, 0u, w, slot, base, nb, tr, any)) {
        report(rs_init, rs_old, 0.0, iters, 1u);
        return;
    }
    report(rs_init, rs_old, tr, iters, 0u);
}

int batch_cg_max_blocks() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    int lo = 1 << 30;
    // The smallest occupancy over the unroll variants, so every variant fits.
    const void* fns[6] = {(const void*)kb_cg<4, true>,  (const void*)kb_cg<2, true>,  (const void*)kb_cg<1, true>,
                          (const void*)kb_cg<4, false>, (const void*)kb_cg<2, false>, (const void*)kb_cg<1, false>};
    for (const void* f : fns) {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, kBCgThreads, 0) != hipSuccess) return 0;
        lo = per_cu < lo ? per_cu : lo;
    }
    return cus * lo;
}

hipError_t launch_batch_cg(const BatchTank* T, const WorkItem* wl, uint32_t nwl, int cg_unroll, const float* b,
                           float* p, float* r, float* d0, float* d1, float* q, const BatchCgWork& w, hipStream_t s) {
    if (nwl == 0) return hipSuccess;
    void* args[] = {(void*)&T, (void*)&wl, (void*)&b, (void*)&p, (void*)&r,
                    (void*)&d0, (void*)&d1, (void*)&q, (void*)&w};
    // FLUIDSTEP_BATCH_DATA_BARRIER=counter takes lockstep_sum after phase B
    // (read at every launch); the default is lockstep_coherent.
    const char* e = std::getenv("FLUIDSTEP_BATCH_DATA_BARRIER");
    const bool coh = !(e && std::strcmp(e, "counter") == 0);
    const void* f = coh ? (cg_unroll == 1   ? (const void*)kb_cg<2, true>
                           : cg_unroll == 2 ? (const void*)kb_cg<1, true>
                                            : (const void*)kb_cg<4, true>)
                        : (cg_unroll == 1   ? (const void*)kb_cg<2, false>
                           : cg_unroll == 2 ? (const void*)kb_cg<1, false>
                                            : (const void*)kb_cg<4, false>);
    return hipLaunchCooperativeKernel(f, dim3(nwl), dim3(kBCgThreads), args, 0, s);
}

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBatchBlock) void kb_gradient_enforce(const BatchTank* __restrict__ T,
                                                                   const WorkItem* __restrict__ wl,
                                                                   const float* __restrict__ pr,
                                                                   const uint8_t* __restrict__ ct, BatchGrid g) {
    const WorkItem wi = wl[blockIdx.x];
    if (wi.tank == kNoTank) return;
    const BatchTank& bt = T[wi.tank];
    const StepParams P = bt.P;
    int a;
    uint32_t f;
    if (!split_face(P, wi.start + threadIdx.x, a, f)) return;
    float* vel = g.vel[a] + bt.face_off[a];
    const uint8_t* ctt = ct + bt.cell_off;
    uint32_t d[3], i, j, k;
    face_dims(P, a, d);
    face_ijk(f, d, i, j, k);
    float v = face_gradient(P, a, pr + bt.cell_off, ctt, i, j, k, vel[f]);
    if (face_solid(P, a, ctt, i, j, k)) v = 0.0f;
    vel[f] = v;
}

hipError_t launch_batch_gradient_enforce(const BatchTank* T, const WorkItem* wl, uint32_t nwl, const float* p,
                                         const uint8_t* ct, const BatchGrid& g, hipStream_t s) {
    if (nwl == 0) return hipSuccess;
    hipLaunchKernelGGL(kb_gradient_enforce, dim3(nwl), dim3(kBatchBlock), 0, s, T, wl, p, ct, g);
    return hipGetLastError();
}

template <int BF16>
__global__ __launch_bounds__(kBatchBlock) void kb_g2p(const BatchTank* __restrict__ T,
                                                      const WorkItem* __restrict__ wl, ParticleSoA sp, ParticleSoA rd,
                                                      const uint32_t* __restrict__ perm, BatchGrid g) {
    const WorkItem wi = wl[blockIdx.x];
    if (wi.tank == kNoTank) return;
    const BatchTank& bt = T[wi.tank];
    const StepParams P = bt.P;
    const uint32_t local = wi.start + threadIdx.x;
    if (local >= P.N) return;
    const uint32_t* fo = bt.face_off;
    const uint32_t idx = bt.part_off + local;
    g2p_particle_from<BF16>(P, rd, perm ? perm[idx] : idx, sp, idx, g.vel[0] + fo[0], g.vel[1] + fo[1],
                            g.vel[2] + fo[2], g.saved[0] + fo[0], g.saved[1] + fo[1], g.saved[2] + fo[2]);
}

hipError_t launch_batch_g2p(const BatchTank* T, const WorkItem* wl, uint32_t nwl, int bf16, const ParticleSoA& sp,
                            const BatchGrid& g, hipStream_t s, const ParticleSoA* from, const uint32_t* perm) {
    if (nwl == 0) return hipSuccess;
    if ((from == nullptr) != (perm == nullptr)) return hipErrorInvalidValue;
    const ParticleSoA rd = from ? *from : sp;
    if (bf16) hipLaunchKernelGGL(kb_g2p<1>, dim3(nwl), dim3(kBatchBlock), 0, s, T, wl, sp, rd, perm, g);
    else hipLaunchKernelGGL(kb_g2p<0>, dim3(nwl), dim3(kBatchBlock), 0, s, T, wl, sp, rd, perm, g);
    return hipGetLastError();
}

}  // namespace fs
